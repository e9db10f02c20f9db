"""Frame outputs that stay on the GPU: the consumer side of the solver's
per-frame results for a renderer or scene writer.

The reference's runtime reads the positions and deformed shapes back to the
host every frame (src/runtime_loop.py:387-402), takes their float64 mean for
the model's velocity and writes the frame only when the largest change since
the last written frame exceeds a threshold (938-946).  ``DeviceFrame`` does the
same through ``xpbd_ext_frame_outputs``: positions, scales and quaternions
land in device buffers a GPU consumer can read in place (their addresses are
``DeviceFrame.*.ptr``), and only the centre and the largest change (four
doubles) come back to the host.  The write-or-skip decision is taken on the
device: the "written" buffer advances to the new positions only when the
change exceeds ``move_eps``.

Device memory comes from the HIP runtime the library itself uses (the
process's ``libamdhip64``), so the buffers are valid for the library's
stream and for any other HIP user of that runtime.
"""
import ctypes

import numpy as np

_HIP = None
_H2D, _D2H = 1, 2


def _hip():
    global _HIP
    if _HIP is None:
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        _HIP = h
    return _HIP


class DeviceBuffer:
    """float32 device array of a fixed shape; ``ptr`` is its device address."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self._p = ctypes.c_void_p()
        rc = _hip().hipMalloc(ctypes.byref(self._p), max(self.nbytes, 4))
        if rc != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed ({rc})")

    @property
    def ptr(self):
        return self._p.value

    def read(self):
        """Host copy (for checks and tools; a GPU consumer reads ``ptr``)."""
        out = np.empty(self.shape, np.float32)
        if self.nbytes and _hip().hipMemcpy(out.ctypes.data, self._p, self.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy to the host failed")
        return out

    def write(self, host):
        host = np.ascontiguousarray(host, np.float32).reshape(self.shape)
        if self.nbytes and _hip().hipMemcpy(self._p, host.ctypes.data, self.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def free(self):
        if self._p.value:
            _hip().hipFree(self._p)
            self._p = ctypes.c_void_p()


class DeviceFrame:
    """Per-frame device outputs of one solver instance (``xpbd_bridge.XPBD``,
    stepped at least once).

    shapes      also update the deformed scales and quaternions each frame
    xyzw        quaternions as (x, y, z, w) (the scene writer's order) instead
                of the solver's (w, x, y, z)
    deform_eps  Gaussians that moved at most this keep their rest shape
    move_eps    a frame counts as written when some coordinate moved more than
                this since the last written frame
    """

    def __init__(self, sim, shapes=True, xyzw=True, deform_eps=1e-2, move_eps=0.0):
        if sim.n <= 0:
            raise RuntimeError("DeviceFrame: create the instance first")
        if not hasattr(sim.lib, "xpbd_ext_frame_outputs"):
            raise RuntimeError("DeviceFrame: the library has no device frame outputs (the HIP library has)")
        self.sim, self.xyzw, self.deform_eps, self.move_eps = sim, bool(xyzw), float(deform_eps), float(move_eps)
        n = sim.n
        self.positions = DeviceBuffer((n, 3))
        self.scales = DeviceBuffer((n, 3)) if shapes else None
        self.quats = DeviceBuffer((n, 4)) if shapes else None
        self.written = DeviceBuffer((n, 3))   # positions of the last frame written
        self.center = None
        self.max_move = None
        self.deformed = 0
        self.frames_written = 0
        self._host = (ctypes.c_double * 4)()
        # the first frame compares against the positions as they are now
        self._call(self.written, None, None, None, -1.0)

    def _call(self, pos, scales, quats, prev, move_eps):
        ptr = lambda b: None if b is None else b.ptr
        k = self.sim.lib.xpbd_ext_frame_outputs(ptr(pos), ptr(scales), ptr(quats), 1 if self.xyzw else 0,
                                                self.deform_eps, ptr(prev), move_eps, self._host, self.sim.n)
        if k < 0:
            self.sim._check(k, "xpbd_ext_frame_outputs")
        return k

    def update(self):
        """This frame's outputs into the device buffers.  Returns True when the
        frame counts as written (largest change > move_eps); ``center``,
        ``max_move`` and ``deformed`` describe the frame."""
        self.deformed = self._call(self.positions, self.scales, self.quats, self.written, self.move_eps)
        h = self._host
        self.center = np.array([h[0], h[1], h[2]]) / self.sim.n
        self.max_move = float(h[3])
        moved = self.max_move > self.move_eps
        self.frames_written += int(moved)
        return moved

    def free(self):
        for b in (self.positions, self.scales, self.quats, self.written):
            if b is not None:
                b.free()
