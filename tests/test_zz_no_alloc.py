"""xpbd_ext_compute_shapes_noalloc, the allocation-free deformed-shape call,
against xpbd_compute_shapes on the same instance: bit-identical scales and
quaternions and the same deformed count, on two calls in a row (the second on
a smaller object, so the shared count array keeps the first call's values
beyond the second's blocks)."""
import ctypes

import numpy as np
import pytest

from harness import XPBD
import scenes

pytestmark = pytest.mark.gpu


def loadgraph_csr(n, edges):
    """Symmetric CSR in input edge order: edge k as a->b, then b->a."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    src = np.stack([e[:, 0], e[:, 1]], 1).ravel()
    dst = np.stack([e[:, 1], e[:, 0]], 1).ravel()
    order = np.argsort(src, kind="stable")
    off = np.zeros(n + 1, np.int32)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off).astype(np.int32), dst[order].astype(np.int32)


class DeviceArray:
    """Device memory for the test's inputs and outputs through the HIP runtime
    the library itself loaded (the entry point under test allocates nothing;
    the caller does)."""

    hip = None

    def __init__(self, host):
        if DeviceArray.hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            DeviceArray.hip = h
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), max(self.host.nbytes, 4)) == 0
        if self.host.nbytes:
            assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0   # host to device

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0                  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def run_object(product_path, n, seed, frames):
    scene = scenes.make_ellipsoid_scene(n, 8, seed=seed)
    script = scenes.GripperScript(scene, seed=seed + 1, attach=min(200, n // 10))
    sim = XPBD(product_path)
    sim.create(scene)
    scenes.config1_solver(sim)
    for f in range(frames):
        if f > 0:                      # attachments need a first step
            script.apply(sim, f)
        for _ in range(10):
            sim.step()
    sc_ref, q_ref, k_ref = sim.shapes(1e-3)
    off, nbr = loadgraph_csr(n, scene["edges"])
    rest, rowoff, nb = DeviceArray(scene["pos"]), DeviceArray(off), DeviceArray(nbr)
    scales, rots = DeviceArray(scene["scales"]), DeviceArray(scene["rots"])
    out_s = DeviceArray(np.full((n, 3), -1.0, np.float32))
    out_q = DeviceArray(np.full((n, 4), -1.0, np.float32))
    fn = sim.lib.xpbd_ext_compute_shapes_noalloc
    fn.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_float] + [ctypes.c_void_p] * 2
    fn.restype = ctypes.c_int
    pos_ptr = sim.positions_device_ptr()
    assert pos_ptr != 0
    k = fn(n, rest.ptr, pos_ptr, rowoff.ptr, nb.ptr, scales.ptr, rots.ptr, 1e-3, out_s.ptr, out_q.ptr)
    assert k >= 0, sim.lib.xpbd_last_error().decode()
    got_s, got_q = out_s.read(), out_q.read()
    for a in (rest, rowoff, nb, scales, rots, out_s, out_q):
        a.free()
    sim.destroy()
    return (sc_ref, q_ref, k_ref), (got_s, got_q, k)


def test_noalloc_shapes_match_compute_shapes(product_path):
    for n, seed, frames in ((6000, 31, 6), (1500, 32, 4)):
        (sr, qr, kr), (sg, qg, kg) = run_object(product_path, n, seed, frames)
        assert kr > 0 and kg == kr, (n, kg, kr)
        assert np.array_equal(sg, sr), f"n={n}: scales differ"
        assert np.array_equal(qg, qr), f"n={n}: quaternions differ"
        print(f"n={n}: {kg} deformed, bit-identical to xpbd_compute_shapes")


def test_noalloc_rejects_bad_sizes(product_path):
    sim = XPBD(product_path)
    fn = sim.lib.xpbd_ext_compute_shapes_noalloc
    fn.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_float] + [ctypes.c_void_p] * 2
    fn.restype = ctypes.c_int
    dummy = ctypes.c_void_p(16)   # never dereferenced: the sizes are refused first
    assert fn(0, dummy, dummy, dummy, dummy, dummy, dummy, 1e-3, dummy, dummy) == -1
    assert fn((1 << 24) + 1, dummy, dummy, dummy, dummy, dummy, dummy, 1e-3, dummy, dummy) == -1
    assert "limit" in sim.lib.xpbd_last_error().decode()
    assert fn(10, None, dummy, dummy, dummy, dummy, dummy, 1e-3, dummy, dummy) == -1
