"""Scripted scene runs shared by the parity tests and the bench."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "3dgs-isaac-sim-xpbd_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import scenes  # noqa: E402
from xpbd_bridge import XPBD  # noqa: E402


def config1_scene(n=20000, k=8, seed=1):
    return scenes.make_ellipsoid_scene(n, k, seed=seed)


def setup_config1(sim, scene):
    sim.create(scene)
    scenes.config1_solver(sim)
    sim.set_colliders([], [], [])


def frame_config1(sim, frame, substeps=10, frames=100):
    scenes.config1_colliders(sim, frame, frames=frames)
    for _ in range(substeps):
        sim.step()


def run_config1(sim, scene, frames, substeps=10):
    """Free run; returns positions after every frame (frames, N, 3)."""
    setup_config1(sim, scene)
    out = []
    for f in range(frames):
        frame_config1(sim, f, substeps, frames)
        out.append(sim.positions())
    return np.stack(out)


def max_abs(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return int(np.max(np.abs(ai - bi))) if a.size else 0
