"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle on small seeded scenes).

CPU: the scene generator still produces the same inputs (digest) and the
oracle reproduces every fixture bit for bit.  GPU: the HIP library reproduces
the integer artefacts bit for bit (hash keys, sorted order, candidate counts
and lists, statistics, colourings, angle pairing) and the positions within
1e-5 of the scene extent per frame (deformed scales within 1e-5 relative,
orientations within 1e-4 rad through the covariances).
"""
import importlib.util
import os

import numpy as np
import pytest

from harness import XPBD, max_abs
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
mg = importlib.util.module_from_spec(spec)
spec.loader.exec_module(mg)

TOL_REL = 1e-5


def cov(scales, quats):
    w, x, y, z = quats.T.astype(np.float64)
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    return np.einsum("nij,nj,nkj->nik", R, scales.astype(np.float64) ** 2, R)


def fixture(name):
    return np.load(os.path.join(HERE, "golden", name), allow_pickle=False)


def test_scene_digests():
    assert str(fixture("plain.npz")["digest"]) == mg.digest(mg.plain_scene())
    assert str(fixture("full.npz")["digest"]) == mg.digest(mg.full_scene()[0])
    assert str(fixture("topology.npz")["digest"]) == mg.digest(
        scenes.make_ellipsoid_scene(mg.N, 8, semi=(0.06, 0.05, 0.04), seed=104))


def test_oracle_reproduces_plain(oracle_path):
    _, pos = mg.run_plain(XPBD(oracle_path))
    np.testing.assert_array_equal(pos, fixture("plain.npz")["positions"])


def test_oracle_reproduces_full(oracle_path):
    _, out = mg.run_full(XPBD(oracle_path))
    fx = fixture("full.npz")
    for k, v in out.items():
        np.testing.assert_array_equal(v, fx[k], err_msg=k)


def test_oracle_reproduces_topology(oracle_path):
    sim = XPBD(oracle_path)
    _, out = mg.run_topology(sim)
    sim.ext_set_gauss_seidel(False, False)
    fx = fixture("topology.npz")
    for k, v in out.items():
        np.testing.assert_array_equal(v, fx[k], err_msg=k)


@pytest.mark.gpu
def test_gpu_plain(product_path):
    scene, pos = mg.run_plain(XPBD(product_path))
    ref = fixture("plain.npz")["positions"]
    ext = scenes.scene_extent(scene)
    errs = [max_abs(a, b) / ext for a, b in zip(pos, ref)]
    print(f"golden plain: worst |dx|/extent {max(errs):.3e}, bit-identical frames "
          f"{sum(np.array_equal(a, b) for a, b in zip(pos, ref))}/{len(ref)}")
    assert max(errs) <= TOL_REL


@pytest.mark.gpu
def test_gpu_full(product_path):
    scene, out = mg.run_full(XPBD(product_path))
    fx = fixture("full.npz")
    ext = scenes.scene_extent(scene)
    errs = [max_abs(a, b) / ext for a, b in zip(out["positions"], fx["positions"])]
    print(f"golden full recipe: worst |dx|/extent {max(errs):.3e}")
    assert max(errs) <= TOL_REL
    for k in ("sc_keys", "sc_order", "sc_counts", "sc_contacts", "sc_stats", "shape_deformed"):
        np.testing.assert_array_equal(out[k], fx[k], err_msg=k)
    np.testing.assert_allclose(out["shape_scales"], fx["shape_scales"], rtol=1e-5)
    # orientations through the covariances R diag(s^2) R^T: where two scales
    # are (nearly) equal the rotation about the third axis is free, so the
    # quaternions themselves are not comparable; 1e-4 rad moves a covariance
    # by at most ~2e-4 of its norm
    Cg, Cc = cov(out["shape_scales"], out["shape_quats"]), cov(fx["shape_scales"], fx["shape_quats"])
    err = np.linalg.norm(Cg - Cc, axis=(1, 2)) / np.linalg.norm(Cc, axis=(1, 2))
    print(f"golden full recipe: deformed shapes max covariance rel err {err.max():.2e}")
    assert err.max() < 2e-4


@pytest.mark.gpu
def test_gpu_topology(product_path):
    sim = XPBD(product_path)
    _, out = mg.run_topology(sim)
    sim.ext_set_gauss_seidel(False, False)
    fx = fixture("topology.npz")
    for k, v in out.items():
        np.testing.assert_array_equal(v, fx[k], err_msg=k)
