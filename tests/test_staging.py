"""Parameter uploads of substeps queued back to back.

Each substep's parameter block goes to the device through a pinned staging
slot that is rewritten only after its previous copy has completed.  Here every
substep runs with different solver settings and the substeps are queued
without waiting (xpbd_batch_step_async); the result must equal the same
sequence run one synchronous substep at a time, bit for bit.
"""
import numpy as np
import pytest

import scenes
import serving
from batch import XPBDBatch

pytestmark = pytest.mark.gpu


def _objects():
    return [scenes.object_request(n, 300 + j) for j, n in enumerate((2500, 4100))]


def _run(objs, queued, substeps=24):
    b = XPBDBatch([o[0] for o in objs])
    for k, (scene, recipe, script) in enumerate(objs):
        scenes.configure_full_batch(b, k, recipe)
        t, p, d = script.colliders(0)
        b.set_colliders(k, t, p, d)
    b.prepare()
    b.step(1)
    for s in range(substeps):
        # damping, under-relaxation and the press of the ground change every substep
        for k in range(len(objs)):
            b.set_solver(k, iters=4, dt=1.0 / 600.0, under_relax=0.3 + 0.02 * (s % 7), vel_damping=0.01 * (s % 5),
                         dist_compliance=1e-6, shape_compliance=0.0, shape_blend=0.1)
            b.set_ground(k, True, (0.0, 0.0, 1.0), 0.002 * (s % 3), friction=0.5, restitution=0.0, gravity=9.81)
        if queued:
            b.step_async(1)
        else:
            b.step(1)
    b.sync()
    out = [b.positions(k) for k in range(len(objs))]
    b.close()
    return out


def test_queued_substeps_keep_their_parameters(product_path):
    objs = _objects()
    ref = _run(objs, queued=False)
    got = _run(objs, queued=True)
    for k in range(len(objs)):
        assert np.array_equal(got[k], ref[k]), f"object {k}: max |dx| {np.abs(got[k] - ref[k]).max()}"
    print("24 substeps with per-substep settings: queued == synchronous, bit for bit")
