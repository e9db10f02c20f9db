"""A splat asset on disk through the HIP library: a viewer graph export
(APGGRPH1, forward.cu:4612-4663) and its cropped PLY are imported to the NPZ
graph (tools/import_graph.py), loaded with ``load_inputs``
(src/xpbd_bridge.py:17-41) and stepped with the full recipe (distance, node +
3NN volume clusters, k-means shape clusters, self-collision, ground) on the
GPU and on the oracle: positions equal bit for bit every frame, and the
deformed shapes agree."""
import numpy as np
import pytest

from harness import XPBD
import graph_io
import scenes
import splat_io
from test_graph_io import viewer_export
from xpbd_bridge import load_inputs

pytestmark = pytest.mark.gpu


def test_imported_asset_steps_like_the_oracle(tmp_path, product_path, oracle_path):
    scene, _ = viewer_export(tmp_path, n=4000, seed=9)
    viewer = graph_io.read_apggrph1(tmp_path / "toy_graph_sibr.bin")
    _, verts = splat_io.read_ply(tmp_path / "toy_crop.ply")
    usd = np.stack([verts["x"], verts["y"], verts["z"]], axis=1)
    graph_io.save_graph_npz(tmp_path / "toy_graph.npz", graph_io.import_graph(viewer, usd))
    inp = load_inputs(str(tmp_path), "toy")
    assert inp["edges"].shape[0] > 0
    recipe = scenes.full_recipe(inp, cluster_size=64, seed=9)
    recipe["table"] = 1 << 14
    gpu, cpu = XPBD(product_path), XPBD(oracle_path)
    for sim in (gpu, cpu):
        sim.create(inp)
        scenes.configure_full(sim, recipe, iters=4, substeps=10)
    moved = 0.0
    for f in range(6):
        for _ in range(10):
            gpu.step()
            cpu.step()
        a, b = gpu.positions(), cpu.positions()
        np.testing.assert_array_equal(a, b, err_msg=f"frame {f}")
        moved = float(np.abs(b - inp["pos"]).max())
    sg, qg, kg = gpu.shapes(1e-3)
    sc, qc, kc = cpu.shapes(1e-3)
    assert kg == kc
    np.testing.assert_allclose(sg, sc, rtol=1e-5)
    print(f"imported asset ({inp['pos'].shape[0]} splats, {inp['edges'].shape[0]} edges): 6 frames bit-exact, "
          f"largest move {moved:.3e}, deformed {kc}")
    assert moved > 1e-3
    gpu.destroy()
    cpu.destroy()
