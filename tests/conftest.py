"""Shared fixtures: library paths, the GPU marker, and scene caching."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "3dgs-isaac-sim-xpbd_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

PRODUCT_LIB = os.path.join(PKG, "libxpbd_isaac.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle_xpbd.so")
HEADER = os.path.join(ROOT, "include", "xpbd_capi.h")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_path():
    if not os.path.isfile(ORACLE_LIB):
        pytest.fail("oracle library missing: run 3dgs-isaac-sim-xpbd_amd/build_lib.py oracle")
    return ORACLE_LIB


@pytest.fixture(scope="session")
def product_path():
    if not os.path.isfile(PRODUCT_LIB):
        pytest.fail("libxpbd_isaac.so missing: run 3dgs-isaac-sim-xpbd_amd/build_lib.py product")
    return PRODUCT_LIB


@pytest.fixture(autouse=True)
def reset_extension_settings():
    """Extension settings persist in each library like the reference's solver
    settings; put them back to their defaults after every test."""
    yield
    import ctypes
    for path in (ORACLE_LIB, PRODUCT_LIB):
        if not os.path.isfile(path):
            continue
        lib = ctypes.CDLL(path)
        lib.xpbd_ext_set_gauss_seidel(0, 0)
        lib.xpbd_ext_set_node_order(1)
        lib.xpbd_ext_set_self_collision_grid(ctypes.c_float(0.0), 0)
        lib.xpbd_ext_set_self_collision_fine(1)
        lib.xpbd_set_self_collision(0, ctypes.c_float(1.5), ctypes.c_float(2.0), 0)
        lib.xpbd_set_object_shape(ctypes.c_float(0.0))
        lib.xpbd_set_colliders(0, None, None, None, ctypes.c_float(0.0), ctypes.c_float(0.0))
        lib.xpbd_press_stop()
        lib.xpbd_press_autolog(1)
        lib.xpbd_set_object_shape_gpu(1)
        lib.xpbd_set_shape_robust(1)
        lib.xpbd_set_contact_shape(0, ctypes.c_float(0.2), ctypes.c_float(0.0))
        lib.xpbd_set_volume(ctypes.c_float(1e-6), 1, 64, 1)
