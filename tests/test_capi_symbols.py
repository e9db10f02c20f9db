"""The C-ABI libraries load and export every symbol include/xpbd_capi.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import re

from conftest import HEADER

DROP_IN = [
    "xpbd_last_error", "xpbd_create", "xpbd_destroy", "xpbd_count", "xpbd_set_solver", "xpbd_set_constraints",
    "xpbd_set_volume", "xpbd_set_ground", "xpbd_set_object_shape", "xpbd_set_object_shape_gpu",
    "xpbd_set_colliders", "xpbd_collider_stats", "xpbd_add_rigid_velocity", "xpbd_set_contact_shape",
    "xpbd_set_attached", "xpbd_set_self_collision", "xpbd_self_collision_stats", "xpbd_set_shape_robust",
    "xpbd_step", "xpbd_positions_device", "xpbd_positions_host", "xpbd_compute_shapes", "xpbd_reset",
    "xpbd_launch", "xpbd_press_start", "xpbd_press_autolog", "xpbd_press_stop", "xpbd_press_state",
    "xpbd_volume_stats",
]


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xpbd_\w+)\s*\(", text)))


def test_header_declares_the_29_reference_symbols():
    names = declared()
    assert len(DROP_IN) == 29
    for s in DROP_IN:
        assert s in names


def test_product_library_exports_header(product_path):
    lib = ctypes.CDLL(product_path)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_oracle_exports_drop_in(oracle_path):
    lib = ctypes.CDLL(oracle_path)
    missing = [s for s in DROP_IN + ["xpbd_ext_get_state", "xpbd_ext_set_state"] if not hasattr(lib, s)]
    assert not missing, missing
