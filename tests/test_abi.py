"""The C-ABI library loads on a machine without a GPU, exports every entry
point that include/radiantium_hip.h declares, and its ctypes mirror has the
header's struct sizes.  No compute calls."""
import ctypes as C
import os
import re

from radiantium_amd import abi, native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "radiantium_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rd_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding's list covers the header
    assert set(names) <= set(native.EXPORTS), set(names) - set(native.EXPORTS)
    # every entry point that takes arguments has its ctypes signature declared
    # (an undeclared one would pass 64-bit handles as 32-bit ints)
    no_args = {"rd_abi_version", "rd_last_error", "rd_image_last_error"}
    undeclared = [n for n in names if n not in no_args and getattr(lib, n).argtypes is None]
    assert not undeclared, undeclared


def test_abi_version_and_struct_sizes():
    lib = native.lib()
    assert lib.rd_abi_version() == abi.ABI_VERSION
    # sizes fixed by the header's field lists (natural alignment, LP64)
    assert C.sizeof(abi.rd_ray) == 32
    assert C.sizeof(abi.rd_hit) == 16
    assert C.sizeof(abi.rd_medium_desc) == 32
    assert C.sizeof(abi.rd_image) == 24
    assert C.sizeof(abi.rd_prim_desc) == 24
    assert C.sizeof(abi.rd_render_params) == 40
    assert C.sizeof(abi.rd_texture_desc) == 40
    assert C.sizeof(abi.rd_material_desc) == 4 * 4 + 16 * 4 + 4 * 4 + 6 * 4
    assert C.sizeof(abi.rd_light_desc) == 4 * 2 + 12 + 4 + 64


def test_host_only_entry_points_work_without_device():
    """rd_bvh_build_host needs no GPU; rd_last_error is always callable."""
    import numpy as np
    lo = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    hi = lo + 1
    nodes, order, depth = native.bvh_build_host(lo, hi, 1)
    assert nodes.shape[0] == 3 and sorted(order.tolist()) == [0, 1] and depth == 1
    native.lib().rd_last_error()
