"""CPU: the persistent tree format's validator (pptk_tree_deserialize).

A small tree is written by hand in the documented layout (include/
pptk_amd.h, pptk_amd/csrc/persist.hip). The intact buffer passes host
validation and only then fails for want of a device (E_DEVICE on a machine
without a GPU); every damaged variant is rejected as invalid (ValueError)
before any device work.
"""
import ctypes

import numpy as np
import pytest

HDR = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"),
                ("dim", "<i4"), ("dtype", "<i4"), ("dist", "<i4"), ("leaf", "<i4"),
                ("depth", "<i4"), ("pad", "<i4"),
                ("n_input", "<i8"), ("n_valid", "<i8"), ("nn", "<i8"), ("nleaves", "<i8"),
                ("nroots", "<i8"),
                ("lo", "<f8", 4), ("hi", "<f8", 4), ("mlo", "<f8", 4), ("mscale", "<f8", 4),
                ("off_coords", "<i8"), ("off_orig", "<i8"), ("off_fwd", "<i8"),
                ("off_nodes", "<i8"), ("off_boxes", "<i8"), ("off_leafof", "<i8"),
                ("total", "<i8")])


def a8(x):
    return (x + 7) & ~7


def make_buffer():
    """6 float32 points in 3-D, one invalid input row (row 2), leaf size 2:
    root [0,5) -> (1: [0,2), 2: [2,5)), node 2 -> (3: [2,3), 4: [3,5))."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [5, 5, 5], [5, 6, 5], [6, 6, 6]], np.float32)
    orig = np.array([0, 1, 3, 4, 5], np.int32)
    fwd = np.array([0, 1, -1, 2, 3, 4], np.int32)
    nodes = np.array([[0, 5, 1, -1], [0, 2, -1, 0], [2, 5, 3, 0], [2, 3, -1, 2],
                      [3, 5, -1, 2]], np.int32)
    boxes = np.zeros((5, 2, 3), np.float32)
    for i, (a, b, _, _) in enumerate(nodes):
        boxes[i, 0] = pts[a:b].min(0)
        boxes[i, 1] = pts[a:b].max(0)
    leafof = np.array([1, 1, 3, 4, 4], np.int32)
    h = np.zeros(1, HDR)
    h["magic"] = b"PPTKTREE"
    h["version"] = 1
    h["header_bytes"] = HDR.itemsize
    h["dim"], h["dtype"], h["dist"], h["leaf"], h["depth"] = 3, 0, 0, 2, 3
    h["n_input"], h["n_valid"], h["nn"], h["nleaves"], h["nroots"] = 6, 5, 5, 3, 1
    h["lo"][0, :3] = pts.min(0)
    h["hi"][0, :3] = pts.max(0)
    h["mlo"][0, :3] = pts.min(0)
    h["mscale"][0, :3] = 1.0
    o = a8(HDR.itemsize)
    sections = [("off_coords", pts), ("off_orig", orig), ("off_fwd", fwd),
                ("off_nodes", nodes), ("off_boxes", boxes), ("off_leafof", leafof)]
    for name, arr in sections:
        h[name] = o
        o = a8(o + arr.nbytes)
    h["total"] = o
    buf = np.zeros(o, np.uint8)
    buf[:HDR.itemsize] = np.frombuffer(h.tobytes(), np.uint8)
    for name, arr in sections:
        off = int(h[name][0])
        buf[off:off + arr.nbytes] = np.frombuffer(arr.tobytes(), np.uint8)
    return buf, {name: int(h[name][0]) for name, _ in sections}


def deserialize(buf):
    from pptk_amd import capi
    h = ctypes.c_void_p()
    return capi.lib().pptk_tree_deserialize(capi._p(buf), buf.size, None, ctypes.byref(h))


def test_header_size_matches_layout():
    assert HDR.itemsize == 264


def test_intact_buffer_passes_validation():
    from pptk_amd import capi
    buf, _ = make_buffer()
    st = deserialize(buf)
    if capi.lib().pptk_device_count() > 0:
        assert st == capi.OK
    else:
        assert st == capi.E_DEVICE  # validated, then no device to upload to


def _box(buf, off, node, side, d, val):
    v = buf[off["off_boxes"]:].view(np.float32)
    v[node * 6 + (3 if side == "hi" else 0) + d] = val


def _i32(buf, off, idx, val):
    v = buf[off:].view(np.int32)
    v[idx] = val


@pytest.mark.parametrize("damage", [
    "magic", "version", "dim", "counts", "short", "node_range", "child_even", "child_parent",
    "leafof", "orig", "fwd", "coords_nan", "cycle", "total",
    "box_shrunk", "box_nan", "box_child", "box_inverted", "depth", "nleaves",
])
def test_damaged_buffers_are_rejected(damage):
    from pptk_amd import capi
    buf, off = make_buffer()
    if damage == "magic":
        buf[0] = ord("X")
    elif damage == "version":
        buf[8] = 2
    elif damage == "dim":
        buf[16] = 7
    elif damage == "counts":
        buf[48:56] = np.frombuffer(np.int64(9).tobytes(), np.uint8)  # n_valid > n_input
    elif damage == "short":
        buf = buf[:-8]
    elif damage == "node_range":
        _i32(buf, off["off_nodes"], 4 * 4 + 1, 9)       # node 4: b = 9 > n_valid
    elif damage == "child_even":
        _i32(buf, off["off_nodes"], 2 * 4 + 2, 4)       # node 2: child 4 (even)
    elif damage == "child_parent":
        _i32(buf, off["off_nodes"], 3 * 4 + 3, 1)       # node 3: parent 1
    elif damage == "leafof":
        _i32(buf, off["off_leafof"], 0, 2)              # position 0 -> inner node
    elif damage == "orig":
        _i32(buf, off["off_orig"], 1, 3)                # two positions claim row 3
    elif damage == "fwd":
        _i32(buf, off["off_fwd"], 2, 4)                 # invalid row mapped
    elif damage == "coords_nan":
        buf[off["off_coords"]:].view(np.float32)[4] = np.nan
    elif damage == "cycle":
        _i32(buf, off["off_nodes"], 1 * 4 + 2, 1)       # node 1 is its own child
    elif damage == "total":
        buf = np.concatenate([buf, np.zeros(8, np.uint8)])
        buf[256:264] = np.frombuffer(np.int64(buf.size + 64).tobytes(), np.uint8)
    elif damage == "box_shrunk":
        _box(buf, off, 4, "hi", 0, 5.5)                 # leaf 4 holds (6, 6, 6)
    elif damage == "box_nan":
        _box(buf, off, 1, "lo", 1, np.nan)
    elif damage == "box_child":
        _box(buf, off, 0, "lo", 2, 1.0)                 # root misses child 1 (z = 0)
    elif damage == "box_inverted":
        _box(buf, off, 3, "lo", 0, 6.0)                 # lo 6 > hi 5
    elif damage == "depth":
        buf[32:36] = np.frombuffer(np.int32(5).tobytes(), np.uint8)
    elif damage == "nleaves":
        buf[64:72] = np.frombuffer(np.int64(4).tobytes(), np.uint8)
    assert deserialize(buf) == capi.E_VALUE, damage
    assert b"invalid tree buffer" in capi.lib().pptk_last_error() or \
        b"unsupported" in capi.lib().pptk_last_error()
