"""The scene compiler's BVH builder (dayflower_amd/csrc/bvh.cpp), built on the
host with g++ into a test helper: structure of the four-wide and two-wide
layouts the kernels walk."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from dayflower_amd import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "native", "bvh_probe.cpp"), os.path.join(HERE, "..", "dayflower_amd", "csrc", "bvh.cpp")]
LIB = os.path.join(HERE, "native", "libbvh_probe.so")
LEAF = 0x80000000
EMPTY = 0xFFFFFFFF


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(LIB) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in SRC):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", LIB] + SRC, check=True)
    lib = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    lib.probe_build_bvh.restype = ctypes.c_longlong
    lib.probe_build_bvh.argtypes = [vp, vp, ctypes.c_longlong, ctypes.c_int, vp, ctypes.c_longlong, vp, vp]
    return lib


def _build(probe, p, t, width):
    n = t.shape[0]
    per = 32 if width == 4 else 16
    cap = 2 * n + 2
    nodes = np.zeros((cap, per), np.float32)
    order = np.zeros(n, np.uint32)
    info = np.zeros(3, np.int32)
    pc = np.ascontiguousarray(p, np.float32)
    tc = np.ascontiguousarray(t, np.int32)
    cnt = probe.probe_build_bvh(pc.ctypes.data, tc.ctypes.data, n, width, nodes.ctypes.data, cap, order.ctypes.data,
                                info.ctypes.data)
    assert cnt > 0
    return nodes[:cnt], order, info


def _children(node, width):
    """(lo[3], hi[3], ref) of each child slot of a node record."""
    f = node
    if width == 4:
        refs = f[24:28].view(np.uint32)
        return [((f[0 + k], f[8 + k], f[16 + k]), (f[4 + k], f[12 + k], f[20 + k]), int(refs[k])) for k in range(4)]
    refs = f[12:14].view(np.uint32)
    return [((f[0], f[2], f[8]), (f[1], f[3], f[9]), int(refs[0])), ((f[4], f[6], f[10]), (f[5], f[7], f[11]), int(refs[1]))]


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("kind", ["soup", "icosphere"])
def test_bvh_structure(probe, width, kind):
    """Every triangle sits in exactly one leaf (<= 8 per leaf); every child box
    contains the triangles of its subtree; every node is reached once; the
    stack bound covers the deepest push sequence of an ordered walk."""
    if kind == "soup":
        p, t = S.triangle_soup(20000, 5)
    else:
        p, t, _ = S.displaced_icosphere(20000, 3)
    nodes, order, info = _build(probe, p, t, width)
    n = t.shape[0]
    tri = p[t]  # (n, 3, 3)
    assert sorted(order.tolist()) == list(range(n))
    seen_tri = np.zeros(n, np.int32)
    seen_node = np.zeros(nodes.shape[0], np.int32)
    deepest = 0
    stack = [(0, 0)]  # node, pushes so far on the walk
    while stack:
        ni, pushes = stack.pop()
        seen_node[ni] += 1
        kids = [c for c in _children(nodes[ni], width) if c[2] != EMPTY]
        assert 1 <= len(kids) <= width
        for lo, hi, ref in kids:
            if ref & LEAF:
                first, count = (ref & 0x7FFFFFFF) >> 3, (ref & 7) + 1
                ids = order[first:first + count]
                seen_tri[ids] += 1
                pts = tri[ids].reshape(-1, 3)
                assert (pts >= np.array(lo) - 1e-6).all() and (pts <= np.array(hi) + 1e-6).all()
            else:
                # the child's box holds every triangle below it: check against its own children
                for clo, chi, cref in _children(nodes[ref], width):
                    if cref != EMPTY:
                        assert (np.array(clo) >= np.array(lo) - 1e-6).all() and (np.array(chi) <= np.array(hi) + 1e-6).all()
                stack.append((ref, pushes + len(kids) - 1))
        deepest = max(deepest, pushes + len(kids) - 1)
    assert (seen_tri == 1).all()
    assert (seen_node == 1).all()
    bound = info[1] if width == 4 else info[0]
    assert deepest <= bound + 1
