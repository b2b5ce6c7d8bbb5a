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


def _bvh_tris(p, t, order):
    """Triangles in BVH order as the compiled scene stores them: a.xyz,
    original id (int bits), b.xyz, primitive 0, c.xyz, 0."""
    tri = p[t[order]].astype(np.float32)  # (n, 3, 3)
    rec = np.zeros((len(order), 12), np.float32)
    rec[:, 0:3], rec[:, 4:7], rec[:, 8:11] = tri[:, 0], tri[:, 1], tri[:, 2]
    rec.view(np.int32)[:, 3] = order.astype(np.int32)
    return rec


def test_oracle_walk_of_the_product_bvh(probe):
    """The oracle's FP64 walk of the compiled four-wide BVH (the counter of
    node visits and triangle tests the roofline prices) finds the hits of the
    oracle's own FP64 intersector on the same soup; any-hit walks visit fewer
    nodes and test fewer triangles on average than closest-hit walks of the
    same rays (per ray the slot order can cost more); a one-leaf tree costs one
    node visit and one test per triangle."""
    from tests import oracle_lib
    from tests.oracle_lib import OracleScene
    p, t = S.triangle_soup(30000, 9)
    nodes, order, _ = _build(probe, p, t, 4)
    tris = _bvh_tris(p, t, order)
    rng = np.random.default_rng(2)
    n = 20000
    rec = np.zeros((n, 12), np.float32)
    rec[:, 0:3] = rng.random((n, 3)) * 1.6 - 0.3
    d = rng.normal(size=(n, 3))
    rec[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 3] = 3.0
    nv, nt, hit, th = oracle_lib.bvh4_walk(nodes, tris, rec)
    sc = S.Scene()
    sc.add_mesh(p, t, sc.add_material(S.matte((0.5, 0.5, 0.5))))
    rays = np.concatenate([rec[:, 0:3], np.full((n, 1), 0.001, np.float32), rec[:, 4:7], rec[:, 3:4]], 1)
    prim64, tri64, t64, nv_own, nt_own = OracleScene(sc).trace_f64(rays, counts=True)
    assert (hit >= 0).mean() > 0.05
    assert (hit == tri64).mean() >= 0.9999
    both = hit >= 0
    np.testing.assert_allclose(th[both], t64[both], rtol=1e-9)
    rec_any = rec.copy()
    rec_any.view(np.int32)[:, 7] = 1
    nv_any, nt_any, hit_any, _ = oracle_lib.bvh4_walk(nodes, tris, rec_any)
    assert ((hit_any >= 0) == (hit >= 0)).all()
    assert nv_any.mean() < nv.mean() and nt_any.mean() < nt.mean()
    print(f"four-wide walk: {nv.mean():.2f} nodes, {nt.mean():.2f} triangles per closest-hit ray "
          f"(the oracle's own binary tree: {nv_own.mean():.2f}, {nt_own.mean():.2f}); any hit {nv_any.mean():.2f}")
    # a tree of a few triangles whose root children are all leaves: a ray
    # costs one node visit plus the triangles of the leaf boxes it enters
    # (rays entering exactly one or no leaf box, found by a float64 slab test)
    q, tt = S.triangle_soup(6, 4, mu=-1.0)
    nodes1, order1, _ = _build(probe, q, tt, 4)
    kids = [c for c in _children(nodes1[0], 4) if c[2] != EMPTY]
    assert nodes1.shape[0] == 1 and all(k[2] & LEAF for k in kids) and len(kids) >= 2
    m = 4000
    few = np.zeros((m, 12), np.float32)
    few[:, 0:3] = rng.random((m, 3)) * 3.0 - 1.0
    dd = rng.normal(size=(m, 3))
    few[:, 4:7] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    few[:, 3] = 100.0
    o64, d64 = few[:, 0:3].astype(np.float64), few[:, 4:7].astype(np.float64)
    inv = 1.0 / np.where(np.abs(d64) < 1e-20, 1e-20, d64)
    entered = []
    for lo, hi, ref in kids:
        t1, t2 = (np.array(lo) - o64) * inv, (np.array(hi) - o64) * inv
        tn = np.maximum(np.minimum(t1, t2).max(1), 0.001)
        tf = np.minimum(np.maximum(t1, t2).min(1), 100.0)
        entered.append((tn <= tf, (ref & 7) + 1))
    hits_per_ray = sum(e.astype(int) for e, _ in entered)
    expect_tris = sum(np.where(e, c, 0) for e, c in entered)
    nv1, nt1, _, _ = oracle_lib.bvh4_walk(nodes1, _bvh_tris(q, tt, order1), few)
    sel = hits_per_ray <= 1
    assert sel.sum() > 1000 and (hits_per_ray == 1).sum() > 50
    assert (nv1[sel] == 1).all()
    np.testing.assert_array_equal(nt1[sel], expect_tris[sel])
