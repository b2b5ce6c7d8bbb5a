"""Mesh instances (Scene.add_instance: several mesh primitives over one
triangle range, each under its own matrix): one object-space copy of the
triangles under a tree of their own and a top-level tree over the instances'
world boxes.  The reference stores a mesh once per distinct TriangleMesh3F
and carries rays into each primitive's object space (Shape3FCache.java
:176-177, 298-300; AbstractSceneKernel.java:1669-1772).  The oracle has no shared ranges: it
renders the same scene with every instance's triangles copied, so the tests
compare against that expansion — hits against the FP64 intersector, paths on
the same streams, moved instances (top-level refit) against a rebuild — and
check the refusals."""
import numpy as np
import pytest

from dayflower_amd import abi
from dayflower_amd import scene as S

pytestmark = pytest.mark.gpu


def _rot(axis, deg):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = np.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


def _affine(A, t):
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = A
    m[:3, 3] = t
    return m


def _field(res, n=24, seed=5, n_tris=2000):
    """A displaced icosphere and n - 1 instances of it under rotations,
    uniform and non-uniform scales, shears and reflections, over a ground
    plane; Plastic / Matte / Glass / Mirror; point light and sky."""
    rng = np.random.default_rng(seed)
    sc = S.Scene()
    mats = [sc.add_material(S.matte((0.7, 0.6, 0.5))), sc.add_material(S.plastic((0.3, 0.5, 0.7), (0.2, 0.2, 0.2), 0.3)),
            sc.add_material(S.glass(1.5)), sc.add_material(S.mirror((0.9, 0.9, 0.9)))]
    sc.add_plane((0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), mats[0])
    p, t, nrm = S.displaced_icosphere(n_tris, 7, radius=0.5, center=(0.0, 0.0, 0.0))
    base = sc.add_mesh(p, t, mats[1], normals=nrm)
    xf = [np.eye(4, dtype=np.float32), _affine(np.eye(3), (0.0, -0.3, 1.0))]
    for k in range(n - 1):
        sc.add_instance(base, int(rng.choice(mats, p=[0.3, 0.4, 0.15, 0.15])))
        A = _rot(rng.normal(size=3), rng.uniform(0, 360)) * rng.uniform(0.3, 0.8)
        if k % 3 == 1:
            A = A @ np.diag(rng.uniform(0.6, 1.4, 3))        # non-uniform scale
        if k % 5 == 2:
            A = A @ np.diag([-1.0, 1.0, 1.0])                 # reflection
        if k % 7 == 3:
            A = A @ np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.0], [0.0, 0.2, 1.0]])  # shear
        xf.append(_affine(A, rng.uniform([-2.5, -0.6, 0.0], [2.5, 1.5, 4.0])))
    sc.add_point_light((1.0, 4.0, -1.0), (40.0, 40.0, 40.0))
    sc.add_perez_light(2.5, (0.5, 1.0, -0.5))
    sc.set_camera(S.camera_record((0.0, 0.8, -4.0), (0.0, 0.3, 1.0), res, res, fov_x_deg=55.0))
    return sc, np.stack(xf), base


def _expanded(sc):
    """The same scene without shared ranges: every mesh primitive whose range
    an earlier primitive already names gets its own copy of the index
    triples, appended.  Returns the scene and, per primitive, the offset that
    takes a triangle id of its shared range to its copy's id."""
    ex = S.Scene()
    ex.positions, ex.normals, ex.has_normals = sc.positions, sc.normals, sc.has_normals
    ex.uvs, ex.has_uvs = sc.uvs, sc.has_uvs
    ex.materials, ex.lights, ex.images, ex.textures = sc.materials, sc.lights, sc.images, sc.textures
    ex.camera = sc.camera
    idx = [sc.indices]
    n = sc.indices.shape[0]
    seen = set()
    offset = np.zeros(len(sc.primitives), np.int64)
    for i, p in enumerate(sc.primitives):
        q = abi.Primitive.from_buffer_copy(p)
        key = (p.first_triangle, p.triangle_count)
        if p.shape == abi.SHAPE_TRIANGLE_MESH and p.triangle_count > 0:
            if key in seen:
                idx.append(sc.indices[p.first_triangle:p.first_triangle + p.triangle_count])
                q.first_triangle = n
                offset[i] = n - p.first_triangle
                n += p.triangle_count
            seen.add(key)
        ex.primitives.append(q)
    ex.indices = np.concatenate(idx)
    return ex, offset


def _hit_parity(prim_g, tri_g, t_g, prim_o, tri_o, t_o, offset):
    """Id agreement and the distance error relative to max(t, 1): an
    instance's world vertices are its object vertices through the matrix in
    float on the device, the expansion's the same product rounded on the
    host, so the two meshes differ by an ulp of the coordinates, which is an
    absolute distance (it dominates the relative error of a hit a few
    thousandths away)."""
    tri_m = np.where((prim_g >= 0) & (tri_g >= 0), tri_g + offset[np.maximum(prim_g, 0)], tri_g)
    same = (prim_g == prim_o) & (tri_m == tri_o)
    both = same & (prim_o >= 0)
    rel = np.abs(t_g[both].astype(np.float64) - t_o[both]) / np.maximum(np.abs(t_o[both]), 1.0)
    q = np.quantile(rel, 0.9999) if rel.size else 0.0
    return same.mean(), q, (rel.max() if rel.size else 0.0), int((prim_o >= 0).sum())


def _rays(rng, n):
    rr = np.zeros((n, 8), np.float32)
    rr[:, 0:3] = rng.uniform([-3, -1, -1], [3, 2, 4], size=(n, 3))
    d = rng.normal(size=(n, 3))
    rr[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rr[:, 3] = 0.001
    rr[:, 7] = rng.uniform(0.05, 6.0, n)
    return rr


def test_instances_match_the_expanded_scene_against_fp64():
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res = 192
    sc, xf, base = _field(res)
    ex, offset = _expanded(sc)
    rr = _rays(np.random.default_rng(4), 100000)
    out = {}
    for name, scene in (("instanced", sc), ("expanded", ex)):
        r = GPURenderer()
        r.setScene(scene)
        r.updateTransforms(xf)
        r.setup()
        rays, prim, tri, t = r.cast_primary(res, res, sample_index=2)
        p2, tr2, t2 = r.trace_rays(rr)
        pa, _, _ = r.trace_rays(rr, any_hit=True)
        r.dispose()
        out[name] = (rays, prim, tri, t, p2, tr2, t2, pa)
    o = OracleScene(ex)
    o.set_transforms(xf)
    rays = out["instanced"][0]
    np.testing.assert_array_equal(rays, out["expanded"][0])
    po, tro, to = o.trace_f64(rays)
    po2, tro2, to2 = o.trace_f64(rr)
    o.close()
    on_inst = int(((po >= 0) & (po != 0)).sum())
    print(f"primary rays on the meshes: {on_inst} of {res * res}", flush=True)
    assert on_inst > res * res // 20
    zero = np.zeros_like(offset)
    for name, off in (("instanced", offset), ("expanded", zero)):
        _, prim, tri, t, p2, tr2, t2, pa = out[name]
        frac, q, mx, nhit = _hit_parity(prim, tri, t, po, tro, to, off)
        frac2, q2, mx2, nhit2 = _hit_parity(p2, tr2, t2, po2, tro2, to2, off)
        anyhit = ((pa >= 0) == (p2 >= 0)).mean()
        print(f"{name}: primary id match {frac:.6f}, rel t 99.99 % {q:.2e} max {mx:.2e}, hits {nhit}; random rays "
              f"{frac2:.6f}, {q2:.2e} max {mx2:.2e}, hits {nhit2}; any hit agrees {anyhit:.6f}", flush=True)
        assert frac >= 0.9999 and frac2 >= 0.9999
        assert q <= 1e-5 and q2 <= 1e-5
        assert anyhit >= 0.9999


@pytest.mark.parametrize("kernel", [0, 1])
def test_instanced_paths_on_the_oracles_streams(kernel):
    """Paths of the instanced scene against the oracle's expansion on the same
    streams, as close as the expanded scene's own GPU render (the instance
    walk rounds a ray into object space where the expansion rounds vertices
    into world space); sample counts exact, 16-spp means within 0.5 %."""
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res, seed = 48, 13
    sc, xf, base = _field(res)
    ex, _ = _expanded(sc)
    o = OracleScene(ex)
    o.set_transforms(xf)
    for spp in (1, 16):
        s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
        ref = s / np.maximum(c, 1)[..., None]
        got = {}
        for name, scene in (("instanced", sc), ("expanded", ex)):
            r = GPURenderer()
            r.setScene(scene)
            r.setSeed(seed)
            r.setMaximumBounce(8)
            r.setKernel(kernel)
            r.updateTransforms(xf)
            r.setup()
            r.render(spp)
            got[name] = r.film(res, res)
            r.dispose()
        for name, (g, k) in got.items():
            np.testing.assert_array_equal(k, c)
        eq = {n: np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean() for n, (g, _) in got.items()}
        rel = {n: abs(g.mean() - ref.mean()) / ref.mean() for n, (g, _) in got.items()}
        print(f"kernel {kernel} spp {spp}: paths equal {eq['instanced']:.4f} (expanded {eq['expanded']:.4f}), "
              f"mean rel {rel['instanced']:.2e} (expanded {rel['expanded']:.2e})", flush=True)
        if spp == 1:
            assert eq["instanced"] >= eq["expanded"] - 0.01 and eq["instanced"] >= 0.9
        else:
            assert rel["instanced"] < 5e-3
    o.close()


def test_moved_instances_refit_the_top_level_tree():
    """Moving every instance after setup refits the top-level tree only: ray
    casting then equals a rebuild's and the FP64 intersector's on the
    expansion under the new matrices."""
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res = 128
    sc, xf, base = _field(res)
    ex, offset = _expanded(sc)
    rng = np.random.default_rng(9)
    xf2 = xf.copy()
    for i in range(1, len(xf2)):
        xf2[i, :3, 3] += rng.uniform(-0.5, 0.5, 3)
        xf2[i, :3, :3] = (_rot(rng.normal(size=3), 40.0) @ xf2[i, :3, :3].astype(np.float64)).astype(np.float32)
    out = []
    for rebuild in (False, True):
        r = GPURenderer()
        r.setScene(sc)
        r.updateTransforms(xf)
        r.setup()
        r.updateTransforms(xf2, rebuild=rebuild)
        out.append(r.cast_primary(res, res, sample_index=1))
        r.dispose()
    (rays, p0, tr0, t0), (_, p1, tr1, t1) = out
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(tr0, tr1)
    o = OracleScene(ex)
    o.set_transforms(xf2)
    po, tro, to = o.trace_f64(rays)
    o.close()
    frac, q, mx, nhit = _hit_parity(p0, tr0, t0, po, tro, to, offset)
    print(f"moved instances: id match {frac:.6f}, rel t 99.99 % {q:.2e} max {mx:.2e}, hits {nhit}", flush=True)
    assert frac >= 0.9999 and q <= 1e-5


def test_shared_ranges_refused_where_they_cannot_work():
    """Ranges that overlap without being equal are an argument error (they
    were silently given to the last primitive before instances existed); an
    area light on a shared mesh is refused (light sampling reads world-space
    triangle records)."""
    from dayflower_amd.renderer import GPURenderer, IllegalArgumentException, UnsupportedOperationException
    sc, xf, base = _field(16, n=3)
    bad = S.Scene()
    bad.__dict__.update(sc.__dict__)
    bad.primitives = [abi.Primitive.from_buffer_copy(p) for p in sc.primitives]
    bad.primitives[-1].first_triangle = sc.primitives[base].first_triangle + 1
    bad.primitives[-1].triangle_count = sc.primitives[base].triangle_count - 1
    r = GPURenderer()
    r.setScene(bad)
    with pytest.raises(IllegalArgumentException, match="overlap"):
        r.setup()
    r.dispose()
    lit = S.Scene()
    lit.__dict__.update(sc.__dict__)
    lit.primitives = [abi.Primitive.from_buffer_copy(p) for p in sc.primitives]
    lit.lights = list(sc.lights)
    L = abi.Light()
    L.type = abi.LIGHT_DIFFUSE_AREA
    L.primitive = base
    L.triangle = sc.primitives[base].first_triangle
    for i in range(3):
        L.radiance[i] = 1.0
    lit.lights.append(L)
    lit.primitives[base].area_light = len(lit.lights) - 1
    r = GPURenderer()
    r.setScene(lit)
    with pytest.raises(UnsupportedOperationException, match="shared"):
        r.setup()
    r.dispose()


def test_compiled_arrays_of_a_shared_mesh_render_as_instances():
    """The reference's arrays keep a shared mesh once (Shape3FCache.java
    :176-177, 298-300); set through dfr_set_compiled_scene they come back as
    instances and render, on the same streams, what the oracle renders for
    the expansion, as closely as the scene set through dfr_set_scene does."""
    from dayflower_amd import renderer as R
    from oracle import scene_compiler as SC
    from tests.oracle_lib import OracleScene
    res, seed, spp = 48, 17, 1
    sc, xf, base = _field(res, n=8)
    arr = SC.compatible(SC.compile_scene(sc, xf))
    t = R.transcode_compiled_scene(arr)
    meshes = [(q.first_triangle, q.triangle_count) for q in t["primitives"] if q.shape == abi.SHAPE_TRIANGLE_MESH]
    assert len(set(meshes)) == 1 and len(meshes) == 8
    ex, _ = _expanded(sc)
    o = OracleScene(ex)
    o.set_transforms(xf)
    s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
    o.close()
    ref = s / np.maximum(c, 1)[..., None]
    films = {}
    for name in ("compiled", "native"):
        r = R.GPURenderer()
        if name == "compiled":
            r.setCompiledScene(arr)
        else:
            r.setScene(sc)
            r.updateTransforms(xf)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setup()
        r.render(spp)
        films[name] = r.film(res, res)
        r.dispose()
    eq = {n: np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean() for n, (g, _) in films.items()}
    print(f"compiled instances: paths equal {eq['compiled']:.4f} (native {eq['native']:.4f})", flush=True)
    np.testing.assert_array_equal(films["compiled"][1], c)
    assert eq["compiled"] >= eq["native"] - 0.01


def test_several_shared_meshes_of_few_triangles():
    """Two shared ranges (a single triangle named by three primitives, a
    five-triangle strip by two) beside an unshared mesh: ray casting equals
    the expansion's against the FP64 intersector."""
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res = 96
    sc = S.Scene()
    m = sc.add_material(S.matte((0.6, 0.6, 0.6)))
    one = sc.add_mesh(np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.0, 0.5, 0]], np.float32),
                      np.array([[0, 2, 1]], np.int32), m)
    xs = np.linspace(-1, 1, 6, dtype=np.float32)
    strip_p = np.array([[x, y, 0.0] for x in xs for y in (-0.2, 0.2)], np.float32)
    strip_t = np.array([[2 * i, 2 * i + 2, 2 * i + 1] for i in range(5)], np.int32)
    strip = sc.add_mesh(strip_p, strip_t, m)
    p, t, n = S.displaced_icosphere(500, 5, radius=0.3, center=(0.0, -0.6, 1.0))
    sc.add_mesh(p, t, m, normals=n)
    extra = [sc.add_instance(one), sc.add_instance(one), sc.add_instance(strip)]
    sc.add_point_light((0.0, 2.0, -2.0), (10.0, 10.0, 10.0))
    sc.set_camera(S.camera_record((0.0, 0.0, -3.0), (0.0, 0.0, 0.5), res, res, fov_x_deg=60.0))
    xf = np.tile(np.eye(4, dtype=np.float32), (len(sc.primitives), 1, 1))
    xf[one][:3, 3] = (-1.2, 0.6, 0.5)
    xf[extra[0]][:3, :3] = _rot((0, 1, 0), 50.0) * 0.8
    xf[extra[0]][:3, 3] = (1.1, 0.7, 0.8)
    xf[extra[1]][:3, :3] = np.diag([1.0, -1.0, 1.0])
    xf[extra[1]][:3, 3] = (0.0, 0.8, 1.5)
    xf[strip][:3, 3] = (0.0, 0.0, 0.2)
    xf[extra[2]][:3, :3] = _rot((1, 0, 0), 60.0)
    xf[extra[2]][:3, 3] = (0.3, -0.3, 2.0)
    ex, offset = _expanded(sc)
    r = GPURenderer()
    r.setScene(sc)
    r.updateTransforms(xf)
    r.setup()
    rays, prim, tri, tt = r.cast_primary(res, res, sample_index=0)
    r.dispose()
    o = OracleScene(ex)
    o.set_transforms(xf)
    po, tro, to = o.trace_f64(rays)
    o.close()
    frac, q, mx, nhit = _hit_parity(prim, tri, tt, po, tro, to, offset)
    on_shared = np.isin(po, [one, strip] + extra).sum()
    print(f"few-triangle instances: id match {frac:.6f}, rel t 99.99 % {q:.2e}, hits {nhit} ({on_shared} on shared meshes)",
          flush=True)
    assert on_shared > 200
    assert frac >= 0.9999 and q <= 1e-5


def test_items_and_instances_in_a_packed_batch():
    """Jobs on a scene with an item tree (hundreds of spheres and
    rectangles) and on a scene with shared meshes render the same alone and
    inside a packed batch with other scenes, where the pre-pass finds each
    ray's scene through its job."""
    from dayflower_amd.renderer import GPURenderer, make_job
    from tests.test_item_bvh import _crowd
    r = GPURenderer()
    crowd = _crowd(32, mesh=False)
    shared, _, _ = _field(32, n=6)           # instances at the identity: coincident copies
    a, b, c = r.addScene(crowd), r.addScene(shared), r.addScene(S.cornell_box(32, 32))
    ja = make_job(a, crowd.camera, 32, 32, 4, seed=5)
    jb = make_job(b, shared.camera, 32, 32, 4, seed=6)
    alone = [r.batchRender([j])[0] for j in (ja, jb)]
    films = r.batchRender([make_job(c, S.cornell_box(32, 32).camera, 24, 24, 4, seed=1), ja,
                           make_job(a, crowd.camera, 16, 16, 8, seed=9), jb])
    r.dispose()
    for got, want in ((films[1], alone[0]), (films[3], alone[1])):
        assert want.mean() > 0
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)
