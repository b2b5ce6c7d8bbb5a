"""GPU parity: the HIP path through the C-ABI against the oracle on the same
seeded inputs.

Bars (stated per test): hit parity of primary rays against the FP64
intersector (primitive/triangle ids on >= 99.99 % of rays, t within 1e-5
relative); path-traced films against the float32 CPU restatement on the same
counter-based random streams (identical paths up to float rounding); and
statistical agreement (3 sigma per pixel, relative RMSE) on independent
streams.
"""
import numpy as np
import pytest

from dayflower_amd import abi
from dayflower_amd import scene as S

pytestmark = pytest.mark.gpu


def _renderer():
    from dayflower_amd.renderer import GPURenderer
    return GPURenderer()


def _hit_parity(prim_g, tri_g, t_g, prim_o, tri_o, t_o):
    hit_o = prim_o >= 0
    same = (prim_g == prim_o) & (tri_g == tri_o)
    frac = same.mean()
    both = same & hit_o
    rel = np.abs(t_g[both].astype(np.float64) - t_o[both]) / np.maximum(np.abs(t_o[both]), 1e-30)
    return frac, (rel.max() if rel.size else 0.0), int(hit_o.sum())


@pytest.mark.parametrize("n_tris,res", [(20000, 256), (200000, 256)])
def test_soup_primary_hits_match_fp64(n_tris, res):
    """Ray casting of a triangle soup: ids >= 99.99 %, t within 1e-5 relative."""
    from tests.oracle_lib import OracleScene
    sc = S.soup_scene(n_tris, seed=11, width=res, height=res)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(3)
    r.setRenderingAlgorithm(abi.ALGO_RAY_CASTING)
    r.setup()
    rays, prim, tri, t = r.cast_primary(res, res, sample_index=0)
    r.dispose()
    o = OracleScene(sc)
    prim_o, tri_o, t_o = o.trace_f64(rays)
    frac, rel, nhit = _hit_parity(prim, tri, t, prim_o, tri_o, t_o)
    print(f"soup {n_tris}: id match {frac:.6f}, max rel t {rel:.2e}, hits {nhit}")
    assert nhit > 0.1 * res * res
    assert frac >= 0.9999
    assert rel <= 1e-5
    # the GPU generates the same camera rays as the oracle's restatement
    rays_o = o.primary_rays(res, res, 3, 0)
    np.testing.assert_allclose(rays[:, 4:7], rays_o[:, 4:7], atol=2e-6)


def test_cornell_primary_hits_match_fp64():
    from tests.oracle_lib import OracleScene
    res = 128
    sc = S.cornell_box(res, res)
    r = _renderer()
    r.setScene(sc)
    r.setup()
    rays, prim, tri, t = r.cast_primary(res, res, sample_index=5)
    r.dispose()
    prim_o, tri_o, t_o = OracleScene(sc).trace_f64(rays)
    frac, rel, nhit = _hit_parity(prim, tri, t, prim_o, tri_o, t_o)
    print(f"cornell: id match {frac:.6f}, max rel t {rel:.2e}")
    assert frac >= 0.9999
    assert rel <= 1e-5


def test_obj_meshes_primary_hits_match_fp64():
    """Meshes read from OBJ text (a fan-triangulated, group-split grid of
    quads under a tetrahedron) through the scene compiler: primary hits
    against the FP64 intersector, same bars as above."""
    from tests.oracle_lib import OracleScene
    from dayflower_amd import obj
    n = 24
    lines = [f"v {x / n * 4 - 2} {0.1 * ((x * 7 + y * 3) % 5) / 5} {y / n * 4 - 2}" for y in range(n + 1) for x in range(n + 1)]
    lines.append("g ground")
    for y in range(n):
        for x in range(n):
            a = y * (n + 1) + x + 1
            lines.append(f"f {a} {a + n + 1} {a + n + 2} {a + 1}")
    base = (n + 1) * (n + 1)
    lines += ["g tetra", "usemtl red", "v 0.6 1.5 0.6", "v 0.6 0.5 -0.6", "v -0.6 1.5 -0.6", "v -0.6 0.5 0.6"]
    lines += [f"f {base + a} {base + b} {base + c}" for a, b, c in ((1, 2, 3), (1, 4, 2), (1, 3, 4), (2, 4, 3))]
    meshes = obj.parse_wavefront_object("\n".join(lines))
    assert [len(m.indices) for m in meshes] == [2 * n * n, 4]
    res = 128
    sc = S.Scene()
    obj.add_to_scene(sc, meshes, {"": sc.add_material(S.matte((0.5, 0.5, 0.5))), "red": sc.add_material(S.plastic())})
    sc.add_point_light((1.0, 4.0, -1.0), (20.0, 20.0, 20.0))
    sc.set_camera(S.camera_record((0.0, 2.5, -4.0), (0.0, 0.5, 0.0), res, res, fov_x_deg=50.0))
    r = _renderer()
    r.setScene(sc)
    r.setup()
    rays, prim, tri, t = r.cast_primary(res, res, sample_index=1)
    r.dispose()
    prim_o, tri_o, t_o = OracleScene(sc).trace_f64(rays)
    frac, rel, nhit = _hit_parity(prim, tri, t, prim_o, tri_o, t_o)
    print(f"obj: id match {frac:.6f}, max rel t {rel:.2e}, hits {nhit}")
    assert nhit > 0.5 * res * res and (prim == 1).sum() > 100
    assert frac >= 0.9999
    assert rel <= 1e-5


def test_soup_5m_full_size_ray_casting_and_ao():
    """The BVH stress configuration at full size: a 5M-triangle soup
    (centroids uniform in the unit cube, lognormal edge lengths, random
    orientation), all 2048 x 2048 primary rays against the FP64 intersector
    (ids on >= 99.99 %, t within 1e-5 relative), and ambient occlusion with
    16 rays of radius 0.05 per hit on a 512 x 512 window against the oracle
    on the same random streams: >= 99.5 % of pixels equal and the window's
    mean within 0.1 %.  At this density rounding alone flips some AO rays
    near triangle edges and at the 0.05 limit: the oracle built with FMA
    contraction agrees with its own non-contracted build on 99.85 % of these
    pixels."""
    from tests.oracle_lib import OracleScene
    res = 2048
    sc = S.soup_scene(5_000_000, seed=7, width=res, height=res)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(3)
    r.setRenderingAlgorithm(abi.ALGO_RAY_CASTING)
    r.setup()
    rays, prim, tri, t = r.cast_primary(res, res, sample_index=0)
    r.setRenderingAlgorithm(abi.ALGO_AMBIENT_OCCLUSION)
    r.setSamples(16)
    r.setMaximumDistance(0.05)
    r.setSeed(8)
    r.setup()
    r.render(1)
    g, _ = r.film(res, res)
    r.dispose()
    o = OracleScene(sc)
    prim_o, tri_o, t_o = o.trace_f64(rays)
    frac, rel, nhit = _hit_parity(prim, tri, t, prim_o, tri_o, t_o)
    print(f"soup 5M at 2048^2: id match {frac:.6f}, max rel t {rel:.2e}, hits {nhit}")
    assert frac >= 0.9999 and rel <= 1e-5
    win = (768, 768, 1280, 1280)
    s_, q_, c_ = o.render(res, res, 1, algorithm=abi.ALGO_AMBIENT_OCCLUSION, samples=16, max_distance=0.05, seed=8,
                          window=win)
    ref = s_ / np.maximum(c_, 1)[..., None]
    x0, y0, x1, y1 = win
    gw, rw = g[y0:y1, x0:x1], ref[y0:y1, x0:x1]
    close = np.isclose(gw, rw, atol=1e-5).all(-1).mean()
    rel = abs(gw.mean() - rw.mean()) / rw.mean()
    print(f"soup 5M AO window: pixels equal {close:.5f}, mean rel diff {rel:.2e}")
    assert close >= 0.995
    assert rel < 1e-3


def test_trace_rays_closest_and_any_hit():
    """Random rays through a soup: closest hit matches the FP64 oracle, and
    any-hit reports a hit exactly when closest-hit does."""
    from tests.oracle_lib import OracleScene
    sc = S.soup_scene(50000, seed=5)
    rng = np.random.default_rng(1)
    n = 200000
    o = rng.random((n, 3)) * 1.4 - 0.2
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 0.001
    rays[:, 4:7] = d
    rays[:, 7] = rng.random(n) * 0.5 + 0.01
    r = _renderer()
    r.setScene(sc)
    r.setup()
    p, tr, t = r.trace_rays(rays)
    pa, _, _ = r.trace_rays(rays, any_hit=True)
    r.dispose()
    po, tro, to = OracleScene(sc).trace_f64(rays)
    frac, rel, nhit = _hit_parity(p, tr, t, po, tro, to)
    print(f"random rays: id match {frac:.6f}, max rel t {rel:.2e}, hits {nhit}")
    assert frac >= 0.9999 and rel <= 1e-5
    assert ((pa >= 0) == (p >= 0)).mean() >= 0.9999


@pytest.mark.parametrize("kernel", [0, 1])
def test_cornell_path_tracing_same_streams(kernel):
    """Wavefront (0) and plain (1) kernels against the oracle on the same
    random streams.

    Same streams give the same paths up to float rounding: a path leaves the
    oracle's only where rounding flips a decision (a Russian-roulette or
    Fresnel draw, a grazing edge).  How often rounding alone does that is
    fixed by the oracle itself: built with FMA contraction it agrees with its
    own non-contracted build on 99.0 % of 1-spp paths of this scene (and on
    84.9 % of 64x64 pixels at 32 spp).  Measured on MI355X: 98.77 % of paths
    equal, image means within 0.04 %.  Bars: >= 98.5 % of 1-spp paths equal
    (rtol 1e-3), and the 32-spp image mean within 0.2 %."""
    from tests.oracle_lib import OracleScene
    res = 128
    sc = S.cornell_box(res, res)
    o = OracleScene(sc)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(42)
    r.setKernel(kernel)
    r.setup()
    r.render(1)
    g1, cnt1 = r.film(res, res)
    s1, _, c1 = o.render(res, res, 1, seed=42)
    paths = np.isclose(g1, s1 / c1[..., None], rtol=1e-3, atol=1e-4).all(-1).mean()
    # 32 spp on a 64x64 film of the same scene
    res2, spp = 64, 32
    sc2 = S.cornell_box(res2, res2)
    r.setScene(sc2)
    r.setup()
    r.render(spp)
    g, cnt = r.film(res2, res2)
    st = r.stats()
    r.dispose()
    assert (cnt1 == 1).all() and (cnt == spp).all()
    s, q, c = OracleScene(sc2).render(res2, res2, spp, seed=42)
    ref = s / c[..., None]
    close = np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
    print(f"kernel {kernel}: 1-spp paths equal {paths:.4f}; 32 spp pixels equal {close:.4f}, "
          f"mean gpu {g.mean():.5f} ref {ref.mean():.5f}, {st}")
    assert paths >= 0.985
    assert abs(g.mean() - ref.mean()) / ref.mean() < 0.002


@pytest.mark.parametrize("name", ["plastic_perez_point", "matte_glass_metal_disney", "glossy_substrate_translucent_uber",
                                  "spot_directional"])
def test_materials_and_lights_same_streams(name):
    """Same-stream parity on the other BSDFs and lights: a Plastic icosphere
    on a Matte ground under a point light and the Perez sun/sky, the
    four-patch Matte/Glass/Metal/Disney icosphere under a point light or
    under a spot and a directional light, and the four-patch
    Glossy/Substrate/Translucent/Uber icosphere on a ClearCoat ground; both
    kernels.  Measured on MI355X (paths equal / mean difference): 99.99 % /
    7e-8, 99.88 % / 2.6e-4, 100 % / 2.2e-5, 99.97 % / 1.1e-5.  Bars just
    below: >= 99.9 % of 1-spp paths equal (99.8 % for the glass scene, whose
    Fresnel draws rounding flips most often; rtol 1e-3) and the 16-spp image
    mean within 0.1 %."""
    from tests.oracle_lib import OracleScene
    from dayflower_amd import workloads as W
    res = 96
    if name == "plastic_perez_point":
        sc = S.icosphere_scene(5000, seed=0x5EED, width=res, height=res, with_perez=True)
    elif name == "matte_glass_metal_disney":
        sc = W.mixed_material_icosphere(5000, 0x5EED, width=res, height=res)
    elif name == "spot_directional":
        sc = W.mixed_material_icosphere(5000, 0x5EED, width=res, height=res)
        sc.lights = []
        sc.add_spot_light((0.5, 3.0, -1.0), (-0.15, -1.0, 0.3), (40.0, 40.0, 40.0), cone_angle=35.0, cone_delta=10.0)
        sc.add_directional_light((-1.0, 1.5, -0.7), (1.5, 1.4, 1.2), radius=10.0)
    else:
        sc = W.more_materials_icosphere(5000, 0x5EED, width=res, height=res)
    o = OracleScene(sc)
    s1, _, c1 = o.render(res, res, 1, seed=21, max_bounce=8)
    s16, _, c16 = o.render(res, res, 16, seed=21, max_bounce=8)
    ref1, ref16 = s1 / c1[..., None], s16 / c16[..., None]
    for kernel in (0, 1):
        r = _renderer()
        r.setScene(sc)
        r.setSeed(21)
        r.setMaximumBounce(8)
        r.setKernel(kernel)
        r.setup()
        r.render(1)
        g1, _ = r.film(res, res)
        r.clear()
        r.render(16)
        g16, cnt = r.film(res, res)
        r.dispose()
        assert (cnt == 16).all()
        paths = np.isclose(g1, ref1, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(g16.mean() - ref16.mean()) / ref16.mean()
        print(f"{name} kernel {kernel}: 1-spp paths equal {paths:.4f}; 16-spp mean rel diff {rel:.2e}")
        assert paths >= (0.998 if name == "matte_glass_metal_disney" else 0.999)
        assert rel < 0.001


def test_cornell_statistical_parity():
    """Converged images on independent streams: the GPU film at 32768 spp
    against the oracle's at 8192 spp (48x48 pixels).

    Bars: >= 99 % of pixel channels and >= 98 % of pixels (all three
    channels) within 3 sigma, relative RMSE <= 1 %.  The oracle against
    itself on two seeds at these sample counts gives 99.54 % / 98.78 % and a
    relative RMSE of 0.58 %; a 1 % brightness bias drops the pixel fraction
    to 82 % and raises the RMSE to 1.14 %.  The test is sensitive enough to
    see a 0.06 % shift: recomputing every hit distance in FP64 before
    spawning the next ray (instead of the reference's float distance) fails
    it, which is why only the ray-query outputs are refined.  (The per-pixel fraction sits
    below 99 % for an unbiased renderer because the glass sphere's caustic
    paths make the sample variance heavy-tailed.)"""
    from tests.oracle_lib import OracleScene, film_agreement
    res, spp_g, spp_o = 48, 32768, 8192
    sc = S.cornell_box(res, res)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(1001)
    r.setup()
    r.render(spp_g)
    g, cnt = r.film(res, res)
    r.dispose()
    assert (cnt == spp_g).all()
    s, q, c = OracleScene(sc).render(res, res, spp_o, seed=2002)
    chan, pix, rrmse = film_agreement(g, spp_g, s, q, c)
    print(f"3-sigma: channels {chan:.4f}, pixels {pix:.4f}; relative RMSE {rrmse:.4f}")
    assert chan >= 0.99 and pix >= 0.98
    assert rrmse <= 0.01


def test_graph_replays_match_eager_launches():
    """The wavefront loop replays a captured HIP graph of four iterations
    between counter polls.  Three renders through the graph (each replaying it
    many times, with host synchronisations in between) equal the same render
    launched kernel by kernel, up to the order of the film's float atomics."""
    res, spp = 64, 16
    sc = S.icosphere_scene(5000, seed=0x5EED, width=res, height=res, with_perez=True)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(5)
    r.setUseGraphs(False)
    r.setup()
    r.render(spp)
    eager, _ = r.film(res, res)
    r.setUseGraphs(True)
    for k in range(3):
        r.clear()
        r.render(spp)
        g, cnt = r.film(res, res)
        assert (cnt == spp).all()
        np.testing.assert_allclose(g, eager, rtol=1e-4, atol=1e-6)
    r.dispose()


def test_ambient_occlusion_matches_oracle():
    from tests.oracle_lib import OracleScene
    res = 96
    sc = S.soup_scene(30000, seed=9, width=res, height=res)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(8)
    r.setRenderingAlgorithm(abi.ALGO_AMBIENT_OCCLUSION)
    r.setSamples(16)
    r.setMaximumDistance(0.05)
    r.setup()
    r.render(1)
    g, _ = r.film(res, res)
    r.dispose()
    s, q, c = OracleScene(sc).render(res, res, 1, algorithm=abi.ALGO_AMBIENT_OCCLUSION, samples=16,
                                     max_distance=0.05, seed=8)
    ref = s / np.maximum(c, 1)[..., None]
    close = np.isclose(g, ref, atol=1e-5).all(-1).mean()
    print(f"AO pixels equal {close:.4f}")
    assert close >= 0.999


def test_batch_alone_equals_in_mix():
    """A job renders the same alone and inside a packed batch of other jobs
    (counter-based RNG keyed by job seed, pixel and sample)."""
    from dayflower_amd.renderer import make_job
    r = _renderer()
    a = r.addScene(S.cornell_box(32, 32))
    b = r.addScene(S.icosphere_scene(5000, with_perez=False, width=48, height=32))
    cam_a = S.cornell_box(32, 32).camera
    cam_b = S.icosphere_scene(20, with_perez=False, width=48, height=32).camera
    job = make_job(a, cam_a, 32, 32, 8, seed=77)
    alone = r.batchRender([job])[0]
    mix = [make_job(b, cam_b, 48, 32, 4, seed=1), job, make_job(a, cam_a, 16, 16, 16, seed=9),
           make_job(b, cam_b, 24, 24, 2, seed=3)]
    films = r.batchRender(mix)
    plain = r.batchRender([job], mode=abi.BATCH_PLAIN)[0]
    r.dispose()
    np.testing.assert_allclose(films[1], alone, rtol=1e-4, atol=1e-5)
    close = np.isclose(plain, alone, rtol=1e-3, atol=1e-4).all(-1).mean()
    assert close > 0.9


def test_server_alone_equals_in_mix():
    """Requests submitted while others are on the GPU are admitted as memory
    frees up; every result equals the same job rendered alone (packed batch of
    one), within float-atomic reordering tolerance."""
    import time
    from dayflower_amd.renderer import Server, make_job
    r = _renderer()
    sid = r.addScene(S.icosphere_scene(20000, with_perez=True, width=64, height=64))
    cam = S.icosphere_scene(20, with_perez=False, width=64, height=64).camera
    jobs = [make_job(sid, cam, 32 + 16 * (i % 3), 32, [1, 4, 8][i % 3], seed=100 + i) for i in range(12)]
    alone = [r.batchRender([j])[0] for j in jobs]
    # budget admits only a few films at a time, so admission waits for frees
    pool = 1 << 16
    pool_bytes = Server.pool_bytes(pool)  # path SoA, queues and counters
    srv = Server(r, hbm_budget_bytes=pool_bytes + 3 * 64 * 32 * 16, max_active=4, pool_paths=pool)
    ids = []
    for j in jobs:
        ids.append(srv.submit(j))
        time.sleep(0.002)
    films = [srv.wait(i) for i in ids]
    lat = [srv.latency(i) for i in ids]
    st = srv.stats()
    srv.close()
    r.dispose()
    for f, a in zip(films, alone):
        np.testing.assert_allclose(f, a, rtol=1e-4, atol=1e-5)
    assert st["completed"] == len(jobs) and st["peak_active"] <= 4
    assert all(x > 0 for x in lat)


def test_rgba8_image_matches_oracle_for_every_tone_mapper():
    """dfr_read_rgba8 against the oracle's image stage applied to the same
    film (dfr_read_film): every tone mapper under path tracing, and a tone
    mapper set under ambient occlusion (which only sRGB-encodes).  The device
    and libm pow/exp may differ in the last ulp, which can move a value
    across a quantisation step: bar >= 99.99 % of bytes equal, none off by
    more than one."""
    from tests import oracle_lib
    res = 128
    sc = S.cornell_box(res, res)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(5)
    r.setup()
    assert r.getToneMapper() == abi.TONE_MAPPER_NONE
    r.render(2)
    g, _ = r.film(res, res)
    print(f"film max {g.max():.3f}, fraction above one {(g > 1).mean():.4f}", flush=True)
    assert (g > 1.0).any()
    images = []
    for op in range(6):
        r.setToneMapper(op)
        img = r.rgba8(res, res)
        ref = oracle_lib.image_rgba8(g, op)
        d = np.abs(img.astype(np.int32) - ref)
        print(f"tone mapper {op}: bytes equal {(d == 0).mean():.6f}, max diff {d.max()}", flush=True)
        assert (d == 0).mean() >= 0.9999 and d.max() <= 1
        images.append(img.tobytes())
    assert len(set(images)) == 6
    r.setRenderingAlgorithm(abi.ALGO_AMBIENT_OCCLUSION)
    r.setToneMapper(abi.TONE_MAPPER_REINHARD)
    r.render(1)
    g, _ = r.film(res, res)
    img = r.rgba8(res, res)
    r.dispose()
    ref = oracle_lib.image_rgba8(g, abi.TONE_MAPPER_REINHARD, abi.ALGO_AMBIENT_OCCLUSION)
    d = np.abs(img.astype(np.int32) - ref)
    assert (d == 0).mean() >= 0.9999 and d.max() <= 1


def test_path_pool_size_does_not_change_the_image():
    """A path's result depends only on its random stream, so pools of 1024
    and 5000 paths (many refills, partial sort tiles) and the default pool
    give the same film up to the order of the film's float additions."""
    res = 64
    sc = S.cornell_box(res, res)
    films = []
    for pool in (1024, 5000, None):
        r = _renderer()
        r.setScene(sc)
        r.setSeed(17)
        if pool:
            r.setWavefront(pool_paths=pool)
        r.setup()
        r.render(4)
        g, cnt = r.film(res, res)
        r.dispose()
        assert (cnt == 4).all()
        films.append(g)
    for g in films[:2]:
        np.testing.assert_allclose(g, films[2], rtol=1e-4, atol=1e-6)


def test_traversal_counts_match_the_oracle_walk():
    """The node visits and triangle tests the roofline prices: a counting
    render records every 16th traced ray with the GPU's per-ray counts; the
    oracle walks the exported BVH over the same rays in FP64 and finds the
    same closest hits (>= 99.99 %) and occlusions.  The GPU's speculative
    traversal visits at least as many nodes as the plain depth-first walk and
    at most 40 % more; it tests the same triangles within 15 %."""
    from tests import oracle_lib
    res = 128
    sc = S.icosphere_scene(200_000, seed=0x5EED, width=res, height=res, with_perez=True)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(3)
    r.setMaximumBounce(8)
    r.setup()
    r.setCountTraversal(True, sample_stride=16)
    r.render(4)
    st = r.stats()
    smp = r.ray_sample()
    nodes, tris = r.bvh()
    r.setCountTraversal(False)
    r.dispose()
    n = len(smp["o"])
    assert n > 5000 and nodes.shape[0] > 1000 and tris.shape[0] == sc.n_triangles
    rec = np.concatenate([smp["o"], smp["tmax"][:, None], smp["d"], smp["any"].view(np.float32)[:, None],
                          np.zeros((n, 4), np.float32)], 1)
    nv, nt, hit, _ = oracle_lib.bvh4_walk(nodes, tris, rec)
    for kind in (0, 1):
        sel = smp["any"] == kind
        assert sel.sum() > 1000
        g_nodes, g_tris = smp["nodes"][sel].mean(), smp["tris"][sel].mean()
        o_nodes, o_tris = nv[sel].mean(), nt[sel].mean()
        print(f"{'any' if kind else 'closest'} hit: {sel.sum()} rays; nodes GPU {g_nodes:.3f} oracle {o_nodes:.3f}; "
              f"triangles GPU {g_tris:.3f} oracle {o_tris:.3f}", flush=True)
        assert o_nodes <= g_nodes * 1.001 and g_nodes <= 1.4 * o_nodes
        assert abs(g_tris - o_tris) <= 0.15 * o_tris
        if kind == 0:
            agree = (hit[sel] == smp["tri"][sel]).mean()
        else:
            agree = ((hit[sel] >= 0) == (smp["prim"][sel] >= 0)).mean()
        print(f"  hit agreement {agree:.6f}")
        assert agree >= 0.9999
    # the sampled rays' GPU counts are a fair sample of the whole run's
    ext = smp["any"] == 0
    assert abs(smp["nodes"][ext].mean() / (st["extend_nodes"] / st["extend_rays"]) - 1.0) < 0.1


def test_benched_frame_full_size_same_streams():
    """The benched frame itself: the 1M-triangle displaced icosphere at
    1920x1080 with the default 64M-path pool (4096 sort tiles), 2 spp, against
    the oracle on a 320x320 window at the image centre on the same streams.
    Bars: >= 99.5 % of pixels equal (rtol 1e-3; measured 99.85 % at 64 spp on
    a 745x745 window) and the window mean within 0.1 %; every pixel counts 2
    samples."""
    from tests.oracle_lib import OracleScene
    w, h, spp = 1920, 1080, 2
    sc = S.icosphere_scene(1_000_000, seed=0x5EED, width=w, height=h, with_perez=True)
    r = _renderer()
    r.setScene(sc)
    r.setSeed(0x5EED)
    r.setMaximumBounce(8)
    r.setMinimumBounceRussianRoulette(5)
    r.setup()
    assert r.getWavefront()["pool_paths"] == 1 << 26
    r.render(spp)
    g, cnt = r.film(w, h)
    r.dispose()
    assert (cnt == spp).all()
    x0, y0, side = w // 2 - 160, h // 2 - 160, 320
    o = OracleScene(sc)
    s, _, c = o.render(w, h, spp, seed=0x5EED, max_bounce=8, min_rr=5, window=(x0, y0, x0 + side, y0 + side))
    o.close()
    ref = (s / np.maximum(c, 1)[..., None])[y0:y0 + side, x0:x0 + side]
    gw = g[y0:y0 + side, x0:x0 + side]
    px = np.isclose(gw, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
    rel = abs(gw.mean() - ref.mean()) / ref.mean()
    print(f"full-size frame: window pixels equal {px:.5f}, mean rel diff {rel:.2e}", flush=True)
    assert px >= 0.995
    assert rel < 0.001


def _jobs_vs_oracle(sc, jobs, films):
    """(pixels equal, relative mean difference) of each job's film against the
    oracle's render of the same streams (one oracle scene for all jobs)."""
    from tests.oracle_lib import OracleScene
    o = OracleScene(sc)
    out = []
    for job, film in zip(jobs, films):
        o.set_camera(np.array(job.camera[:], np.float32))
        s, _, c = o.render(job.width, job.height, job.spp, seed=job.seed, max_bounce=job.maximum_bounce,
                           min_rr=job.minimum_bounce_rr)
        ref = s / np.maximum(c, 1)[..., None]
        px = np.isclose(film, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(film.mean() - ref.mean()) / max(ref.mean(), 1e-12)
        print(f"job {job.width}x{job.height} {job.spp} spp (scene {job.scene}): pixels equal {px:.5f}, "
              f"mean rel diff {rel:.2e}", flush=True)
        out.append((px, rel))
    o.close()
    return out


def test_batched_jobs_match_the_oracle():
    """The 64-job batched workload (4 mixed-material icospheres of 10k, 100k,
    500k and 2M triangles; Matte, Glass, Metal and Disney patches): the first
    two jobs of each scene, at their own resolutions with spp capped at 2,
    rendered together in one packed batch; each job's film against the
    oracle on the same streams.  Calibration by the oracle against its own
    FMA-contracted build on the serving workload's jobs (rounding alone):
    98.3-99.4 % of 2-spp pixels equal, means within 0.17 %; measured on
    MI355X: 98.7-99.4 % and 2.2e-4..7.7e-4.  Bars: >= 98 % of pixels equal
    (rtol 1e-3) and the mean within 0.3 %."""
    from dayflower_amd import workloads as W
    scenes = W.batched_scenes()
    r = _renderer()
    ids = [r.addScene(sc) for sc in scenes]
    jobs = W.batched_jobs(ids)
    pick = []
    for sid in ids:
        pick += [j for j in jobs if j.scene == sid][:2]
    for j in pick:
        j.spp = min(j.spp, 2)
    films = r.batchRender(pick)
    r.dispose()
    assert sorted({j.scene for j in pick}) == sorted(ids)
    for k, sid in enumerate(ids):
        sel = [i for i, j in enumerate(pick) if j.scene == sid]
        for px, rel in _jobs_vs_oracle(scenes[k], [pick[i] for i in sel], [films[i] for i in sel]):
            assert px >= 0.98 and rel < 0.003


def test_served_requests_match_the_oracle():
    """The serving workload's first 20 requests (2M-triangle mixed-material
    icosphere, cameras on a radius-3 sphere, resolutions 256^2..1024^2) with
    spp capped at 2, submitted to the server together: every served film
    against the oracle on the same streams.  Measured on MI355X: 98.2-99.5 %
    of pixels equal, means within 1.2e-3.  Bars as for the batch."""
    from dayflower_amd import workloads as W
    from dayflower_amd.renderer import Server
    sc = W.mixed_material_icosphere(2_000_000, 0x5EED)
    r = _renderer()
    sid = r.addScene(sc)
    stream = W.request_stream(sid, duration=0.2)[:20]
    assert len(stream) == 20
    jobs = [j for _, j in stream]
    for j in jobs:
        j.spp = min(j.spp, 2)
    srv = Server(r, hbm_budget_bytes=8 << 30, max_active=8, pool_paths=1 << 22)
    ids = [srv.submit(j) for j in jobs]
    films = [srv.wait(i) for i in ids]
    srv.close()
    r.dispose()
    for px, rel in _jobs_vs_oracle(sc, jobs, films):
        assert px >= 0.98 and rel < 0.003


def test_server_reuses_slots_and_leaves_the_context_free():
    """A server with 3 job slots serves 10 requests (slots of finished
    requests are reused); every film equals the same job rendered alone.
    While it serves, the context renders its own frame and compiles another
    scene (the server holds the context only while it enqueues a round), and
    a scene with requests on it cannot be removed."""
    import threading
    from dayflower_amd.renderer import IllegalStateException, Server, make_job
    r = _renderer()
    sc = S.icosphere_scene(20000, with_perez=True, width=48, height=48)
    sid = r.addScene(sc)
    cam = sc.camera
    jobs = [make_job(sid, cam, 32 + 8 * (i % 3), 24 + 8 * (i % 2), [2, 4, 8][i % 3], seed=300 + i) for i in range(10)]
    alone = [r.batchRender([j])[0] for j in jobs]
    r.setScene(S.cornell_box(32, 32))
    r.setSeed(1)
    r.setup()
    srv = Server(r, hbm_budget_bytes=Server.pool_bytes(1 << 16) + (64 << 20), max_active=3, pool_paths=1 << 16,
                 slot_capacity=3)
    ids = [srv.submit(j) for j in jobs]
    with pytest.raises(IllegalStateException):
        r.removeScene(sid)
    # the context's own work interleaves with the serving rounds
    r.render(8)
    own, cnt = r.film(32, 32)
    extra = r.addScene(S.icosphere_scene(2000, with_perez=False, width=16, height=16))
    films = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    r.removeScene(sid)
    r.removeScene(extra)
    r.dispose()
    assert (cnt == 8).all() and own.mean() > 0
    assert st["completed"] == len(jobs) and st["peak_active"] <= 3
    for f, a in zip(films, alone):
        np.testing.assert_allclose(f, a, rtol=1e-4, atol=1e-5)
