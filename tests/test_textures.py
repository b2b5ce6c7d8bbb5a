"""Textures on material colours (M/scene/texture/*Texture.java getColor; the
materials' computeBSDF read every colour through its texture) and the
texture coordinates of hits (Triangle3F interpolation, Sphere3F, Rectangle3F
.java:448-475, Plane3F.java:280-290).

Oracle pins (CPU): a Matte quad lit by a directional light along the view
renders kd / pi x E x cos per pixel, deterministically, so the texture's value
at every hit is read back and compared with the closed form of the texture at
the hit's coordinates: the interpolated vertex coordinates of a triangle
pair, a sphere's (phi / 2 pi, (theta - pi) / -pi), an axis-aligned
rectangle's linear map, a checkerboard's cells, an image's bilinear texels,
a blend.  A texture that is a constant renders bit-identically to the
constant itself.

GPU (through the C-ABI): textured Matte, Plastic, Mirror and Glass over
triangles, spheres, rectangles and planes on the same random streams as the
oracle, both kernels; the textures the kernels cannot follow are refused.
"""
import numpy as np
import pytest

from dayflower_amd import abi
from dayflower_amd import scene as S

E = 2.0   # directional light radiance


def _lit_scene(res, add_geometry):
    sc = S.Scene()
    add_geometry(sc)
    sc.add_directional_light((0.0, 0.0, -1.0), (E, E, E), radius=10.0)
    sc.set_camera(S.camera_record((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), res, res, fov_x_deg=40.0))
    return sc


def _direct(sc, res, seed=7):
    """Per-pixel radiance with the direct light only (max bounce 1: the
    continuation ray leaves the scene and sees nothing), the primary rays and
    their hits."""
    from tests.oracle_lib import OracleScene
    o = OracleScene(sc)
    s, _, c = o.render(res, res, 1, seed=seed, max_bounce=1, min_rr=50)
    rays = o.primary_rays(res, res, seed, 0)
    prim, _, t = o.trace_f32(rays)
    o.close()
    img = (s / np.maximum(c, 1)[..., None]).reshape(-1, 3)
    p = rays[:, 0:3] + rays[:, 4:7] * t[:, None]
    return img, rays, prim, p


def _kd_from(img, rays, normal):
    """kd = L pi / (E cos), cos between the surface normal and the light
    direction (the light shines along +z onto a surface facing -z)."""
    cos = np.abs(normal[..., 2]) if normal.ndim > 1 else abs(normal[2])
    return img * np.pi / (E * cos if np.ndim(cos) == 0 else (E * cos)[:, None])


def _quad(sc, material, uv=True):
    p = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.int32)   # facing -z, toward the camera
    uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32) if uv else None
    sc.add_mesh(p, t, material, normals=np.tile([0, 0, -1], (4, 1)).astype(np.float32), uvs=uvs)


def test_uv_texture_on_triangles_reads_the_interpolated_coordinates():
    res = 32

    def geo(sc):
        m = sc.add_material(S.matte((1, 1, 1)))
        sc.set_texture(m, abi.SLOT_KD, sc.uv_texture())
        _quad(sc, m)
    img, rays, prim, p = _direct(_lit_scene(res, geo), res)
    hit = prim >= 0
    kd = _kd_from(img[hit], rays[hit], np.array([0, 0, -1.0]))
    expect = np.stack([(p[hit, 0] + 1) / 2, (p[hit, 1] + 1) / 2, np.zeros(hit.sum())], -1)
    print(f"hits {hit.sum()}; max |kd - (u, v, 0)| {np.abs(kd - expect).max():.2e}")
    assert hit.sum() > res * res // 4
    np.testing.assert_allclose(kd, expect, atol=2e-5)


def test_checkerboard_cells():
    res = 48

    def geo(sc):
        m = sc.add_material(S.matte((1, 1, 1)))
        a = sc.constant_texture((0.9, 0.2, 0.1))
        b = sc.constant_texture((0.1, 0.3, 0.8))
        sc.set_texture(m, abi.SLOT_KD, sc.checkerboard_texture(a, b, 0.0, (4.0, 4.0)))
        _quad(sc, m)
    img, rays, prim, p = _direct(_lit_scene(res, geo), res)
    hit = prim >= 0
    kd = _kd_from(img[hit], rays[hit], np.array([0, 0, -1.0]))
    u, v = (p[hit, 0] + 1) / 2, (p[hit, 1] + 1) / 2
    fu, fv = np.mod(u * 4, 1.0), np.mod(v * 4, 1.0)
    away = (np.abs(fu - 0.5) > 1e-3) & (np.abs(fv - 0.5) > 1e-3) & (fu > 1e-3) & (fv > 1e-3)
    is_a = (fu > 0.5) ^ (fv > 0.5)
    expect = np.where(is_a[:, None], [0.9, 0.2, 0.1], [0.1, 0.3, 0.8])
    print(f"cells away from the edges: {away.sum()} of {hit.sum()}; cells of a: {is_a[away].mean():.3f}")
    np.testing.assert_allclose(kd[away], expect[away], rtol=1e-5)
    assert 0.3 < is_a[away].mean() < 0.7


def test_image_texture_bilinear_and_blend():
    """A 2 x 2 image over the quad: u * 2 - 0.5 and v * 2 - 0.5 are the image
    coordinates, wrapped; a blend of it with a constant mixes per component."""
    res = 32
    img2 = np.array([[[0.2, 0.4, 0.6], [0.8, 0.1, 0.3]], [[0.5, 0.9, 0.0], [0.1, 0.6, 0.7]]], np.float32)
    t3 = (0.25, 0.5, 0.75)

    def geo(sc):
        m = sc.add_material(S.matte((1, 1, 1)))
        im = sc.image_texture(img2)
        k = sc.constant_texture((1.0, 1.0, 1.0))
        sc.set_texture(m, abi.SLOT_KD, sc.blend_texture(im, k, t3))
        _quad(sc, m)
    img, rays, prim, p = _direct(_lit_scene(res, geo), res)
    hit = prim >= 0
    kd = _kd_from(img[hit], rays[hit], np.array([0, 0, -1.0]))
    u, v = (p[hit, 0] + 1) / 2, (p[hit, 1] + 1) / 2
    x, y = np.mod(u * 2 - 0.5, 2), np.mod(v * 2 - 0.5, 2)
    x0, y0 = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - x0, y - y0
    T = lambda a, b: img2[b % 2, a % 2]
    c = ((1 - fy)[:, None] * ((1 - fx)[:, None] * T(x0, y0) + fx[:, None] * T(x0 + 1, y0)) +
         fy[:, None] * ((1 - fx)[:, None] * T(x0, y0 + 1) + fx[:, None] * T(x0 + 1, y0 + 1)))
    expect = (1 - np.array(t3)) * c + np.array(t3) * 1.0
    print(f"max |kd - blend(bilinear, 1)| {np.abs(kd - expect).max():.2e}")
    np.testing.assert_allclose(kd, expect, atol=5e-5)


def test_sphere_and_rectangle_coordinates():
    """UV texture on a sphere: (phi / 2 pi, (theta - pi) / -pi) of the hit
    relative to the centre; on the rectangle a(-1,-1,0), b(-1,1,0),
    c(1,1,0) (d = a + c - b): the projection formula gives u = (y + 1) / 2,
    v = (x + 1) / 2."""
    res = 32
    # sphere
    def geo_s(sc):
        m = sc.add_material(S.matte((1, 1, 1)))
        sc.set_texture(m, abi.SLOT_KD, sc.uv_texture())
        sc.add_sphere((0.0, 0.0, 0.0), 1.0, m)
    img, rays, prim, p = _direct(_lit_scene(res, geo_s), res)
    hit = prim >= 0
    n = p[hit] / np.linalg.norm(p[hit], axis=1, keepdims=True)
    kd = img[hit] * np.pi / (E * np.abs(n[:, 2]))[:, None]
    phi = np.mod(np.arctan2(p[hit, 1], p[hit, 0]), 2 * np.pi)
    th = np.arccos(np.clip(p[hit, 2] / np.linalg.norm(p[hit], axis=1), -1, 1))
    expect = np.stack([phi / (2 * np.pi), (th - np.pi) / -np.pi, np.zeros(hit.sum())], -1)
    ok = np.abs(n[:, 2]) > 0.2
    print(f"sphere: max |kd - uv| {np.abs(kd - expect)[ok].max():.2e} over {ok.sum()} hits")
    np.testing.assert_allclose(kd[ok], expect[ok], atol=5e-4)
    # rectangle
    def geo_r(sc):
        m = sc.add_material(S.matte((1, 1, 1)))
        sc.set_texture(m, abi.SLOT_KD, sc.uv_texture())
        sc.add_rectangle((-1, -1, 0), (-1, 1, 0), (1, 1, 0), m)
    img, rays, prim, p = _direct(_lit_scene(res, geo_r), res)
    hit = prim >= 0
    kd = _kd_from(img[hit], rays[hit], np.array([0, 0, -1.0]))
    expect = np.stack([(p[hit, 1] + 1) / 2, (p[hit, 0] + 1) / 2, np.zeros(hit.sum())], -1)
    print(f"rectangle: max |kd - uv| {np.abs(kd - expect).max():.2e} over {hit.sum()} hits")
    np.testing.assert_allclose(kd, expect, atol=5e-5)


def test_constant_texture_equals_the_constant():
    from tests.oracle_lib import OracleScene
    res = 24
    a = S.cornell_box(res, res)
    b = S.cornell_box(res, res)
    for m in range(len(b.materials)):
        mat = b.materials[m]
        slots = {abi.MAT_MATTE: [(abi.SLOT_KD, mat.kd)], abi.MAT_MIRROR: [(abi.SLOT_KR, mat.kr)],
                 abi.MAT_GLASS: [(abi.SLOT_KR, mat.kr), (abi.SLOT_KT, mat.kt)]}.get(mat.type, [])
        for slot, col in slots:
            b.set_texture(m, slot, b.constant_texture(tuple(col)))
    out = []
    for sc in (a, b):
        o = OracleScene(sc)
        s, _, c = o.render(res, res, 4, seed=5, max_bounce=6)
        o.close()
        out.append((s, c))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _textured_scene(res):
    sc = S.Scene()
    white = sc.constant_texture((0.8, 0.8, 0.75))
    dark = sc.constant_texture((0.1, 0.12, 0.15))
    floor = sc.add_material(S.matte((0.8, 0.8, 0.8)))
    sc.set_texture(floor, abi.SLOT_KD, sc.checkerboard_texture(white, dark, 15.0, (6.0, 6.0)))
    sc.add_rectangle((-3, 0, -3), (-3, 0, 3), (3, 0, 3), floor)
    # an icosphere with spherical texture coordinates and an image texture on Plastic kd
    p, t, n = S.displaced_icosphere(2000, 0x5EED, radius=0.5, center=(-0.6, 0.55, 0.2))
    q = p - np.array([-0.6, 0.55, 0.2], np.float32)
    uv = np.stack([np.mod(np.arctan2(q[:, 2], q[:, 0]) / (2 * np.pi), 1.0),
                   np.arccos(np.clip(q[:, 1] / np.linalg.norm(q, axis=1), -1, 1)) / np.pi], -1)
    rng = np.random.default_rng(9)
    image = rng.random((16, 32, 3)).astype(np.float32)
    image[:, ::4] = 0.0          # black columns: the Lambertian lobe drops out there
    plast = sc.add_material(S.plastic((0.8, 0.8, 0.8), (0.25, 0.25, 0.25), 0.3))
    sc.set_texture(plast, abi.SLOT_KD, sc.image_texture(image, 10.0, (2.0, 1.0)))
    sc.add_mesh(p, t, plast, normals=n, uvs=uv)
    # a sphere with polka dots on Plastic kd and ks
    pd = sc.add_material(S.plastic((0.8, 0.8, 0.8), (0.3, 0.3, 0.3), 0.2))
    red, blue = sc.constant_texture((0.8, 0.1, 0.1)), sc.constant_texture((0.1, 0.2, 0.7))
    sc.set_texture(pd, abi.SLOT_KD, sc.polka_dot_texture(red, blue, 30.0, 8.0, 0.3))
    sc.set_texture(pd, abi.SLOT_KS, sc.dot_product_texture())
    sc.add_sphere((0.6, 0.4, -0.2), 0.4, pd)
    # glass with a blended transmittance, a mirror with a bullseye and a UV-coloured plane behind
    gl = sc.add_material(S.glass(1.5))
    sc.set_texture(gl, abi.SLOT_KT, sc.blend_texture(sc.surface_normal_texture(), white, (0.3, 0.6, 0.9)))
    sc.add_sphere((0.1, 0.25, -0.8), 0.25, gl)
    mi = sc.add_material(S.mirror((1, 1, 1)))
    sc.set_texture(mi, abi.SLOT_KR, sc.bullseye_texture(white, dark, (0.0, 1.0, 1.5), 3.0))
    sc.add_rectangle((-1.5, 0.1, 1.4), (-1.5, 1.6, 1.6), (0.0, 1.6, 1.6), mi)
    back = sc.add_material(S.matte((0.8, 0.8, 0.8)))
    sc.set_texture(back, abi.SLOT_KD, sc.uv_texture())
    sc.add_plane((0.0, 0.0, 3.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, -1.0), back)
    sc.add_point_light((1.5, 3.0, -1.5), (18.0, 18.0, 18.0))
    sc.add_perez_light(2.5, (0.5, 1.0, -0.7))
    sc.set_camera(S.camera_record((0.0, 1.5, -3.2), (0.0, 0.4, 0.0), res, res, fov_x_deg=50.0))
    return sc


def _same_streams(sc, res, seed, kernel, spp_list=(1, 16)):
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    o = OracleScene(sc)
    refs = []
    for spp in spp_list:
        s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
        refs.append((s / np.maximum(c, 1)[..., None], c))
    o.close()
    out = []
    for spp, (ref, c) in zip(spp_list, refs):
        r = GPURenderer()
        r.setScene(sc)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setKernel(kernel)
        r.setup()
        r.render(spp)
        g, k = r.film(res, res)
        r.dispose()
        np.testing.assert_array_equal(k, c)
        gc, rc = np.minimum(g, CLAMP), np.minimum(ref, CLAMP)
        diff = (gc - rc).reshape(-1)
        z = diff.mean() / max(diff.std() / np.sqrt(diff.size), 1e-30)
        out.append((np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean(), abs(g.mean() - ref.mean()) / ref.mean(),
                    abs(gc.mean() - rc.mean()) / rc.mean(), z))
    return out


# Per-channel clamp of the mean comparison.  The untextured scene's pixels stay
# below 2.0 at 16 spp, so the clamp leaves it alone.  In the textured scene the
# UV texture on the infinite back plane gives it albedos above 1 (a plane's
# texture coordinates are unbounded projections, Plane3F.java:288, and
# UVTexture returns them as red and green), and the glass sphere focuses the
# sun onto the floor: one pixel holds about half of the image's total (a
# single sample of radiance near 1.6e5), so a rounding-level difference in one
# diverged path moves the raw mean by 1e-3 while saying nothing about the
# textures.
CLAMP = 8.0


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_textured_scene_same_streams(kernel):
    """The glass and mirror send rounding-level differences into other paths
    (the same scene with every texture removed measures 95.3 % of 1-spp
    paths equal), so the textured scene is held to that baseline: its paths
    agree as often (measured 95.4 %; a wrong texture value moves every path
    that sees it by more than the rtol 1e-3 of path equality).  What the
    diverged paths add to the 16-spp image is noise, not bias: the paired
    per-pixel difference of the clamped images has a mean within 4 standard
    errors of zero (measured |z| 0.9 textured, 0.2 without) and within 1 %."""
    res, seed = 48, 51
    sc = _textured_scene(res)
    (p1, _, _, _), (_, rel, crel, z) = _same_streams(sc, res, seed, kernel)
    plain = _textured_scene(res)
    for m in plain.materials:
        for k in range(4):
            m.textures[k] = 0
    (q1, _, _, _), (_, qrel, qcrel, qz) = _same_streams(plain, res, seed, kernel)
    print(f"kernel {kernel}: 1-spp paths equal {p1:.5f} textured, {q1:.5f} without textures "
          f"(bar: no more than 0.005 below, and 0.94); 16-spp clamped mean rel diff {crel:.2e} textured, "
          f"{qcrel:.2e} without (bar 1e-2), paired z {z:+.2f} textured, {qz:+.2f} without (bar |z| < 4); "
          f"raw means {rel:.2e}, {qrel:.2e}", flush=True)
    assert p1 >= 0.94 and p1 >= q1 - 0.005
    assert abs(z) < 4 and abs(qz) < 4
    assert crel < 1e-2 and qcrel < 1e-3


@pytest.mark.gpu
def test_textures_the_kernels_cannot_follow_are_refused():
    from dayflower_amd.renderer import GPURenderer, IllegalArgumentException, UnsupportedOperationException
    cases = []
    sc = S.cornell_box(16, 16)                     # a texture on a metal
    m = sc.add_material(S.metal())
    sc.set_texture(m, abi.SLOT_KD, sc.uv_texture())
    cases.append((sc, UnsupportedOperationException))
    sc = S.cornell_box(16, 16)                     # a blend of blends
    k = sc.constant_texture((1, 1, 1))
    b1 = sc.blend_texture(k, k)
    sc.set_texture(0, abi.SLOT_KD, sc.blend_texture(b1, k))
    cases.append((sc, UnsupportedOperationException))
    sc = S.cornell_box(16, 16)                     # a checkerboard of itself
    c = sc.checkerboard_texture(0, 0)
    sc.textures[c].a = c
    sc.set_texture(0, abi.SLOT_KD, c)
    cases.append((sc, IllegalArgumentException))
    r = GPURenderer()
    for s, exc in cases:
        with pytest.raises(exc):
            r.setScene(s)
    r.dispose()


def _glowing_quad(res, kind):
    """A black Matte quad whose emission is a texture (textureEmission):
    UV colours or a checkerboard of two constants."""
    def geo(sc):
        m = sc.add_material(S.matte((0.0, 0.0, 0.0)))
        if kind == "uv":
            t = sc.uv_texture()
        else:
            t = sc.checkerboard_texture(sc.constant_texture((3.0, 0.5, 0.25)), sc.constant_texture((0.2, 1.5, 0.0)),
                                        scale=(4.0, 4.0))
        sc.set_emission_texture(m, t)
        _quad(sc, m)
    return _lit_scene(res, geo)


@pytest.mark.parametrize("kind", ["uv", "checkerboard"])
def test_emission_texture_is_what_a_camera_ray_sees(kind):
    """With no bounce the image is the emission at each primary hit: the quad's
    UV coordinates ((x + 1) / 2, (y + 1) / 2, 0) or the checkerboard's two
    colours by cell."""
    from tests.oracle_lib import OracleScene
    res, seed = 48, 3
    sc = _glowing_quad(res, kind)
    o = OracleScene(sc)
    s, _, c = o.render(res, res, 1, seed=seed, max_bounce=0)
    rays = o.primary_rays(res, res, seed, 0)
    prim, _, t = o.trace_f32(rays)
    o.close()
    img = (s / np.maximum(c, 1)[..., None]).reshape(-1, 3)
    hit = prim >= 0
    p = (rays[:, 0:3] + rays[:, 4:7] * t[:, None])[hit]
    u, v = (p[:, 0] + 1) / 2, (p[:, 1] + 1) / 2
    if kind == "uv":
        want = np.stack([u, v, np.zeros_like(u)], 1)
        err = np.abs(img[hit] - want).max()
        assert err < 2e-6, err
    else:
        cu, cv = np.modf(u * 4)[0] > 0.5, np.modf(v * 4)[0] > 0.5
        away = (np.abs(np.modf(u * 4)[0] - 0.5) > 0.01) & (np.abs(np.modf(v * 4)[0] - 0.5) > 0.01) & \
               (np.modf(u * 4)[0] > 0.01) & (np.modf(v * 4)[0] > 0.01)
        want = np.where((cu ^ cv)[:, None], [3.0, 0.5, 0.25], [0.2, 1.5, 0.0])
        assert np.array_equal(img[hit][away], want[away].astype(np.float32))
    assert np.all(img[~hit] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_emission_texture_on_the_gpu(kernel):
    """The glowing quads on both kernels against the oracle per pixel, and
    the same scene through the reference's compiled arrays."""
    from dayflower_amd import renderer as R
    from oracle import scene_compiler as SC
    from tests.oracle_lib import OracleScene
    res, seed = 48, 3
    for kind in ("uv", "checkerboard"):
        sc = _glowing_quad(res, kind)
        o = OracleScene(sc)
        s, _, c = o.render(res, res, 4, seed=seed, max_bounce=2)
        o.close()
        ref = s / np.maximum(c, 1)[..., None]
        for compiled in (False, True):
            r = R.GPURenderer()
            if compiled:
                r.setCompiledScene(SC.compatible(SC.compile_scene(sc)))
            else:
                r.setScene(sc)
            r.setSeed(seed)
            r.setMaximumBounce(2)
            r.setKernel(kernel)
            r.setup()
            r.render(4)
            g, k = r.film(res, res)
            r.dispose()
            np.testing.assert_array_equal(k, c)
            eq = np.isclose(g, ref, rtol=1e-4, atol=1e-5).all(-1).mean()
            print(f"emission {kind} kernel {kernel} compiled {compiled}: pixels equal {eq:.4f}", flush=True)
            assert eq >= 0.999


def _rough_scene(res, constant_textures=False, constants=False):
    """Metal and Plastic spheres whose roughness (and Metal's eta and k) come
    from textures: a checkerboard of two roughnesses, UV-coloured eta, a
    blended k; with constant_textures the same parameters from constant
    textures of the values the constants would have (constants: no
    textures at all)."""
    sc = S.Scene()
    floor = sc.add_material(S.matte((0.6, 0.6, 0.6)))
    metal = sc.add_material(S.metal(roughness=0.25))
    plastic = sc.add_material(S.plastic((0.2, 0.4, 0.7), (0.5, 0.5, 0.5), 0.5))
    sc.materials[metal].remap_roughness = 1
    M = sc.materials[metal]
    if constant_textures:
        sc.set_texture(metal, abi.SLOT_ROUGHNESS_U, sc.constant_texture((0.25, 0.25, 0.25)))
        sc.set_texture(metal, abi.SLOT_ROUGHNESS_V, sc.constant_texture((M.roughness_v,) * 3))
        sc.set_texture(metal, abi.SLOT_ETA, sc.constant_texture(tuple(M.eta)))
        sc.set_texture(metal, abi.SLOT_K, sc.constant_texture(tuple(M.k)))
        sc.set_texture(plastic, abi.SLOT_ROUGHNESS_U, sc.constant_texture((0.5, 0.5, 0.5)))
    elif not constants:
        lo, hi = sc.constant_texture((0.05, 0.05, 0.05)), sc.constant_texture((0.6, 0.6, 0.6))
        sc.set_texture(metal, abi.SLOT_ROUGHNESS_U, sc.checkerboard_texture(lo, hi, scale=(6.0, 3.0)))
        sc.set_texture(metal, abi.SLOT_ETA, sc.uv_texture())
        sc.set_texture(metal, abi.SLOT_K, sc.blend_texture(sc.constant_texture((3.0, 2.5, 2.0)), sc.uv_texture()))
        sc.set_texture(plastic, abi.SLOT_ROUGHNESS_U, sc.checkerboard_texture(hi, lo, scale=(4.0, 4.0)))
    sc.add_plane((0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), floor)
    sc.add_sphere((-0.6, 0.0, 0.5), 0.6, metal)
    sc.add_sphere((0.7, -0.2, 0.3), 0.5, plastic)
    sc.add_point_light((1.0, 3.0, -1.5), (30.0, 30.0, 30.0))
    sc.add_perez_light(2.5, (0.5, 1.0, -0.5))
    sc.set_camera(S.camera_record((0.0, 0.4, -3.0), (0.0, 0.0, 0.5), res, res, fov_x_deg=50.0))
    return sc


def test_constant_textures_on_roughness_eta_and_k_equal_the_constants():
    """Roughness, eta and k read through constant textures (colour average
    of equal components: the value itself) render bit-identically to the
    constants."""
    from tests.oracle_lib import OracleScene
    res = 32
    films = []
    for kw in (dict(constants=True), dict(constant_textures=True)):
        o = OracleScene(_rough_scene(res, **kw))
        s, _, c = o.render(res, res, 4, seed=5, max_bounce=6)
        o.close()
        films.append((s, c))
    np.testing.assert_array_equal(films[0][1], films[1][1])
    np.testing.assert_array_equal(films[0][0], films[1][0])


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_textured_roughness_eta_and_k_on_the_gpu(kernel):
    """Checkerboard roughness on Metal and Plastic, UV eta and blended k on
    Metal: same streams as the oracle (>= 98 % of 1-spp paths, 16-spp mean
    within 0.5 %, counts exact); constant textures render like the
    constants; the reference's compiled arrays like dfr_set_scene."""
    from dayflower_amd import renderer as R
    from oracle import scene_compiler as SC
    from tests.oracle_lib import OracleScene
    res, seed = 48, 23
    sc = _rough_scene(res)
    o = OracleScene(sc)

    def gpu(setter, spp):
        r = R.GPURenderer()
        setter(r)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setKernel(kernel)
        r.setup()
        r.render(spp)
        g, k = r.film(res, res)
        r.dispose()
        return g, k
    for spp in (1, 16):
        s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
        ref = s / np.maximum(c, 1)[..., None]
        g, k = gpu(lambda r: r.setScene(sc), spp)
        np.testing.assert_array_equal(k, c)
        eq = np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(g.mean() - ref.mean()) / ref.mean()
        print(f"textured roughness kernel {kernel} spp {spp}: paths equal {eq:.4f}, mean rel {rel:.2e}", flush=True)
        if spp == 1:
            assert eq >= 0.98
        else:
            assert rel < 5e-3
    o.close()
    a, _ = gpu(lambda r: r.setScene(_rough_scene(res, constants=True)), 4)
    b, _ = gpu(lambda r: r.setScene(_rough_scene(res, constant_textures=True)), 4)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
    arr = SC.compatible(SC.compile_scene(sc))
    x, _ = gpu(lambda r: r.setCompiledScene(arr), 4)
    y, _ = gpu(lambda r: r.setScene(sc), 4)
    eq = np.isclose(x, y, rtol=1e-4, atol=1e-5).all(-1).mean()
    print(f"compiled textured roughness: pixels equal to dfr_set_scene's {eq:.4f}", flush=True)
    assert eq >= 0.97


def _ext_textured_scene(res):
    """Substrate (checkerboard kd, UV ks, checkerboard roughness v), Glossy
    (UV kr, checkerboard roughness) and ClearCoat (checkerboard kd,
    constant-texture ks) spheres, and smooth and rough Glass spheres with a
    checkerboard eta, over a Matte floor."""
    sc = S.Scene()
    floor = sc.add_material(S.matte((0.6, 0.6, 0.6)))
    sub = sc.add_material(S.substrate((0.5, 0.3, 0.2), (0.1, 0.1, 0.1), 0.3))
    glo = sc.add_material(S.glossy((0.8, 0.6, 0.4), 0.2))
    cc = sc.add_material(S.clear_coat((0.4, 0.5, 0.6), (0.2, 0.2, 0.2)))
    red, blue = sc.constant_texture((0.8, 0.1, 0.1)), sc.constant_texture((0.1, 0.2, 0.9))
    sc.set_texture(sub, abi.SLOT_KD, sc.checkerboard_texture(red, blue, scale=(8.0, 4.0)))
    sc.set_texture(sub, abi.SLOT_KS, sc.uv_texture())
    sc.set_texture(glo, abi.SLOT_KR, sc.blend_texture(sc.uv_texture(), sc.constant_texture((0.9, 0.9, 0.9))))
    sc.set_texture(cc, abi.SLOT_KD, sc.checkerboard_texture(blue, red, scale=(6.0, 6.0)))
    # smooth and rough glass whose eta is a checkerboard of 1.3 and 1.7
    e13, e17 = sc.constant_texture((1.3, 1.3, 1.3)), sc.constant_texture((1.7, 1.7, 1.7))
    smooth = sc.add_material(S.glass(1.5))
    rough = sc.add_material(S.glass(1.5))
    sc.materials[rough].roughness_u = sc.materials[rough].roughness_v = 0.2
    sc.set_texture(smooth, abi.SLOT_ETA, sc.checkerboard_texture(e13, e17, scale=(6.0, 3.0)))
    sc.set_texture(rough, abi.SLOT_ETA, sc.checkerboard_texture(e17, e13, scale=(4.0, 4.0)))
    sc.add_sphere((-0.5, 0.5, 1.6), 0.45, smooth)
    sc.add_sphere((0.6, 0.5, 1.7), 0.45, rough)
    rough_lo, rough_hi = sc.constant_texture((0.05, 0.05, 0.05)), sc.constant_texture((0.5, 0.5, 0.5))
    sc.set_texture(sub, abi.SLOT_ROUGHNESS_V, sc.checkerboard_texture(rough_lo, rough_hi, scale=(5.0, 5.0)))
    sc.set_texture(glo, abi.SLOT_ROUGHNESS_U, sc.checkerboard_texture(rough_hi, rough_lo, scale=(4.0, 8.0)))
    sc.set_texture(cc, abi.SLOT_KS, sc.constant_texture((0.3, 0.3, 0.3)))
    sc.add_plane((0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), floor)
    sc.add_sphere((-1.0, -0.4, 0.8), 0.55, sub)
    sc.add_sphere((0.1, -0.4, 0.5), 0.55, glo)
    sc.add_sphere((1.1, -0.4, 0.9), 0.55, cc)
    sc.add_point_light((1.0, 3.0, -1.5), (30.0, 30.0, 30.0))
    sc.add_perez_light(2.5, (0.5, 1.0, -0.5))
    sc.set_camera(S.camera_record((0.0, 0.4, -3.0), (0.0, -0.2, 0.6), res, res, fov_x_deg=55.0))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_textured_substrate_glossy_and_clear_coat(kernel):
    """Same streams as the oracle (>= 98 % of 1-spp paths, 16-spp mean
    within 0.5 %, counts exact) and the compiled arrays like dfr_set_scene."""
    from dayflower_amd import renderer as R
    from oracle import scene_compiler as SC
    from tests.oracle_lib import OracleScene
    res, seed = 48, 31
    sc = _ext_textured_scene(res)
    o = OracleScene(sc)

    def gpu(setter, spp):
        r = R.GPURenderer()
        setter(r)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setKernel(kernel)
        r.setup()
        r.render(spp)
        g, k = r.film(res, res)
        r.dispose()
        return g, k
    for spp in (1, 16):
        s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
        ref = s / np.maximum(c, 1)[..., None]
        g, k = gpu(lambda r: r.setScene(sc), spp)
        np.testing.assert_array_equal(k, c)
        eq = np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(g.mean() - ref.mean()) / ref.mean()
        print(f"textured substrate/glossy/clear coat kernel {kernel} spp {spp}: paths equal {eq:.4f}, "
              f"mean rel {rel:.2e}", flush=True)
        if spp == 1:
            assert eq >= 0.98
        else:
            assert rel < 5e-3
    o.close()
    x, _ = gpu(lambda r: r.setCompiledScene(SC.compatible(SC.compile_scene(sc))), 4)
    y, _ = gpu(lambda r: r.setScene(sc), 4)
    eq = np.isclose(x, y, rtol=1e-4, atol=1e-5).all(-1).mean()
    print(f"compiled: pixels equal to dfr_set_scene's {eq:.4f}", flush=True)
    assert eq >= 0.97
