"""ImageLight (M/scene/light/ImageLight.java:134-150, 165-240, 348-372,
544-584): an image on the sphere of directions, looked up bilinearly with
wrap-around, sampled through a 32 x 32 distribution of luminance x sin(theta).

Oracle pins (CPU): texel lookups at directions whose image coordinates are
known in closed form (the light's object frame, the half-texel shift, the
wrap-around at the seam and the pole, rotation and scale); a constant image
gives constant radiance and a pdf that integrates to one; samples carry the
evaluated pdf and the lookup's radiance, and follow the pdf (chi-square);
a Lambertian plane under a constant sky renders to albedo x radiance.

GPU (through the C-ABI): path tracing with an image light on the same random
streams as the oracle, both kernels; ray casting sees the image on a miss.
"""
import numpy as np
import pytest

from dayflower_amd import abi
from dayflower_amd import scene as S

# 4 x 2 image with distinct texels (value = 10 * row + column + 1 in red)
_IMG = np.zeros((2, 4, 3), np.float32)
for _r in range(2):
    for _c in range(4):
        _IMG[_r, _c] = (10 * _r + _c + 1, 0.5, 0.25)


def _texel_mean(cells):
    return np.mean([_IMG[r % 2, c % 4] for r, c in cells], axis=0)


@pytest.mark.parametrize("world_dir,cells", [
    # world +z is the light's object +x: u = 0, v = 1/2 -> image (3.5, 0.5): columns 3 and 0 (the seam)
    ((0.0, 0.0, 1.0), [(0, 3), (0, 0), (1, 3), (1, 0)]),
    # world +x is object +y: u = 1/4 -> x = 0.5
    ((1.0, 0.0, 0.0), [(0, 0), (0, 1), (1, 0), (1, 1)]),
    # world +y is the object pole: v = 0 -> y = -0.5, wrapped to 1.5: rows 1 and 0
    ((0.0, 1.0, 0.0), [(1, 3), (1, 0), (0, 3), (0, 0)]),
])
def test_lookup_at_known_image_coordinates(world_dir, cells):
    from tests.oracle_lib import image_light
    rad, _, _, _, _ = image_light(_IMG, [world_dir])
    np.testing.assert_allclose(rad[0], _texel_mean(cells), rtol=1e-6)


def test_rotation_and_scale_move_the_lookup():
    """Rotated by 90 degrees, (u, v) -> (u cos - v sin, v cos + u sin) =
    (-v, u): world +z (u 0, v 1/2) reads image (1.5, 1.5); scaled by (2, 1),
    world +x (u 1/4) reads x = 2 * 1/4 * 4 - 0.5 = 1.5."""
    from tests.oracle_lib import image_light
    rad, _, _, _, _ = image_light(_IMG, [(0.0, 0.0, 1.0)], angle=90.0)
    np.testing.assert_allclose(rad[0], _texel_mean([(1, 1), (1, 2), (0, 1), (0, 2)]), rtol=1e-5)
    rad, _, _, _, _ = image_light(_IMG, [(1.0, 0.0, 0.0)], scale=(2.0, 1.0))
    np.testing.assert_allclose(rad[0], _texel_mean([(0, 1), (0, 2), (1, 1), (1, 2)]), rtol=1e-6)


def _sphere_grid(nu, nv):
    """World directions at the centres of an (nu x nv) grid over the light's
    spherical coordinates, with their solid angles."""
    u = (np.arange(nu) + 0.5) / nu
    v = (np.arange(nv) + 0.5) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    phi, th = 2 * np.pi * U, np.pi * V
    obj = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], -1).reshape(-1, 3)
    world = np.stack([obj[:, 1], obj[:, 2], obj[:, 0]], -1)   # object (x, y, z) -> world (y, z, x)
    dw = (2 * np.pi / nu) * (np.pi / nv) * np.sin(th).reshape(-1)
    return world, dw, U.reshape(-1), V.reshape(-1)


def test_constant_image_pdf_integrates_to_one():
    from tests.oracle_lib import image_light
    img = np.full((16, 32, 3), 0.7, np.float32)
    world, dw, _, _ = _sphere_grid(256, 128)
    rad, pdf, _, _, _ = image_light(img, world)
    np.testing.assert_allclose(rad, 0.7, rtol=1e-6)
    total = float((pdf * dw).sum())
    print(f"integral of the pdf over the sphere: {total:.6f}")
    assert abs(total - 1.0) < 1e-3


def _spot_image():
    """A dim gradient with a bright spot: the sampler must find the spot."""
    h, w = 32, 64
    img = np.zeros((h, w, 3), np.float32)
    img[..., 0] = np.linspace(0.05, 0.2, w)[None, :]
    img[..., 1] = np.linspace(0.1, 0.02, h)[:, None]
    img[..., 2] = 0.08
    img[8:12, 40:46] = (30.0, 28.0, 22.0)
    return img


def test_samples_carry_the_pdf_and_follow_it():
    from tests.oracle_lib import image_light
    img = _spot_image()
    rng = np.random.default_rng(5)
    uv = rng.random((200000, 2), dtype=np.float32)
    _, _, sd, sp, sr = image_light(img, np.zeros((0, 3)), uv)
    ok = sp > 0
    assert ok.mean() > 0.999
    rad, pdf, _, _, _ = image_light(img, sd[ok])
    np.testing.assert_allclose(pdf, sp[ok], rtol=2e-4)
    np.testing.assert_allclose(rad, sr[ok], rtol=1e-5, atol=1e-6)
    # chi-square over 16 x 8 bins of the spherical coordinates
    world, dw, U, V = _sphere_grid(256, 128)
    _, pdf_grid, _, _, _ = image_light(img, world)
    mass = pdf_grid * dw
    bu, bv = (U * 16).astype(int), (V * 8).astype(int)
    expect = np.zeros((16, 8))
    np.add.at(expect, (bu, bv), mass)
    expect *= ok.sum() / expect.sum()
    obj = np.stack([sd[ok][:, 2], sd[ok][:, 0], sd[ok][:, 1]], -1)   # world -> object
    phi = np.mod(np.arctan2(obj[:, 1], obj[:, 0]), 2 * np.pi)
    th = np.arccos(np.clip(obj[:, 2], -1, 1))
    got = np.zeros((16, 8))
    np.add.at(got, (np.minimum((phi / (2 * np.pi) * 16).astype(int), 15), np.minimum((th / np.pi * 8).astype(int), 7)), 1)
    m = expect > 20
    chi2 = float((((got - expect) ** 2) / expect)[m].sum())
    dof = int(m.sum()) - 1
    print(f"chi-square {chi2:.1f} over {dof} degrees of freedom; spot bins hold {got.max() / ok.sum():.3f} of the samples")
    assert chi2 < dof + 5 * np.sqrt(2 * dof)


def _floor_under_sky(res, sky):
    sc = S.Scene()
    m = sc.add_material(S.matte((0.5, 0.5, 0.5)))
    sc.add_rectangle((-10, 0, -10), (-10, 0, 10), (10, 0, 10), m)
    sc.add_image_light(sky)
    sc.set_camera(S.camera_record((0.0, 3.0, -1.0), (0.0, 0.0, 0.5), res, res, fov_x_deg=40.0))
    return sc


def test_lambertian_plane_under_a_constant_sky():
    """Radiance c from every direction of the upper hemisphere: a Lambertian
    plane of albedo 0.5 reflects 0.5 c toward the camera."""
    from tests.oracle_lib import OracleScene
    c = np.array([0.8, 0.6, 0.4], np.float32)
    sky = np.broadcast_to(c, (8, 16, 3)).copy()
    res, spp = 32, 64
    o = OracleScene(_floor_under_sky(res, sky))
    s, q, n = o.render(res, res, spp, seed=3, max_bounce=1, min_rr=50)
    o.close()
    mean = s / n[..., None]
    var = q / n[..., None] - mean ** 2
    err = np.sqrt(var.sum((0, 1)) / n.sum()) / np.sqrt(n.size)
    got = mean.mean((0, 1))
    print(f"mean radiance {got}, expected {0.5 * c}, standard error {err}")
    assert np.all(np.abs(got - 0.5 * c) < 4 * np.maximum(err, 1e-6))


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _env_scene(res):
    sc = S.Scene()
    g = sc.add_material(S.plastic((0.3, 0.4, 0.5), (0.2, 0.2, 0.2), 0.3))
    m = sc.add_material(S.matte((0.6, 0.55, 0.5)))
    p, t, n = S.displaced_icosphere(2000, 0x5EED, radius=0.6, center=(0.0, 0.6, 0.0))
    sc.add_mesh(p, t, g, normals=n)
    sc.add_rectangle((-3, 0, -3), (-3, 0, 3), (3, 0, 3), m)
    sc.add_image_light(_spot_image(), angle_deg=30.0, scale=(1.0, 1.0))
    sc.set_camera(S.camera_record((0.0, 1.6, -2.8), (0.0, 0.5, 0.0), res, res, fov_x_deg=50.0))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_image_light_same_streams(kernel):
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res, seed = 48, 43
    sc = _env_scene(res)
    o = OracleScene(sc)
    s1, _, c1 = o.render(res, res, 1, seed=seed, max_bounce=8)
    s16, _, c16 = o.render(res, res, 16, seed=seed, max_bounce=8)
    o.close()
    ref1 = s1 / np.maximum(c1, 1)[..., None]
    ref16 = s16 / np.maximum(c16, 1)[..., None]
    out = []
    for spp in (1, 16):
        r = GPURenderer()
        r.setScene(sc)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setKernel(kernel)
        r.setup()
        r.render(spp)
        out.append(r.film(res, res))
        r.dispose()
    (g1, k1), (g16, k16) = out
    np.testing.assert_array_equal(k1, c1)
    np.testing.assert_array_equal(k16, c16)
    paths = np.isclose(g1, ref1, rtol=1e-3, atol=1e-4).all(-1).mean()
    rel = abs(g16.mean() - ref16.mean()) / ref16.mean()
    print(f"kernel {kernel}: 1-spp paths equal {paths:.5f} (bar 0.99); 16-spp mean rel diff {rel:.2e} (bar 3e-3)",
          flush=True)
    assert paths >= 0.99 and rel < 3e-3


@pytest.mark.gpu
def test_ray_casting_sees_the_image_on_a_miss():
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res, seed = 48, 44
    sc = _env_scene(res)
    o = OracleScene(sc)
    s, _, c = o.render(res, res, 2, seed=seed, algorithm=abi.ALGO_RAY_CASTING)
    prim, _, _ = o.trace_f32(o.primary_rays(res, res, seed, 0))
    o.close()
    ref = s / np.maximum(c, 1)[..., None]
    r = GPURenderer()
    r.setScene(sc)
    r.setSeed(seed)
    r.setRenderingAlgorithm(abi.ALGO_RAY_CASTING)
    r.setup()
    r.render(2)
    g, k = r.film(res, res)
    r.dispose()
    np.testing.assert_array_equal(k, c)
    eq = np.isclose(g, ref, rtol=1e-4, atol=1e-6).all(-1).mean()
    sky = (prim < 0).mean()
    print(f"ray casting: pixels equal {eq:.5f} (bar 0.999); camera rays that miss {sky:.3f}", flush=True)
    assert eq >= 0.999 and sky > 0.2
