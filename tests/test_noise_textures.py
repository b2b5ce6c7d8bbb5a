"""Marble and simplex fBm textures (M/scene/texture/MarbleTexture.java
:131-146, SimplexFractionalBrownianMotionTexture.java:127-137) and the noise
under them (M/noise/PerlinNoiseF.java:113-250, SimplexNoiseF.java:246-507).

Oracle pins (CPU):
* Perlin noise against Ken Perlin's published improved-noise algorithm
  ("Improving Noise", 2002) evaluated here in float64 over the library's
  lattice table: equal wherever none of a point's eight corner hashes is 13
  or 14, where the reference chooses other gradients (u = x for hashes 12
  and 13, v = y for 12 and 13); those points follow the reference's rule;
  zero on the integer lattice, periodic over 256;
* simplex noise against Gustavson's published 3D simplex noise ("Simplex
  noise demystified", 2005) in float64;
* turbulence and fBm as the closed sums of their octaves;
* the textures' colours read back from a lit quad equal the closed forms of
  MarbleTexture and SimplexFractionalBrownianMotionTexture over the oracle's
  noise at the world-space hit points.

GPU (through the C-ABI): the same quad against the oracle per pixel, path
tracing on the same streams on both kernels, and the reference's compiled
arrays of the textures.
"""
import os
import re

import numpy as np
import pytest

from dayflower_amd import abi
from dayflower_amd import scene as S
from tests.test_textures import E, _lit_scene, _quad

TABLES = os.path.join(os.path.dirname(__file__), "..", "dayflower_amd", "csrc", "noise_tables.inc")


def _tables():
    text = open(TABLES).read()
    perm = [int(v) for v in re.search(r"NOISE_PERM_256 \\(.*?)#define", text, re.S).group(1).replace("\\", "").split(",")]
    grad = [float(v.strip().rstrip("f")) for v in
            re.search(r"NOISE_GRAD3_36 \\(.*)", text, re.S).group(1).replace("\\", "").split(",") if v.strip()]
    return np.array(perm + perm, np.int64), np.array(grad, np.float64).reshape(12, 3)


def _perlin_textbook(p, P, reference_gradients=False):
    """Perlin's improved noise (2002) in float64; with reference_gradients the
    reference's choice of u and v for hashes 12 to 15.  Also returns the
    eight corner hashes of each point."""
    x, y, z = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64), p[:, 2].astype(np.float64)
    X, Y, Z = (np.floor(c).astype(np.int64) & 255 for c in (x, y, z))
    x, y, z = x - np.floor(x), y - np.floor(y), z - np.floor(z)
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
    u, v, w = fade(x), fade(y), fade(z)
    A, B = P[X] + Y, P[X + 1] + Y
    AA, AB, BA, BB = P[A] + Z, P[A + 1] + Z, P[B] + Z, P[B + 1] + Z

    def grad(h, gx, gy, gz):
        if reference_gradients:
            uu = np.where((h < 8) | (h == 12) | (h == 13), gx, gy)
            vv = np.where((h < 4) | (h == 12) | (h == 13), gy, gz)
        else:
            uu = np.where(h < 8, gx, gy)
            vv = np.where(h < 4, gy, np.where((h == 12) | (h == 14), gx, gz))
        return np.where(h & 1, -uu, uu) + np.where(h & 2, -vv, vv)

    hashes = [P[AA] & 15, P[BA] & 15, P[AB] & 15, P[BB] & 15, P[AA + 1] & 15, P[BA + 1] & 15, P[AB + 1] & 15,
              P[BB + 1] & 15]
    lerp = lambda t, a, b: a + t * (b - a)
    n = lerp(w, lerp(v, lerp(u, grad(hashes[0], x, y, z), grad(hashes[1], x - 1, y, z)),
                     lerp(u, grad(hashes[2], x, y - 1, z), grad(hashes[3], x - 1, y - 1, z))),
             lerp(v, lerp(u, grad(hashes[4], x, y, z - 1), grad(hashes[5], x - 1, y, z - 1)),
                  lerp(u, grad(hashes[6], x, y - 1, z - 1), grad(hashes[7], x - 1, y - 1, z - 1))))
    return n, np.stack(hashes, 1)


def _simplex_textbook(p, P, G):
    """Gustavson's 3D simplex noise (2005) in float64."""
    out = np.empty(len(p))
    F3, G3 = 1.0 / 3.0, 1.0 / 6.0
    for n, (x, y, z) in enumerate(p.astype(np.float64)):
        s = (x + y + z) * F3
        i, j, k = int(np.floor(x + s)), int(np.floor(y + s)), int(np.floor(z + s))
        t = (i + j + k) * G3
        x0, y0, z0 = x - (i - t), y - (j - t), z - (k - t)
        if x0 >= y0:
            o = ((1, 0, 0), (1, 1, 0)) if y0 >= z0 else ((1, 0, 0), (1, 0, 1)) if x0 >= z0 else ((0, 0, 1), (1, 0, 1))
        else:
            o = ((0, 0, 1), (0, 1, 1)) if y0 < z0 else ((0, 1, 0), (0, 1, 1)) if x0 < z0 else ((0, 1, 0), (1, 1, 0))
        corners = [(0, 0, 0), o[0], o[1], (1, 1, 1)]
        total = 0.0
        for c, (a, b, d) in enumerate(corners):
            q = np.array([x0 - a + c * G3, y0 - b + c * G3, z0 - d + c * G3])
            r = 0.6 - q @ q
            if r > 0:
                gi = P[(i & 255) + a + P[(j & 255) + b + P[(k & 255) + d]]] % 12
                total += r ** 4 * (G[gi] @ q)
        out[n] = 32.0 * total
    return out


def test_perlin_noise_is_the_textbook_algorithm_with_the_references_gradients():
    from tests.oracle_lib import noise
    P, _ = _tables()
    rng = np.random.default_rng(1)
    p = rng.uniform(-300, 300, size=(20000, 3)).astype(np.float32)
    got = noise("perlin", p).astype(np.float64)
    book, hashes = _perlin_textbook(p, P)
    ref, _ = _perlin_textbook(p, P, reference_gradients=True)
    plain = ~np.isin(hashes, (13, 14)).any(1)
    print(f"perlin: {plain.mean():.3f} of points with no corner hash 13 or 14; max |oracle - textbook| there "
          f"{np.abs(got - book)[plain].max():.2e}, elsewhere vs the reference rule "
          f"{np.abs(got - ref)[~plain].max():.2e}; textbook differs elsewhere up to {np.abs(book - ref).max():.2f}")
    assert 0.3 < plain.mean() < 0.7
    assert np.abs(got - book)[plain].max() < 5e-6
    assert np.abs(got - ref).max() < 5e-6
    assert np.abs(book - ref)[~plain].max() > 0.1          # the rule matters
    lattice = rng.integers(-1000, 1000, size=(2000, 3)).astype(np.float32)
    assert np.all(noise("perlin", lattice) == 0.0)
    # periodic over 256 (points on a 1/64 grid, which the shift keeps exact)
    q = (np.round(rng.uniform(0, 40, size=(2000, 3)) * 64) / 64).astype(np.float32)
    assert np.array_equal(noise("perlin", q), noise("perlin", q + np.float32(256.0)))


def test_simplex_noise_is_gustavsons():
    from tests.oracle_lib import noise
    P, G = _tables()
    rng = np.random.default_rng(2)
    p = rng.uniform(-50, 50, size=(5000, 3)).astype(np.float32)
    got = noise("simplex", p).astype(np.float64)
    book = _simplex_textbook(p, P, G)
    err = np.abs(got - np.clip(book, -1, 1))
    print(f"simplex: max |oracle - textbook| {err.max():.2e}, range [{got.min():.3f}, {got.max():.3f}]")
    # float32 offsets from cells up to 50 away carry ulps of 4e-6 into a sum
    # scaled by 32
    assert err.max() < 5e-5
    assert got.min() < -0.5 and got.max() > 0.5


def test_turbulence_and_fbm_are_the_sums_of_their_octaves():
    from tests.oracle_lib import noise
    rng = np.random.default_rng(3)
    p = rng.uniform(-20, 20, size=(3000, 3)).astype(np.float32)
    n = lambda k: noise("perlin", p * np.float32(k)).astype(np.float64)
    assert np.array_equal(noise("turbulence", p, octaves=1), np.abs(noise("perlin", p)))
    assert np.array_equal(noise("turbulence", p, octaves=0), np.abs(noise("perlin", p)))
    t3 = np.abs(n(1)) + np.abs(n(2)) / 2 + np.abs(n(4)) / 4
    assert np.allclose(noise("turbulence", p, octaves=3), t3, rtol=0, atol=1e-6)
    f, g = np.float32(0.7), 0.5
    s = lambda k: noise("simplex", p * np.float32(f * k)).astype(np.float64)
    assert np.allclose(noise("fbm", p, frequency=f, gain=g, octaves=1), (s(1) + 1) / 2, rtol=0, atol=1e-6)
    f3 = ((s(1) + g * s(2) + g * g * s(4)) / (1 + g + g * g) + 1) / 2
    assert np.allclose(noise("fbm", p, frequency=f, gain=g, octaves=3), f3, rtol=0, atol=1e-6)
    assert np.all(np.isnan(noise("fbm", p[:10], frequency=f, gain=g, octaves=0)))   # 0 / 0, as the reference


def _marble_scene(res, fbm=False):
    def geo(sc):
        m = sc.add_material(S.matte((0.5, 0.5, 0.5)))
        if fbm:
            t = sc.simplex_fbm_texture((0.8, 0.6, 1.0), frequency=2.0, gain=0.6, octaves=5)
        else:
            t = sc.marble_texture((0.8, 0.8, 0.8), (0.4, 0.2, 0.1), (0.06, 0.04, 0.02), scale=5.0, stripes=0.5,
                                  octaves=6)
        sc.set_texture(m, abi.SLOT_KD, t)
        _quad(sc, m)
    return _lit_scene(res, geo)


def _closed_form(sc, p, fbm):
    from tests.oracle_lib import noise
    T = sc.textures[0]
    if fbm:
        return np.outer(noise("fbm", p, frequency=T.scale[0], gain=T.scale[1], octaves=int(T.cell)), T.color)
    q = (p * np.float32(T.scale[0])).astype(np.float32)
    r = np.float32(T.scale[1]) * noise("turbulence", q, octaves=int(T.cell))
    s = 2.0 * np.abs(np.sin((q[:, 0] + r).astype(np.float32).astype(np.float64)))
    t = np.where(s < 1, s, s - 1)[:, None]
    A, B, C = np.array(T.color), np.array(T.origin), np.array(T.t)
    return np.where((s < 1)[:, None], (1 - t) * C + t * B, (1 - t) * B + t * A)


@pytest.mark.parametrize("fbm", [False, True])
def test_texture_colours_read_back_from_a_lit_quad(fbm):
    from tests.test_textures import _direct, _kd_from
    res = 48
    sc = _marble_scene(res, fbm)
    img, rays, prim, p = _direct(sc, res)
    hit = prim >= 0
    kd = _kd_from(img, rays, np.array([0.0, 0.0, -1.0]))[hit]
    want = _closed_form(sc, p[hit].astype(np.float32), fbm)
    err = np.abs(kd - want).max()
    print(f"{'fbm' if fbm else 'marble'}: {hit.sum()} hits, max |kd - closed form| {err:.2e}, "
          f"kd range [{kd.min():.3f}, {kd.max():.3f}]")
    assert hit.sum() > res * res // 3 and err < 1e-4
    assert kd.max() - kd.min() > 0.1                      # the texture varies over the quad


@pytest.mark.gpu
@pytest.mark.parametrize("fbm", [False, True])
def test_noise_textures_on_the_gpu_match_the_oracle(fbm):
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res, seed = 64, 7
    sc = _marble_scene(res, fbm)
    o = OracleScene(sc)
    s, _, c = o.render(res, res, 1, seed=seed, max_bounce=1, min_rr=50)
    o.close()
    ref = s / np.maximum(c, 1)[..., None]
    r = GPURenderer()
    r.setScene(sc)
    r.setSeed(seed)
    r.setMaximumBounce(1)
    r.setMinimumBounceRussianRoulette(50)
    r.setup()
    r.render(1)
    g, k = r.film(res, res)
    r.dispose()
    np.testing.assert_array_equal(k, c)
    eq = np.isclose(g, ref, rtol=1e-4, atol=1e-5).all(-1).mean()
    print(f"{'fbm' if fbm else 'marble'} on the GPU: pixels equal {eq:.4f}, max |diff| {np.abs(g - ref).max():.2e}")
    assert eq >= 0.999


def _noise_scene(res):
    sc = S.Scene()
    mats = [sc.add_material(S.matte((0.5, 0.5, 0.5))), sc.add_material(S.plastic((0.3, 0.5, 0.7), (0.2, 0.2, 0.2), 0.3)),
            sc.add_material(S.glass(1.5))]
    sc.set_texture(mats[0], abi.SLOT_KD, sc.marble_texture())
    sc.set_texture(mats[1], abi.SLOT_KD, sc.simplex_fbm_texture())
    # colours on the 8-bit grid, so the reference's packed form holds them exactly
    sc.set_texture(mats[2], abi.SLOT_KT, sc.simplex_fbm_texture((204 / 255, 1.0, 153 / 255), 3.0, 0.5, 4))
    sc.add_plane((0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), mats[0])
    sc.add_sphere((-0.8, 0.0, 1.0), 0.7, mats[1])
    sc.add_sphere((0.8, 0.0, 0.6), 0.6, mats[2])
    p, t, n = S.displaced_icosphere(2000, 7, radius=0.5, center=(0.0, 0.6, 2.2))
    sc.add_mesh(p, t, mats[0], normals=n)
    sc.add_point_light((1.0, 4.0, -1.0), (40.0, 40.0, 40.0))
    sc.add_perez_light(2.5, (0.5, 1.0, -0.5))
    sc.set_camera(S.camera_record((0.0, 0.5, -3.0), (0.0, 0.0, 1.0), res, res, fov_x_deg=55.0))
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
def test_noise_textured_scene_same_streams(kernel):
    """Marble on a plane and a mesh, fBm on Plastic and on Glass
    transmittance: >= 95 % of 1-spp paths equal to the oracle's (the glass
    sphere makes some paths rounding-sensitive), counts exact, 16-spp mean
    within 0.5 %."""
    from dayflower_amd.renderer import GPURenderer
    from tests.oracle_lib import OracleScene
    res, seed = 48, 29
    sc = _noise_scene(res)
    o = OracleScene(sc)
    for spp in (1, 16):
        s, _, c = o.render(res, res, spp, seed=seed, max_bounce=8)
        ref = s / np.maximum(c, 1)[..., None]
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
        eq = np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(g.mean() - ref.mean()) / ref.mean()
        print(f"noise textures kernel {kernel} spp {spp}: paths equal {eq:.4f}, mean rel {rel:.2e}", flush=True)
        if spp == 1:
            assert eq >= 0.95
        else:
            assert rel < 5e-3
    o.close()


@pytest.mark.gpu
def test_compiled_arrays_carry_the_noise_textures():
    """The reference's arrays of the same scene (colours on the 8-bit grid of
    Color3F.pack) render what dfr_set_scene renders."""
    from dayflower_amd import renderer as R
    from oracle import scene_compiler as SC
    res, seed = 32, 5
    sc = _noise_scene(res)
    for T in sc.textures:       # the 8-bit grid for every colour
        for f in (T.color, T.origin, T.t):
            for i in range(3):
                f[i] = round(f[i] * 255) / 255
    arr = SC.compatible(SC.compile_scene(sc))
    t = R.transcode_compiled_scene(arr)
    kinds = sorted({x.type for x in t["textures"]})
    assert abi.TEX_MARBLE in kinds and abi.TEX_SIMPLEX_FBM in kinds
    films = []
    for compiled in (True, False):
        r = R.GPURenderer()
        if compiled:
            r.setCompiledScene(arr)
        else:
            r.setScene(sc)
        r.setSeed(seed)
        r.setMaximumBounce(8)
        r.setup()
        r.render(4)
        films.append(r.film(res, res))
        r.dispose()
    eq = np.isclose(films[0][0], films[1][0], rtol=1e-4, atol=1e-5).all(-1).mean()
    print(f"compiled noise textures: pixels equal to dfr_set_scene's {eq:.4f}")
    assert eq >= 0.97
