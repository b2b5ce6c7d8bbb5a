"""Seeded workload generators for the batched and serving measurements.

* `batched_scenes` / `batched_jobs`: 64 render jobs; resolution drawn from
  {128^2, 256^2, 512^2, 1024^2}, spp from {1, 4, 16, 64, 256}, each job on one
  of 4 noise-displaced icospheres (about 10k, 100k, 500k and 2M triangles)
  whose surface is split into Matte, Glass, Metal and Disney patches;
  maxBounce 8.
* `request_stream`: Poisson arrivals at `rate` requests/s over `duration`
  seconds against one resident 2M-triangle icosphere; camera on a sphere of
  radius 3 looking at the origin; resolution from {256^2, 512^2, 1024^2},
  spp from {4, 16, 64}.
"""
import math

import numpy as np

from . import scene as S
from .renderer import make_job

BATCH_TRIANGLES = (10_000, 100_000, 500_000, 2_000_000)
BATCH_RES = (128, 256, 512, 1024)
BATCH_SPP = (1, 4, 16, 64, 256)
SERVE_RES = (256, 512, 1024)
SERVE_SPP = (4, 16, 64)


def mixed_material_icosphere(n_triangles, seed, with_perez=False, width=512, height=512, materials=None,
                             ground=None):
    """Displaced icosphere split into four contiguous patches (five
    icosahedron faces each) of Matte, Glass, Metal and Disney (or the four
    `materials` given), on a Matte ground plane (or `ground`), lit by a point
    light (and optionally the Perez sky)."""
    sc = S.Scene()
    if materials is None:
        materials = [S.matte((0.7, 0.6, 0.5)), S.glass(1.5, roughness=0.0), S.metal(roughness=0.25),
                     S.disney((0.3, 0.5, 0.8), metallic=0.2, roughness=0.4, clearcoat=0.5, sheen=0.3)]
    assert len(materials) == 4
    mats = [sc.add_material(m) for m in materials]
    ground = sc.add_material(ground if ground is not None else S.matte((0.5, 0.5, 0.5)))
    p, t, nrm = S.displaced_icosphere(n_triangles, seed)
    nt = t.shape[0]
    cuts = [0, nt // 4, nt // 2, 3 * nt // 4, nt]
    # one vertex array; every patch references it through its own index range
    first = sc.add_mesh(p, t[cuts[0]:cuts[1]], mats[0], normals=nrm)
    base = sc.positions.shape[0] - p.shape[0]
    for k in range(1, 4):
        idx = t[cuts[k]:cuts[k + 1]]
        # reuse the vertex block already appended by the first patch
        prim = S.abi.Primitive()
        prim.shape = S.abi.SHAPE_TRIANGLE_MESH
        prim.material = mats[k]
        prim.area_light = -1
        prim.first_triangle = sc.indices.shape[0]
        prim.triangle_count = idx.shape[0]
        sc.indices = np.concatenate([sc.indices, idx + base])
        sc.primitives.append(prim)
    del first
    sc.add_plane((0.0, -1.1, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), ground)
    sc.add_point_light((2.0, 4.0, 2.0), (50.0, 50.0, 50.0))
    if with_perez:
        sc.add_perez_light(2.0, (1.0, 1.0, -1.0))
    sc.set_camera(S.camera_record((0.0, 0.4, -3.6), (0.0, -0.1, 0.0), width, height, fov_x_deg=40.0))
    return sc


def orbit_camera(rng, width, height, radius=3.0):
    """Camera at a uniform direction on a sphere of `radius` (upper
    hemisphere, so the ground plane stays below) looking at the origin."""
    z = rng.uniform(0.05, 1.0)
    phi = rng.uniform(0, 2 * math.pi)
    s = math.sqrt(1 - z * z)
    eye = (radius * s * math.cos(phi), radius * z, radius * s * math.sin(phi))
    return S.camera_record(eye, (0.0, 0.0, 0.0), width, height, fov_x_deg=40.0)


def batched_scenes(seed=0x5EED, triangles=BATCH_TRIANGLES):
    return [mixed_material_icosphere(n, seed + i) for i, n in enumerate(triangles)]


def batched_jobs(scene_ids, seed=0x5EED, n_jobs=64, max_bounce=8):
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        res = int(rng.choice(BATCH_RES))
        spp = int(rng.choice(BATCH_SPP))
        sid = int(scene_ids[rng.integers(len(scene_ids))])
        cam = orbit_camera(rng, res, res)
        jobs.append(make_job(sid, cam, res, res, spp, seed=(seed << 8) ^ (j * 0x9E3779B1), maximum_bounce=max_bounce))
    return jobs


def request_stream(scene_id, seed=0x5EED, rate=200.0, duration=60.0, max_bounce=8):
    """[(arrival_seconds, job)] with exponential inter-arrival times."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    i = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        res = int(rng.choice(SERVE_RES))
        spp = int(rng.choice(SERVE_SPP))
        cam = orbit_camera(rng, res, res)
        out.append((t, make_job(scene_id, cam, res, res, spp, seed=(seed << 16) ^ i, maximum_bounce=max_bounce)))
        i += 1
    return out


def more_materials_icosphere(n_triangles, seed, with_perez=False, width=512, height=512):
    """The four-patch icosphere in Glossy, Substrate, Translucent and Uber
    (partly transparent) on a ClearCoat ground plane."""
    return mixed_material_icosphere(
        n_triangles, seed, with_perez=with_perez, width=width, height=height,
        materials=[S.glossy((0.8, 0.6, 0.3), roughness=0.15),
                   S.substrate((0.5, 0.3, 0.2), (0.2, 0.2, 0.2), roughness=0.1),
                   S.translucent((0.6, 0.6, 0.4), (0.3, 0.3, 0.3), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), roughness=0.2),
                   S.uber((0.5, 0.4, 0.3), (0.2, 0.2, 0.2), kr=(0.1, 0.1, 0.1), kt=(0.1, 0.1, 0.1),
                          opacity=(0.8, 0.7, 0.8), eta=1.5, roughness=0.15)],
        ground=S.clear_coat((0.4, 0.4, 0.5), (1.0, 1.0, 1.0)))
