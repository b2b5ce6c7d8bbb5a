"""TEST INFRASTRUCTURE ONLY: a numpy restatement of the reference's
SceneCompiler (M/scene/compiler/SceneCompiler.java:81-164 and the
Compiled*Cache.to* methods), writing a scene of dayflower_amd.scene into the
array set doSetupScene uploads (M/renderer/gpu/AbstractSceneKernel.java
:1777-1855).  The tests feed these arrays to dfr_set_compiled_scene /
dfr_transcode_compiled_scene; nothing in the product imports this module.

M/ = src/main/java/org/dayflower/.  Layout constants come from
tests/golden/compiled_layout.json, which tests/golden/gen_compiled_layout.py
extracted from the reference's Compiled*Cache.java and type-ID declarations,
so a record written here sits where the reference's compiler puts it.

What the restatement does as the reference does:
* a primitive is 8 ints (instance ID, Ints.pack(area light ID, offset),
  bounding volume ID and offset, material ID and offset, shape ID and offset)
  and 32 floats (object to world, world to object, row major)
  (CompiledPrimitiveCache.java:323-393);
* shapes and materials are addressed by element offset, textures, lights
  and bounding volumes by index, LDR image textures and image / Perez lights
  through offset arrays (the mapDistinctToOffsets calls of
  Shape3FCache.java:320-360, MaterialCache.java:345-397,
  TextureCache.java:242-270, LightCache.java:213-233);
* Sphere3F is the unit sphere and Plane3F the z = 0 plane in object space
  (AbstractShape3FKernel.java:1048-1068, 2025-2038): a sphere's centre and
  radius and a plane's frame go into the primitive's matrix;
* a triangle mesh is a threaded BVH of int records (leaf: ID 1, bounding
  volume, next, count, triangle offsets, padded to 8; tree: ID 2, bounding
  volume, next, left, 4 ints of padding), nodes in depth-first order with
  `left` the first child and `next` the next node not below this one
  (CompiledShape3FCache.java:2617-2733); the tree itself is a plain median
  split (the reference's swept SAH only changes which leaves exist);
* every material parameter is a texture: a colour or value becomes a
  ConstantTexture, a textured slot its texture tree; material words pack
  (ID, offset) byte pairs, texture records (ID, offset) 16-bit pairs
  (CompiledMaterialCache.java:2632-3580, CompiledTextureCache.java:1551-1980);
* a flat triangle's vertex normals are its face normal
  normalize(cross(normalize(B - A), normalize(C - A))) (Triangle3F.java:102).
"""
import json
import math
import os

import numpy as np

from dayflower_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LAYOUT = json.load(open(os.path.join(os.path.dirname(HERE), "tests", "golden", "compiled_layout.json")))
C = LAYOUT["constants"]
ID = LAYOUT["ids"]


def pack16(a, b):
    """Ints.pack(a, b) (M/utility/Ints.java:237-242)."""
    assert 0 <= a <= 65535 and 0 <= b <= 65535
    return ((a & 0xFFFF) << 16) | (b & 0xFFFF)


def pack8(a, b, c, d):
    """Ints.pack(a, b, c, d) (M/utility/Ints.java:258-265), as a signed int."""
    for x in (a, b, c, d):
        assert 0 <= x <= 255, x
    v = (a & 0xFF) | ((b & 0xFF) << 8) | ((c & 0xFF) << 16) | ((d & 0xFF) << 24)
    return v - (1 << 32) if v >= (1 << 31) else v


def padding(n, block=8):
    return 0 if n % block == 0 else block - n % block


def _face_normal(a, b, c):
    a, b, c = (np.asarray(x, np.float32) for x in (a, b, c))
    e1 = (b - a) / np.float32(np.linalg.norm(b - a))
    e2 = (c - a) / np.float32(np.linalg.norm(c - a))
    n = np.cross(e1, e2).astype(np.float32)
    return n / np.float32(np.linalg.norm(n))


def _argb(rgb):
    """Color3F.pack(): 0xFFRRGGBB as int bits in a float (exact: alpha 255
    keeps the value within 2^24 of zero); components rounded from x * 255,
    so colours that are multiples of 1/255 pack exactly whatever the
    reference's rounding rule (Ints.saturateAndScaleToInt, absent)."""
    q = [int(np.clip(np.round(np.float32(x) * 255.0), 0, 255)) for x in rgb]
    v = (255 << 24) | (q[0] << 16) | (q[1] << 8) | q[2]
    return float(v - (1 << 32) if v >= (1 << 31) else v)


class _Distinct:
    """Distinct records of one type in first-use order (the NodeCache's
    distinct lists) with their offsets."""

    def __init__(self, length=None):
        self.keys, self.records, self.length = {}, [], length

    def add(self, key, make):
        if key in self.keys:
            return self.keys[key]
        rec = make()
        index = len(self.records)
        self.records.append(rec)
        if self.length is None:              # indexed
            off = index
        else:                                # element offset
            off = sum(len(r) for r in self.records[:-1])
        self.keys[key] = off
        return off

    def array(self, dtype):
        if not self.records:
            return np.zeros(0, dtype)
        return np.concatenate([np.asarray(r, dtype) for r in self.records])


class SceneCompiler:
    def __init__(self, scene, object_to_world=None):
        self.sc = scene
        self.m = object_to_world
        n = len(scene.primitives)
        if self.m is not None:
            self.m = np.asarray(self.m, np.float64).reshape(n, 4, 4)
        self.const = _Distinct()
        self.tex = {k: _Distinct() for k in ("blend", "bullseye", "checkerboard", "polka_dot", "marble", "simplex_fbm")}
        self.ldr = _Distinct()
        self.mats = {k: _Distinct(length=True) for k in ("clear_coat", "disney", "glass", "glossy", "matte", "metal",
                                                         "mirror", "plastic", "substrate", "translucent", "uber")}
        self.triangles = []
        self.meshes = []          # int records, concatenated
        self.mesh_offsets = {}    # (first, count) -> offset of its mesh record
        self.aabbs = []
        self.rects = []

    # -- textures -----------------------------------------------------------
    def constant(self, rgb):
        rgb = tuple(float(np.float32(x)) for x in rgb)
        return ID["ConstantTexture"], self.const.add(rgb, lambda: [rgb[0], rgb[1], rgb[2], 0.0])

    def value(self, v):
        return self.constant((v, v, v))

    def texture(self, t):
        """(ID, offset) of the scene's texture index t."""
        T = self.sc.textures[t]
        if T.type == abi.TEX_CONSTANT:
            return self.constant(tuple(T.color))
        if T.type == abi.TEX_UV:
            return ID["UVTexture"], 0
        if T.type == abi.TEX_SURFACE_NORMAL:
            return ID["SurfaceNormalTexture"], 0
        if T.type == abi.TEX_DOT_PRODUCT:
            return ID["DotProductTexture"], 0
        if T.type == abi.TEX_IMAGE:
            im = np.asarray(self.sc.images[T.image], np.float32)
            h, w = im.shape[:2]
            q = np.clip(np.round(im * 255.0), 0, 255).astype(np.int64)
            packed = (255 << 24) | (q[..., 0] << 16) | (q[..., 1] << 8) | q[..., 2]
            packed = np.where(packed >= (1 << 31), packed - (1 << 32), packed).reshape(-1)
            rad = math.radians(T.angle)

            def make():
                rec = [rad, T.scale[0], T.scale[1], float(w), float(h)] + [float(v) for v in packed]
                return rec + [0.0] * padding(len(rec))
            return ID["LDRImageTexture"], self.ldr.add(("img", t), make)
        if T.type == abi.TEX_MARBLE:
            # CompiledTextureCache.toArray(MarbleTexture): colours as Color3F.pack()
            rec = [0.0] * C["MARBLE_TEXTURE_LENGTH"]
            rec[C["MARBLE_TEXTURE_OFFSET_COLOR_A"]] = _argb(T.color)
            rec[C["MARBLE_TEXTURE_OFFSET_COLOR_B"]] = _argb(T.origin)
            rec[C["MARBLE_TEXTURE_OFFSET_COLOR_C"]] = _argb(T.t)
            rec[C["MARBLE_TEXTURE_OFFSET_FREQUENCY"]] = T.scale[0]
            rec[C["MARBLE_TEXTURE_OFFSET_SCALE"]] = T.scale[1]
            rec[C["MARBLE_TEXTURE_OFFSET_STRIPES"]] = float(np.float32(T.scale[0]) / np.float32(math.pi))
            rec[C["MARBLE_TEXTURE_OFFSET_OCTAVES"]] = T.cell
            return ID["MarbleTexture"], self.tex["marble"].add(("mb", t), lambda: rec)
        if T.type == abi.TEX_SIMPLEX_FBM:
            rec = [0.0] * C["SIMPLEX_FRACTIONAL_BROWNIAN_MOTION_TEXTURE_LENGTH"]
            rec[C["SIMPLEX_FRACTIONAL_BROWNIAN_MOTION_TEXTURE_OFFSET_COLOR"]] = _argb(T.color)
            rec[C["SIMPLEX_FRACTIONAL_BROWNIAN_MOTION_TEXTURE_OFFSET_FREQUENCY"]] = T.scale[0]
            rec[C["SIMPLEX_FRACTIONAL_BROWNIAN_MOTION_TEXTURE_OFFSET_GAIN"]] = T.scale[1]
            rec[C["SIMPLEX_FRACTIONAL_BROWNIAN_MOTION_TEXTURE_OFFSET_OCTAVES"]] = T.cell
            return ID["SimplexFractionalBrownianMotionTexture"], self.tex["simplex_fbm"].add(("fbm", t), lambda: rec)
        a = pack16(*self.texture(T.a))
        b = pack16(*self.texture(T.b))
        if T.type == abi.TEX_CHECKERBOARD:
            rec = [T.angle, math.radians(T.angle), float(a), float(b), T.scale[0], T.scale[1], 0.0, 0.0]
            return ID["CheckerboardTexture"], self.tex["checkerboard"].add(("cb", t), lambda: rec)
        if T.type == abi.TEX_BLEND:
            rec = [float(a), float(b), T.t[0], T.t[1], T.t[2], 0.0, 0.0, 0.0]
            return ID["BlendTexture"], self.tex["blend"].add(("bl", t), lambda: rec)
        if T.type == abi.TEX_BULLSEYE:
            rec = [T.origin[0], T.origin[1], T.origin[2], float(a), float(b), T.scale[0], 0.0, 0.0]
            return ID["BullseyeTexture"], self.tex["bullseye"].add(("be", t), lambda: rec)
        if T.type == abi.TEX_POLKA_DOT:
            rec = [T.angle, math.radians(T.angle), float(a), float(b), T.cell, T.radius, 0.0, 0.0]
            return ID["PolkaDotTexture"], self.tex["polka_dot"].add(("pd", t), lambda: rec)
        raise ValueError(f"texture type {T.type}")

    def slot(self, M, k, color):
        return self.texture(M.textures[k] - 1) if M.textures[k] else self.constant(tuple(color))

    def float_slot(self, M, k, value):
        return self.texture(M.textures[k] - 1) if M.textures[k] else self.value(value)

    # -- materials ----------------------------------------------------------
    def material(self, i):
        M = self.sc.materials[i]
        emission = self.texture(M.emission_texture - 1) if M.emission_texture else self.constant(tuple(M.emission))
        em = pack8(*emission, ID["NoOpModifier"], 0)
        t = M.type
        if t == abi.MAT_MATTE:
            w = [em, pack8(*self.value(M.angle), *self.slot(M, abi.SLOT_KD, M.kd))]
            name = "matte"
        elif t == abi.MAT_MIRROR:
            w = [em, pack8(*self.slot(M, abi.SLOT_KR, M.kr), 0, 0)]
            name = "mirror"
        elif t == abi.MAT_GLASS:
            w = [em, pack8(*self.float_slot(M, abi.SLOT_ETA, M.eta[0]), 0, 0),
                 pack8(*self.slot(M, abi.SLOT_KR, M.kr), *self.slot(M, abi.SLOT_KT, M.kt)),
                 pack8(*self.value(M.roughness_u), *self.value(M.roughness_v)), int(M.remap_roughness), 0, 0, 0]
            name = "glass"
        elif t == abi.MAT_METAL:
            w = [em, pack8(*self.slot(M, abi.SLOT_ETA, M.eta), *self.slot(M, abi.SLOT_K, M.k)),
                 pack8(*self.float_slot(M, abi.SLOT_ROUGHNESS_U, M.roughness_u),
                       *self.float_slot(M, abi.SLOT_ROUGHNESS_V, M.roughness_v)), int(M.remap_roughness)]
            name = "metal"
        elif t == abi.MAT_PLASTIC:
            w = [em, pack8(*self.slot(M, abi.SLOT_KD, M.kd), *self.slot(M, abi.SLOT_KS, M.ks)),
                 pack8(*self.float_slot(M, abi.SLOT_ROUGHNESS_U, M.roughness_u), 0, 0), int(M.remap_roughness)]
            name = "plastic"
        elif t == abi.MAT_DISNEY:
            v = self.value
            w = [em, pack8(*v(M.disney_anisotropic), 0, 0),
                 pack8(*v(M.disney_clearcoat), *v(M.disney_clearcoat_gloss)),
                 pack8(*self.constant(tuple(M.kd)), *v(M.disney_diff_trans)),
                 pack8(*v(M.eta[0]), *v(M.disney_flatness)),
                 pack8(*v(M.disney_metallic), *v(M.disney_roughness)),
                 pack8(*self.constant((0.0, 0.0, 0.0)), *v(M.disney_sheen)),
                 pack8(*v(M.disney_sheen_tint), *v(M.disney_specular_tint)),
                 pack8(*v(M.disney_spec_trans), int(M.thin), 0)] + [0] * 7
            name = "disney"
        elif t == abi.MAT_GLOSSY:
            w = [em, pack8(*self.slot(M, abi.SLOT_KR, M.kr), *self.float_slot(M, abi.SLOT_ROUGHNESS_U, M.roughness_u))]
            name = "glossy"
        elif t == abi.MAT_SUBSTRATE:
            w = [em, pack8(*self.slot(M, abi.SLOT_KD, M.kd), *self.slot(M, abi.SLOT_KS, M.ks)),
                 pack8(*self.float_slot(M, abi.SLOT_ROUGHNESS_U, M.roughness_u),
                       *self.float_slot(M, abi.SLOT_ROUGHNESS_V, M.roughness_v)), int(M.remap_roughness)]
            name = "substrate"
        elif t == abi.MAT_TRANSLUCENT:
            w = [em, pack8(*self.constant(tuple(M.kd)), *self.constant(tuple(M.ks))),
                 pack8(*self.constant(tuple(M.kr)), *self.constant(tuple(M.kt))),
                 pack8(*self.value(M.roughness_u), int(M.remap_roughness), 0)]
            name = "translucent"
        elif t == abi.MAT_CLEAR_COAT:
            w = [em, pack8(*self.slot(M, abi.SLOT_KD, M.kd), *self.slot(M, abi.SLOT_KS, M.ks))]
            name = "clear_coat"
        elif t == abi.MAT_UBER:
            w = [em, pack8(*self.constant(tuple(M.kd)), *self.constant(tuple(M.kr))),
                 pack8(*self.constant(tuple(M.ks)), *self.constant(tuple(M.kt))),
                 pack8(*self.value(M.eta[0]), *self.constant(tuple(M.k))),
                 pack8(*self.value(M.roughness_u), *self.value(M.roughness_v)), int(M.remap_roughness), 0, 0]
            name = "uber"
        else:
            raise ValueError(f"material type {t}")
        key = {"matte": "MatteMaterial", "mirror": "MirrorMaterial", "glass": "GlassMaterial",
               "metal": "MetalMaterial", "plastic": "PlasticMaterial", "disney": "DisneyMaterial",
               "glossy": "GlossyMaterial", "substrate": "SubstrateMaterial", "translucent": "TranslucentMaterial",
               "clear_coat": "ClearCoatMaterial", "uber": "UberMaterial"}[name]
        return ID[key], self.mats[name].add(("m", i), lambda: w)

    # -- shapes -------------------------------------------------------------
    def triangle(self, t):
        """Offset of the scene's triangle t as a Triangle3F record."""
        sc = self.sc
        ia, ib, ic = sc.indices[t]
        P = [sc.positions[j] for j in (ia, ib, ic)]
        if sc.has_normals:
            N = [sc.normals[j] for j in (ia, ib, ic)]
        else:
            fn = _face_normal(*P)
            N = [fn, fn, fn]
        T = [sc.uvs[j] for j in (ia, ib, ic)]
        rec = [float(x) for p in P for x in p] + [float(x) for uv in T for x in uv] + [float(x) for n in N for x in n]
        assert len(rec) == C["TRIANGLE_3_F_LENGTH"]
        off = len(self.triangles) * C["TRIANGLE_3_F_LENGTH"]
        self.triangles.append(rec)
        return off

    def mesh(self, first, count):
        """Element offset of a TriangleMesh3F over triangles [first, first + count)
        (compiled once per distinct mesh: primitives that share a triangle
        range share one TriangleMesh3F, as primitives sharing a mesh object
        share one record in the reference's distinct lists, Shape3FCache.java
        :176-177, 298-300)."""
        if (first, count) in self.mesh_offsets:
            return self.mesh_offsets[(first, count)]
        sc = self.sc
        tris = list(range(first, first + count))
        offs = {t: self.triangle(t) for t in tris}
        cen = sc.positions[sc.indices[first:first + count]].mean(axis=1)
        nodes = []       # (depth, kind, payload) in depth-first order

        def build(ts, depth):
            p = sc.positions[sc.indices[ts]].reshape(-1, 3)
            self.aabbs.append(list(p.max(0)) + list(p.min(0)) + [0.0, 0.0])
            bv = len(self.aabbs) - 1
            if len(ts) <= 4:
                nodes.append((depth, ID["LeafBVHNode3F"], bv, [offs[t] for t in ts]))
                return
            c = cen[np.asarray(ts) - first]
            axis = int(np.argmax(c.max(0) - c.min(0)))
            order = [ts[k] for k in np.argsort(c[:, axis], kind="stable")]
            half = len(order) // 2
            nodes.append((depth, ID["TreeBVHNode3F"], bv, None))
            build(order[:half], depth + 1)
            build(order[half:], depth + 1)

        build(tris, 0)
        lengths = [4 + len(n[3]) + padding(4 + len(n[3])) if n[1] == ID["LeafBVHNode3F"] else 8 for n in nodes]
        starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(int)
        out = []
        for i, (depth, kind, bv, payload) in enumerate(nodes):
            nxt = next((int(starts[j]) for j in range(i + 1, len(nodes)) if nodes[j][0] <= depth), -1)
            if kind == ID["LeafBVHNode3F"]:
                rec = [kind, bv, nxt, len(payload)] + payload
                rec += [0] * padding(len(rec))
            else:
                left = next((int(starts[j]) for j in range(i + 1, len(nodes)) if nodes[j][0] == depth + 1), -1)
                rec = [kind, bv, nxt, left, 0, 0, 0, 0]
            out += rec
        base = sum(len(m) for m in self.meshes)
        self.meshes.append(out)
        self.mesh_offsets[(first, count)] = base
        return base

    # -- the whole scene ----------------------------------------------------
    def compile(self):
        sc = self.sc
        lights = []           # (ID, offset) per scene light
        point, spot, directional, diffuse, perez, image = [], [], [], [], [], []
        perez_offs, image_offs = [], []
        prim_of_light = {}
        for pi, p in enumerate(sc.primitives):
            if p.area_light >= 0:
                prim_of_light[p.area_light] = pi
        for li, L in enumerate(sc.lights):
            if L.type == abi.LIGHT_DIFFUSE_AREA:
                prim_of_light.setdefault(li, L.primitive)

        # primitives first (their shapes give the area lights' shape offsets)
        prim_recs, matrices, shape_of_prim = [], [], []
        for pi, p in enumerate(sc.primitives):
            M = np.eye(4)
            if p.shape == abi.SHAPE_TRIANGLE_MESH:
                lit = p.area_light >= 0 or any(L.type == abi.LIGHT_DIFFUSE_AREA and L.primitive == pi
                                               for L in sc.lights)
                if p.triangle_count == 1 and lit:
                    sid, soff = ID["Triangle3F"], self.triangle(p.first_triangle)
                else:
                    sid, soff = ID["TriangleMesh3F"], self.mesh(p.first_triangle, p.triangle_count)
            elif p.shape == abi.SHAPE_SPHERE:
                sid, soff = ID["Sphere3F"], 0
                c, r = np.array(p.data[0:3], np.float64), float(p.data[3])
                M = np.array([[r, 0, 0, c[0]], [0, r, 0, c[1]], [0, 0, r, c[2]], [0, 0, 0, 1]], np.float64)
            elif p.shape == abi.SHAPE_PLANE:
                sid, soff = ID["Plane3F"], 0
                o, u, v, w = (np.array(p.data[3 * k:3 * k + 3], np.float64) for k in range(4))
                M = np.eye(4)
                M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = u, v, w, o
            elif p.shape == abi.SHAPE_RECTANGLE:
                a, b, c = (np.array(p.data[3 * k:3 * k + 3], np.float32) for k in range(3))
                n = _face_normal(a, b, c)
                rec = list(a) + list(b - a) + list(c - b) + list(n) + [0.0] * 4
                sid, soff = ID["Rectangle3F"], len(self.rects) * C["RECTANGLE_3_F_LENGTH"]
                self.rects.append([float(x) for x in rec])
            else:
                raise ValueError(f"shape {p.shape}")
            if self.m is not None:
                M = self.m[pi] @ M
            shape_of_prim.append((sid, soff))
            mid, moff = self.material(p.material)
            matrices.append(np.concatenate([M.reshape(-1), np.linalg.inv(M).reshape(-1)]).astype(np.float32))
            prim_recs.append([0, 0, ID["AxisAlignedBoundingBox3F"], 0, mid, moff, sid, soff])

        for li, L in enumerate(sc.lights):
            if L.type == abi.LIGHT_POINT:
                lights.append((ID["PointLight"], len(point)))
                point.append(list(L.radiance) + list(L.position) + [0.0, 0.0])
            elif L.type == abi.LIGHT_DIRECTIONAL:
                lights.append((ID["DirectionalLight"], len(directional)))
                d = np.array(L.direction, np.float64)
                d = d / np.linalg.norm(d)
                directional.append(list(L.radiance) + list(d) + [L.radius, 0.0])
            elif L.type == abi.LIGHT_SPOT:
                lights.append((ID["SpotLight"], len(spot)))
                # the cone axis is the object-space +z: world to object maps
                # the axis to z (an orthonormal frame around it, then the position)
                z = np.array(L.direction, np.float64)
                z /= np.linalg.norm(z)
                x = np.cross([0.0, 1.0, 0.0] if abs(z[1]) < 0.9 else [1.0, 0.0, 0.0], z)
                x /= np.linalg.norm(x)
                y = np.cross(z, x)
                o2w = np.eye(4)
                o2w[:3, 0], o2w[:3, 1], o2w[:3, 2], o2w[:3, 3] = x, y, z, np.array(L.position, np.float64)
                w2o = np.linalg.inv(o2w)
                c1 = float(np.float32(math.cos(math.radians(L.cone_angle))))
                c2 = float(np.float32(math.cos(math.radians(L.cone_angle - L.cone_delta))))
                spot.append(list(w2o.reshape(-1)) + list(L.radiance) + list(L.position) + [c1, c2])
            elif L.type == abi.LIGHT_PEREZ:
                lights.append((ID["PerezLight"], len(perez)))
                rec = [0.0] * C["PEREZ_LIGHT_OFFSET_DISTRIBUTION"]
                d = np.array(L.direction, np.float64)
                d = d / np.linalg.norm(d)
                rec[C["PEREZ_LIGHT_OFFSET_SUN_DIRECTION_WORLD_SPACE"]:C["PEREZ_LIGHT_OFFSET_SUN_DIRECTION_WORLD_SPACE"] + 3] = list(d)
                rec[C["PEREZ_LIGHT_OFFSET_TURBIDITY"]] = L.turbidity
                perez_offs.append(sum(len(r) for r in perez))
                perez.append(rec)
            elif L.type == abi.LIGHT_IMAGE:
                lights.append((ID["ImageLight"], len(image)))
                im = np.asarray(sc.images[L.primitive], np.float32)
                h, w = im.shape[:2]
                rec = [0.0] * C["IMAGE_LIGHT_OFFSET_IMAGE"]
                rec[C["IMAGE_LIGHT_OFFSET_ANGLE_RADIANS"]] = math.radians(L.cone_angle)
                su, sv = (L.direction[0], L.direction[1]) if (L.direction[0] or L.direction[1]) else (1.0, 1.0)
                rec[C["IMAGE_LIGHT_OFFSET_SCALE"]:C["IMAGE_LIGHT_OFFSET_SCALE"] + 2] = [su, sv]
                rec[C["IMAGE_LIGHT_OFFSET_RADIUS"]] = L.radius if L.radius > 0 else 10.0
                rec[C["IMAGE_LIGHT_OFFSET_RESOLUTION_X"]] = float(w)
                rec[C["IMAGE_LIGHT_OFFSET_RESOLUTION_Y"]] = float(h)
                rec += [float(x) for x in im.reshape(-1)]
                image_offs.append(sum(len(r) for r in image))
                image.append(rec)
            elif L.type == abi.LIGHT_DIFFUSE_AREA:
                lights.append((ID["DiffuseAreaLight"], len(diffuse)))
                pi = prim_of_light[li]
                sid, soff = shape_of_prim[pi]
                rec = [0.0] * C["DIFFUSE_AREA_LIGHT_LENGTH"]
                M = matrices[pi]
                rec[0:32] = [float(x) for x in M]
                rec[C["DIFFUSE_AREA_LIGHT_OFFSET_RADIANCE_EMITTED"]:C["DIFFUSE_AREA_LIGHT_OFFSET_RADIANCE_EMITTED"] + 3] = list(L.radiance)
                rec[C["DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_ID"]] = float(sid)
                rec[C["DIFFUSE_AREA_LIGHT_OFFSET_SHAPE_OFFSET"]] = float(soff)
                rec[C["DIFFUSE_AREA_LIGHT_OFFSET_IS_TWO_SIDED"]] = 1.0 if L.two_sided else 0.0
                diffuse.append(rec)
                prim_recs[pi][C["PRIMITIVE_OFFSET_AREA_LIGHT_ID_AND_OFFSET"]] = pack16(ID["DiffuseAreaLight"], len(diffuse) - 1)
            else:
                raise ValueError(f"light type {L.type}")

        f32 = lambda recs: np.asarray([x for r in recs for x in r], np.float32)
        i32 = lambda recs: np.asarray([x for r in recs for x in r], np.int32)
        out = {
            "axis_aligned_bounding_boxes": f32(self.aabbs), "triangles": f32(self.triangles),
            "triangle_meshes": i32(self.meshes), "rectangles": f32(self.rects),
            "constant_textures": self.const.array(np.float32),
            "checkerboard_textures": self.tex["checkerboard"].array(np.float32),
            "blend_textures": self.tex["blend"].array(np.float32),
            "bullseye_textures": self.tex["bullseye"].array(np.float32),
            "polka_dot_textures": self.tex["polka_dot"].array(np.float32),
            "marble_textures": self.tex["marble"].array(np.float32),
            "simplex_fbm_textures": self.tex["simplex_fbm"].array(np.float32),
            "ldr_image_textures": self.ldr.array(np.float32),
            "ldr_image_texture_offsets": np.asarray(
                np.concatenate([[0], np.cumsum([len(r) for r in self.ldr.records])[:-1]]) if self.ldr.records else [],
                np.int32),
            "light_ids_and_offsets": np.asarray([pack16(a, b) for a, b in lights], np.int32),
            "point_lights": f32(point), "spot_lights": f32(spot), "directional_lights": f32(directional),
            "diffuse_area_lights": f32(diffuse), "perez_lights": f32(perez),
            "perez_light_offsets": np.asarray(perez_offs, np.int32), "image_lights": f32(image),
            "image_light_offsets": np.asarray(image_offs, np.int32),
            "camera": np.asarray(sc.camera, np.float32),
            "primitives": i32(prim_recs), "primitive_matrices": f32(matrices),
        }
        for name, d in self.mats.items():
            out[f"{name}_materials"] = d.array(np.int32)
        out["primitive_count"] = len(sc.primitives)
        out["light_count"] = len(lights)
        return out


def compile_scene(scene, object_to_world=None):
    """The compiled arrays of `scene` (dict of numpy arrays plus the two
    counts), as SceneCompiler.compile would hand them to doSetupScene."""
    return SceneCompiler(scene, object_to_world).compile()


def compatible(arrays):
    """doGetCompatibleArray: every empty array becomes one zero
    (AbstractSceneKernel.java:1857-1866)."""
    out = {}
    for name, kind in abi.COMPILED_SCENE_ARRAYS:
        a = arrays.get(name)
        dt = np.float32 if kind == "f" else np.int32
        a = np.zeros(0, dt) if a is None else np.asarray(a, dt)
        out[name] = a if a.size > 0 else np.zeros(1, dt)
    out["primitive_count"] = arrays["primitive_count"]
    out["light_count"] = arrays["light_count"]
    return out
