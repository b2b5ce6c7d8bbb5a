"""Scene description and synthetic scene generators.

A `Scene` holds world-space geometry (one concatenated triangle array for all
meshes plus analytic spheres, rectangles and planes), materials, lights and a
camera, and converts itself into the `dfr_scene_desc` record of the C-ABI.

The reference builds scenes from Java snippets compiled at run time
(M/scene/loader/JavaSceneLoader.java:252-278); that needs a JDK, so the
benchmark scenes are generated here from seeds instead:

* `cornell_box`     closed room of 5 rectangles, area light, glass + mirror spheres
* `displaced_icosphere`  geodesic icosphere with seeded fBm simplex displacement
* `triangle_soup`   random equilateral triangles with log-normal edge lengths

M/ = src/main/java/org/dayflower/ of the reference tree.
"""
import ctypes
import math

import numpy as np

from . import abi

# ---------------------------------------------------------------------------
# Camera (M/scene/Camera.java:75-124, 243-291; 19-float record of
# M/scene/compiler/CompiledCameraCache.java:158-197)
# ---------------------------------------------------------------------------


def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def camera_record(eye, look_at, width, height, fov_x_deg=40.0, up=(0.0, 1.0, 0.0), lens=abi.LENS_THIN,
                  focal_distance=30.0, aperture=0.0):
    """Camera looking from `eye` at `look_at`.

    The basis is OrthonormalBasis33F(w, v): w = view direction, u =
    normalize(cross(v, w)), v = cross(w, u) (M/geometry/OrthonormalBasis33F.java
    :110-114).  The vertical field of view follows AngleF.fieldOfViewY
    (M/geometry/AngleF.java:312-314) from the horizontal one.
    """
    w = _normalize(np.subtract(look_at, eye))
    u = _normalize(np.cross(_normalize(up), w))
    v = np.cross(w, u)
    fx = math.radians(fov_x_deg)
    fy = 2.0 * math.atan(math.tan(fx * 0.5) * (height / width))
    rec = np.zeros(abi.CAMERA_LENGTH, dtype=np.float32)
    rec[0] = math.tan(+fx * 0.5)
    rec[1] = math.tan(-fy * 0.5)
    rec[2] = lens
    rec[3:6] = u
    rec[6:9] = v
    rec[9:12] = w
    rec[12:15] = eye
    rec[15] = aperture
    rec[16] = focal_distance
    rec[17] = width
    rec[18] = height
    return rec


# ---------------------------------------------------------------------------
# Materials and lights
# ---------------------------------------------------------------------------


def _mat(type_, **kw):
    m = abi.Material()
    m.type = type_
    for k, v in kw.items():
        if isinstance(v, (tuple, list, np.ndarray)):
            arr = getattr(m, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        else:
            setattr(m, k, v)
    return m


def matte(kd, angle=0.0, emission=(0, 0, 0)):
    """MatteMaterial(kd, emission, angle) (M/scene/material/MatteMaterial.java:144)."""
    return _mat(abi.MAT_MATTE, kd=kd, angle=float(angle), emission=emission)


def mirror(kr=(1, 1, 1)):
    return _mat(abi.MAT_MIRROR, kr=kr)


def glass(eta=1.5, kr=(1, 1, 1), kt=(1, 1, 1), roughness=0.0, remap=True):
    return _mat(abi.MAT_GLASS, kr=kr, kt=kt, eta=(eta, 0, 0), roughness_u=float(roughness),
                roughness_v=float(roughness), remap_roughness=int(remap))


def metal(eta=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14), roughness=0.2, remap=True):
    return _mat(abi.MAT_METAL, eta=eta, k=k, roughness_u=float(roughness), roughness_v=float(roughness),
                remap_roughness=int(remap))


def plastic(kd=(0.5, 0.5, 0.5), ks=(0.25, 0.25, 0.25), roughness=0.2, remap=True):
    return _mat(abi.MAT_PLASTIC, kd=kd, ks=ks, roughness_u=float(roughness), roughness_v=float(roughness),
                remap_roughness=int(remap))


def disney(color=(0.8, 0.5, 0.3), metallic=0.0, roughness=0.5, specular_tint=0.0, anisotropic=0.0, sheen=0.0,
           sheen_tint=0.5, clearcoat=0.0, clearcoat_gloss=1.0, spec_trans=0.0, eta=1.5):
    return _mat(abi.MAT_DISNEY, kd=color, eta=(eta, 0, 0), disney_metallic=float(metallic),
                disney_roughness=float(roughness), disney_specular_tint=float(specular_tint),
                disney_anisotropic=float(anisotropic), disney_sheen=float(sheen),
                disney_sheen_tint=float(sheen_tint), disney_clearcoat=float(clearcoat),
                disney_clearcoat_gloss=float(clearcoat_gloss), disney_spec_trans=float(spec_trans))


def glossy(kr=(0.5, 0.5, 0.5), roughness=0.2):
    """GlossyMaterial: TorranceSparrowBRDF(KR) with ConductorFresnel(KR, 1, 1);
    the roughness is used as the microfacet alpha (no remapping)."""
    return _mat(abi.MAT_GLOSSY, kr=kr, roughness_u=float(roughness), roughness_v=float(roughness))


def substrate(kd=(0.5, 0.5, 0.5), ks=(0.5, 0.5, 0.5), roughness=0.1, remap=True):
    """SubstrateMaterial: FresnelBlendBRDF(KD, KS)."""
    return _mat(abi.MAT_SUBSTRATE, kd=kd, ks=ks, roughness_u=float(roughness), roughness_v=float(roughness),
                remap_roughness=int(remap))


def translucent(kd=(0.25, 0.25, 0.25), ks=(0.25, 0.25, 0.25), reflectance=(0.5, 0.5, 0.5),
                transmittance=(0.5, 0.5, 0.5), roughness=0.1, remap=True):
    """TranslucentMaterial: Lambertian BRDF/BTDF and Torrance-Sparrow
    BRDF/BTDF (1 -> 1.5); reflectance and transmittance travel in kr, kt."""
    return _mat(abi.MAT_TRANSLUCENT, kd=kd, ks=ks, kr=reflectance, kt=transmittance, roughness_u=float(roughness),
                roughness_v=float(roughness), remap_roughness=int(remap))


def clear_coat(kd=(0.5, 0.5, 0.5), ks=(1.0, 1.0, 1.0)):
    """ClearCoatMaterial: per hit, a specular (KS) or Lambertian (KD) layer
    chosen by Russian roulette on the 1 -> 1.5 dielectric reflectance."""
    return _mat(abi.MAT_CLEAR_COAT, kd=kd, ks=ks)


def uber(kd=(0.25, 0.25, 0.25), ks=(0.25, 0.25, 0.25), kr=(0.0, 0.0, 0.0), kt=(0.0, 0.0, 0.0),
         opacity=(1.0, 1.0, 1.0), eta=1.5, roughness=0.1, remap=True):
    """UberMaterial; the opacity travels in k."""
    return _mat(abi.MAT_UBER, kd=kd, ks=ks, kr=kr, kt=kt, k=opacity, eta=(eta, 0, 0), roughness_u=float(roughness),
                roughness_v=float(roughness), remap_roughness=int(remap))


class Scene:
    """World-space scene; see the module docstring."""

    def __init__(self):
        self.positions = np.zeros((0, 3), np.float32)
        self.normals = np.zeros((0, 3), np.float32)
        self.has_normals = False
        self.uvs = np.zeros((0, 2), np.float32)
        self.has_uvs = False
        self.indices = np.zeros((0, 3), np.int32)
        self.primitives = []
        self.materials = []
        self.lights = []
        self.images = []
        self.textures = []
        self.camera = camera_record((0, 0, -3), (0, 0, 0), 800, 800)
        self._keep = None

    # -- building ---------------------------------------------------------
    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_mesh(self, positions, indices, material, normals=None, area_light=-1, uvs=None):
        positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        base_v = self.positions.shape[0]
        first = self.indices.shape[0]
        if normals is None:
            normals = _flat_vertex_normals(positions, indices)
        else:
            self.has_normals = True
        if uvs is None:
            uvs = np.zeros((positions.shape[0], 2), np.float32)
        else:
            uvs = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
            if uvs.shape[0] != positions.shape[0]:
                raise ValueError("one (u, v) per vertex")
            self.has_uvs = True
        self.uvs = np.concatenate([self.uvs, uvs])
        self.positions = np.concatenate([self.positions, positions])
        self.normals = np.concatenate([self.normals, np.ascontiguousarray(normals, np.float32).reshape(-1, 3)])
        self.indices = np.concatenate([self.indices, indices + base_v])
        p = abi.Primitive()
        p.shape = abi.SHAPE_TRIANGLE_MESH
        p.material = material
        p.area_light = area_light
        p.first_triangle = first
        p.triangle_count = indices.shape[0]
        self.primitives.append(p)
        return len(self.primitives) - 1

    def add_instance(self, prim, material=None):
        """Another primitive over mesh primitive `prim`'s triangles (the same
        triangle range), with `material` or prim's.  Place it with a
        per-primitive object-to-world matrix (GPURenderer.updateTransforms);
        the library keeps one object-space copy of the triangles and a
        top-level tree over the instances.  Returns the new primitive."""
        src = self.primitives[prim]
        if src.shape != abi.SHAPE_TRIANGLE_MESH:
            raise ValueError("instances are of mesh primitives")
        p = abi.Primitive()
        p.shape = abi.SHAPE_TRIANGLE_MESH
        p.material = src.material if material is None else material
        p.area_light = -1
        p.first_triangle = src.first_triangle
        p.triangle_count = src.triangle_count
        self.primitives.append(p)
        return len(self.primitives) - 1

    def _add_analytic(self, shape, material, data):
        p = abi.Primitive()
        p.shape = shape
        p.material = material
        p.area_light = -1
        p.first_triangle = 0
        p.triangle_count = 0
        for i, x in enumerate(data):
            p.data[i] = float(x)
        self.primitives.append(p)
        return len(self.primitives) - 1

    def add_sphere(self, center, radius, material):
        return self._add_analytic(abi.SHAPE_SPHERE, material, list(center) + [radius])

    def add_rectangle(self, a, b, c, material):
        return self._add_analytic(abi.SHAPE_RECTANGLE, material, list(a) + list(b) + list(c))

    def add_plane(self, origin, u, v, w, material):
        return self._add_analytic(abi.SHAPE_PLANE, material, list(origin) + list(u) + list(v) + list(w))

    def add_triangle_light(self, a, b, c, radiance, material, two_sided=False):
        """DiffuseAreaLight over a Triangle3F (M/scene/light/DiffuseAreaLight.java:190)."""
        li = len(self.lights)
        prim = self.add_mesh(np.array([a, b, c], np.float32), np.array([[0, 1, 2]], np.int32), material,
                             area_light=li)
        L = abi.Light()
        L.type = abi.LIGHT_DIFFUSE_AREA
        L.primitive = prim
        L.triangle = self.primitives[prim].first_triangle
        L.two_sided = int(two_sided)
        for i in range(3):
            L.radiance[i] = float(radiance[i])
        self.lights.append(L)
        return li

    def add_sphere_light(self, center, radius, radiance, material, two_sided=False):
        """DiffuseAreaLight over a Sphere3F (M/scene/light/DiffuseAreaLight.java
        :190; sampled by M/geometry/shape/Sphere3F.java:156-226).  Returns the
        light index; the sphere primitive names it as its area light."""
        li = len(self.lights)
        prim = self.add_sphere(center, radius, material)
        self.primitives[prim].area_light = li
        L = abi.Light()
        L.type = abi.LIGHT_DIFFUSE_AREA
        L.primitive = prim
        L.triangle = -1
        L.two_sided = int(two_sided)
        for i in range(3):
            L.radiance[i] = float(radiance[i])
        self.lights.append(L)
        return li

    def add_point_light(self, position, intensity):
        """PointLight(intensity, position) (M/scene/light/PointLight.java:100)."""
        L = abi.Light()
        L.type = abi.LIGHT_POINT
        L.primitive = -1
        L.triangle = -1
        for i in range(3):
            L.radiance[i] = float(intensity[i])
            L.position[i] = float(position[i])
        self.lights.append(L)
        return len(self.lights) - 1

    def add_spot_light(self, position, direction, intensity, cone_angle=30.0, cone_delta=5.0):
        """SpotLight(coneAngle, coneAngleDelta, intensity, position, direction)
        (M/scene/light/SpotLight.java:184); angles in degrees."""
        L = abi.Light()
        L.type = abi.LIGHT_SPOT
        L.primitive = -1
        L.triangle = -1
        L.cone_angle = float(cone_angle)
        L.cone_delta = float(cone_delta)
        for i in range(3):
            L.radiance[i] = float(intensity[i])
            L.position[i] = float(position[i])
            L.direction[i] = float(direction[i])
        self.lights.append(L)
        return len(self.lights) - 1

    def add_directional_light(self, direction, radiance, radius=10.0):
        """DirectionalLight(transform, radiance, direction, radius)
        (M/scene/light/DirectionalLight.java:150); `direction` points towards
        the light, `radius` bounds the scene."""
        L = abi.Light()
        L.type = abi.LIGHT_DIRECTIONAL
        L.primitive = -1
        L.triangle = -1
        L.radius = float(radius)
        for i in range(3):
            L.radiance[i] = float(radiance[i])
            L.direction[i] = float(direction[i])
        self.lights.append(L)
        return len(self.lights) - 1

    # -- textures (M/scene/texture/*Texture.java) ---------------------------
    def _add_texture(self, type_, **kw):
        t = abi.Texture()
        t.type = type_
        t.a = t.b = t.image = -1
        t.scale[0] = t.scale[1] = 1.0
        for k, v in kw.items():
            f = getattr(t, k)
            if isinstance(f, (int, float)):
                setattr(t, k, v)
            else:
                for i, x in enumerate(v):
                    f[i] = float(x)
        self.textures.append(t)
        return len(self.textures) - 1

    def constant_texture(self, color):
        return self._add_texture(abi.TEX_CONSTANT, color=color)

    def checkerboard_texture(self, a, b, angle_deg=0.0, scale=(1.0, 1.0)):
        return self._add_texture(abi.TEX_CHECKERBOARD, a=a, b=b, angle=float(angle_deg), scale=scale)

    def uv_texture(self):
        return self._add_texture(abi.TEX_UV)

    def surface_normal_texture(self):
        return self._add_texture(abi.TEX_SURFACE_NORMAL)

    def dot_product_texture(self):
        return self._add_texture(abi.TEX_DOT_PRODUCT)

    def bullseye_texture(self, a, b, origin=(0.0, 0.0, 0.0), scale=1.0):
        return self._add_texture(abi.TEX_BULLSEYE, a=a, b=b, origin=origin, scale=(scale, scale))

    def polka_dot_texture(self, a, b, angle_deg=0.0, cell=10.0, radius=0.25):
        return self._add_texture(abi.TEX_POLKA_DOT, a=a, b=b, angle=float(angle_deg), cell=float(cell),
                                 radius=float(radius))

    def blend_texture(self, a, b, t=(0.5, 0.5, 0.5)):
        return self._add_texture(abi.TEX_BLEND, a=a, b=b, t=t)

    def marble_texture(self, a=(0.8, 0.8, 0.8), b=(0.4, 0.2, 0.1), c=(0.06, 0.04, 0.02), scale=5.0, stripes=0.15,
                       octaves=8):
        """MarbleTexture(colorA, colorB, colorC, scale, stripes, octaves)
        (M/scene/texture/MarbleTexture.java:109-117; defaults as its
        no-argument constructor): the record's frequency is PI * stripes in
        float."""
        frequency = np.float32(np.float32(np.pi) * np.float32(stripes))
        return self._add_texture(abi.TEX_MARBLE, color=a, origin=b, t=c, scale=(float(frequency), scale),
                                 cell=float(octaves))

    def simplex_fbm_texture(self, color=(0.75, 0.5, 0.75), frequency=5.0, gain=0.5, octaves=16):
        """SimplexFractionalBrownianMotionTexture(color, frequency, gain,
        octaves) (defaults as its no-argument constructor)."""
        return self._add_texture(abi.TEX_SIMPLEX_FBM, color=color, scale=(frequency, gain), cell=float(octaves))

    def image_texture(self, image, angle_deg=0.0, scale=(1.0, 1.0)):
        """LDRImageTexture: `image` (H, W, 3) linear RGB in [0, 1], row 0 first."""
        im = np.ascontiguousarray(image, np.float32)
        if im.ndim != 3 or im.shape[2] != 3:
            raise ValueError("image must be (H, W, 3)")
        self.images.append(im)
        return self._add_texture(abi.TEX_IMAGE, image=len(self.images) - 1, angle=float(angle_deg), scale=scale)

    def set_emission_texture(self, material, texture):
        """The material's emission from `texture` (textureEmission)."""
        self.materials[material].emission_texture = texture + 1

    def set_texture(self, material, slot, texture):
        """Replaces a colour slot (abi.SLOT_KD, KS, KR, KT) of a material by a texture."""
        self.materials[material].textures[slot] = texture + 1

    def add_image_light(self, image, angle_deg=0.0, scale=(1.0, 1.0), radius=10.0):
        """ImageLight(image, angle, scale) (M/scene/light/ImageLight.java:134):
        `image` is an (H, W, 3) array of linear RGB, row 0 first."""
        im = np.ascontiguousarray(image, np.float32)
        if im.ndim != 3 or im.shape[2] != 3 or im.shape[0] < 1 or im.shape[1] < 1:
            raise ValueError("image must be (H, W, 3)")
        self.images.append(im)
        li = len(self.lights)
        L = abi.Light()
        L.type = abi.LIGHT_IMAGE
        L.primitive = len(self.images) - 1
        L.triangle = -1
        L.cone_angle = float(angle_deg)
        L.direction[0], L.direction[1] = float(scale[0]), float(scale[1])
        L.radius = float(radius)
        self.lights.append(L)
        return li

    def add_perez_light(self, turbidity=2.0, sun_direction=(1.0, 1.0, -1.0)):
        """PerezLight(turbidity, sunDirectionWorldSpace) (M/scene/light/PerezLight.java)."""
        L = abi.Light()
        L.type = abi.LIGHT_PEREZ
        L.primitive = -1
        L.triangle = -1
        L.turbidity = float(turbidity)
        for i in range(3):
            L.direction[i] = float(sun_direction[i])
        self.lights.append(L)
        return len(self.lights) - 1

    def set_camera(self, rec):
        self.camera = np.asarray(rec, np.float32).copy()

    # -- C-ABI record -------------------------------------------------------
    @property
    def n_triangles(self):
        return int(self.indices.shape[0])

    def desc(self):
        """dfr_scene_desc pointing into this object's arrays (kept alive by it)."""
        pos = np.ascontiguousarray(self.positions, np.float32)
        nrm = np.ascontiguousarray(self.normals, np.float32)
        idx = np.ascontiguousarray(self.indices, np.int32)
        prims = (abi.Primitive * max(1, len(self.primitives)))(*self.primitives)
        mats = (abi.Material * max(1, len(self.materials)))(*self.materials)
        lights = (abi.Light * max(1, len(self.lights)))(*self.lights)
        d = abi.SceneDesc()
        d.n_vertices = pos.shape[0]
        fp = ctypes.POINTER(ctypes.c_float)
        d.positions = pos.ctypes.data_as(fp)
        d.normals = nrm.ctypes.data_as(fp) if self.has_normals else None
        uv = np.ascontiguousarray(self.uvs, np.float32)
        d.uvs = uv.ctypes.data_as(fp) if self.has_uvs else None
        d.n_triangles = idx.shape[0]
        d.indices = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.n_primitives = len(self.primitives)
        d.primitives = prims
        d.n_materials = len(self.materials)
        d.materials = mats
        d.n_lights = len(self.lights)
        d.lights = lights
        for i in range(abi.CAMERA_LENGTH):
            d.camera[i] = float(self.camera[i])
        imgs = [np.ascontiguousarray(im, np.float32) for im in self.images]
        recs = (abi.Image * max(1, len(imgs)))()
        for k, im in enumerate(imgs):
            recs[k].height, recs[k].width = int(im.shape[0]), int(im.shape[1])
            recs[k].rgb = im.ctypes.data_as(fp)
        d.n_images = len(imgs)
        d.images = recs if imgs else None
        texs = (abi.Texture * max(1, len(self.textures)))(*self.textures)
        d.n_textures = len(self.textures)
        d.textures = texs if self.textures else None
        self._keep = (pos, nrm, idx, prims, mats, lights, imgs, recs, uv, texs)
        return d


def _flat_vertex_normals(positions, indices):
    """Per-vertex normals for flat meshes: the face normal of the (unshared)
    vertices; shared vertices get the normalised sum of their faces."""
    p = positions.astype(np.float64)
    a, b, c = p[indices[:, 0]], p[indices[:, 1]], p[indices[:, 2]]
    fn = np.cross(b - a, c - a)
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, indices[:, k], fn)
    ln = np.linalg.norm(vn, axis=1, keepdims=True)
    ln[ln == 0] = 1.0
    return (vn / ln).astype(np.float32)


# ---------------------------------------------------------------------------
# Seeded simplex noise (3D, after Gustavson's "Simplex noise demystified")
# ---------------------------------------------------------------------------
_GRAD3 = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
                   [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], np.float64)


def _simplex_order(x0, y0, z0):
    """Offsets (i1,j1,k1), (i2,j2,k2) of the second and third simplex corners."""
    a = np.stack([x0, y0, z0], -1)
    order = np.argsort(-a, axis=-1, kind="stable")
    first = np.zeros(a.shape, np.int64)
    second = np.zeros(a.shape, np.int64)
    np.put_along_axis(first, order[..., :1], 1, -1)
    np.put_along_axis(second, order[..., :1], 1, -1)
    np.put_along_axis(second, order[..., 1:2], 1, -1)
    return first, second


def simplex3_ordered(x, y, z, perm):
    """3D simplex noise with the corner order derived by sorting the offsets."""
    F3, G3 = 1.0 / 3.0, 1.0 / 6.0
    s = (x + y + z) * F3
    i, j, k = np.floor(x + s), np.floor(y + s), np.floor(z + s)
    t = (i + j + k) * G3
    x0, y0, z0 = x - (i - t), y - (j - t), z - (k - t)
    o1, o2 = _simplex_order(x0, y0, z0)
    ii, jj, kk = (i.astype(np.int64) & 255), (j.astype(np.int64) & 255), (k.astype(np.int64) & 255)
    total = 0.0
    for off, g_off in ((np.zeros_like(o1), 0.0), (o1, G3), (o2, 2 * G3), (np.ones_like(o1), 3 * G3)):
        xc = x0 - off[..., 0] + g_off
        yc = y0 - off[..., 1] + g_off
        zc = z0 - off[..., 2] + g_off
        g = perm[ii + off[..., 0] + perm[jj + off[..., 1] + perm[kk + off[..., 2]]]] % 12
        tt = 0.6 - xc * xc - yc * yc - zc * zc
        gr = _GRAD3[g]
        total = total + (tt > 0) * (np.maximum(tt, 0) ** 4) * (gr[..., 0] * xc + gr[..., 1] * yc + gr[..., 2] * zc)
    return 32.0 * total


def fbm3(p, seed, octaves=3, frequency=2.0, gain=0.5, lacunarity=2.0):
    """Fractal sum of `octaves` simplex octaves at points p (N,3)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm]).astype(np.int64)
    out = np.zeros(p.shape[0])
    amp, f = 1.0, frequency
    for _ in range(octaves):
        out += amp * simplex3_ordered(p[:, 0] * f, p[:, 1] * f, p[:, 2] * f, perm)
        amp *= gain
        f *= lacunarity
    return out


# ---------------------------------------------------------------------------
# Geodesic icosphere
# ---------------------------------------------------------------------------


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_icosphere(frequency):
    """Icosahedron with every face split into frequency^2 triangles and the
    vertices projected to the unit sphere: 20*f^2 triangles, 10*f^2+2 shared
    vertices (closed, consistently wound outward)."""
    n = int(frequency)
    v, f = _icosahedron()
    # lattice points (i, j) with i + j <= n on every face
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    mask = ii + jj <= n
    ii, jj = ii[mask], jj[mask]
    local = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(ii, jj))}
    A, B, C = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    pts = (A[:, None, :] + (B - A)[:, None, :] * (ii[None, :, None] / n) + (C - A)[:, None, :] * (jj[None, :, None] / n))
    pts /= np.linalg.norm(pts, axis=2, keepdims=True)
    # triangles of the lattice: (i,j),(i+1,j),(i,j+1) and (i+1,j),(i+1,j+1),(i,j+1)
    tris = []
    for a in range(n):
        for b in range(n - a):
            tris.append((local[(a, b)], local[(a + 1, b)], local[(a, b + 1)]))
            if a + b + 1 < n:
                tris.append((local[(a + 1, b)], local[(a + 1, b + 1)], local[(a, b + 1)]))
    tris = np.array(tris, np.int64)
    npf = ii.shape[0]
    allp = pts.reshape(-1, 3)
    allt = (tris[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    # merge duplicated edge / corner vertices
    q = np.round(allp * 2**30).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    verts = allp[first]
    allt = inv.reshape(-1)[allt]
    if verts.shape[0] != 10 * n * n + 2:
        raise RuntimeError("icosphere vertex merge failed")
    # orient outward
    a, b, c = verts[allt[:, 0]], verts[allt[:, 1]], verts[allt[:, 2]]
    fn = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", fn, a + b + c) < 0
    allt[flip] = allt[flip][:, [0, 2, 1]]
    return verts, allt.astype(np.int32)


def displaced_icosphere(n_triangles, seed=0x5EED, amplitude=0.1, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Geodesic icosphere of about `n_triangles` triangles (frequency
    round(sqrt(n/20))), each vertex moved along its radius by
    amplitude * fBm3(simplex, 3 octaves, frequency 2, gain 0.5, lacunarity 2).
    Returns (positions, indices, vertex normals)."""
    freq = max(1, int(round(math.sqrt(n_triangles / 20.0))))
    verts, tris = geodesic_icosphere(freq)
    d = fbm3(verts, seed)
    p = verts * (radius * (1.0 + amplitude * d))[:, None] + np.asarray(center)[None, :]
    normals = _smooth_normals(p, tris)
    return p.astype(np.float32), tris, normals.astype(np.float32)


def _smooth_normals(p, tris):
    a, b, c = p[tris[:, 0]], p[tris[:, 1]], p[tris[:, 2]]
    fn = np.cross(b - a, c - a)
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, tris[:, k], fn)
    return vn / np.linalg.norm(vn, axis=1, keepdims=True)


# ---------------------------------------------------------------------------
# Triangle soup
# ---------------------------------------------------------------------------


def random_rotations(n, rng):
    """Uniform random rotation matrices from uniform unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def triangle_soup(n, seed=0x5EED, mu=-5.0, sigma=0.5):
    """n triangles: centroid ~ U[0,1]^3, edge length ~ LogNormal(mu, sigma),
    uniform random orientation, equilateral template.  Returns (positions
    (3n,3), indices (n,3))."""
    rng = np.random.default_rng(seed)
    cen = rng.random((n, 3))
    edge = rng.lognormal(mu, sigma, n)
    tmpl = np.array([[1.0, 0.0, 0.0], [-0.5, math.sqrt(3) / 2, 0.0], [-0.5, -math.sqrt(3) / 2, 0.0]])
    tmpl /= math.sqrt(3.0)  # unit edge length (circumradius 1/sqrt(3))
    out = np.empty((n, 3, 3), np.float32)
    chunk = 1 << 20
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        R = random_rotations(e - s, rng)
        local = np.einsum("nij,kj->nki", R, tmpl) * edge[s:e, None, None]
        out[s:e] = (local + cen[s:e, None, :]).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return out.reshape(-1, 3), idx


# ---------------------------------------------------------------------------
# Benchmark scenes
# ---------------------------------------------------------------------------


def cornell_box(width=400, height=400):
    """Closed Cornell-style room (5 Rectangle3F walls, Matte 0.75 white / red
    (0.75,0.1,0.1) / green (0.1,0.75,0.1)), a 0.5 x 0.5 ceiling light made of
    two Triangle3F DiffuseAreaLights of radiance 12, a glass (eta 1.5) and a
    mirror sphere of radius 0.35, pinhole camera with fov 40 degrees."""
    s = Scene()
    white = s.add_material(matte((0.75, 0.75, 0.75)))
    red = s.add_material(matte((0.75, 0.1, 0.1)))
    green = s.add_material(matte((0.1, 0.75, 0.1)))
    black = s.add_material(matte((0.0, 0.0, 0.0)))
    glass_m = s.add_material(glass(1.5, roughness=0.0))
    mirror_m = s.add_material(mirror((1.0, 1.0, 1.0)))
    # Rectangle3F(a, b, c): normal = normalize(cross(b - a, c - a)) faces the room
    s.add_rectangle((-1, 0, -1), (-1, 0, 1), (1, 0, 1), white)        # floor, normal +y
    s.add_rectangle((-1, 2, -1), (1, 2, -1), (1, 2, 1), white)        # ceiling, normal -y
    s.add_rectangle((-1, 0, 1), (-1, 2, 1), (1, 2, 1), white)         # back wall z=1, normal -z
    s.add_rectangle((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), red)        # left wall x=-1, normal +x
    s.add_rectangle((1, 0, -1), (1, 0, 1), (1, 2, 1), green)          # right wall x=1, normal -x
    y = 1.999
    a, b, c, d = (-0.25, y, -0.25), (0.25, y, -0.25), (0.25, y, 0.25), (-0.25, y, 0.25)
    s.add_triangle_light(a, b, c, (12.0, 12.0, 12.0), black)  # normal -y
    s.add_triangle_light(a, c, d, (12.0, 12.0, 12.0), black)
    s.add_sphere((-0.45, 0.35, -0.3), 0.35, glass_m)
    s.add_sphere((0.45, 0.35, 0.3), 0.35, mirror_m)
    s.set_camera(camera_record((0.0, 1.0, -3.4), (0.0, 1.0, 0.0), width, height, fov_x_deg=40.0))
    return s


def icosphere_scene(n_triangles, seed=0x5EED, width=1920, height=1080, mesh_material=None, with_perez=True):
    """Noise-displaced icosphere on a ground Plane at y = -1.1; Plastic
    (roughness 0.2) mesh over Matte ground; a point light (50 at (2,4,2)) and
    optionally a Perez sun/sky (turbidity 2, sun (1,1,-1))."""
    s = Scene()
    mm = s.add_material(mesh_material if mesh_material is not None else plastic())
    ground = s.add_material(matte((0.5, 0.5, 0.5)))
    p, t, nrm = displaced_icosphere(n_triangles, seed)
    s.add_mesh(p, t, mm, normals=nrm)
    s.add_plane((0.0, -1.1, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), ground)
    s.add_point_light((2.0, 4.0, 2.0), (50.0, 50.0, 50.0))
    if with_perez:
        s.add_perez_light(2.0, (1.0, 1.0, -1.0))
    s.set_camera(camera_record((0.0, 0.4, -3.6), (0.0, -0.1, 0.0), width, height, fov_x_deg=40.0))
    return s


def soup_scene(n_triangles, seed=0x5EED, width=2048, height=2048):
    """Triangle soup in the unit cube, matte, camera outside looking at its
    centre with fov 40 degrees."""
    s = Scene()
    m = s.add_material(matte((0.75, 0.75, 0.75)))
    p, t = triangle_soup(n_triangles, seed)
    s.add_mesh(p, t, m)
    s.set_camera(camera_record((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), width, height, fov_x_deg=40.0))
    return s
