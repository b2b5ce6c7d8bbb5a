"""Host-side mirror of the reference renderer API over the HIP C-ABI.

`GPURenderer` keeps the method names, defaults and error behaviour of the
reference's GPURenderer (M/renderer/gpu/AbstractGPURenderer.java:143-556,
M/renderer/gpu/GPURenderer.java:55-57; M/ = src/main/java/org/dayflower/):
setScene, setImage, setMaximumBounce, setMinimumBounceRussianRoulette,
setRenderingAlgorithm, setSamples, setMaximumDistance, setup, render(n),
clear, renderShutdown, dispose.  Every call goes through
`libdayflower_hip.so`; there is no CPU fallback.
"""
import ctypes
import enum

import numpy as np

from . import abi


class NullPointerException(TypeError):
    pass


class IllegalArgumentException(ValueError):
    pass


class IllegalStateException(RuntimeError):
    pass


class DeviceException(RuntimeError):
    pass


class UnsupportedOperationException(NotImplementedError):
    """A scene feature the compiled path does not carry (DFR_E_UNSUPPORTED)."""


class RenderingAlgorithm(enum.IntEnum):
    """M/renderer/RenderingAlgorithm.java:33-53 (ordinals as on the GPU)."""
    AMBIENT_OCCLUSION = abi.ALGO_AMBIENT_OCCLUSION
    DEPTH_CAMERA = abi.ALGO_DEPTH_CAMERA
    PATH_TRACING = abi.ALGO_PATH_TRACING
    RAY_CASTING = abi.ALGO_RAY_CASTING
    RAY_TRACING = abi.ALGO_RAY_TRACING


class ToneMapper(enum.IntEnum):
    """AbstractGPURenderer.java:50-79 (TONE_MAPPER_* ordinals; NONE is the default)."""
    FILMIC_CURVE_ACES_MODIFIED_VERSION_1 = abi.TONE_MAPPER_FILMIC_CURVE_ACES_MODIFIED_VERSION_1
    NONE = abi.TONE_MAPPER_NONE
    REINHARD = abi.TONE_MAPPER_REINHARD
    REINHARD_MODIFIED_VERSION_1 = abi.TONE_MAPPER_REINHARD_MODIFIED_VERSION_1
    REINHARD_MODIFIED_VERSION_2 = abi.TONE_MAPPER_REINHARD_MODIFIED_VERSION_2
    UNREAL_3 = abi.TONE_MAPPER_UNREAL_3


def _check(lib, ctx, rc):
    if rc >= 0:
        return rc
    msg = lib.dfr_last_error(ctx)
    msg = msg.decode() if msg else ""
    if rc == abi.E_NULL:
        raise NullPointerException(msg)
    if rc == abi.E_ARGUMENT:
        raise IllegalArgumentException(msg)
    if rc == abi.E_STATE:
        raise IllegalStateException(msg)
    if rc == abi.E_STOPPED:
        return rc
    if rc == abi.E_UNSUPPORTED:
        raise UnsupportedOperationException(msg)
    raise DeviceException(f"{msg} (code {rc})")


class GPURenderer:
    """Drop-in for the reference GPURenderer (one context per instance)."""

    KERNEL_WAVEFRONT = 0
    KERNEL_PLAIN = 1

    def __init__(self, renderer_observer=None, device=-1):
        self.lib = abi.load()
        self.ctx = self.lib.dfr_create(device)
        if not self.ctx:
            raise DeviceException("dfr_create failed (no HIP device?)")
        self.renderer_observer = renderer_observer
        self._scene = None
        self._desc = None

    # -- lifecycle ----------------------------------------------------------
    def dispose(self):
        if self.ctx:
            self.lib.dfr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

    def _c(self, rc):
        return _check(self.lib, self.ctx, rc)

    # -- setters (AbstractGPURenderer.java:447-526) --------------------------
    def setScene(self, scene):
        if scene is None:
            raise NullPointerException("scene == null")
        self._scene = scene
        self._desc = scene.desc()
        self._c(self.lib.dfr_set_scene(self.ctx, ctypes.byref(self._desc)))
        w, h = int(scene.camera[abi.CAMERA_LENGTH - 2]), int(scene.camera[abi.CAMERA_LENGTH - 1])
        self._c(self.lib.dfr_set_resolution(self.ctx, w, h))

    def updateTransforms(self, object_to_world, rebuild=False):
        """updateMatrix44Fs (AbstractSceneKernel.java:292-300): one 4x4
        object-to-world matrix per primitive of the scene, shape (n, 4, 4) or
        (n, 16), row major.  After setup the geometry moves at once (the BVH
        is refitted, or rebuilt with rebuild=True)."""
        import numpy as np
        m = np.ascontiguousarray(np.asarray(object_to_world, dtype=np.float32).reshape(-1, 16))
        self._xforms = m
        self._c(self.lib.dfr_update_transforms(self.ctx, m.ctypes.data, m.shape[0], 1 if rebuild else 0))

    def setCompiledScene(self, arrays):
        """setScene for a caller holding SceneCompiler.compile's arrays
        (AbstractSceneKernel.java:1777-1855): `arrays` maps the field names of
        dfr_compiled_scene to float32 / int32 arrays, plus "primitive_count"
        and "light_count".  The resolution comes from the compiled camera."""
        if arrays is None:
            raise NullPointerException("scene == null")
        cs, keep = compiled_scene_struct(arrays)
        self._compiled = (cs, keep)
        self._scene = None
        self._c(self.lib.dfr_set_compiled_scene(self.ctx, ctypes.byref(cs)))
        cam = keep["camera"]
        self._c(self.lib.dfr_set_resolution(self.ctx, int(cam[abi.CAMERA_LENGTH - 2]), int(cam[abi.CAMERA_LENGTH - 1])))

    def updateMatrix44Fs(self, matrices):
        """updateMatrix44Fs with CompiledPrimitiveCache.getMatrix44Fs()'s
        array: 32 floats per primitive (object to world, world to object)."""
        import numpy as np
        m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float32).reshape(-1))
        self._c(self.lib.dfr_update_matrix44fs(self.ctx, m.ctypes.data, m.size))

    def setImage(self, width=None, height=None):
        """setImage(): resolution from the scene camera; setImage(w, h) explicit."""
        if width is None:
            if self._scene is None:
                raise NullPointerException("scene == null")
            width = int(self._scene.camera[abi.CAMERA_LENGTH - 2])
            height = int(self._scene.camera[abi.CAMERA_LENGTH - 1])
        self._c(self.lib.dfr_set_resolution(self.ctx, int(width), int(height)))

    def _seti(self, key, v):
        self._c(self.lib.dfr_set_i32(self.ctx, key, int(v)))

    def _geti(self, key):
        out = ctypes.c_int32()
        self._c(self.lib.dfr_get_i32(self.ctx, key, ctypes.byref(out)))
        return out.value

    def setMaximumBounce(self, v):
        self._seti(abi.KEY_MAXIMUM_BOUNCE, v)

    def getMaximumBounce(self):
        return self._geti(abi.KEY_MAXIMUM_BOUNCE)

    def setMinimumBounceRussianRoulette(self, v):
        self._seti(abi.KEY_MINIMUM_BOUNCE_RR, v)

    def getMinimumBounceRussianRoulette(self):
        return self._geti(abi.KEY_MINIMUM_BOUNCE_RR)

    def setRenderingAlgorithm(self, algorithm):
        if algorithm is None:
            raise NullPointerException("renderingAlgorithm == null")
        self._seti(abi.KEY_RENDERING_ALGORITHM, int(algorithm))

    def getRenderingAlgorithm(self):
        return RenderingAlgorithm(self._geti(abi.KEY_RENDERING_ALGORITHM))

    def setSamples(self, v):
        self._seti(abi.KEY_SAMPLES, v)

    def getSamples(self):
        return self._geti(abi.KEY_SAMPLES)

    def setMaximumDistance(self, v):
        self._c(self.lib.dfr_set_f32(self.ctx, abi.KEY_MAXIMUM_DISTANCE, float(v)))

    def getMaximumDistance(self):
        out = ctypes.c_float()
        self._c(self.lib.dfr_get_f32(self.ctx, abi.KEY_MAXIMUM_DISTANCE, ctypes.byref(out)))
        return out.value

    def setToneMapper(self, v):
        self._seti(abi.KEY_TONE_MAPPER, int(v))

    def getToneMapper(self):
        return self._geti(abi.KEY_TONE_MAPPER)

    def setPreviewMode(self, v):
        self._seti(abi.KEY_PREVIEW_MODE, int(bool(v)))

    # build-specific settings
    def setSeed(self, seed):
        seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        lo, hi = seed & 0xFFFFFFFF, seed >> 32
        self._seti(abi.KEY_SEED, lo - (1 << 32) if lo >= 1 << 31 else lo)
        self._seti(abi.KEY_SEED_HI, hi - (1 << 32) if hi >= 1 << 31 else hi)

    def setKernel(self, kernel):
        self._seti(abi.KEY_KERNEL, kernel)

    def setLightPdfMode(self, mode):
        self._seti(abi.KEY_LIGHT_PDF_MODE, mode)

    def setSamplesPerPass(self, spp):
        self._seti(abi.KEY_SPP_PER_PASS, spp)

    def setProfile(self, on):
        self._seti(abi.KEY_PROFILE, int(bool(on)))

    def setCountTraversal(self, on, sample_stride=0):
        """Count BVH node visits and triangle tests of the next renders; with
        `sample_stride` > 0 also record every n-th traced ray (ray_sample())."""
        self._seti(abi.KEY_COUNT_TRAVERSAL, int(bool(on)))
        self._seti(abi.KEY_RAY_SAMPLE, int(sample_stride) if on else 0)

    def ray_sample(self):
        """Rays recorded by the last counting render: (origin (n,3), far limit
        at BVH entry (n,), direction (n,3), any-hit flag (n,), node visits,
        triangle tests, hit primitive, hit triangle (n,) each)."""
        n = ctypes.c_int64()
        self._c(self.lib.dfr_read_ray_sample(self.ctx, None, 0, ctypes.byref(n)))
        rec = np.zeros((n.value, 12), np.float32)
        self._c(self.lib.dfr_read_ray_sample(self.ctx, rec.ctypes.data, n.value, ctypes.byref(n)))
        ints = rec.view(np.int32)
        return {"o": rec[:, 0:3], "tmax": rec[:, 3], "d": rec[:, 4:7], "any": ints[:, 7], "nodes": ints[:, 8],
                "tris": ints[:, 9], "prim": ints[:, 10], "tri": ints[:, 11]}

    def bvh(self):
        """The compiled four-wide BVH: nodes (n, 32) float32 and triangles in
        BVH order (m, 12) float32 (layouts in dfr_read_bvh's header comment)."""
        nn, nt = ctypes.c_int64(0), ctypes.c_int64(0)
        self._c(self.lib.dfr_read_bvh(self.ctx, None, ctypes.byref(nn), None, ctypes.byref(nt)))
        nodes = np.zeros((nn.value, 32), np.float32)
        tris = np.zeros((nt.value, 12), np.float32)
        self._c(self.lib.dfr_read_bvh(self.ctx, nodes.ctypes.data, ctypes.byref(nn), tris.ctypes.data,
                                      ctypes.byref(nt)))
        return nodes, tris

    def setWavefront(self, block=None, pool_paths=None, sort_keys=None):
        """Wavefront engine configuration (the values dfr_autotune picks)."""
        for key, v in ((abi.KEY_WF_BLOCK, block), (abi.KEY_WF_POOL_PATHS, pool_paths),
                       (abi.KEY_WF_SORT_KEYS, sort_keys)):
            if v is not None:
                self._seti(key, v)

    def setUseGraphs(self, on):
        """Replay captured HIP graphs of the wavefront iterations (default on)."""
        self._seti(abi.KEY_USE_GRAPHS, int(bool(on)))

    def getBvhWidth(self):
        """Children per node of the BVH the library builds (2 or 4)."""
        return self._geti(abi.KEY_BVH_WIDTH)

    def getWavefront(self):
        return {"block": self._geti(abi.KEY_WF_BLOCK), "pool_paths": self._geti(abi.KEY_WF_POOL_PATHS),
                "sort_keys": self._geti(abi.KEY_WF_SORT_KEYS)}

    # -- rendering -----------------------------------------------------------
    def setup(self, compile_scene=True):
        self._c(self.lib.dfr_setup(self.ctx, 1 if compile_scene else 0))

    def render(self, passes=1):
        """render(n): n < 1 raises IllegalArgumentException; returns True."""
        rc = self._c(self.lib.dfr_render(self.ctx, int(passes)))
        return rc != abi.E_STOPPED

    def clear(self):
        self._c(self.lib.dfr_clear(self.ctx))

    def getRenderPass(self):
        return self.lib.dfr_get_render_pass(self.ctx)

    def renderShutdown(self):
        self.lib.dfr_request_stop(self.ctx)

    def updateCamera(self, camera):
        cam = np.ascontiguousarray(camera, np.float32)
        self._c(self.lib.dfr_update_camera(self.ctx, cam.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    # -- readback ------------------------------------------------------------
    def film(self, width, height):
        """Mean radiance (H, W, 3) float32 and per-pixel sample counts."""
        rgb = np.zeros((height, width, 3), np.float32)
        cnt = np.zeros((height, width), np.int32)
        self._c(self.lib.dfr_read_film(self.ctx, rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return rgb, cnt

    def rgba8(self, width, height):
        out = np.zeros((height, width, 4), np.uint8)
        self._c(self.lib.dfr_read_rgba8(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def setFilmFilter(self, on):
        """Keeps CPURenderer's Gaussian XYZ film (PixelImageF with
        GaussianFilter2F) beside the running-mean film; a change clears both."""
        self._seti(abi.KEY_FILM_FILTER, 1 if on else 0)

    def film_xyz(self, width, height):
        """CPURenderer's XYZ film: (H, W, 4) float32 of X, Y, Z and the filter
        weight sum."""
        out = np.zeros((height, width, 4), np.float32)
        self._c(self.lib.dfr_read_film_xyz(self.ctx, out.ctypes.data))
        return out

    def film_filtered(self, width, height):
        """filmRender() of CPURenderer's film: (H, W, 3) gamma-encoded sRGB."""
        out = np.zeros((height, width, 3), np.float32)
        self._c(self.lib.dfr_read_film_filtered(self.ctx, out.ctypes.data))
        return out

    def save(self, path, width, height, filtered=False):
        """Writes the image as PNG: GPURenderer's tone-mapped RGBA8 image, or
        with filtered=True CPURenderer's filtered film in bytes."""
        from dayflower_amd.png import float_to_bytes, write_png
        if filtered:
            write_png(path, float_to_bytes(self.film_filtered(width, height)))
        else:
            write_png(path, self.rgba8(width, height))

    def stats(self):
        st = abi.Stats()
        self._c(self.lib.dfr_get_stats(self.ctx, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in abi.Stats._fields_}

    # -- ray queries -----------------------------------------------------------
    def trace_rays(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        tri = np.empty(n, np.int64)
        t = np.empty(n, np.float32)
        self._c(self.lib.dfr_trace_rays(self.ctx, rays.ctypes.data, n, int(any_hit), prim.ctypes.data,
                                        tri.ctypes.data, t.ctypes.data, 0))
        return prim, tri, t

    def trace_rays_device(self, rays_ptr, n, prim_ptr, tri_ptr, t_ptr, any_hit=False):
        """Device-resident variant: every pointer is a device address."""
        self._c(self.lib.dfr_trace_rays(self.ctx, rays_ptr, n, int(any_hit), prim_ptr, tri_ptr, t_ptr, 1))

    def trace_rays_device_noalloc(self, rays_ptr, n, prim_ptr, tri_ptr, t_ptr, any_hit=False):
        """Device-resident variant that allocates nothing (static scratch)."""
        self._c(self.lib.dfr_trace_rays_noalloc(self.ctx, rays_ptr, n, int(any_hit), prim_ptr, tri_ptr, t_ptr))

    def cast_primary(self, width, height, sample_index=0):
        n = width * height
        rays = np.empty((n, 8), np.float32)
        prim = np.empty(n, np.int32)
        tri = np.empty(n, np.int64)
        t = np.empty(n, np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._c(self.lib.dfr_cast_primary(self.ctx, sample_index, rays.ctypes.data_as(fp),
                                          prim.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), t.ctypes.data_as(fp)))
        return rays, prim, tri, t

    # -- resident scenes / batching --------------------------------------------
    def addScene(self, scene):
        desc = scene.desc()
        sid = self._c(self.lib.dfr_scene_add(self.ctx, ctypes.byref(desc)))
        return sid

    def removeScene(self, sid):
        self._c(self.lib.dfr_scene_remove(self.ctx, int(sid)))

    def autotune(self, jobs):
        """Pick block size, path-pool capacity and material-sort keys for the
        packed launch on `jobs`; returns (block, capacity, sort_keys)."""
        n = len(jobs)
        arr = (abi.Job * n)(*jobs)
        out = (ctypes.c_int32 * 3)()
        self._c(self.lib.dfr_autotune(self.ctx, arr, n, out))
        return tuple(out)

    def batchRender(self, jobs, mode=abi.BATCH_PACKED, readback=True):
        """Render a list of abi.Job records; returns per-job (H, W, 3) films."""
        n = len(jobs)
        arr = (abi.Job * n)(*jobs)
        films = [np.zeros((j.height, j.width, 3), np.float32) for j in jobs] if readback else []
        fp = ctypes.POINTER(ctypes.c_float)
        ptrs = (fp * n)(*[f.ctypes.data_as(fp) for f in films]) if readback else None
        self._c(self.lib.dfr_batch_render(self.ctx, arr, n, int(mode), ptrs))
        return films


def compiled_scene_struct(arrays):
    """A dfr_compiled_scene pointing into `arrays` (see setCompiledScene) and
    the contiguous copies it points at, which the caller keeps alive."""
    import numpy as np
    cs = abi.CompiledScene()
    keep = {}
    for name, kind in abi.COMPILED_SCENE_ARRAYS:
        a = arrays.get(name)
        if a is None:
            continue
        dt, ct = (np.float32, ctypes.c_float) if kind == "f" else (np.int32, ctypes.c_int32)
        a = np.ascontiguousarray(a, dt).reshape(-1)
        keep[name] = a
        f = getattr(cs, name)
        f.data = a.ctypes.data_as(ctypes.POINTER(ct))
        f.length = a.size
    cs.primitive_count = int(arrays.get("primitive_count", 0))
    cs.light_count = int(arrays.get("light_count", 0))
    return cs, keep


def transcode_compiled_scene(arrays):
    """The scene the library reads from compiled arrays, without a device:
    a dict of the dfr_scene_desc's arrays and records (numpy copies), with
    "transforms" (n, 4, 4).  Raises like the renderer does."""
    import numpy as np
    lib = abi.load()
    cs, keep = compiled_scene_struct(arrays)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    rc = lib.dfr_transcode_compiled_scene(ctypes.byref(cs), ctypes.byref(h), err, 512)
    if rc != abi.E_OK:
        msg = err.value.decode(errors="replace")
        if rc == abi.E_UNSUPPORTED:
            raise UnsupportedOperationException(msg)
        if rc == abi.E_NULL:
            raise NullPointerException(msg)
        raise IllegalArgumentException(msg)
    try:
        d = lib.dfr_transcoded_desc(h).contents
        nv, nt = d.n_vertices, d.n_triangles
        arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n else np.zeros(0, dt)
        out = {
            "positions": arr(d.positions, 3 * nv, np.float32).reshape(-1, 3),
            "normals": arr(d.normals, 3 * nv, np.float32).reshape(-1, 3),
            "uvs": arr(d.uvs, 2 * nv, np.float32).reshape(-1, 2),
            "indices": arr(d.indices, 3 * nt, np.int32).reshape(-1, 3),
            "primitives": [abi.Primitive.from_buffer_copy(d.primitives[i]) for i in range(d.n_primitives)],
            "materials": [abi.Material.from_buffer_copy(d.materials[i]) for i in range(d.n_materials)],
            "lights": [abi.Light.from_buffer_copy(d.lights[i]) for i in range(d.n_lights)],
            "textures": [abi.Texture.from_buffer_copy(d.textures[i]) for i in range(d.n_textures)],
            "images": [np.ctypeslib.as_array(d.images[i].rgb, shape=(d.images[i].height, d.images[i].width, 3)).copy()
                       for i in range(d.n_images)],
            "camera": np.array(d.camera[:], np.float32),
        }
        xf = lib.dfr_transcoded_transforms(h)
        out["transforms"] = arr(xf, 16 * d.n_primitives, np.float32).reshape(-1, 4, 4)
    finally:
        lib.dfr_transcoded_free(h)
    return out


def make_job(scene_id, camera, width, height, spp, seed, maximum_bounce=8, minimum_bounce_rr=5):
    j = abi.Job()
    j.scene = int(scene_id)
    j.width, j.height, j.spp = int(width), int(height), int(spp)
    j.maximum_bounce, j.minimum_bounce_rr = int(maximum_bounce), int(minimum_bounce_rr)
    j.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    cam = np.asarray(camera, np.float32)
    for i in range(abi.CAMERA_LENGTH):
        j.camera[i] = float(cam[i])
    j.camera[17] = float(width)
    j.camera[18] = float(height)
    return j


class Server:
    """Serving loop over a GPURenderer's resident scenes (continuous
    batching; see dfr_server_* in include/dayflower_hip.h)."""

    def __init__(self, renderer, hbm_budget_bytes=64 << 30, max_active=256, pool_paths=1 << 20, slot_capacity=0):
        self.r = renderer
        self.lib = renderer.lib
        o = abi.ServerOpts()
        o.hbm_budget_bytes, o.max_active, o.pool_paths, o.slot_capacity = int(hbm_budget_bytes), int(max_active), \
            int(pool_paths), int(slot_capacity)
        self.h = self.lib.dfr_server_create(renderer.ctx, ctypes.byref(o))
        if not self.h:
            msg = self.lib.dfr_last_error(renderer.ctx)
            raise DeviceException("dfr_server_create failed: " + (msg.decode() if msg else ""))
        self._jobs = {}

    @staticmethod
    def pool_bytes(pool_paths):
        """Device bytes of the wavefront engine for `pool_paths` paths."""
        return int(abi.load().dfr_wavefront_bytes(int(pool_paths)))

    def submit(self, job):
        rid = self.lib.dfr_server_submit(self.h, ctypes.byref(job))
        if rid < 0:
            raise IllegalArgumentException(f"submit failed ({rid})")
        self._jobs[rid] = job
        return rid

    def poll(self, rid):
        return self.lib.dfr_server_poll(self.h, rid)

    def wait(self, rid):
        j = self._jobs[rid]
        film = np.zeros((j.height, j.width, 3), np.float32)
        rc = self.lib.dfr_server_wait(self.h, rid, film.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        if rc != 0:
            raise DeviceException(f"request {rid} failed ({rc})")
        return film

    def latency(self, rid):
        out = ctypes.c_double()
        rc = self.lib.dfr_server_latency(self.h, rid, ctypes.byref(out))
        if rc != 0:
            raise IllegalStateException(f"request {rid} not done ({rc})")
        return out.value

    def stats(self):
        st = abi.ServerStats()
        self.lib.dfr_server_get_stats(self.h, ctypes.byref(st))
        return {name: getattr(st, name) for name, _ in abi.ServerStats._fields_}

    def close(self):
        if self.h:
            self.lib.dfr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
