"""The no-allocation ray query (dfr_trace_rays_noalloc) against
dfr_trace_rays on the same device-resident rays: two calls in a row, the
second with fewer rays, each bit-identical to the allocating entry point,
and the device's free memory unchanged across the no-allocation call."""
import numpy as np
import pytest

from dayflower_amd import scene as S

pytestmark = pytest.mark.gpu


def test_noalloc_trace_matches_allocating_trace():
    import torch
    from dayflower_amd.renderer import GPURenderer
    sc = S.soup_scene(200_000, seed=5, width=64, height=64)
    sc.add_plane((0.0, -0.2, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 0)
    r = GPURenderer()
    r.setScene(sc)
    r.setup()
    rng = np.random.default_rng(3)
    dev = torch.device("cuda:0")
    for n, any_hit in ((300_000, False), (70_001, True), (50_000, False)):
        o = rng.random((n, 3)) * 1.4 - 0.2
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, 5.0
        rays_d = torch.from_numpy(rays).to(dev)
        outs = []
        for variant in ("alloc", "noalloc"):
            prim = torch.full((n,), -7, dtype=torch.int32, device=dev)
            tri = torch.full((n,), -7, dtype=torch.int64, device=dev)
            t = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
            torch.cuda.synchronize()
            if variant == "alloc":
                r.trace_rays_device(rays_d.data_ptr(), n, prim.data_ptr(), tri.data_ptr(), t.data_ptr(), any_hit)
            else:
                free0 = torch.cuda.mem_get_info(0)[0]
                r.trace_rays_device_noalloc(rays_d.data_ptr(), n, prim.data_ptr(), tri.data_ptr(), t.data_ptr(),
                                            any_hit)
                assert torch.cuda.mem_get_info(0)[0] == free0
            torch.cuda.synchronize()
            outs.append((prim.cpu().numpy(), tri.cpu().numpy(), t.cpu().numpy()))
        (pa, ta, da), (pn, tn, dn) = outs
        hits = (pa >= 0).mean()
        print(f"{n} rays any_hit={any_hit}: hit fraction {hits:.3f}", flush=True)
        assert 0.05 < hits < 0.95
        np.testing.assert_array_equal(pn, pa)
        np.testing.assert_array_equal(tn, ta)
        np.testing.assert_array_equal(dn, da)
    r.dispose()
