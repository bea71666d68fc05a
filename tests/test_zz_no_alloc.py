"""raw_ao_trace_ao_dev_noalloc, the entry point that allocates nothing, against
raw_ao_trace_ao_dev (which it varies): the same AO bit for bit on two calls in a row, the
second on a smaller image, at 64 spp (the refill kernel) and below (the per-pixel counts in
the library's scratch array); images beyond its maximum are refused with an error."""
import numpy as np
import pytest
import torch

import raw_ao
import raw_ao.host as H

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def frame(w, h, seed):
    vpos, vnrm, idx, _ = H.scene_seeded()
    _, pos, nrm, mask = H.rasterize(vpos, vnrm, idx, H.orbit_camera(seed), (0, 2, 0), (0, 1, 0), 0.9,
                                    float(np.float32(w) / np.float32(h)), 0.1, 200.0, w, h)
    return H.mesh_tris(vpos, idx), pos, nrm, mask


def dev(*arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrays]


@pytest.mark.parametrize("spp", [64, 16])
def test_noalloc_trace_equals_trace_ao_dev(spp):
    print("no-alloc trace, spp", spp, flush=True)
    with raw_ao.Context(0) as c:
        tris, pos, nrm, mask = frame(640, 360, 7)
        c.build_ploc(tris, 16)
        for w, h in ((640, 360), (320, 180)):
            if (w, h) != (640, 360):
                _, pos, nrm, mask = frame(w, h, 8)
            P, N, M = dev(pos, nrm, mask)
            want = torch.full((h, w), -1.0, device="cuda:0")
            got = torch.full((h, w), -2.0, device="cuda:0")
            c.compute_rtao_dev(P, N, M, w, h, spp, 1.0, want)
            c.compute_rtao_dev_noalloc(P, N, M, w, h, spp, 1.0, got)
            torch.cuda.synchronize()
            a, b = want.cpu().numpy(), got.cpu().numpy()
            assert np.array_equal(bits(a), bits(b)), (w, h)
            assert 0.0 < float(a[mask != 0].mean()) < 1.0


def test_noalloc_trace_refuses_too_large_images():
    with raw_ao.Context(0) as c:
        tris, pos, nrm, mask = frame(64, 64, 3)
        c.build_ploc(tris, 16)
        z = torch.zeros(16, device="cuda:0")
        with pytest.raises(raw_ao.RawAoError) as e:
            c.compute_rtao_dev_noalloc(z, z, z, 4096, 1025, 64, 1.0, z)
        assert "RAW_AO_NOALLOC_MAX_PIXELS" in str(e.value)
