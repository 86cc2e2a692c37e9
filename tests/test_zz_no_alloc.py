"""The allocation-free hologram stage (mtd_holo_to_crop_noalloc) against the
allocating one (mtd_holo_to_crop) on the same holograms, through the C ABI:
two calls in a row, the second with a smaller input (fewer holograms, a
smaller hologram side, crop and dtype), crops bit-identical. The first call
spans several scratch chunks (40 x 1024^2 uint16: 2 MiB of row intermediate
per hologram, 31 per 63 MiB chunk); a repeated call leaves the device's
free memory unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_noalloc_matches_allocating(gpu_ctx):
    import torch
    import mtd_transmission_amd as m
    from cases import LAM, N0, NA
    cases = [(1024, 128, 40, torch.uint16), (512, 64, 7, torch.float32)]
    gen = torch.Generator().manual_seed(5)
    for n, nx, A, dt in cases:
        _say(f"noalloc hologram stage N={n} crop={2 * nx} A={A} {dt}")
        pix = nx * LAM / (NA * n)
        g = m.geom_derive(n, nx, 4 * nx, n // 2 - round(0.3 * n), n // 2 + round(0.3 * n),
                          NA, N0, LAM, pix)
        if dt == torch.uint16:
            h = torch.randint(0, 60000, (A, n, n), generator=gen, dtype=torch.int32).to(dt)
        else:
            h = torch.rand((A, n, n), generator=gen, dtype=dt) * 4095
        h = h.cuda()
        C = g.C
        want = torch.empty((A, C, C), dtype=torch.complex64, device="cuda")
        got = torch.full((A, C, C), float("nan"), dtype=torch.complex64, device="cuda")
        gpu_ctx.holo_to_crop(g, h, want)
        torch.cuda.synchronize()
        gpu_ctx.holo_to_crop_noalloc(g, h, got)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(got), torch.view_as_real(want))
        # again on poisoned outputs: same bits, and the device's free memory
        # is unchanged by the call
        got.fill_(float("nan"))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        gpu_ctx.holo_to_crop_noalloc(g, h, got)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        assert torch.equal(torch.view_as_real(got), torch.view_as_real(want))


def test_noalloc_rejects(gpu_ctx):
    """Host stacks (the allocating path stages them) are refused with a clear
    error, as are invalid geometries."""
    import torch
    import mtd_transmission_amd as m
    from cases import LAM, N0, NA
    n, nx = 256, 32
    g = m.geom_derive(n, nx, 4 * nx, n // 2 - round(0.3 * n), n // 2 + round(0.3 * n),
                      NA, N0, LAM, nx * LAM / (NA * n))
    h = torch.zeros((2, n, n), dtype=torch.float32)
    out = torch.empty((2, g.C, g.C), dtype=torch.complex64, device="cuda")
    with pytest.raises(m.MTDError, match="device memory"):
        m.lib()  # loaded
        rc = m.lib().mtd_holo_to_crop_noalloc(gpu_ctx.h, m.ctypes.byref(g),
                                              m.ctypes.c_void_p(h.data_ptr()), m.MTD_F32, 2,
                                              m.ctypes.c_void_p(out.data_ptr()))
        gpu_ctx._check(rc, "holo_to_crop_noalloc")
    g.circle_cx = 1
    with pytest.raises(m.MTDError, match="crop window"):
        gpu_ctx.holo_to_crop_noalloc(g, h.cuda(), out)
