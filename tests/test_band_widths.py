"""Band-pipeline kernel choice by window width: at 512^2 the 16-lane row and
column transforms take windows of |k| < 32; a clip of 512 pixels at 8 nm (period
4096 nm, Nx = 34) is wider and takes the general two-stage kernels (row DFT,
mask columns, FI columns, image rows).  Both give the FP64 oracle's image to
1e-5 (max |diff| / max |ref|), and the 768^2 clip at 4 nm (Nx = 25) takes the
16-lane kernels throughout."""
import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA, max_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,pix", [(512, 8), (768, 4)])
def test_window_widths_match_oracle(gpu, orc, base_source, N, pix):
    src, s = base_source
    period = pix * N
    Nx = orc.compute_nx(BASE_NA, period, s, BASE_LAMBDA)
    print(f"[case] band widths N={N} pixel={pix} nm Nx={Nx}", flush=True)
    assert (Nx > 31) == (N == 512)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, period, period, Nx, Nx, 8)
    masks = np.stack([gpu.manhattan_mask(N, 500 + i) for i in range(2)])
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.socs_image_batch(bank, dm, 2, N, di)
        img = ctx.d2h(np.zeros_like(masks), di)
        _, stages = ctx.last_timing()
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
    assert "socs_band" in stages
    for g, m in zip(img, masks):
        ref = orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32),
                            Nx, Nx, nthreads=8)
        e = max_rel(g, ref)
        print(f"[band widths N={N} Nx={Nx}] max rel err vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5
