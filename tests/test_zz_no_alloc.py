"""litho_socs_image_batch_noalloc (the entry point that allocates nothing:
its workspaces sit in the library's fixed per-GPU scratch array) against
litho_socs_image_batch, whose workspaces the context allocates.  Both run the
same launches, so the images must be identical bit for bit; two calls in a
row, the second on a smaller batch, show that the shared scratch carries
nothing from one call to the next.  One image is also checked against the
FP64 oracle (1e-5 max rel err)."""
import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA, max_rel

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def _kernels(gpu, N, nk):
    src = gpu.createSource(41, "Annular", 0.6, 0.9)
    s = gpu.getOuterSigma(src)
    Nx = gpu.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx,
                                       nk)
    return Nx, sc, K


def _run(gpu, ctx, fn, bank, masks, N):
    B = masks.shape[0]
    dm = ctx.alloc(masks.nbytes)
    di = ctx.alloc(masks.nbytes)
    try:
        ctx.h2d(dm, masks)
        fn(bank, dm, B, N, di)
        ctx.sync()
        return ctx.d2h(np.zeros_like(masks), di)
    finally:
        ctx.free(dm)
        ctx.free(di)


@pytest.mark.parametrize("N,nk,batches,path", [
    (512, 10, (6, 3), 0),      # band pipeline, 81^2 conv grid
    (1024, 24, (8, 2), 0),     # production clip, 144^2 band grid
    (512, 10, (3, 1), 1),      # the reference's nextPow2 grid (128)
])
def test_noalloc_matches_batch(gpu, orc, N, nk, batches, path):
    Nx, sc, K = _kernels(gpu, N, nk)
    ctx = gpu.Context()
    bank = ctx.bank(K, sc, Nx, Nx)
    try:
        ctx.set_path(path)
        for call, B in enumerate(batches):
            case(f"noalloc N={N} Nx={Nx} K={nk} path={path} call {call} B={B}")
            masks = np.stack([gpu.manhattan_mask(N, 100 * call + b + 1) for b in range(B)])
            a = _run(gpu, ctx, ctx.socs_image_batch_noalloc, bank, masks, N)
            b = _run(gpu, ctx, ctx.socs_image_batch, bank, masks, N)
            assert np.array_equal(a, b)
            if call == 0 and N == 512:
                ref = orc.calc_socs(orc.mask_spectrum(masks[0]), K.astype(np.complex128),
                                    sc.astype(np.float32), Nx, Nx)
                assert max_rel(a[0], ref) < 1e-5
    finally:
        bank.close()
        ctx.close()


def test_noalloc_rejects_batches_beyond_the_scratch(gpu):
    case("noalloc rejects a batch beyond 64 MiB of workspaces")
    N, nk = 1024, 24
    Nx, sc, K = _kernels(gpu, N, nk)
    ctx = gpu.Context()
    bank = ctx.bank(K, sc, Nx, Nx)
    try:
        d = ctx.alloc(4)  # never touched: the size check comes first
        rc = gpu.lib().litho_socs_image_batch_noalloc(ctx.h, bank, d, 200, N, d)
        assert rc == gpu.LITHO_ESCRATCH
        assert b"scratch" in gpu.lib().litho_status_string(rc)
        ctx.free(d)
    finally:
        bank.close()
        ctx.close()
