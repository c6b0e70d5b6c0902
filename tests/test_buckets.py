"""Mixed clip sizes in one step (the batching form): buckets of 256, 512, 768
and 1024^2 clips at 4 nm, each with its own K=16 kernel set; FP32 buckets
equal the single-size batch path bit for bit and the FP64 oracle to 1e-5
(768 is a 16 x 48 FFT size); the autotuner's plans keep that parity; the
matrix-core form stays within 1e-3."""
import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA, max_rel

pytestmark = pytest.mark.gpu

SIZES = (256, 512, 768, 1024)
COUNTS = (5, 4, 3, 2)


@pytest.fixture(scope="module")
def setup(gpu, orc, base_source):
    src, s = base_source
    ctx = gpu.Context()
    buckets, host, banks, dev = [], [], [], []
    for i, (N, cnt) in enumerate(zip(SIZES, COUNTS)):
        Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
        sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx,
                                           Nx, 16)
        bank = ctx.bank(K, sc, Nx, Nx)
        masks = np.stack([gpu.manhattan_mask(N, 300 + 10 * i + j) for j in range(cnt)])
        dm = ctx.alloc(masks.nbytes)
        dm8 = ctx.alloc(masks.size)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.h2d(dm8, (masks * 255).astype(np.uint8))
        host.append((N, Nx, sc, K, masks))
        banks.append(bank)
        dev += [dm, dm8, di]
        buckets.append((bank, N, cnt, dm, di, dm8))
    yield ctx, buckets, host
    for p in dev:
        ctx.free(p)
    for b in banks:
        b.close()
    ctx.close()


def oracle_image(orc, N, Nx, sc, K, m):
    return orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32), Nx,
                         Nx, nthreads=8)


def test_buckets_fp32_equal_single_size_path(gpu, orc, setup):
    ctx, buckets, host = setup
    print("[case] mixed-size buckets FP32", flush=True)
    bl = [(b, N, c, dm, di) for b, N, c, dm, di, _ in buckets]
    for graph in (False, True, True):
        ctx.socs_image_buckets(bl, 0, graph)
        ctx.sync()
        got = [ctx.d2h(np.zeros((c, N, N), np.float32), di) for _, N, c, _, di in bl]
        for (b, N, c, dm, di), g in zip(bl, got):
            ref = ctx.alloc(c * N * N * 4)
            ctx.socs_image_batch(b, dm, c, N, ref)
            assert np.array_equal(ctx.d2h(np.zeros((c, N, N), np.float32), ref), g), (N, graph)
            ctx.free(ref)
    t = None
    ctx.socs_image_buckets(bl, 0, False)
    t = ctx.bucket_times(len(bl))
    assert len(t) == 4 and np.all(t > 0), t
    for (N, Nx, sc, K, masks), g in zip(host, got):
        e = max_rel(g[-1], oracle_image(orc, N, Nx, sc, K, masks[-1]))
        print(f"[buckets N={N} Nx={Nx}] max rel err vs FP64 {e:.2e}; bucket end {t}", flush=True)
        assert e < 1e-5


def test_autotuned_plans_keep_parity(gpu, orc, setup):
    ctx, buckets, host = setup
    print("[case] autotune", flush=True)
    bl = [(b, N, c, dm, di) for b, N, c, dm, di, _ in buckets]
    n = ctx.autotune(bl, 0, reps=2)
    assert n == 4
    assert ctx.autotune(bl, 0, reps=2) == 0  # already tuned
    for (N, Nx, sc, K, masks) in host:
        G, shape, lm, tuned = ctx.plan(N, Nx, 16, 0)
        print(f"[plan N={N} Nx={Nx}] grid {G} socs shape {shape} lean_fft {lm}", flush=True)
        assert tuned and G >= 4 * Nx + 1
    ctx.socs_image_buckets(bl, 0, True)
    ctx.sync()
    for (N, Nx, sc, K, masks), (_, _, c, _, di) in zip(host, bl):
        g = ctx.d2h(np.zeros((c, N, N), np.float32), di)
        assert max_rel(g[0], oracle_image(orc, N, Nx, sc, K, masks[0])) < 1e-5


def test_every_plan_matches_oracle(gpu, orc, setup):
    """Each conv grid and SOCS block shape (incl. the wave-synced transforms,
    and shape 3: each clip's kernel entries over two blocks, the two band
    spectra added by the FI column pass) gives the FP64 image to 1e-5 on every
    clip size."""
    ctx, buckets, host = setup
    print("[case] every band plan", flush=True)
    for (b, N, c, dm, di, _), (_, Nx, sc, K, masks) in zip(buckets, host):
        ref = oracle_image(orc, N, Nx, sc, K, masks[0])
        for G in (36, 81, 144):
            for shape in (0, 1, 2, 3):
                for lean in (0, 1):
                    try:
                        ctx.set_plan(N, Nx, 16, 0, G, shape, lean)
                    except gpu.LithoError:
                        continue  # not a built combination for this shape
                    ctx.socs_image_buckets([(b, N, c, dm, di)], 0, False)
                    g = ctx.d2h(np.zeros((c, N, N), np.float32), di)
                    e = max_rel(g[0], ref)
                    print(f"[plan N={N} G={G} shape={shape} lean={lean}] {e:.2e}", flush=True)
                    assert e < 1e-5, (N, G, shape, lean)


def test_buckets_matrix_core_form(gpu, orc, setup):
    ctx, buckets, host = setup
    print("[case] mixed-size buckets matrix-core form", flush=True)
    prec = gpu.LITHO_PREC_COMPACT
    bl = [(b, N, c, dm8, di) for b, N, c, dm, di, dm8 in buckets]
    ctx.socs_image_buckets(bl, prec, True)
    ctx.sync()
    for (N, Nx, sc, K, masks), (_, _, c, _, di) in zip(host, bl):
        g = ctx.d2h(np.zeros((c, N, N), np.float16), di).astype(np.float32)
        e = max_rel(g[1], oracle_image(orc, N, Nx, sc, K, masks[1]))
        print(f"[buckets compact N={N}] max rel err vs FP64 {e:.2e}", flush=True)
        assert e < 1e-3


def test_buckets_one_bit_masks(gpu, orc, setup):
    """The compact form on binary masks packed one bit per pixel equals the same
    form on 8-bit levels to one fp16 step, and FP64 to 1e-3."""
    ctx, buckets, host = setup
    print("[case] mixed-size buckets, 1-bit masks", flush=True)
    prec = (gpu.LITHO_PREC_COMPACT & ~gpu.LITHO_PREC_MASK_U8) | gpu.LITHO_PREC_MASK_1BIT
    dev, bl1 = [], []
    for (b, N, c, dm, di, dm8), (_, _, _, _, masks) in zip(buckets, host):
        bits = np.packbits(masks.astype(bool), axis=-1, bitorder="little")
        d1 = ctx.alloc(bits.nbytes)
        ctx.h2d(d1, bits)
        dev.append(d1)
        bl1.append((b, N, c, d1, di))
    ctx.socs_image_buckets(bl1, prec, True)
    ctx.sync()
    got = [ctx.d2h(np.zeros((c, N, N), np.float16), di).astype(np.float32)
           for _, N, c, _, di in bl1]
    bl8 = [(b, N, c, dm8, di) for b, N, c, dm, di, dm8 in buckets]
    ctx.socs_image_buckets(bl8, gpu.LITHO_PREC_COMPACT, False)
    ctx.sync()
    for (N, Nx, sc, K, masks), (_, _, c, _, di), g in zip(host, bl8, got):
        via_u8 = ctx.d2h(np.zeros((c, N, N), np.float16), di).astype(np.float32)
        assert max_rel(g, via_u8) < 1e-3
        e = max_rel(g[0], oracle_image(orc, N, Nx, sc, K, masks[0]))
        print(f"[buckets 1-bit N={N}] max rel err vs FP64 {e:.2e}", flush=True)
        assert e < 1e-3
    for d in dev:
        ctx.free(d)


def test_bucket_argument_checks(gpu, setup):
    ctx, buckets, _ = setup
    b, N, c, dm, di, _ = buckets[0]
    with pytest.raises(gpu.LithoError):
        ctx.socs_image_buckets([(b, 640, c, dm, di)], 0, False)  # unsupported size
    with pytest.raises(gpu.LithoError):
        ctx.socs_image_buckets([(b, N, -1, dm, di)], 0, False)
    with pytest.raises(gpu.LithoError):
        ctx.socs_image_buckets([(b, N, c, dm, di)], 256, False)  # unknown precision flag
    with pytest.raises(gpu.LithoError):  # two bank formats at once
        ctx.socs_image_buckets([(b, N, c, dm, di)],
                               gpu.LITHO_PREC_BANK_BF16 | gpu.LITHO_PREC_BANK_FP16, False)
    ctx.socs_image_buckets([(b, N, 0, 0, 0)], 0, False)  # empty bucket: nothing to do
