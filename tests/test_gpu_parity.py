"""Parity of the HIP path (through the C-ABI) with the FP64 oracle on the same
seeded inputs, at sizes the oracle finishes in seconds, plus the reference's
own fixtures and size-independent properties at the full 1024^2 shape.

Tolerances: TCC and eigenvalues are FP64 on the GPU (1e-12 relative at zero
defocus; 1e-7 with defocus, where device cosf/sinf may differ from glibc by
an ulp); images are FP32 on the GPU and must be within 1e-5 of the FP64
oracle in max|a-b|/max|b| (validation/golden/run_verification.py L203-204).
"""
import numpy as np
import pytest

from conftest import (BASE_LAMBDA, BASE_NA, FIX_DEFOCUS, FIX_NA, golden, max_rel)

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def clip512(orc, base_source):
    """512^2 at 4 nm, Nx = 17, K = 10: oracle kernels via the Gram route."""
    src, s = base_source
    N = 512
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    sc, K, split = orc.calc_kernels_gram(A, S, 10, Nx, Nx)
    assert not split
    return dict(N=N, Nx=Nx, src=src, s=s, A=A, S=S, sc=sc, K=K)


@pytest.fixture(scope="module")
def fixture_tcc(orc, fix_source):
    src, s = fix_source
    return orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 512, 512, 4, 4)


# ------------------------------------------------------------ Module 4 -------
def test_tcc_build_fixture_config(gpu, orc, fix_source, fixture_tcc):
    case("tcc 101-source Lx=512 Nx=4 defocus 0.2")
    src, s = fix_source
    T = gpu.calcTCC(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 512, 512, 4, 4)
    assert max_rel(T, fixture_tcc) < 1e-7
    tr = np.fromfile(golden("tcc_expected_r.bin"), np.float32).reshape(81, 81)
    ti = np.fromfile(golden("tcc_expected_i.bin"), np.float32).reshape(81, 81)
    assert max_rel(T, tr + 1j * ti) < 1e-6
    assert np.array_equal(T, T.conj().T)


def test_tcc_build_baseline_c1(gpu, orc, clip512):
    case("tcc 41-source N=512 Nx=17 defocus 0")
    T = gpu.calcTCC(clip512["src"], clip512["s"], BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048, clip512["Nx"], clip512["Nx"])
    To = orc.calc_tcc(clip512["src"], clip512["s"], BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048, clip512["Nx"], clip512["Nx"])
    assert max_rel(T, To) < 1e-12
    assert not T.imag.any()


# ------------------------------------------------------------ Module 6 -------
def test_kernels_topk_fixture(gpu, orc, fixture_tcc):
    case("kernels_topk n=81 nk=12")
    sc, K, fl = gpu.calcKernels(fixture_tcc, 12, 4, 4)
    so, Ko, split = orc.calc_kernels(fixture_tcc, 12, 4, 4)
    assert np.abs(sc - so).max() / so[0] < 1e-10
    assert np.abs(K - Ko).max() < 1e-6  # canonical basis + phase rule, f32 output
    assert bool(fl & gpu.LITHO_FLAG_CLUSTER_SPLIT) == split


def test_kernels_topk_split_cluster_is_flagged(gpu, fixture_tcc):
    case("kernels_topk n=81 nk=10 (splits a pair)")
    _, _, fl = gpu.calcKernels(fixture_tcc, 10, 4, 4)
    assert fl & gpu.LITHO_FLAG_CLUSTER_SPLIT


@pytest.mark.parametrize("L,fname", [(256, "eig_resolution_256x256.txt"),
                                     (1024, "eig_resolution_1024x1024.txt")])
def test_eigenvalues_match_reference_runs(gpu, orc, fix_source, L, fname):
    case(f"kernels_from_source 101-source L={L}")
    src, s = fix_source
    Nx = orc.compute_nx(FIX_NA, L, s, BASE_LAMBDA)
    ref = np.loadtxt(golden(fname))
    k = min(16, ref.size, (2 * Nx + 1) ** 2)
    sc, _, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, L, L, Nx, Nx, k)
    big = ref[:k] > 1e-6 * ref[0]
    assert np.all(np.abs(sc[big] - ref[:k][big]) / ref[:k][big] < 6e-6)


def test_kernels_gram_route_c1(gpu, orc, clip512):
    case("kernels_from_source 41-source N=512 nk=10 (Gram)")
    sc, K, fl = gpu.kernels_from_source(clip512["src"], clip512["s"], BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048,
                                        clip512["Nx"], clip512["Nx"], 10)
    assert fl & gpu.LITHO_FLAG_GRAM_ROUTE
    assert not fl & gpu.LITHO_FLAG_CLUSTER_SPLIT
    assert np.abs(sc - clip512["sc"]).max() / clip512["sc"][0] < 1e-10
    assert np.abs(K - clip512["K"]).max() < 1e-6


# ------------------------------------------------------------ Module 3 / 7 ---
def test_mask_spectrum(gpu, orc):
    case("mask spectrum N=512")
    m = gpu.manhattan_mask(512, 3)
    assert max_rel(gpu.FT_r2c_padded(m), orc.mask_spectrum(m)) < IMG_TOL


def test_socs_tmpimg_hls_entry(gpu, orc, clip512):
    case("socs tmpimg 512^2 K=10")
    m = gpu.manhattan_mask(clip512["N"], 4)
    mf = orc.mask_spectrum(m)
    t = gpu.calc_socs_hls(mf, clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
    # Nx = 17 is above v18's MAX_KERNEL_SIZE range: grid max(128, nextPow2(69)) = 128
    to = orc.socs_tmpimg(mf, clip512["K"], clip512["sc"].astype(np.float32), clip512["Nx"],
                         clip512["Nx"], grid=(128, 128))
    assert t.shape == (128, 128)
    assert max_rel(t, to) < IMG_TOL


@pytest.mark.parametrize("N", [512, 4096])
def test_fi(gpu, orc, clip512, N):
    """FI of a 128^2 tmpImgp to N^2; at 4096 it runs the 4096-point column and
    image-row kernels on row-major U (the layout of litho_fi)."""
    case(f"FI 128 -> {N}")
    m = gpu.manhattan_mask(clip512["N"], 5)
    to = orc.socs_tmpimg(orc.mask_spectrum(m), clip512["K"], clip512["sc"].astype(np.float32), clip512["Nx"],
                         clip512["Nx"])
    assert max_rel(gpu.FI(to, N, 2 * clip512["Nx"]), orc.fi(to, N, N)) < IMG_TOL


def test_calc_socs_from_spectrum(gpu, orc, clip512):
    case("calcSOCS from spectrum 512^2 K=10")
    m = gpu.manhattan_mask(clip512["N"], 6)
    mf = orc.mask_spectrum(m)
    img = gpu.calcSOCS(mf, clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
    ref = orc.calc_socs(mf, clip512["K"], clip512["sc"].astype(np.float32), clip512["Nx"], clip512["Nx"])
    assert max_rel(img, ref) < IMG_TOL


@pytest.mark.parametrize("seed", [1, 2])
def test_socs_image_c1(gpu, orc, clip512, seed):
    case(f"socs_image 512^2 K=10 seed {seed}")
    m = gpu.manhattan_mask(clip512["N"], seed)
    img = gpu.socs_image(m, clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
    ref = orc.calc_socs(orc.mask_spectrum(m), clip512["K"], clip512["sc"].astype(np.float32), clip512["Nx"],
                        clip512["Nx"])
    assert max_rel(img, ref) < IMG_TOL


def test_socs_image_reference_fixture_full_rank(gpu, orc, fixture_tcc):
    case("socs_image fixture mask, K = n = 81")
    m = np.fromfile(golden("mask_256x256.bin"), np.float32).reshape(256, 256)
    M = np.zeros((512, 512), np.float32)
    M[128:384, 128:384] = m
    sc, K, _ = gpu.calcKernels(fixture_tcc, 81, 4, 4)
    img = gpu.socs_image(M, K, sc, 4, 4)
    gold = np.fromfile(golden("image_expected_512x512.bin"), np.float32).reshape(512, 512)
    # SOCS with every kernel equals the TCC-direct image the fixture holds
    assert max_rel(img, gold) < IMG_TOL


def test_socs_image_c2_full_size(gpu, orc, base_source):
    case("socs_image N=1024 Nx=34 K=24")
    src, s = base_source
    N = 1024
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    assert Nx == 34
    sc, K, fl = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 24)
    assert not fl & gpu.LITHO_FLAG_CLUSTER_SPLIT
    m = gpu.manhattan_mask(N, 11)
    img = gpu.socs_image(m, K, sc, Nx, Nx)
    ref = orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32), Nx,
                        Nx, nthreads=8)
    assert max_rel(img, ref) < IMG_TOL
    # open frame: intensity of a clear mask is sum(lambda_k |K_k(0)|^2), uniform
    clear = gpu.socs_image(np.ones((N, N), np.float32), K, sc, Nx, Nx)
    c = (2 * Nx + 1) ** 2 // 2
    expect = float(np.sum(sc.astype(np.float32) * np.abs(K[:, c]) ** 2))
    assert np.abs(clear - expect).max() < 1e-5 * expect


def test_batch_equals_single(gpu, orc, clip512):
    case("batch of 3 vs single")
    ctx = gpu.Context()
    try:
        N = clip512["N"]
        masks = np.stack([gpu.manhattan_mask(N, 20 + i) for i in range(3)])
        bank = ctx.bank(clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.socs_image_batch(bank, dm, 3, N, di)
        out = ctx.d2h(np.zeros_like(masks), di)
        for i in range(3):
            single = gpu.socs_image(masks[i], clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
            assert np.array_equal(out[i], single)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("N", [512, 1024])
def test_band_and_pow2_grids_agree(gpu, orc, base_source, N):
    """The image path's 81^2 / 144^2 band grid and the reference's nextPow2
    grid give the same image (FI is exact for any grid >= 4Nx+1)."""
    case(f"band vs pow2 grid N={N}")
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    nk = 10 if N == 512 else 24
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, nk)
    masks = np.stack([gpu.manhattan_mask(N, 40 + i) for i in range(2)])
    outs = []
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        for path in (gpu.LITHO_PATH_AUTO, gpu.LITHO_PATH_POW2):
            ctx.set_path(path)
            ctx.socs_image_batch(bank, dm, 2, N, di)
            outs.append(ctx.d2h(np.zeros_like(masks), di))
            _, stages = ctx.last_timing()
            assert ("socs_band" in stages) == (path == gpu.LITHO_PATH_AUTO)
            # without per-stage events: the same images, and no timing to report
            ctx.set_timing(False)
            ctx.socs_image_batch(bank, dm, 2, N, di)
            assert np.array_equal(ctx.d2h(np.zeros_like(masks), di), outs[-1])
            with pytest.raises(gpu.LithoError):
                ctx.last_timing()
            ctx.set_timing(True)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
    ref = orc.calc_socs(orc.mask_spectrum(masks[1]), K.astype(np.complex128),
                        sc.astype(np.float32), Nx, Nx, nthreads=8)
    assert max_rel(outs[0][1], ref) < IMG_TOL
    assert max_rel(outs[1][1], ref) < IMG_TOL
    assert max_rel(outs[0], outs[1]) < IMG_TOL


def test_large_clip_576_and_pow2_grids_agree(gpu, base_source):
    """4096^2 clips take the 576 = 24^2 conv grid on the automatic path and the
    reference's nextPow2 grid (1024) on LITHO_PATH_POW2: the same images to
    FP32 round-off (the intensity band 4Nx+1 = 549 fits both), through the
    batch entry point and the pipelined one with a 16-bit bank."""
    case("576 vs 1024 conv grid at 4096^2, K=16")
    N, nk = 4096, 16
    src, s = base_source
    Nx = gpu.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    assert 4 * Nx + 1 == 549
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx,
                                       nk)
    masks = np.stack([gpu.manhattan_mask(N, 60 + i) for i in range(2)])
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        outs = {}
        for path in (gpu.LITHO_PATH_AUTO, gpu.LITHO_PATH_POW2):
            ctx.set_path(path)
            ctx.socs_image_batch(bank, dm, 2, N, di)
            outs[("batch", path)] = ctx.d2h(np.zeros_like(masks), di)
            ctx.socs_image_batch_pipelined(bank, dm, 2, N, di, gpu.LITHO_PREC_BANK_FP16, 1, True)
            outs[("fp16", path)] = ctx.d2h(np.zeros_like(masks), di)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
    a, p = gpu.LITHO_PATH_AUTO, gpu.LITHO_PATH_POW2
    print("576 vs 1024 grid: fp32", max_rel(outs[("batch", a)], outs[("batch", p)]),
          "fp16 bank", max_rel(outs[("fp16", a)], outs[("fp16", p)]))
    assert max_rel(outs[("batch", a)], outs[("batch", p)]) < IMG_TOL
    assert max_rel(outs[("fp16", a)], outs[("fp16", p)]) < IMG_TOL


@pytest.mark.parametrize("N,B", [(2048, 11), (4096, 9)])
def test_large_grid_batch_matches_single_clips(gpu, orc, base_source, N, B):
    """The large-grid column pass takes several clips per block: a batch whose
    size is not a multiple of that count gives every clip the image it gets
    alone (bit for bit), in FP32 and with the fp16 bank, and the last clip
    matches the FP64 oracle."""
    case(f"large-grid batch N={N} B={B}")
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 16)
    masks = np.stack([gpu.manhattan_mask(N, 300 + i) for i in range(B)])
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm, di = ctx.alloc(masks.nbytes), ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        one = masks[0].nbytes
        for flag in (0, gpu.LITHO_PREC_BANK_FP16):
            ctx.socs_image_batch_pipelined(bank, dm, B, N, di, flag, 1, False)
            batch = ctx.d2h(np.zeros_like(masks), di)
            for b in (0, 7, B - 1):
                ctx.socs_image_batch_pipelined(bank, type(dm)(dm.value + b * one), 1, N, di, flag,
                                               1, False)
                alone = ctx.d2h(np.zeros_like(masks[0]), di)
                assert np.array_equal(batch[b], alone), (flag, b)
            if flag == 0:
                last = batch[B - 1]
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
    ref = orc.calc_socs(orc.mask_spectrum(masks[B - 1]), K.astype(np.complex128),
                        sc.astype(np.float32), Nx, Nx, nthreads=16)
    assert max_rel(last, ref) < IMG_TOL


@pytest.mark.parametrize("N", [256, 2048])
def test_other_clip_sizes(gpu, orc, base_source, N):
    case(f"socs_image N={N}")
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 16)
    m = gpu.manhattan_mask(N, 31)
    img = gpu.socs_image(m, K, sc, Nx, Nx)
    ref = orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32), Nx,
                        Nx, nthreads=8)
    assert max_rel(img, ref) < IMG_TOL


def test_single_kernel_and_empty_mask(gpu, orc, clip512):
    case("K=1 and empty mask")
    m = gpu.manhattan_mask(clip512["N"], 9)
    img = gpu.socs_image(m, clip512["K"][:1], clip512["sc"][:1], clip512["Nx"], clip512["Nx"])
    ref = orc.calc_socs(orc.mask_spectrum(m), clip512["K"][:1], clip512["sc"][:1].astype(np.float32),
                        clip512["Nx"], clip512["Nx"])
    assert max_rel(img, ref) < IMG_TOL
    z = gpu.socs_image(np.zeros((clip512["N"], clip512["N"]), np.float32), clip512["K"], clip512["sc"], clip512["Nx"],
                       clip512["Nx"])
    assert not z.any()


@pytest.mark.parametrize("nk,B", [(1, 1), (3, 3), (24, 5)])
def test_entry_split_socs(gpu, orc, base_source, nk, B):
    """Launches of few 1024^2 clips split each clip's kernel entries over two
    blocks (the second half's band spectrum in a second plane, added by the FI
    column pass): a single kernel (one block of the pair gets no entry), an
    odd entry count and an odd batch, against the FP64 oracle and against the
    pinned one-block plan (shape 2) of the pipelined entry."""
    case(f"entry-split SOCS 1024^2 K={nk} B={B}")
    N = 1024
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 24)
    K, sc = K[:nk], sc[:nk]
    masks = np.stack([gpu.manhattan_mask(N, 70 + i) for i in range(B)])
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.socs_image_batch(bank, dm, B, N, di)  # few clips: split by default
        split = ctx.d2h(np.zeros_like(masks), di)
        ctx.set_plan(N, Nx, nk, 0, 144, 2, False)  # one block per clip
        ctx.socs_image_batch_pipelined(bank, dm, B, N, di, 0, 1, False)
        one = ctx.d2h(np.zeros_like(masks), di)
        ctx.set_plan(N, Nx, nk, 0, 144, 3, False)  # the split, pinned
        ctx.socs_image_batch_pipelined(bank, dm, B, N, di, 0, 1, False)
        pinned = ctx.d2h(np.zeros_like(masks), di)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
    assert np.array_equal(split, pinned)
    ref = orc.calc_socs(orc.mask_spectrum(masks[-1]), K.astype(np.complex128),
                        sc.astype(np.float32), Nx, Nx, nthreads=8)
    e = max_rel(split[-1], ref)
    print(f"entry split K={nk} B={B}: {e:.2e} vs FP64, {max_rel(split, one):.2e} vs one block")
    assert e < IMG_TOL
    assert max_rel(split, one) < IMG_TOL


def test_invalid_requests_raise(gpu, clip512):
    case("invalid shapes")
    with pytest.raises(gpu.LithoError):
        gpu.socs_image(np.zeros((100, 100), np.float32), clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
    with pytest.raises(gpu.LithoError):
        gpu.socs_image(np.zeros((64, 64), np.float32), clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])


@pytest.mark.parametrize("precision,graph", [(0, False), (0, True), (5, True)])
def test_pipelined_batch_forms(gpu, orc, base_source, precision, graph):
    """The pipelined / graph-captured batch equals the plain path (FP32) or
    stays within 1e-3 of FP64 (mixed storage), including a replayed graph and
    a batch that does not divide into the sub-batches."""
    case(f"pipelined batch precision={precision} graph={graph}")
    src, s = base_source
    N = 1024
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 24)
    B = 5
    masks = np.stack([gpu.manhattan_mask(N, 60 + i) for i in range(B)])
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(masks.nbytes)
        d1 = ctx.alloc(masks.nbytes)
        d2 = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.socs_image_batch(bank, dm, B, N, d1)
        plain = ctx.d2h(np.zeros_like(masks), d1)
        for _ in range(3):  # capture once, replay
            ctx.socs_image_batch_pipelined(bank, dm, B, N, d2, precision, 2, graph)
            ctx.sync()
            got = ctx.d2h(np.zeros_like(masks), d2)
            if precision == 0:
                assert np.array_equal(got, plain)
        ref = orc.calc_socs(orc.mask_spectrum(masks[4]), K.astype(np.complex128),
                            sc.astype(np.float32), Nx, Nx, nthreads=8)
        tol = IMG_TOL if precision == 0 else 1e-3
        assert max_rel(got[4], ref) < tol
        if precision:
            # report the error of every storage flag on its own
            for flag in (1, 2, 4, 64):
                ctx.socs_image_batch_pipelined(bank, dm, B, N, d2, flag, 2, False)
                ctx.sync()
                e = max_rel(ctx.d2h(np.zeros_like(masks), d2)[4], ref)
                print(f"[precision flag {flag}] max rel err vs FP64 {e:.3e}", flush=True)
                if flag == 64:
                    # bf16 hi + lo bank: 16 significant bits per kernel value
                    assert e < 1e-4
        for p_ in (dm, d1, d2):
            ctx.free(p_)
        bank.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("N,prec", [(1024, 32), (1024, 8 | 32 | 1 | 4), (1024, 8 | 32 | 4 | 16),
                                    (512, 8 | 32 | 1 | 4), (256, 8 | 32), (1024, 8 | 4 | 16),
                                    (768, 8 | 16), (512, 8 | 1 | 4 | 16),
                                    (1024, 8 | 32 | 1 | 4 | 16), (512, 8 | 32 | 1 | 4 | 16),
                                    (1024, 8 | 32 | 1 | 64 | 16), (1024, 128 | 32 | 1 | 4 | 16),
                                    (512, 128 | 32 | 1 | 4 | 16), (768, 128 | 32 | 16)])
def test_matrix_core_forms(gpu, orc, base_source, N, prec):
    """u8 / f32 masks and f32 / fp16 images, with the band DFTs on the fp16
    matrix cores (flag 32) or as FP32 FFTs, stay within 1e-3 of FP64 on binary
    and 8-bit grey masks; graph replays reproduce the eager result bit for bit."""
    case(f"matrix-core form N={N} precision={prec}")
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    K_n = 24 if N == 1024 else 16
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx,
                                       K_n)
    B = 3
    u8 = bool(prec & gpu.LITHO_PREC_MASK_U8)
    bits = bool(prec & gpu.LITHO_PREC_MASK_1BIT)
    out16 = bool(prec & gpu.LITHO_PREC_IMAGE_FP16)
    rng = np.random.default_rng(N + prec)
    masks = np.stack([gpu.manhattan_mask(N, 70 + i) for i in range(B)])
    if bits:
        # binary masks, one bit per pixel (pixel 8j + i at bit i of byte j)
        masks[2] = (rng.random((N, N)) < 0.5).astype(np.float32)  # a random pattern too
        host_in = np.packbits(masks.astype(bool), axis=-1, bitorder="little")
        assert host_in.shape == (B, N, N // 8)
    elif u8:
        levels = (masks * 255).astype(np.uint8)
        # last clip: random grey levels 0..255 (mask = level / 255)
        levels[2] = rng.integers(0, 256, (N, N), dtype=np.uint8)
        masks = levels.astype(np.float32) / np.float32(255.0)
        host_in = levels
    else:
        host_in = masks
    odt = np.float16 if out16 else np.float32
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm = ctx.alloc(host_in.nbytes)
        di = ctx.alloc(B * N * N * 4)
        ctx.h2d(dm, host_in)
        ctx.socs_image_batch_pipelined(bank, dm, B, N, di, prec, 1, False)
        ctx.sync()
        eager = ctx.d2h(np.zeros((B, N, N), odt), di)
        for _ in range(2):
            ctx.socs_image_batch_pipelined(bank, dm, B, N, di, prec, 2, True)
            ctx.sync()
            assert np.array_equal(ctx.d2h(np.zeros((B, N, N), odt), di), eager)
        for i in (0, 2):
            ref = orc.calc_socs(orc.mask_spectrum(masks[i]), K.astype(np.complex128),
                                sc.astype(np.float32), Nx, Nx, nthreads=8)
            e = max_rel(eager[i].astype(np.float32), ref)
            print(f"[matrix-core N={N} prec={prec} clip {i}] max rel err vs FP64 {e:.3e}",
                  flush=True)
            assert e < 1e-3
        if bits:
            # the same binary masks as u8 levels: the same images to FP32 round-off
            lv = (masks * 255).astype(np.uint8)
            du = ctx.alloc(lv.nbytes)
            ctx.h2d(du, lv)
            p8 = (prec & ~gpu.LITHO_PREC_MASK_1BIT) | gpu.LITHO_PREC_MASK_U8
            ctx.socs_image_batch_pipelined(bank, du, B, N, di, p8, 1, False)
            ctx.sync()
            via_u8 = ctx.d2h(np.zeros((B, N, N), odt), di).astype(np.float32)
            d = max_rel(eager.astype(np.float32), via_u8)
            print(f"[matrix-core N={N}] bits vs u8 levels {d:.3e}", flush=True)
            assert d < (1e-3 if out16 else 1e-5)
            ctx.free(du)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()


def test_matrix_core_flag_rules(gpu, clip512):
    """Unknown flags and two bank formats at once fail."""
    ctx = gpu.Context()
    try:
        bank = ctx.bank(clip512["K"], clip512["sc"], clip512["Nx"], clip512["Nx"])
        d = ctx.alloc(4 * 512 * 512)
        for bad in (256, 512, 2 | 4, 2 | 4 | 32, 4 | 64, 2 | 64, 128, 128 | 8 | 32, 128 | 1):
            with pytest.raises(gpu.LithoError):
                ctx.socs_image_batch_pipelined(bank, d, 1, 512, d, bad, 1, False)
        ctx.free(d)
        bank.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("defocus", [0.0, 0.2])
def test_split_bf16_tcc(gpu, base_source, defocus):
    """Split-BF16 matrix-core TCC (explicit route, n <= n_src) against the FP64
    MFMA TCC: real pupil at zero defocus, complex pupil with defocus."""
    case(f"split-bf16 TCC defocus={defocus}")
    src, s = base_source
    N = 256
    Nx = gpu.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    args = (src, s, BASE_LAMBDA, BASE_NA, defocus, 4 * N, 4 * N, Nx, Nx)
    t64 = gpu.calcTCC(*args)
    t16 = gpu.calcTCC(*args, tcc_prec=gpu.LITHO_TCC_SPLIT_BF16)
    e = np.abs(t16 - t64).max() / np.abs(t64).max()
    print(f"[split-bf16 TCC n={t64.shape[0]} defocus={defocus}] max err / max|TCC| {e:.3e}",
          flush=True)
    assert np.allclose(t16, t16.conj().T)
    assert e < 5e-5


def test_split_bf16_kernels_and_image(gpu, orc, base_source):
    """1024^2, K=24 precision form: split-BF16 Gram contraction + FP64 eigensolve + FP64
    lift; eigenvalues, kernels and the FP32 SOCS image against FP64."""
    case("split-bf16 Gram route 1024^2 K=24")
    src, s = base_source
    N, nk = 1024, 24
    Nx = gpu.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    args = (src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, nk)
    sc64, K64, f64 = gpu.kernels_from_source(*args)
    sc16, K16, f16 = gpu.kernels_from_source(*args, tcc_prec=gpu.LITHO_TCC_SPLIT_BF16)
    assert f64 == f16
    eig = np.abs(sc16 - sc64).max() / sc64[0]
    kerr = max(np.abs(K16[k] - K64[k]).max() / np.abs(K64[k]).max() for k in range(nk))
    m = gpu.manhattan_mask(N, 77)
    img = gpu.socs_image(m, K16, sc16, Nx, Nx)
    ref = orc.calc_socs(orc.mask_spectrum(m), K64.astype(np.complex128), sc64.astype(np.float32),
                        Nx, Nx, nthreads=8)
    ierr = max_rel(img, ref)
    print(f"[split-bf16 1024^2 K=24] eigenvalues {eig:.3e} kernels {kerr:.3e} image {ierr:.3e} (vs FP64)",
          flush=True)
    assert eig < 1e-4
    assert kerr < 1e-3  # canonical basis in the mirror-degenerate pairs as well
    assert ierr < 1e-4


def test_tcc_direct_image_reference_fixture(gpu, orc, fixture_tcc):
    """GPU calcImage reproduces the reference's image_expected_512x512.bin."""
    case("tcc-direct image fixture")
    m = np.fromfile(golden("mask_256x256.bin"), np.float32).reshape(256, 256)
    M = np.zeros((512, 512), np.float32)
    M[128:384, 128:384] = m
    img = gpu.calcImage(M, fixture_tcc, 4, 4)
    gold = np.fromfile(golden("image_expected_512x512.bin"), np.float32).reshape(512, 512)
    assert max_rel(img, gold) < IMG_TOL


def test_tcc_direct_image_c1_and_full_rank_socs(gpu, orc, clip512):
    """TCC-direct image on a 512^2 production-optics clip vs the oracle, and the
    SOCS image with every kernel (K = n via the explicit route) converging to it."""
    case("tcc-direct image 512^2")
    T = orc.calc_tcc(clip512["src"], clip512["s"], BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048, clip512["Nx"], clip512["Nx"])
    m = gpu.manhattan_mask(clip512["N"], 77)
    img = gpu.calcImage(m, T, clip512["Nx"], clip512["Nx"])
    ref = orc.calc_image_tcc(orc.mask_spectrum(m), T, clip512["Nx"], clip512["Nx"])
    assert max_rel(img, ref) < IMG_TOL
    # rank of the TCC is 572 (< n = 1225): the top 572 kernels span it
    sc, K, _ = gpu.kernels_from_source(clip512["src"], clip512["s"], BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048,
                                       clip512["Nx"], clip512["Nx"], 572)
    socs = gpu.socs_image(m, K, sc, clip512["Nx"], clip512["Nx"])
    assert max_rel(socs, img) < 1e-4


@pytest.mark.parametrize("shape,nxy,period", [((128, 128), (4, 4), (512, 512)),
                                              ((64, 128), (4, 2), (512, 256)),
                                              ((256, 64), (3, 8), (256, 1024)),
                                              ((2048, 2048), (4, 4), (448, 448)),
                                              ((256, 4096), (8, 4), (896, 448)),
                                              ((4096, 256), (4, 8), (448, 896))])
def test_tcc_direct_image_general_shapes(gpu, orc, shape, nxy, period):
    """calcImage on the general power-of-two path: square sizes outside the band
    grids (128^2, 2048^2), rectangular Ly x Lx clips and Nx != Ny, against the
    oracle's calcImage + FT_c2r_padded on the same spectrum.  The TCC depends on
    the period (nm) and Nx, Ny only; the 2048^2 case takes a 448 nm period so
    that Nx = 4 covers the pupil (n = 81).  The 4096-point sides run the
    register-core row, column and image-row transforms."""
    Ly, Lx = shape
    Nx, Ny = nxy
    case(f"tcc-direct image {Ly}x{Lx}, Nx={Nx}, Ny={Ny}")
    src = orc.source_annular(41, 0.6, 0.9)
    s = orc.outer_sigma(src)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, BASE_NA, 0.0, period[0], period[1], Nx, Ny)
    m = gpu.manhattan_mask(max(Lx, Ly), 91)[:Ly, :Lx]
    img = gpu.calcImage(m, T, Nx, Ny)
    ref = orc.calc_image_tcc(orc.mask_spectrum(m), T, Nx, Ny)
    assert img.shape == (Ly, Lx)
    assert max_rel(img, ref) < IMG_TOL


@pytest.mark.parametrize("shape,nxy,period", [((64, 128), (4, 2), (512, 256)),
                                              ((256, 64), (3, 8), (256, 1024)),
                                              ((512, 1024), (17, 8), (2048, 1024)),
                                              ((256, 4096), (8, 4), (896, 448)),
                                              ((4096, 256), (4, 8), (448, 896))])
def test_socs_image_rect(gpu, orc, shape, nxy, period):
    """calcSOCS on a rectangular Ly x Lx clip with Nx != Ny: mask spectrum
    (rectangular FT_r2c_padded) against the oracle, then the image against the
    oracle's calcSOCS (its own nextPow2 grid per axis) fed the same kernels."""
    Ly, Lx = shape
    Nx, Ny = nxy
    case(f"socs_image_rect {Ly}x{Lx}, Nx={Nx}, Ny={Ny}")
    src = orc.source_annular(41, 0.6, 0.9)
    s = orc.outer_sigma(src)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, BASE_NA, 0.0, period[0], period[1], Nx, Ny)
    sc, K, _ = orc.calc_kernels(T, 12, Nx, Ny)
    m = gpu.manhattan_mask(max(Lx, Ly), 93)[:Ly, :Lx]
    mf = orc.mask_spectrum(m)
    spec = gpu.FT_r2c_padded(m)
    assert spec.shape == (Ly, Lx)
    assert np.abs(spec - mf).max() < 1e-5 * np.abs(mf).max()
    img = gpu.socs_image(m, K, sc, Nx, Ny)
    ref = orc.calc_socs(mf, K, np.asarray(sc, np.float32), Nx, Ny)
    assert img.shape == (Ly, Lx)
    assert max_rel(img, ref) < IMG_TOL


def test_large_clip_implicit_tcc(gpu, orc, base_source):
    """4096^2 clip at 4 nm (period 16384 nm): Nx = 137, a 75625-dim TCC that is
    never formed; the source-space Gram route gives K = 64 kernels (checked
    against LAPACK on the Gram and by their residual through B B^H), and the
    image matches the FP64 oracle fed the same kernels."""
    case("large clip N=4096 Nx=137 K=64")
    src, s = base_source
    N = 4096
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    assert Nx == 137
    sc, K, fl = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 64)
    assert fl & gpu.LITHO_FLAG_GRAM_ROUTE
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    B = A * np.sqrt(S)[None, :]
    w = np.linalg.eigvalsh(B.conj().T @ B)[::-1]
    assert np.abs(sc - w[:64]).max() / w[0] < 1e-10
    V = K[:, ::-1].T.astype(np.complex128)  # v = flip(K)
    R = B @ (B.conj().T @ V) - V * sc
    assert np.abs(R).max() / sc[0] < 1e-5  # f32 kernel storage
    del A, B
    m = gpu.manhattan_mask(N, 99)
    img = gpu.socs_image(m, K, sc, Nx, Nx)
    ref = orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32), Nx,
                        Nx, nthreads=16)
    assert max_rel(img, ref) < IMG_TOL
    # the same clip with 16-bit kernel storage (FP32 SOCS loop, the nextPow2
    # grid of 1024 points): bf16 is config 5's storage, fp16 and bf16 hi + lo
    # are reported beside it
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        dm, di = ctx.alloc(m.nbytes), ctx.alloc(m.nbytes)
        ctx.h2d(dm, m)
        for flag, tol in ((0, IMG_TOL), (gpu.LITHO_PREC_BANK_BF16, 1e-2),
                          (gpu.LITHO_PREC_BANK_FP16, 2e-3), (gpu.LITHO_PREC_BANK_BF16X2, 1e-4)):
            for graph in (False, True):
                ctx.socs_image_batch_pipelined(bank, dm, 1, N, di, flag, 1, graph)
                ctx.sync()
                got = ctx.d2h(np.zeros_like(m), di)
                e = max_rel(got, ref)
                print(f"[C5 bank flag {flag} graph {graph}] max rel err vs FP64 {e:.3e}",
                      flush=True)
                assert e < tol, (flag, e)
            if flag == 0:
                assert np.array_equal(got, img)
        with pytest.raises(gpu.LithoError):  # only bank formats on this path
            ctx.socs_image_batch_pipelined(bank, dm, 1, N, di, gpu.LITHO_PREC_COMPACT, 1, False)
        ctx.free(dm)
        ctx.free(di)
        bank.close()
    finally:
        ctx.close()
