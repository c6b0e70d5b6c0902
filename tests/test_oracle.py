"""Pins of the FP64 CPU oracle (CPU only).

Every oracle function is checked against something it was not derived from:
the reference's own fixtures (tests/golden, copied from
source/TCC_HLS/hls/tb/golden and output/Different_resolution_tests), LAPACK
zheevr through scipy (the routine calcKernels calls, litho_thread.cpp
L621-640), numpy's pocketfft, brute-force sums on tiny grids and closed forms.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import (BASE_LAMBDA, BASE_NA, FIX_DEFOCUS, FIX_NA, golden, max_rel)


# ---------------------------------------------------------------- FFT -------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 64, 137, 768, 1024])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    assert max_rel(orc.fft1d(x, -1), ref) < 1e-13
    assert max_rel(orc.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-13


@pytest.mark.parametrize("n", [6, 15])
def test_fft_matches_direct_sum(orc, n):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    direct = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
    assert max_rel(orc.fft1d(x, -1), direct) < 1e-13


def test_fft2d_matches_numpy(orc):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))
    assert max_rel(orc.fft2d(x, -1), np.fft.fft2(x)) < 1e-13


# ---------------------------------------------------------------- source ----
def test_annular_counts_and_weights(orc):
    # integer radii int(0.6*20)=12, int(0.9*20)=18 (common/source.cpp L6-7)
    src = orc.source_annular(41, 0.6, 0.9)
    nz = src[src != 0]
    assert nz.size == 572
    assert np.all(nz == np.float32(1.0) / np.float32(572.0))
    assert orc.source_annular(101, 0.6, 0.9).astype(bool).sum() == 3552
    assert orc.outer_sigma(src) == pytest.approx(18.0 / 20.0, abs=1e-7)
    assert orc.count_live_sources(src, orc.outer_sigma(src)) == 572


def test_nx_values(orc):
    src = orc.source_annular(41, 0.6, 0.9)
    s = orc.outer_sigma(src)
    # 4 nm pixels, NA 0.85: N = 512 / 1024 / 4096 -> Nx 17 / 34 / 137
    assert [orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA) for N in (256, 512, 768, 1024, 4096)] \
        == [8, 17, 25, 34, 137]
    # reference runs (output/Different_resolution_tests/*/fft_meta.txt: Nx 2..32)
    fs = orc.outer_sigma(orc.source_annular(101, 0.6, 0.9))
    assert [orc.compute_nx(FIX_NA, L, fs, BASE_LAMBDA) for L in (256, 512, 1024, 2048, 4096)] \
        == [2, 4, 8, 16, 32]


def test_other_sources(orc):
    d = orc.source_dipole(41, 0.2, 0.5, True)
    q = orc.source_quadrupole(41, 0.2, 0.5)
    p = orc.source_point(41, 0.0, 0.0)
    for s in (d, q, p):
        assert s.sum() == pytest.approx(1.0, abs=1e-6)
    # dipole circles are mirror images across the centre column/row
    assert np.array_equal(d.astype(bool), d.astype(bool)[::-1, ::-1])
    assert np.array_equal(q.astype(bool), q.astype(bool).T)
    assert (p != 0).sum() == 1


# ---------------------------------------------------------------- mask FFT --
def test_mask_spectrum_direct_formula(orc):
    # out[N/2+ky][N/2+kx] = (1/N^2) sum mask e^{-2 pi i (kx (x-N/2) + ky (y-N/2))/N}
    N = 8
    rng = np.random.default_rng(1)
    m = (rng.random((N, N)) > 0.5).astype(np.float32)
    k = np.arange(N) - N // 2
    x = np.arange(N) - N // 2
    E = np.exp(-2j * np.pi * np.outer(k, x) / N)
    direct = E @ m.astype(np.float64) @ E.T / N ** 2
    assert max_rel(orc.mask_spectrum(m), direct) < 1e-13


def test_mask_spectrum_dc_hermitian_parseval(orc):
    N = 64
    rng = np.random.default_rng(2)
    m = rng.random((N, N)).astype(np.float32)
    S = orc.mask_spectrum(m)
    assert S[N // 2, N // 2].real == pytest.approx(m.astype(np.float64).mean(), rel=1e-13)
    # Hermitian about the centre: S[N/2+k] = conj(S[N/2-k]) for |k| < N/2
    c = S[1:, 1:]
    assert max_rel(c, np.conj(c[::-1, ::-1])) < 1e-13
    assert np.sum(np.abs(S) ** 2) * N * N == pytest.approx(np.sum(m.astype(np.float64) ** 2),
                                                          rel=1e-12)


# ---------------------------------------------------------------- TCC -------
@pytest.fixture(scope="module")
def fix_tcc(orc, fix_source):
    src, s = fix_source
    Nx = orc.compute_nx(FIX_NA, 512, s, BASE_LAMBDA)
    return orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 512, 512, Nx, Nx), Nx


def test_tcc_matches_reference_fixture_bit_exact(fix_tcc):
    # source/TCC_HLS/hls/tb/golden/tcc_expected_{r,i}.bin (81x81 f32)
    T, Nx = fix_tcc
    assert Nx == 4
    tr = np.fromfile(golden("tcc_expected_r.bin"), np.float32).reshape(81, 81)
    ti = np.fromfile(golden("tcc_expected_i.bin"), np.float32).reshape(81, 81)
    assert np.array_equal(T.real.astype(np.float32), tr)
    assert np.array_equal(T.imag.astype(np.float32), ti)


def test_tcc_equals_pupil_product_and_invariants(orc, base_source):
    src, s = base_source
    N = 256
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    # zero defocus: the pupil is a 0/1 indicator, TCC = S * (integer overlap counts)
    assert set(np.unique(A.real)) <= {0.0, 1.0} and not A.imag.any()
    assert max_rel(T, (A * S) @ A.conj().T) < 1e-14
    assert np.array_equal(T, T.conj().T)
    c = T.shape[0] // 2
    # every source point lies inside the unit pupil, so TCC(0,0) = sum S = 1
    assert T[c, c].real == pytest.approx(1.0, abs=1e-6)
    assert np.trace(T).real == pytest.approx(float((A != 0).sum(0) @ S), rel=1e-13)
    assert np.linalg.eigvalsh(T).min() > -1e-12


# ---------------------------------------------------------------- eigen -----
def test_jacobi_matches_lapack_zheevr(orc, fix_tcc):
    T, _ = fix_tcc
    w, V = orc.eigh(T)
    wl, Vl = scipy.linalg.eigh(T, driver="evr")
    assert np.abs(w - wl).max() / wl.max() < 1e-13
    assert np.abs(T @ V - V * w).max() < 1e-13
    assert np.abs(V.conj().T @ V - np.eye(T.shape[0])).max() < 1e-12


@pytest.mark.parametrize("L,fname", [(256, "eig_resolution_256x256.txt"),
                                     (512, "eig_resolution_512x512.txt"),
                                     (1024, "eig_resolution_1024x1024.txt")])
def test_eigenvalues_match_reference_runs(orc, fix_source, L, fname):
    # output/Different_resolution_tests/*/kernels/kernel_info.txt (6 significant digits)
    src, s = fix_source
    Nx = orc.compute_nx(FIX_NA, L, s, BASE_LAMBDA)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, L, L, Nx, Nx)
    ref = np.loadtxt(golden(fname))
    k = min(ref.size, T.shape[0], 16)
    sc, _, _ = orc.calc_kernels(T, k, Nx, Nx)
    big = ref[:k] > 1e-6 * ref[0]
    assert np.all(np.abs(sc[:k][big] - ref[:k][big]) / ref[:k][big] < 6e-6)


def test_eigenvalues_match_matlab(orc, fix_source):
    # data/matlab/out_validate/validation_metrics.csv: FP64 pupil, so ~1e-7
    src, s = fix_source
    vals = dict(l.strip().split(",") for l in open(golden("matlab_validation_nx8.csv")))
    T = orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 1024, 1024, 8, 8)
    assert np.trace(T).real == pytest.approx(float(vals["tcc_trace"]), rel=1e-6)
    sc, _, _ = orc.calc_kernels(T, 10, 8, 8)
    ref = np.array([float(vals[f"eig_{i}"]) for i in range(10)])
    assert np.abs(sc - ref).max() / ref[0] < 1e-6


def test_kernels_are_flipped_eigenvectors(orc, fix_tcc):
    T, Nx = fix_tcc
    sc, K, split = orc.calc_kernels(T, 12, Nx, Nx)
    V = K[:, ::-1].T  # v[j] = K[n-1-j]
    assert np.abs(T @ V - V * sc).max() < 1e-12
    assert np.all(np.diff(sc) <= 0)
    # phase rule: first near-maximal component is real positive
    for v in V.T:
        a = np.abs(v)
        k = int(np.argmax(a >= (1 - 1e-6) * a.max()))
        assert v[k].imag == 0 and v[k].real > 0


def test_canonical_basis_of_degenerate_pair(orc, fix_tcc):
    T, Nx = fix_tcc
    sc, K, split = orc.calc_kernels(T, 3, Nx, Nx)  # 1 and 2 form the 2-D irrep pair
    assert not split and abs(sc[1] - sc[2]) < 1e-12 * sc[0]
    W = 2 * Nx + 1
    V = K[:, ::-1].T
    mirror = np.array([(i // W) * W + (W - 1 - i % W) for i in range(W * W)])
    mu = [np.vdot(V[:, j], V[mirror, j]).real for j in (1, 2)]
    assert mu[0] == pytest.approx(1.0, abs=1e-10) and mu[1] == pytest.approx(-1.0, abs=1e-10)
    # a split cluster is reported
    _, _, split10 = orc.calc_kernels(T, 10, Nx, Nx)
    assert split10


def test_gram_route_equals_explicit(orc):
    src = orc.source_annular(21, 0.6, 0.9)  # 148 live points, fewer than n = 289
    s = orc.outer_sigma(src)
    Nx = 8
    T = orc.calc_tcc(src, s, BASE_LAMBDA, BASE_NA, 0.0, 1024, 1024, Nx, Nx)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 1024, 1024, Nx, Nx)
    assert A.shape[1] < T.shape[0]
    se, Ke, _ = orc.calc_kernels(T, 12, Nx, Nx)
    sg, Kg, _ = orc.calc_kernels_gram(A, S, 12, Nx, Nx)
    assert np.abs(se - sg).max() / se[0] < 1e-13
    assert np.abs(Ke - Kg).max() < 1e-9


# ---------------------------------------------------------------- images ----
def test_tcc_direct_image_matches_reference_fixture_bit_exact(orc, fix_tcc):
    # source/TCC_HLS/hls/tb/golden/image_expected_512x512.bin from mask_256x256.bin
    T, Nx = fix_tcc
    m = np.fromfile(golden("mask_256x256.bin"), np.float32).reshape(256, 256)
    M = np.zeros((512, 512), np.float32)
    M[128:384, 128:384] = m  # createMask centring (common/mask.cpp L54-59)
    img = orc.calc_image_tcc(orc.mask_spectrum(M), T, Nx, Nx)
    gold = np.fromfile(golden("image_expected_512x512.bin"), np.float32).reshape(512, 512)
    assert np.array_equal(img, gold)


def test_socs_full_rank_equals_tcc_direct(orc, fix_tcc):
    T, Nx = fix_tcc
    m = np.fromfile(golden("mask_256x256.bin"), np.float32).reshape(256, 256)
    M = np.zeros((512, 512), np.float32)
    M[128:384, 128:384] = m
    mf = orc.mask_spectrum(M)
    sc, K, _ = orc.calc_kernels(T, T.shape[0], Nx, Nx)
    socs = orc.calc_socs(mf, K, sc.astype(np.float32), Nx, Nx)
    ref = orc.calc_image_tcc(mf, T, Nx, Nx)
    assert max_rel(socs, ref) < 1e-6


def test_mask_spectrum_rectangular_direct_formula(orc):
    # Ly x Lx: out[Ly/2+ky][Lx/2+kx] = (1/(Lx Ly)) sum mask e^{-2 pi i (kx x'/Lx + ky y'/Ly)}
    Ly, Lx = 8, 16
    m = (np.random.default_rng(2).random((Ly, Lx)) > 0.5).astype(np.float32)

    def E(n):
        k = np.arange(n) - n // 2
        return np.exp(-2j * np.pi * np.outer(k, k) / n)

    direct = E(Ly) @ m.astype(np.float64) @ E(Lx).T / (Lx * Ly)
    assert max_rel(orc.mask_spectrum(m), direct) < 1e-13


def test_rectangular_clip_full_rank_and_transpose(orc):
    """A 64 x 128 clip (period 256 x 512 nm, Nx = 4, Ny = 2): SOCS with every
    kernel equals the TCC-direct image, and the clip transposed with the
    periods and Nx, Ny swapped gives the transposed image (the annulus and the
    in-focus pupil are symmetric under x <-> y), which an Lx/Ly or Nx/Ny mix-up
    anywhere in the rectangular path breaks."""
    src = orc.source_annular(41, 0.6, 0.9)
    s = orc.outer_sigma(src)
    m = np.zeros((64, 128), np.float32)
    m[16:48, 24:104] = orc.linespace(80, 32, 6, 10)
    T = orc.calc_tcc(src, s, 193.0, 0.8, 0.0, 512, 256, 4, 2)
    mf = orc.mask_spectrum(m)
    img = orc.calc_image_tcc(mf, T, 4, 2)
    assert img.shape == (64, 128) and img.max() > 0
    sc, K, _ = orc.calc_kernels(T, T.shape[0], 4, 2)
    assert max_rel(orc.calc_socs(mf, K, sc.astype(np.float32), 4, 2), img) < 1e-6
    Tt = orc.calc_tcc(src, s, 193.0, 0.8, 0.0, 256, 512, 2, 4)
    imgt = orc.calc_image_tcc(orc.mask_spectrum(np.ascontiguousarray(m.T)), Tt, 2, 4)
    assert max_rel(imgt, img.T) < 1e-6


def test_fi_is_exact_on_band_limited_grids(orc):
    G, N, b = 32, 96, 7
    rng = np.random.default_rng(5)
    ky, kx = np.meshgrid(np.arange(-b, b + 1), np.arange(-b, b + 1), indexing="ij")
    c = rng.standard_normal(ky.shape) + 1j * rng.standard_normal(ky.shape)
    c = c + np.conj(c[::-1, ::-1])  # Hermitian -> real signal

    def evaluate(n):
        y = np.arange(n)[:, None, None, None]
        x = np.arange(n)[None, :, None, None]
        ph = np.exp(2j * np.pi * (ky * (y - n // 2) + kx * (x - n // 2)) / n)
        return (ph * c).sum(axis=(2, 3)).real

    out = orc.fi(evaluate(G), N, N)
    assert max_rel(out, evaluate(N)) < 1e-12


def test_socs_is_independent_of_conv_grid(orc, base_source):
    src, s = base_source
    N = 256
    Nx = 8
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    T = (A * S) @ A.conj().T
    sc, K, _ = orc.calc_kernels(T, 16, Nx, Nx)
    rng = np.random.default_rng(9)
    m = (rng.random((N, N)) > 0.6).astype(np.float32)
    mf = orc.mask_spectrum(m)
    imgs = []
    for G in (orc.conv_grid(Nx), 2 * orc.conv_grid(Nx), 4 * Nx + 1 + 1):
        lib = orc.lib()
        t = np.zeros(G * G)
        rc = lib.orc_socs_tmpimg(t, G, G, np.ascontiguousarray(mf.real).ravel(),
                                 np.ascontiguousarray(mf.imag).ravel(), N, N,
                                 np.ascontiguousarray(K.real).ravel(),
                                 np.ascontiguousarray(K.imag).ravel(),
                                 sc.astype(np.float32), 16, Nx, Nx, 1)
        assert rc == 0
        imgs.append(orc.fi(t.reshape(G, G), N, N))
    assert max_rel(imgs[1], imgs[0]) < 1e-12
    assert max_rel(imgs[2], imgs[0]) < 1e-12


def test_socs_single_flat_kernel_is_lowpass_intensity(orc):
    # one kernel equal to 1 on the whole window, lambda = 1: the image is
    # |lowpass(mask)|^2 with the window as the pass band
    N, Nx = 128, 6
    rng = np.random.default_rng(11)
    m = (rng.random((N, N)) > 0.5).astype(np.float32)
    W = 2 * Nx + 1
    K = np.ones((1, W * W), np.complex128)
    img = orc.calc_socs(orc.mask_spectrum(m), K, np.ones(1, np.float32), Nx, Nx)
    f = np.fft.fftfreq(N, 1.0 / N)
    keep = (np.abs(f)[:, None] <= Nx) & (np.abs(f)[None, :] <= Nx)
    low = np.fft.ifft2(np.fft.fft2(m.astype(np.float64)) * keep)
    assert max_rel(img, np.abs(low) ** 2) < 1e-6


def test_linespace_mask(orc):
    m = orc.linespace(16, 4, 3, 5, horizontal=False)
    assert np.array_equal(m[0], (np.arange(16) % 8 < 3).astype(np.float32))
    assert np.array_equal(m[0], m[3])
