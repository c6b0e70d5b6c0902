"""The accelerator-shaped entry litho_socs_tmpimg against the HLS v18 top's
contract (source/SOCS_HLS/src/socs_2048_stream_refactored_v18_hls_fft.h
L34-35, L43, L86-97; .cpp L287-388): a fixed 128 x 128 fftshifted tmpImgp
written to both output buffers for nx, ny <= 8, nx and ny taken separately.

The oracle runs the reference's calcSOCS loop (litho_thread.cpp L897-970) on
the same grid; |IFFT|^2 does not depend on where the window is embedded
(v18 uses (128 (64 - W)) / 64, litho_thread.cpp the centre), so the two agree
to FP32 round-off.  Tolerance 1e-5 in max|a-b|/max|b|.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import BASE_LAMBDA, FIX_DEFOCUS, FIX_NA, max_rel

pytestmark = pytest.mark.gpu

TOL = 1e-5


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def golden_nx8(orc, fix_source):
    """The reference's golden_1024 optics: 1 nm pixels, period 1024 x 1024,
    NA 0.8, defocus 0.2, 101^2 annulus 0.6-0.9 -> Nx = Ny = 8 (v18's range)."""
    src, s = fix_source
    Nx = orc.compute_nx(FIX_NA, 1024, s, BASE_LAMBDA)
    assert Nx == 8
    T = orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 1024, 1024, Nx, Nx)
    sc, K, _ = orc.calc_kernels(T, 10, Nx, Nx)
    return Nx, sc, K


@pytest.fixture(scope="module")
def golden_8x5(orc, fix_source):
    """Same optics on a 1024 x 640 nm period: Nx = 8, Ny = 5 (kernels 11 x 17)."""
    src, s = fix_source
    nx = orc.compute_nx(FIX_NA, 1024, s, BASE_LAMBDA)
    ny = orc.compute_nx(FIX_NA, 640, s, BASE_LAMBDA)
    assert (nx, ny) == (8, 5)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, FIX_NA, FIX_DEFOCUS, 1024, 640, nx, ny)
    sc, K, _ = orc.calc_kernels(T, 12, nx, ny)
    return nx, ny, sc, K


def _mask(gpu, rows, cols, seed):
    m = gpu.manhattan_mask(max(rows, cols), seed)
    return np.ascontiguousarray(m[:rows, :cols])


def _call_v18(gpu, mskf, K, sc, nx, ny, nbuf):
    """Call the C-ABI with two distinct caller buffers of nbuf floats each
    (as the v18 testbench allocates tmpImg_ddr and output)."""
    Ly, Lx = mskf.shape
    mr, mi = gpu._planes(mskf)
    kr, ki = gpu._planes(K)
    sc = np.ascontiguousarray(sc, np.float32)
    t = np.full(nbuf, -1.0, np.float32)
    o = np.full(nbuf, -1.0, np.float32)
    rc = gpu.lib().litho_socs_tmpimg(mr, mi, kr, ki, sc, t, o, sc.size, nx, ny, Lx, Ly)
    assert rc == 0, rc
    return t, o


def test_v18_grid_rule(gpu):
    case("litho_tmpimg_grid")
    assert [gpu.tmpimg_grid(n) for n in (0, 2, 4, 8)] == [128] * 4
    assert gpu.tmpimg_grid(17) == 128 and gpu.tmpimg_grid(31) == 128
    assert gpu.tmpimg_grid(32) == 256 and gpu.tmpimg_grid(34) == 256
    assert gpu.tmpimg_grid(137) == 1024


def test_v18_nx8_golden_optics(gpu, orc, golden_nx8):
    case("v18 tmpImgp 128^2, golden 1 nm optics, Nx = 8, K = 10")
    Nx, sc, K = golden_nx8
    m = _mask(gpu, 1024, 1024, 11)
    mf = orc.mask_spectrum(m)
    t, o = _call_v18(gpu, mf, K, sc, Nx, Nx, 128 * 128)
    ref = orc.socs_tmpimg(mf, K, sc.astype(np.float32), Nx, Nx, grid=(128, 128))
    assert max_rel(t.reshape(128, 128), ref) < TOL
    assert np.array_equal(t, o)
    # the centre of tmpImgp is the zero-offset intensity (fftshift by 64)
    assert ref.shape == (128, 128)
    # and the same grid through the Python front end
    assert max_rel(gpu.calc_socs_hls(mf, K, sc, Nx, Nx), ref) < TOL


def test_v18_nx_ne_ny(gpu, orc, golden_8x5):
    case("v18 tmpImgp 128^2, Nx = 8, Ny = 5, 640 x 1024 spectrum")
    nx, ny, sc, K = golden_8x5
    m = _mask(gpu, 640, 1024, 12)
    mf = orc.mask_spectrum(m)
    t, o = _call_v18(gpu, mf, K, sc, nx, ny, 128 * 128)
    ref = orc.socs_tmpimg(mf, K, sc.astype(np.float32), nx, ny, grid=(128, 128))
    assert max_rel(t.reshape(128, 128), ref) < TOL
    assert np.array_equal(t, o)


def test_tmpimg_on_grid_unequal_axes(gpu, orc, golden_8x5):
    case("tmpImgp on a 128 x 64 grid (sampled from the square grid)")
    nx, ny, sc, K = golden_8x5
    m = _mask(gpu, 640, 1024, 13)
    mf = orc.mask_spectrum(m)
    for grid in ((128, 64), (32, 128), (64, 64)):
        got = gpu.calc_socs_hls(mf, K, sc, nx, ny, grid=grid)
        ref = orc.socs_tmpimg(mf, K, sc.astype(np.float32), nx, ny, grid=grid)
        assert got.shape == grid
        assert max_rel(got, ref) < TOL, grid


def test_tmpimg_reference_grid_and_fi(gpu, orc, golden_nx8):
    case("calcSOCS: nextPow2 grid (64 at Nx = 8) + FI equals the v18 grid + FI")
    Nx, sc, K = golden_nx8
    m = _mask(gpu, 1024, 1024, 14)
    mf = orc.mask_spectrum(m)
    img = gpu.calcSOCS(mf, K, sc, Nx, Nx)
    ref = orc.calc_socs(mf, K, sc.astype(np.float32), Nx, Nx)
    assert max_rel(img, ref) < TOL
    # FI of the 128 grid gives the same image (grid independence)
    img128 = gpu.FI(gpu.calc_socs_hls(mf, K, sc, Nx, Nx), 1024, 2 * Nx)
    assert max_rel(img128, ref) < TOL


def test_tmpimg_dev_matches_host_entry(gpu, orc, golden_8x5):
    case("litho_socs_tmpimg_dev (bank + device planes) vs litho_socs_tmpimg")
    nx, ny, sc, K = golden_8x5
    m = _mask(gpu, 640, 1024, 15)
    mf = orc.mask_spectrum(m).astype(np.complex64)
    host = gpu.calc_socs_hls(mf, K, sc, nx, ny)
    ctx = gpu.Context()
    bank = ctx.bank(K, sc, nx, ny)
    try:
        mr, mi = gpu._planes(mf)
        dr, di = ctx.alloc(mr.nbytes), ctx.alloc(mi.nbytes)
        dt, do = ctx.alloc(128 * 128 * 4), ctx.alloc(128 * 128 * 4)
        ctx.h2d(dr, mr)
        ctx.h2d(di, mi)
        for grid in ((128, 128), (128, 64)):
            Gy, Gx = grid
            rc = gpu.lib().litho_socs_tmpimg_dev(ctx.h, bank, dr, di, 1024, 640, Gx, Gy, dt, do)
            assert rc == 0, rc
            ctx.sync()
            t = ctx.d2h(np.zeros(Gy * Gx, np.float32), dt).reshape(grid)
            o = ctx.d2h(np.zeros(Gy * Gx, np.float32), do).reshape(grid)
            assert np.array_equal(t, o)
            if grid == (128, 128):
                assert np.array_equal(t, host)
            ref = orc.socs_tmpimg(mf, K, sc.astype(np.float32), nx, ny, grid=grid)
            assert max_rel(t, ref) < TOL
        # a bank of other window sizes is rejected for a grid it does not fit
        assert gpu.lib().litho_socs_tmpimg_dev(ctx.h, bank, dr, di, 1024, 640, 16, 128, dt,
                                               do) == -1
        for p in (dr, di, dt, do):
            ctx.free(p)
    finally:
        bank.close()
        ctx.close()
