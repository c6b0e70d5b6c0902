"""numpy front end of the FP64 CPU oracle (liblitho_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke check and the CPU
baseline leg of bench.py; the product package never imports it.  Each wrapper
names the reference function (validation/golden/src/litho_thread.cpp unless
stated) that the C function restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OracleError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblitho_oracle.so")
        if not os.path.exists(path):
            raise OracleError(f"{path} not built (run make -C oracle)")
        L = C.CDLL(path)
        i, f, d = C.c_int, C.c_float, C.c_double
        sig = {
            "orc_fft1d": [i, _f64p, _f64p, i],
            "orc_fft2d": [i, i, _f64p, _f64p, i],
            "orc_source_annular": [_f32p, i, f, f],
            "orc_source_dipole": [_f32p, i, f, f, i],
            "orc_source_quadrupole": [_f32p, i, f, f],
            "orc_source_point": [_f32p, i, f, f],
            "orc_count_live_sources": [_f32p, i, f],
            "orc_linespace": [_f32p, i, i, i, i, i],
            "orc_mask_spectrum": [_f32p, i, i, _f64p, _f64p],
            "orc_calc_tcc": [_f64p, _f64p, i, _f32p, i, f, f, f, f, i, i, i, i],
            "orc_pupil_matrix": [_f64p, _f64p, _f64p, C.POINTER(i), i, _f32p, i, f, f, f, f,
                                 i, i, i, i],
            "orc_eigh": [i, _f64p, _f64p, _f64p, _f64p, _f64p],
            "orc_calc_kernels": [_f64p, _f64p, _f64p, i, _f64p, _f64p, i, i, i, C.POINTER(i)],
            "orc_calc_kernels_gram": [_f64p, _f64p, _f64p, i, _f64p, _f64p, _f64p, i, i, i, i,
                                      C.POINTER(i)],
            "orc_canonicalize": [_f64p, _f64p, _f64p, i, i, i, i, d],
            "orc_conv_grid": [i],
            "orc_socs_tmpimg": [_f64p, i, i, _f64p, _f64p, i, i, _f64p, _f64p, _f32p, i, i, i, i],
            "orc_fi": [_f64p, i, i, _f64p, i, i],
            "orc_calc_socs": [_f32p, _f64p, _f64p, i, i, _f64p, _f64p, _f32p, i, i, i, i],
            "orc_calc_image_tcc": [_f32p, _f64p, _f64p, i, i, _f64p, _f64p, i, i, i],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.orc_outer_sigma.argtypes = [_f32p, i]
        L.orc_outer_sigma.restype = C.c_float
        L.orc_compute_nx.argtypes = [f, i, f, f]
        L.orc_compute_nx.restype = C.c_int
        L.orc_normalize.argtypes = [_f32p, i]
        L.orc_normalize.restype = None
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with code {rc}")


def fft1d(x, sign):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    _check(lib().orc_fft1d(x.size, re, im, sign), "orc_fft1d")
    return re + 1j * im


def fft2d(x, sign):
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    _check(lib().orc_fft2d(x.shape[0], x.shape[1], re, im, sign), "orc_fft2d")
    return re + 1j * im


def source_annular(size, inner, outer):
    """createAnnular + normalizeMatrix (common/source.cpp L5-23, L89-103)."""
    s = np.zeros(size * size, np.float32)
    _check(lib().orc_source_annular(s, size, inner, outer), "annular")
    return s.reshape(size, size)


def source_dipole(size, radius, offset, on_x=True):
    s = np.zeros(size * size, np.float32)
    _check(lib().orc_source_dipole(s, size, radius, offset, int(on_x)), "dipole")
    return s.reshape(size, size)


def source_quadrupole(size, radius, offset):
    s = np.zeros(size * size, np.float32)
    _check(lib().orc_source_quadrupole(s, size, radius, offset), "quadrupole")
    return s.reshape(size, size)


def source_point(size, px, py):
    s = np.zeros(size * size, np.float32)
    _check(lib().orc_source_point(s, size, px, py), "point")
    return s.reshape(size, size)


def outer_sigma(src):
    """getOuterSigma (L427-439), clamped to 1 as main (L1344)."""
    src = np.ascontiguousarray(src, np.float32)
    v = lib().orc_outer_sigma(src.ravel(), src.shape[0])
    return min(float(np.float32(v)), 1.0)


def compute_nx(NA, period, sigma, lam):
    """Nx = floor(NA*period*(1+sigma)/lambda) in float (L1346)."""
    return lib().orc_compute_nx(NA, period, sigma, lam)


def count_live_sources(src, sigma):
    src = np.ascontiguousarray(src, np.float32)
    return lib().orc_count_live_sources(src.ravel(), src.shape[0], sigma)


def linespace(sx, sy, line, space, horizontal=False):
    m = np.zeros(sx * sy, np.float32)
    _check(lib().orc_linespace(m, sx, sy, line, space, int(horizontal)), "linespace")
    return m.reshape(sy, sx)


def mask_spectrum(mask):
    """FT_r2c_padded (L255-331): full centred spectrum / N^2, complex128."""
    mask = np.ascontiguousarray(mask, np.float32)
    sy, sx = mask.shape
    re = np.zeros(sx * sy)
    im = np.zeros(sx * sy)
    _check(lib().orc_mask_spectrum(mask.ravel(), sy, sx, re, im), "mask_spectrum")
    return (re + 1j * im).reshape(sy, sx)


def calc_tcc(src, sigma, lam, NA, defocus, Lx, Ly, Nx, Ny):
    """calcTCC (L441-532): n x n complex128 TCC, float pupil arithmetic."""
    src = np.ascontiguousarray(src, np.float32)
    n = (2 * Nx + 1) * (2 * Ny + 1)
    re = np.zeros(n * n)
    im = np.zeros(n * n)
    _check(lib().orc_calc_tcc(re, im, n, src.ravel(), src.shape[0], sigma, lam, NA, defocus,
                              Lx, Ly, Nx, Ny), "calc_tcc")
    return (re + 1j * im).reshape(n, n)


def pupil_matrix(src, sigma, lam, NA, defocus, Lx, Ly, Nx, Ny):
    """Pupil of the live sources (calcTCC L457-484): A (n x nsrc) and S (nsrc)."""
    src = np.ascontiguousarray(src, np.float32)
    ns = C.c_int(0)
    dummy = np.zeros(1)
    L = lib()
    # count first (A_re may not be NULL through ndpointer; use a raw call)
    fn = L.orc_pupil_matrix
    argt = fn.argtypes
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p] + argt[3:]
    try:
        _check(fn(None, None, None, C.byref(ns), 0, src.ravel(), src.shape[0], sigma, lam, NA,
                  defocus, Lx, Ly, Nx, Ny), "pupil count")
    finally:
        fn.argtypes = argt
    n = (2 * Nx + 1) * (2 * Ny + 1)
    m = ns.value
    Ar = np.zeros(n * m)
    Ai = np.zeros(n * m)
    S = np.zeros(max(m, 1))
    _check(fn(Ar, Ai, S, C.byref(ns), m, src.ravel(), src.shape[0], sigma, lam, NA, defocus,
              Lx, Ly, Nx, Ny), "pupil_matrix")
    del dummy
    return (Ar + 1j * Ai).reshape(n, m), S[:m]


def eigh(H):
    """Full Hermitian eigendecomposition (own Jacobi): w ascending, V columns."""
    H = np.asarray(H, np.complex128)
    n = H.shape[0]
    hr = np.ascontiguousarray(H.real).ravel().copy()
    hi = np.ascontiguousarray(H.imag).ravel().copy()
    w = np.zeros(n)
    vr = np.zeros(n * n)
    vi = np.zeros(n * n)
    _check(lib().orc_eigh(n, hr, hi, w, vr, vi), "eigh")
    V = (vr + 1j * vi).reshape(n, n).T  # column j = eigenvector j
    return w, V


def calc_kernels(tcc, nk, Nx, Ny):
    """calcKernels (L597-663): scales (FP64, descending) and flipped kernels
    (nk x n complex128), canonical basis in degenerate clusters.  Returns
    (scales, kernels, cluster_split)."""
    tcc = np.asarray(tcc, np.complex128)
    n = tcc.shape[0]
    kr = np.zeros(nk * n)
    ki = np.zeros(nk * n)
    sc = np.zeros(nk)
    split = C.c_int(0)
    _check(lib().orc_calc_kernels(kr, ki, sc, nk, np.ascontiguousarray(tcc.real).ravel(),
                                  np.ascontiguousarray(tcc.imag).ravel(), n, Nx, Ny,
                                  C.byref(split)), "calc_kernels")
    return sc, (kr + 1j * ki).reshape(nk, n), bool(split.value)


def calc_kernels_gram(A, S, nk, Nx, Ny):
    """Kernels through the source-space Gram route (same result as calc_kernels)."""
    A = np.asarray(A, np.complex128)
    n, m = A.shape
    kr = np.zeros(nk * n)
    ki = np.zeros(nk * n)
    sc = np.zeros(nk)
    split = C.c_int(0)
    _check(lib().orc_calc_kernels_gram(kr, ki, sc, nk, np.ascontiguousarray(A.real).ravel(),
                                       np.ascontiguousarray(A.imag).ravel(),
                                       np.ascontiguousarray(S, np.float64), n, m, Nx, Ny,
                                       C.byref(split)), "calc_kernels_gram")
    return sc, (kr + 1j * ki).reshape(nk, n), bool(split.value)


def conv_grid(Nx):
    return lib().orc_conv_grid(Nx)


def socs_tmpimg(mskf, krn, scales, Nx, Ny, nthreads=1, grid=None):
    """calcSOCS accumulate + fftshift (L897-970): tmpImgp (float64) on the
    nextPow2(4N+1) grid of L833-834, or on grid = (Gy, Gx) when given (the
    v18 top's fixed 128 x 128, socs_2048_stream_refactored_v18_hls_fft.h L34-35)."""
    mskf = np.asarray(mskf, np.complex128)
    fy, fx = mskf.shape
    krn = np.asarray(krn, np.complex128)
    nk = krn.shape[0]
    Gx, Gy = conv_grid(Nx), conv_grid(Ny)
    if grid is not None:
        Gy, Gx = grid
    out = np.zeros(Gx * Gy)
    _check(lib().orc_socs_tmpimg(out, Gx, Gy, np.ascontiguousarray(mskf.real).ravel(),
                                 np.ascontiguousarray(mskf.imag).ravel(), fx, fy,
                                 np.ascontiguousarray(krn.real).ravel(),
                                 np.ascontiguousarray(krn.imag).ravel(),
                                 np.ascontiguousarray(scales, np.float32), nk, Nx, Ny, nthreads),
           "socs_tmpimg")
    return out.reshape(Gy, Gx)


def fi(img, oy, ox):
    """FI (L692-772): band-limited interpolation to oy x ox (float64)."""
    img = np.ascontiguousarray(img, np.float64)
    iy, ix = img.shape
    out = np.zeros(ox * oy)
    _check(lib().orc_fi(img.ravel(), ix, iy, out, ox, oy), "fi")
    return out.reshape(oy, ox)


def calc_socs(mskf, krn, scales, Nx, Ny, nthreads=1):
    """calcSOCS complete (L813-1040): float32 aerial image, size of mskf."""
    mskf = np.asarray(mskf, np.complex128)
    fy, fx = mskf.shape
    krn = np.asarray(krn, np.complex128)
    img = np.zeros(fx * fy, np.float32)
    _check(lib().orc_calc_socs(img, np.ascontiguousarray(mskf.real).ravel(),
                               np.ascontiguousarray(mskf.imag).ravel(), fx, fy,
                               np.ascontiguousarray(krn.real).ravel(),
                               np.ascontiguousarray(krn.imag).ravel(),
                               np.ascontiguousarray(scales, np.float32), krn.shape[0], Nx, Ny,
                               nthreads), "calc_socs")
    return img.reshape(fy, fx)


def calc_image_tcc(mskf, tcc, Nx, Ny):
    """calcImage + FT_c2r_padded (L539-583, L371-425): TCC-direct image."""
    mskf = np.asarray(mskf, np.complex128)
    fy, fx = mskf.shape
    tcc = np.asarray(tcc, np.complex128)
    img = np.zeros(fx * fy, np.float32)
    _check(lib().orc_calc_image_tcc(img, np.ascontiguousarray(mskf.real).ravel(),
                                    np.ascontiguousarray(mskf.imag).ravel(), fx, fy,
                                    np.ascontiguousarray(tcc.real).ravel(),
                                    np.ascontiguousarray(tcc.imag).ravel(), tcc.shape[0], Nx, Ny),
           "calc_image_tcc")
    return img.reshape(fy, fx)
