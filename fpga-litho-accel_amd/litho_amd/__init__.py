"""Python mirror of the reference's SOCS call surface over the MI355X C-ABI.

The functions keep the names and argument meaning of the reference's CPU
oracle (validation/golden/src/litho_thread.cpp): FT_r2c_padded (L255),
calcTCC (L441), calcKernels (L597), FI (L692), calcSOCS (L813), plus the HLS
accelerator top calc_socs_2048_hls_stream_refactored_v18
(source/SOCS_HLS/src/socs_2048_stream_refactored_v18_hls_fft.cpp L287).
Everything runs in liblitho_socs.so on the GPU; a missing library or GPU
raises LithoError -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(os.path.dirname(_HERE), "liblitho_socs.so")
_LIB = None

LITHO_FLAG_CLUSTER_SPLIT = 1
LITHO_FLAG_GRAM_ROUTE = 2

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

# every symbol include/litho_socs.h declares
EXPORTS = [
    "litho_status_string", "litho_version", "litho_ctx_create", "litho_ctx_destroy",
    "litho_ctx_sync", "litho_dev_alloc", "litho_dev_free", "litho_memcpy_h2d",
    "litho_memcpy_d2h", "litho_memset_dev", "litho_memcpy_d2d", "litho_source_annular", "litho_source_quadrupole",
    "litho_source_dipole", "litho_source_point", "litho_outer_sigma", "litho_compute_nx",
    "litho_manhattan_mask", "litho_mask_spectrum", "litho_tcc_build", "litho_kernels_topk",
    "litho_kernels_from_source", "litho_setup_timing", "litho_tmpimg_grid", "litho_socs_tmpimg",
    "litho_socs_tmpimg_on_grid", "litho_socs_tmpimg_dev", "litho_socs_image", "litho_fi",
    "litho_image_tcc", "litho_image_tcc_rect", "litho_mask_spectrum_rect", "litho_socs_image_rect",
    "litho_bank_create", "litho_bank_destroy", "litho_bank_info", "litho_socs_image_batch",
    "litho_socs_image_batch_noalloc",
    "litho_ctx_set_path", "litho_ctx_set_timing",
    "litho_last_timing", "litho_stage_name", "litho_socs_image_batch_pipelined",
    "litho_tcc_build_prec", "litho_kernels_from_source_prec",
    "litho_socs_image_buckets", "litho_bucket_times", "litho_autotune", "litho_plan_get",
    "litho_plan_set",
    "litho_fno_create", "litho_fno_destroy", "litho_fno_param_count", "litho_fno_get_params",
    "litho_fno_set_params", "litho_fno_forward", "litho_fno_forward_fp32", "litho_fno_loss_grad",
    "litho_fno_adam", "litho_fno_verify",
]
LITHO_TCC_FP64 = 0
LITHO_TCC_SPLIT_BF16 = 1
LITHO_PREC_FP32 = 0
LITHO_PREC_INTER_FP16 = 1
LITHO_PREC_BANK_BF16 = 2
LITHO_PREC_BANK_FP16 = 4
LITHO_PREC_MASK_U8 = 8
LITHO_PREC_IMAGE_FP16 = 16
LITHO_PREC_DFT_MFMA = 32
LITHO_PREC_BANK_BF16X2 = 64
LITHO_PREC_MASK_1BIT = 128
LITHO_PREC_MIXED = LITHO_PREC_INTER_FP16 | LITHO_PREC_BANK_FP16
LITHO_PREC_COMPACT = (LITHO_PREC_INTER_FP16 | LITHO_PREC_BANK_FP16 | LITHO_PREC_MASK_U8
                      | LITHO_PREC_DFT_MFMA | LITHO_PREC_IMAGE_FP16)
LITHO_PATH_AUTO = 0
LITHO_PATH_POW2 = 1
LITHO_ESCRATCH = -7
LITHO_NOALLOC_SCRATCH_BYTES = 64 << 20


class LithoError(RuntimeError):
    def __init__(self, status, what):
        msg = lib().litho_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: status {status} ({msg})")
        self.status = status


class LithoOptics(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("NA", C.c_float), ("defocus", C.c_float),
                ("outerSigma", C.c_float), ("Lx", C.c_int), ("Ly", C.c_int),
                ("Nx", C.c_int), ("Ny", C.c_int)]


class LithoBucket(C.Structure):
    """One clip size of a mixed-size step (litho_socs_image_buckets)."""
    _fields_ = [("bank", C.c_void_p), ("N", C.c_int), ("count", C.c_int),
                ("d_masks", C.c_void_p), ("d_images", C.c_void_p)]


def library_path():
    return _LIBPATH


def lib():
    """Load liblitho_socs.so (built in-tree); raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError(f"{_LIBPATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(_LIBPATH)
        i, f, vp, sz = C.c_int, C.c_float, C.c_void_p, C.c_size_t
        L.litho_status_string.argtypes = [i]
        L.litho_status_string.restype = C.c_char_p
        L.litho_version.restype = i
        L.litho_ctx_create.argtypes = [C.POINTER(vp), i]
        L.litho_ctx_destroy.argtypes = [vp]
        L.litho_ctx_sync.argtypes = [vp]
        L.litho_dev_alloc.argtypes = [vp, C.POINTER(vp), sz]
        L.litho_dev_free.argtypes = [vp, vp]
        L.litho_memcpy_h2d.argtypes = [vp, vp, vp, sz]
        L.litho_memcpy_d2h.argtypes = [vp, vp, vp, sz]
        L.litho_memset_dev.argtypes = [vp, vp, i, sz]
        L.litho_memcpy_d2d.argtypes = [vp, vp, vp, sz]
        L.litho_source_annular.argtypes = [_f32p, i, f, f]
        L.litho_source_quadrupole.argtypes = [_f32p, i, f, f]
        L.litho_source_dipole.argtypes = [_f32p, i, f, f, i]
        L.litho_source_point.argtypes = [_f32p, i, f, f]
        L.litho_outer_sigma.argtypes = [_f32p, i]
        L.litho_outer_sigma.restype = f
        L.litho_compute_nx.argtypes = [f, i, f, f]
        L.litho_manhattan_mask.argtypes = [_f32p, i, C.c_ulonglong, i, i, f]
        L.litho_mask_spectrum.argtypes = [_f32p, i, _f32p, _f32p]
        L.litho_mask_spectrum_rect.argtypes = [_f32p, i, i, _f32p, _f32p]
        L.litho_socs_image_rect.argtypes = [_f32p, i, i, _f32p, _f32p, _f32p, i, i, i, _f32p]
        L.litho_tcc_build.argtypes = [C.POINTER(LithoOptics), _f32p, i, _f64p, _f64p]
        L.litho_kernels_topk.argtypes = [_f64p, _f64p, i, i, i, i, _f64p, _f32p, _f32p,
                                         C.POINTER(i)]
        L.litho_kernels_from_source.argtypes = [C.POINTER(LithoOptics), _f32p, i, i, _f64p,
                                                _f32p, _f32p, C.POINTER(i)]
        L.litho_fno_create.argtypes = [vp, i, C.c_ulonglong, C.POINTER(vp)]
        L.litho_fno_destroy.argtypes = [vp]
        L.litho_fno_param_count.argtypes = [vp]
        L.litho_fno_param_count.restype = C.c_longlong
        L.litho_fno_get_params.argtypes = [vp, _f32p, C.c_longlong, i]
        L.litho_fno_set_params.argtypes = [vp, _f32p, C.c_longlong]
        L.litho_fno_forward.argtypes = [vp, vp, i, vp]
        L.litho_fno_forward_fp32.argtypes = [vp, vp, i, vp]
        L.litho_fno_verify.argtypes = [vp, vp, i, vp, i, f, _f32p]
        L.litho_fno_loss_grad.argtypes = [vp, vp, vp, i, C.POINTER(f)]
        L.litho_fno_adam.argtypes = [vp, f, f, f, f]
        bp = C.POINTER(LithoBucket)
        L.litho_socs_image_buckets.argtypes = [vp, bp, i, i, i]
        L.litho_bucket_times.argtypes = [vp, _f32p, i]
        L.litho_autotune.argtypes = [vp, bp, i, i, i]
        L.litho_plan_get.argtypes = [vp, i, i, i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        L.litho_plan_set.argtypes = [vp, i, i, i, i, i, i, i]
        L.litho_tcc_build_prec.argtypes = [C.POINTER(LithoOptics), _f32p, i, i, _f64p, _f64p]
        L.litho_kernels_from_source_prec.argtypes = [C.POINTER(LithoOptics), _f32p, i, i, i,
                                                     _f64p, _f32p, _f32p, C.POINTER(i)]
        L.litho_tmpimg_grid.argtypes = [i]
        L.litho_setup_timing.argtypes = [_f64p, i]
        L.litho_socs_tmpimg.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                        i, i, i, i, i]
        L.litho_socs_tmpimg_on_grid.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                                _f32p, i, i, i, i, i, i, i]
        L.litho_socs_tmpimg_dev.argtypes = [vp, vp, vp, vp, i, i, i, i, vp, vp]
        L.litho_socs_image.argtypes = [_f32p, i, _f32p, _f32p, _f32p, i, i, i, _f32p]
        L.litho_fi.argtypes = [_f32p, i, i, i, _f32p]
        L.litho_image_tcc.argtypes = [_f32p, i, _f64p, _f64p, i, i, _f32p]
        L.litho_image_tcc_rect.argtypes = [_f32p, i, i, _f64p, _f64p, i, i, _f32p]
        L.litho_bank_create.argtypes = [vp, _f32p, _f32p, _f32p, i, i, i, C.POINTER(vp)]
        L.litho_bank_destroy.argtypes = [vp]
        L.litho_bank_info.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
        L.litho_socs_image_batch.argtypes = [vp, vp, vp, i, i, vp]
        L.litho_socs_image_batch_noalloc.argtypes = [vp, vp, vp, i, i, vp]
        L.litho_last_timing.argtypes = [vp, C.POINTER(f), C.POINTER(f), i]
        L.litho_ctx_set_path.argtypes = [vp, i]
        L.litho_ctx_set_timing.argtypes = [vp, i]
        L.litho_socs_image_batch_pipelined.argtypes = [vp, vp, vp, i, i, vp, i, i, i]
        L.litho_stage_name.argtypes = [vp, i]
        L.litho_stage_name.restype = C.c_char_p
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("litho_status_string", "litho_outer_sigma", "litho_stage_name"):
                fn.restype = C.c_int
        _LIB = L
    return _LIB


def _check(status, what):
    if status != 0:
        raise LithoError(status, what)


def _planes(z):
    z = np.asarray(z)
    return (np.ascontiguousarray(z.real, dtype=np.float32).ravel(),
            np.ascontiguousarray(z.imag, dtype=np.float32).ravel())


# ---- Module 1 ---------------------------------------------------------------
def createSource(size, kind="Annular", inner=0.6, outer=0.9, radius=0.5, offset=0.3,
                 on_x=True, px=0.0, py=0.0):
    """createSource (common/source.cpp L105-141) incl. normalizeMatrix."""
    s = np.zeros(size * size, np.float32)
    L = lib()
    if kind == "Annular":
        rc = L.litho_source_annular(s, size, inner, outer)
    elif kind == "CrossQuadrupole":
        rc = L.litho_source_quadrupole(s, size, radius, offset)
    elif kind == "Dipole":
        rc = L.litho_source_dipole(s, size, radius, offset, int(on_x))
    elif kind == "Point":
        rc = L.litho_source_point(s, size, px, py)
    else:
        raise ValueError(kind)
    _check(rc, "createSource")
    return s.reshape(size, size)


def getOuterSigma(src):
    """getOuterSigma (L427-439), clamped to 1 as main does (L1344)."""
    src = np.ascontiguousarray(src, np.float32)
    return float(lib().litho_outer_sigma(src.ravel(), src.shape[0]))


def compute_nx(NA, period, outer_sigma, wavelength):
    """Nx = floor(NA*period*(1+sigma)/lambda) in float arithmetic (L1346)."""
    return int(lib().litho_compute_nx(NA, int(period), outer_sigma, wavelength))


def manhattan_mask(N, seed, wmin=12, wmax=45, fill=0.40):
    """Seeded synthetic Manhattan mask (PCG32), float32 0/1, N x N."""
    m = np.zeros(N * N, np.float32)
    _check(lib().litho_manhattan_mask(m, N, seed, wmin, wmax, fill), "litho_manhattan_mask")
    return m.reshape(N, N)


def optics(wavelength, NA, defocus, outer_sigma, Lx, Ly, Nx, Ny):
    return LithoOptics(wavelength, NA, defocus, outer_sigma, int(Lx), int(Ly), int(Nx), int(Ny))


# ---- Module 3 ---------------------------------------------------------------
def FT_r2c_padded(mask):
    """Full centred mask spectrum / (Lx Ly) (L255-331), complex64 Ly x Lx."""
    mask = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = mask.shape
    r = np.zeros(Lx * Ly, np.float32)
    i = np.zeros(Lx * Ly, np.float32)
    if Lx == Ly:
        _check(lib().litho_mask_spectrum(mask.ravel(), Lx, r, i), "litho_mask_spectrum")
    else:
        _check(lib().litho_mask_spectrum_rect(mask.ravel(), Lx, Ly, r, i),
               "litho_mask_spectrum_rect")
    return (r + 1j * i).astype(np.complex64).reshape(Ly, Lx)


# ---- Module 4 ---------------------------------------------------------------
def calcTCC(src, outerSigma, wavelength, NA, defocus, Lx, Ly, Nx, Ny, tcc_prec=LITHO_TCC_FP64):
    """TCC = A diag(S) A^H (L441-532), complex128 n x n, built on the GPU
    (FP64 MFMA, or split BF16 with tcc_prec=LITHO_TCC_SPLIT_BF16)."""
    src = np.ascontiguousarray(src, np.float32)
    n = (2 * Nx + 1) * (2 * Ny + 1)
    re = np.zeros(n * n)
    im = np.zeros(n * n)
    op = optics(wavelength, NA, defocus, outerSigma, Lx, Ly, Nx, Ny)
    _check(lib().litho_tcc_build_prec(C.byref(op), src.ravel(), src.shape[0], tcc_prec, re, im),
           "litho_tcc_build_prec")
    return (re + 1j * im).reshape(n, n)


# ---- Module 6 ---------------------------------------------------------------
def calcKernels(tcc, nk, Nx, Ny):
    """Top-nk kernels (L597-663): (scales float64 desc, kernels complex64 nk x n, flags)."""
    tcc = np.asarray(tcc, np.complex128)
    n = tcc.shape[0]
    sc = np.zeros(nk)
    kr = np.zeros(nk * n, np.float32)
    ki = np.zeros(nk * n, np.float32)
    fl = C.c_int(0)
    _check(lib().litho_kernels_topk(np.ascontiguousarray(tcc.real).ravel(),
                                    np.ascontiguousarray(tcc.imag).ravel(), n, Nx, Ny, nk, sc,
                                    kr, ki, C.byref(fl)), "litho_kernels_topk")
    return sc, (kr + 1j * ki).astype(np.complex64).reshape(nk, n), fl.value


def kernels_from_source(src, outerSigma, wavelength, NA, defocus, Lx, Ly, Nx, Ny, nk,
                        tcc_prec=LITHO_TCC_FP64):
    """Source -> top-nk kernels on the GPU (explicit or Gram route); the
    contraction in FP64 or split BF16 (tcc_prec), the eigensolve in FP64."""
    src = np.ascontiguousarray(src, np.float32)
    n = (2 * Nx + 1) * (2 * Ny + 1)
    sc = np.zeros(nk)
    kr = np.zeros(nk * n, np.float32)
    ki = np.zeros(nk * n, np.float32)
    fl = C.c_int(0)
    op = optics(wavelength, NA, defocus, outerSigma, Lx, Ly, Nx, Ny)
    _check(lib().litho_kernels_from_source_prec(C.byref(op), src.ravel(), src.shape[0], nk,
                                                tcc_prec, sc, kr, ki, C.byref(fl)),
           "litho_kernels_from_source_prec")
    return sc, (kr + 1j * ki).astype(np.complex64).reshape(nk, n), fl.value


def setup_timing():
    """Breakdown of this thread's last Module 4 / 6 call: pupil, contraction
    and eigensolve (ms), the eigensolver (jacobi / subspace) and its
    iterations."""
    t = np.zeros(5)
    _check(min(lib().litho_setup_timing(t, 5), 0), "litho_setup_timing")
    return {"pupil_ms": float(t[0]), "contraction_ms": float(t[1]), "eigensolve_ms": float(t[2]),
            "solver": "subspace" if t[3] == 1 else "jacobi", "iterations": int(t[4])}


# ---- Module 7 ---------------------------------------------------------------
def tmpimg_grid(n):
    """Conv grid of the accelerator-shaped entry along an axis of window
    half-width n: v18's fixed 128 (v18 .h L34-35) up to n = 8, at least
    nextPow2(4n+1) above (litho_thread.cpp L833-834)."""
    g = lib().litho_tmpimg_grid(int(n))
    _check(min(g, 0), "litho_tmpimg_grid")
    return g


def conv_grid(n):
    """calcSOCS's conv grid nextPow2(4n+1) (litho_thread.cpp L833-834)."""
    g = 1
    while g < 4 * n + 1:
        g *= 2
    return g


def calc_socs_hls(mskf, krns, scales, nx, ny, grid=None):
    """HLS-v18-shaped SOCS (v18 .cpp L287-388): the fftshifted tmpImgp,
    Gy x Gx = tmpimg_grid(ny) x tmpimg_grid(nx) (128 x 128 for nx, ny <= 8,
    as v18), or on grid = (Gy, Gx) when given."""
    mskf = np.asarray(mskf)
    Ly, Lx = mskf.shape
    mr, mi = _planes(mskf)
    kr, ki = _planes(krns)
    sc = np.ascontiguousarray(scales, np.float32)
    nk = sc.size
    if grid is None:
        Gy, Gx = tmpimg_grid(ny), tmpimg_grid(nx)
        out = np.zeros(Gx * Gy, np.float32)
        _check(lib().litho_socs_tmpimg(mr, mi, kr, ki, sc, out, out, nk, nx, ny, Lx, Ly),
               "litho_socs_tmpimg")
    else:
        Gy, Gx = grid
        out = np.zeros(Gx * Gy, np.float32)
        _check(lib().litho_socs_tmpimg_on_grid(mr, mi, kr, ki, sc, out, out, nk, nx, ny, Lx, Ly,
                                               Gx, Gy), "litho_socs_tmpimg_on_grid")
    return out.reshape(Gy, Gx)


def FI(tmpImgp, N, band):
    """Fourier interpolation (L692-772) of a centred G x G grid to N x N."""
    t = np.ascontiguousarray(tmpImgp, np.float32)
    G = t.shape[0]
    out = np.zeros(N * N, np.float32)
    _check(lib().litho_fi(t.ravel(), G, N, band, out), "litho_fi")
    return out.reshape(N, N)


def calcSOCS(mskf, krns, scales, Nx, Ny):
    """calcSOCS (L813-1040) from a full centred spectrum: SOCS loop on the
    nextPow2(4Nx+1) grid (L833) + FI."""
    mskf = np.asarray(mskf)
    G = conv_grid(Nx)
    t = calc_socs_hls(mskf, krns, scales, Nx, Ny, grid=(G, G))
    return FI(t, mskf.shape[0], 2 * Nx)


def calcImage(mask, tcc, Nx, Ny):
    """TCC-direct image (calcImage + FT_c2r_padded, L539-583 / L371-425) on the GPU;
    mask [Ly][Lx] with power-of-two sides."""
    mask = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = mask.shape
    tcc = np.asarray(tcc, np.complex128)
    tr, ti = np.ascontiguousarray(tcc.real).ravel(), np.ascontiguousarray(tcc.imag).ravel()
    out = np.zeros(Lx * Ly, np.float32)
    if Lx == Ly:
        _check(lib().litho_image_tcc(mask.ravel(), Lx, tr, ti, Nx, Ny, out), "litho_image_tcc")
    else:
        _check(lib().litho_image_tcc_rect(mask.ravel(), Lx, Ly, tr, ti, Nx, Ny, out),
               "litho_image_tcc_rect")
    return out.reshape(Ly, Lx)


def socs_image(mask, krns, scales, Nx, Ny):
    """Mask [Ly][Lx] -> aerial image on the GPU (Modules 3 + 7 fused pipeline)."""
    mask = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = mask.shape
    kr, ki = _planes(krns)
    sc = np.ascontiguousarray(scales, np.float32)
    out = np.zeros(Lx * Ly, np.float32)
    if Lx == Ly:
        _check(lib().litho_socs_image(mask.ravel(), Lx, kr, ki, sc, sc.size, Nx, Ny, out),
               "litho_socs_image")
    else:
        _check(lib().litho_socs_image_rect(mask.ravel(), Lx, Ly, kr, ki, sc, sc.size, Nx, Ny,
                                           out), "litho_socs_image_rect")
    return out.reshape(Ly, Lx)


# ---- device-resident batch path ----------------------------------------------
class Bank:
    """One kernel set in HBM (litho_bank).  Passed to the C-ABI as its handle;
    close() (also called by the owning Context's close and on garbage
    collection) frees the device memory once."""

    def __init__(self, handle):
        self.h = handle

    @property
    def _as_parameter_(self):
        if not self.h:
            raise LithoError(-1, "bank used after close()")
        return self.h

    @property
    def value(self):
        return self._as_parameter_.value

    def close(self):
        if self.h:
            lib().litho_bank_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP stream + workspaces (litho_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._banks = []
        _check(lib().litho_ctx_create(C.byref(self.h), device), "litho_ctx_create")

    def alloc(self, nbytes):
        p = C.c_void_p()
        _check(lib().litho_dev_alloc(self.h, C.byref(p), nbytes), "litho_dev_alloc")
        return p

    def free(self, p):
        _check(lib().litho_dev_free(self.h, p), "litho_dev_free")

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().litho_memcpy_h2d(self.h, dst, arr.ctypes.data, arr.nbytes), "h2d")

    def d2h(self, arr, src):
        _check(lib().litho_memcpy_d2h(self.h, arr.ctypes.data, src, arr.nbytes), "d2h")
        return arr

    def d2d(self, dst, src, nbytes):
        _check(lib().litho_memcpy_d2d(self.h, dst, src, nbytes), "d2d")

    def sync(self):
        _check(lib().litho_ctx_sync(self.h), "litho_ctx_sync")

    def bank(self, krns, scales, Nx, Ny):
        kr, ki = _planes(krns)
        sc = np.ascontiguousarray(scales, np.float32)
        b = C.c_void_p()
        _check(lib().litho_bank_create(self.h, kr, ki, sc, sc.size, Nx, Ny, C.byref(b)),
               "litho_bank_create")
        bank = Bank(b)
        self._banks.append(bank)
        return bank

    @staticmethod
    def bank_info(bank):
        """(entries, pairs) of the bank's mask -> image form: kernels whose
        fields are real on a real mask share one transform two at a time."""
        e, p = C.c_int(0), C.c_int(0)
        _check(lib().litho_bank_info(bank, C.byref(e), C.byref(p)), "litho_bank_info")
        return e.value, p.value

    def socs_image_batch(self, bank, d_mask, B, N, d_image):
        _check(lib().litho_socs_image_batch(self.h, bank, d_mask, B, N, d_image),
               "litho_socs_image_batch")

    def socs_image_batch_noalloc(self, bank, d_mask, B, N, d_image):
        """socs_image_batch with its workspaces in the library's fixed
        per-GPU scratch (LITHO_NOALLOC_SCRATCH_BYTES): allocates nothing."""
        _check(lib().litho_socs_image_batch_noalloc(self.h, bank, d_mask, B, N, d_image),
               "litho_socs_image_batch_noalloc")

    def socs_image_batch_pipelined(self, bank, d_mask, B, N, d_image, precision=5, chunks=1,
                                   graph=True):
        """Batched form (flags LITHO_PREC_*): d_mask holds f32 masks, or uint8
        levels with LITHO_PREC_MASK_U8; d_image receives f32, or fp16 with
        LITHO_PREC_IMAGE_FP16."""
        _check(lib().litho_socs_image_batch_pipelined(self.h, bank, d_mask, B, N, d_image,
                                                      precision, chunks, int(graph)),
               "litho_socs_image_batch_pipelined")

    @staticmethod
    def _buckets(buckets):
        """[(bank, N, count, d_masks, d_images), ...] -> ctypes array."""
        arr = (LithoBucket * len(buckets))()
        for k, (bank, N, count, dm, di) in enumerate(buckets):
            arr[k] = LithoBucket(getattr(bank, "value", bank), int(N), int(count),
                                 getattr(dm, "value", dm),
                                 getattr(di, "value", di))
        return arr

    def socs_image_buckets(self, buckets, precision=0, graph=True):
        """Mixed clip sizes in one step: buckets = [(bank, N, count, d_masks,
        d_images), ...]; precision flags as socs_image_batch_pipelined."""
        arr = self._buckets(buckets)
        _check(lib().litho_socs_image_buckets(self.h, arr, len(buckets), precision, int(graph)),
               "litho_socs_image_buckets")

    def bucket_times(self, nb):
        """Per-bucket device end times (ms) of the last eager bucket step."""
        t = np.zeros(nb, np.float32)
        n = lib().litho_bucket_times(self.h, t, nb)
        _check(min(n, 0), "litho_bucket_times")
        return t[:n]

    def autotune(self, buckets, precision=0, reps=3):
        n = lib().litho_autotune(self.h, self._buckets(buckets), len(buckets), precision, reps)
        _check(min(n, 0), "litho_autotune")
        return n

    def set_plan(self, N, Nx, nk, precision, G, socs_shape, lean_fft):
        _check(lib().litho_plan_set(self.h, N, Nx, nk, precision, G, socs_shape, int(lean_fft)),
               "litho_plan_set")

    def plan(self, N, Nx, nk, precision=0):
        """(conv grid, SOCS block shape 0/1/2/3, lean FFT block, tuned?) for a shape."""
        g, ls, lm = C.c_int(), C.c_int(), C.c_int()
        t = lib().litho_plan_get(self.h, N, Nx, nk, precision, C.byref(g), C.byref(ls),
                                 C.byref(lm))
        _check(min(t, 0), "litho_plan_get")
        return g.value, int(ls.value), bool(lm.value), bool(t)

    def set_path(self, path):
        _check(lib().litho_ctx_set_path(self.h, path), "litho_ctx_set_path")

    def set_timing(self, on):
        """Per-stage HIP events around batch launches (on by default)."""
        _check(lib().litho_ctx_set_timing(self.h, 1 if on else 0), "litho_ctx_set_timing")

    def last_timing(self, nstage=8):
        """(total ms, {stage name: ms}) of the last batch call."""
        tot = C.c_float()
        st = (C.c_float * nstage)()
        n = lib().litho_last_timing(self.h, C.byref(tot), st, nstage)
        if n < 0:
            _check(n, "litho_last_timing")
        names = [lib().litho_stage_name(self.h, i).decode() for i in range(n)]
        return tot.value, dict(zip(names, list(st)[:n]))

    def close(self):
        for b in self._banks:
            b.close()
        self._banks = []
        if self.h:
            lib().litho_ctx_destroy(self.h)
            self.h = None
