"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, validates arguments before touching the device, and its host
helpers (source sigma, Nx, synthetic masks) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "litho_socs.h")).read()
    return sorted(set(re.findall(r"\b(litho_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gpu):
    L = gpu.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(gpu.EXPORTS) == syms


def test_status_strings(gpu):
    L = gpu.lib()
    assert L.litho_status_string(0) == b"ok"
    for code in (-1, -2, -4, -5, -6, -7):
        assert L.litho_status_string(code) != b"unknown status"
    assert L.litho_version() >= 100


def test_arguments_rejected_before_device_work(gpu):
    L = gpu.lib()
    m = np.zeros(100 * 100, np.float32)
    r = np.zeros_like(m)
    assert L.litho_mask_spectrum(m, 100, r, r) == -1  # not a power of two
    t = np.zeros(9 * 9)
    k = np.zeros(9 * 9, np.float32)
    sc = np.zeros(100)
    assert L.litho_kernels_topk(t, t, 81, 4, 4, 100, sc, k, k, None) == -1  # nk > n
    assert L.litho_kernels_topk(t, t, 80, 4, 4, 5, sc, k, k, None) == -1  # n != W^2
    assert L.litho_fi(m, 48, 100, 3, m) == -1
    tn = np.eye(9).ravel()
    tn[10] = np.nan  # a non-finite TCC is refused before the eigensolver
    assert L.litho_kernels_topk(tn, np.zeros(81), 9, 1, 1, 2, sc, k, k, None) == -1
    assert L.litho_socs_image_batch(None, None, None, 1, 512, None) == -1
    # TCC-direct image: sides must be powers of two in 32..4096 holding the band
    img = np.zeros(64 * 96, np.float32)
    te = np.eye(45).ravel()
    z = np.zeros(45 * 45)
    assert L.litho_image_tcc_rect(img, 96, 64, te, z, 4, 2, img) == -1  # 96 not a power of two
    assert L.litho_image_tcc_rect(img, 64, 16, te, z, 4, 2, img) == -1  # below 32
    assert L.litho_image_tcc_rect(img, 64, 32, te, z, 16, 2, img) == -1  # 4Nx+1 > Lx
    te[3] = np.inf
    assert L.litho_image_tcc_rect(img, 64, 32, te, z, 4, 2, img) == -1  # non-finite TCC
    assert L.litho_image_tcc(img, 64, te, z, 4, 2, img) == -1


def test_tmpimg_grid_and_arguments(gpu):
    """v18's fixed 128 grid up to n = 8 (v18 .h L34-35, L43), at least
    nextPow2(4n+1) above (litho_thread.cpp L833-834); grids that cannot hold
    the window are rejected before any device work."""
    L = gpu.lib()
    assert [L.litho_tmpimg_grid(n) for n in (0, 8, 17, 31, 32, 34, 137, 255)] == \
        [128, 128, 128, 128, 256, 256, 1024, 1024]
    assert L.litho_tmpimg_grid(-1) == -1 and L.litho_tmpimg_grid(256) == -1
    x = np.zeros(17 * 17 * 4, np.float32)
    sc = np.ones(1, np.float32)
    args = (x, x, x, x, sc, x, x, 1, 8, 5, 32, 32)
    assert L.litho_socs_tmpimg_on_grid(*args, 16, 128) == -1    # 16 < 2*8+1
    assert L.litho_socs_tmpimg_on_grid(*args, 128, 8) == -1     # 8 < 2*5+1
    assert L.litho_socs_tmpimg_on_grid(*args, 96, 128) == -1    # not a power of two
    assert L.litho_socs_tmpimg_on_grid(*args, 2048, 128) == -1  # beyond 1024
    assert L.litho_socs_tmpimg(x, x, x, x, sc, x, x, 1, 300, 5, 32, 32) == -1
    assert L.litho_socs_tmpimg_dev(None, None, None, None, 32, 32, 128, 128, None, None) == -1
    assert L.litho_socs_image_batch_noalloc(None, None, None, 1, 512, None) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_context_reports_missing_gpu(gpu):
    h = C.c_void_p()
    assert gpu.lib().litho_ctx_create(C.byref(h), 0) == -5


def test_host_helpers_match_oracle(gpu, orc):
    for size, a, b in ((41, 0.6, 0.9), (101, 0.6, 0.9), (41, 0.3, 0.5)):
        src = orc.source_annular(size, a, b)
        assert gpu.getOuterSigma(src) == orc.outer_sigma(src)
        s = orc.outer_sigma(src)
        for NA, P in ((0.85, 2048), (0.85, 4096), (0.8, 512), (0.85, 16384)):
            assert gpu.compute_nx(NA, P, s, 193.0) == orc.compute_nx(NA, P, s, 193.0)


def test_source_generators_match_oracle(gpu, orc):
    assert np.array_equal(gpu.createSource(41), orc.source_annular(41, 0.6, 0.9))
    assert np.array_equal(gpu.createSource(101), orc.source_annular(101, 0.6, 0.9))
    assert np.array_equal(gpu.createSource(41, "Dipole", radius=0.2, offset=0.5),
                          orc.source_dipole(41, 0.2, 0.5, True))
    assert np.array_equal(gpu.createSource(41, "CrossQuadrupole", radius=0.2, offset=0.5),
                          orc.source_quadrupole(41, 0.2, 0.5))
    assert np.array_equal(gpu.createSource(41, "Point", px=0.3, py=-0.2),
                          orc.source_point(41, 0.3, -0.2))


def test_manhattan_mask_is_seeded_and_filled(gpu):
    a = gpu.manhattan_mask(512, 7)
    b = gpu.manhattan_mask(512, 7)
    c = gpu.manhattan_mask(512, 8)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert set(np.unique(a)) == {0.0, 1.0}
    assert 0.40 <= a.mean() < 0.45
    with pytest.raises(gpu.LithoError):
        gpu.manhattan_mask(64, 1, wmin=0)
