"""calcKernels (litho_thread.cpp L597-663) on explicit TCCs beyond the Jacobi
range: litho_kernels_topk switches to block subspace iteration at n >= 1024.

Checked against the FP64 oracle's source-space Gram route (orc_calc_kernels_gram,
pinned to LAPACK zheevr and to the oracle's explicit route in test_oracle.py):
eigenvalues within 1e-10 of the largest, kernels (canonical cluster basis and
phase rule, f32 output) within 1e-6.  C1 is 512^2 (Nx = 17, n = 1225), C2 the
production clip 1024^2 (Nx = 34, n = 4761); both at 4 nm, NA 0.85, zero
defocus, 41^2 annulus 0.6-0.9.
"""
import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA

pytestmark = pytest.mark.gpu

EIG_TOL = 1e-10
KRN_TOL = 1e-6


def case(msg):
    print(f"[case] {msg}", flush=True)


def _oracle_gram(orc, base_source, N, nk):
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    sc, K, split = orc.calc_kernels_gram(A, S, nk, Nx, Nx)
    return Nx, sc, K, split


@pytest.fixture(scope="module")
def c2_oracle(orc, base_source):
    case("oracle Gram-route kernels at C2 (572^2 Jacobi, about 10 s)")
    return _oracle_gram(orc, base_source, 1024, 24)


def _check(sc, K, fl, so, Ko, split, gpu):
    assert np.abs(sc - so).max() / so[0] < EIG_TOL
    assert np.abs(K - Ko).max() < KRN_TOL
    assert bool(fl & gpu.LITHO_FLAG_CLUSTER_SPLIT) == split


def test_explicit_topk_c1(gpu, orc, base_source):
    case("kernels_topk on the explicit C1 TCC, n = 1225, nk = 10")
    src, s = base_source
    Nx, so, Ko, split = _oracle_gram(orc, base_source, 512, 10)
    T = orc.calc_tcc(src, s, BASE_LAMBDA, BASE_NA, 0.0, 2048, 2048, Nx, Nx)
    sc, K, fl = gpu.calcKernels(T, 10, Nx, Nx)
    t = gpu.setup_timing()
    print(f"[info] C1 explicit top-k: {t}", flush=True)
    assert t["solver"] == "subspace"
    _check(sc, K, fl, so, Ko, split, gpu)


def test_explicit_topk_c2(gpu, orc, base_source, c2_oracle):
    case("kernels_topk on the explicit C2 TCC, n = 4761, nk = 24")
    src, s = base_source
    Nx, so, Ko, split = c2_oracle
    T = gpu.calcTCC(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4096, 4096, Nx, Nx)
    tt = gpu.setup_timing()
    print(f"[info] C2 TCC build: {tt}", flush=True)
    sc, K, fl = gpu.calcKernels(T, 24, Nx, Nx)
    t = gpu.setup_timing()
    print(f"[info] C2 explicit top-k: {t}", flush=True)
    assert t["solver"] == "subspace"
    _check(sc, K, fl, so, Ko, split, gpu)


def test_gram_route_c2(gpu, orc, base_source, c2_oracle):
    case("kernels_from_source at C2 (Gram route, 572^2) vs the oracle")
    src, s = base_source
    Nx, so, Ko, split = c2_oracle
    sc, K, fl = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4096, 4096, Nx, Nx, 24)
    t = gpu.setup_timing()
    print(f"[info] C2 Gram route: {t}", flush=True)
    assert fl & gpu.LITHO_FLAG_GRAM_ROUTE
    _check(sc, K, fl, so, Ko, split, gpu)


def test_subspace_falls_back_on_an_indefinite_matrix(gpu):
    case("kernels_topk, n = 1089 with a large negative eigenvalue: Jacobi fallback")
    rng = np.random.default_rng(3)
    n, Nx = 33 * 33, 16
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    w = np.concatenate([np.geomspace(1.0, 1e-3, n - 1), [-5.0]])
    H = (Q * w) @ Q.conj().T
    sc, K, _ = gpu.calcKernels(H, 8, Nx, Nx)
    assert gpu.setup_timing()["solver"] == "jacobi"
    assert np.abs(sc - w[:8]).max() < 1e-10
