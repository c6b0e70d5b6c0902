"""Kernel pairing in the mask -> image paths.

For a real mask the window spectrum M satisfies M(-f) = conj(M(f)) about the
window centre.  A kernel with K(-f) = e^{ib} conj(K(f)) then has a field
IFFT(K M) that is real up to the phase ramp every field of the clip shares, so
two such kernels a, b share one transform: the entry sqrt(la) Ka + i sqrt(lb) Kb
with weight 1 gives |IFFT|^2 = la |Ea|^2 + lb |Eb|^2.  Zero defocus with a
point-symmetric source gives real kernels of definite parity, all of which
pair; defocus breaks the symmetry and nothing pairs.

The CPU test checks the identity on the oracle's FP64 SOCS intensity (a numpy
restatement of the pairing, not the library's code); the GPU tests check which
kernels the library pairs and that its images equal the oracle's unpaired
calcSOCS to 1e-5 (max |diff| / max |ref|).
"""
import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA, max_rel


def random_mask(N, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((N, N), np.float32)
    for _ in range(40):
        y, x = rng.integers(0, N - 40, 2)
        h, w = rng.integers(6, 40, 2)
        m[y:y + h, x:x + w] = 1.0
    return m


def conj_symmetric_part(K):
    """Phase that maximises the part with K(-f) = conj(K(f)), then that part,
    and the energy fraction left out."""
    Km = K[::-1]
    n = np.sum(np.abs(K) ** 2)
    s = np.sum(K * Km)
    a = np.exp(-0.5j * np.angle(s))
    Kp = a * K
    return 0.5 * (Kp + np.conj(Kp[::-1])), (1 - abs(s) / n) / 2


def pairing_plan(K, sc):
    """The pairing rule: a kernel with positive weight whose conjugate-symmetric
    part (after the best phase) leaves at most 1e-7 of its energy out pairs with
    the next unpaired such kernel of its class (real- or imaginary-dominant
    symmetric part).  Returns [(k, q or None)] in entry order."""
    n = len(K)
    cls = [-1] * n
    syms = [None] * n
    for k in range(n):
        if not sc[k] > 0:
            continue
        v, r = conj_symmetric_part(K[k])
        if r > 1e-7:
            continue
        syms[k] = v
        cls[k] = int(np.sum(v.imag ** 2) > np.sum(v.real ** 2))
    used = [False] * n
    plan = []
    for k in range(n):
        if used[k]:
            continue
        used[k] = True
        q = None
        if cls[k] >= 0:
            q = next((r for r in range(k + 1, n) if not used[r] and cls[r] == cls[k]), None)
        if q is not None:
            used[q] = True
        plan.append((k, q))
    return plan, syms


def q16(E):
    """fp16 bank: one scale for the real and one for the imaginary parts."""
    out = []
    for e in E:
        sx = max(np.abs(e.real).max(), 1e-30)
        sy = max(np.abs(e.imag).max(), 1e-30)
        out.append((e.real / sx).astype(np.float16).astype(np.float64) * sx +
                   1j * (e.imag / sy).astype(np.float16).astype(np.float64) * sy)
    return np.array(out)


def test_pairing_identity_on_oracle(orc, base_source):
    src, s = base_source
    N = 256
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    sc, K, _ = orc.calc_kernels_gram(A, S, 8, Nx, Nx)
    mskf = orc.mask_spectrum(random_mask(N, 5))
    ref = orc.socs_tmpimg(mskf, K, sc.astype(np.float32), Nx, Nx)  # FP64 intensity
    entries = []
    for a in range(0, 8, 2):
        Ka, ra = conj_symmetric_part(K[a])
        Kb, rb = conj_symmetric_part(K[a + 1])
        assert ra < 1e-12 and rb < 1e-12  # real kernels of definite parity
        entries.append(np.sqrt(float(np.float32(sc[a]))) * Ka +
                       1j * np.sqrt(float(np.float32(sc[a + 1]))) * Kb)
    got = orc.socs_tmpimg(mskf, np.array(entries), np.ones(4, np.float32), Nx, Nx)
    assert max_rel(got, ref) < 1e-12
    # a kernel without the symmetry does not pair: the same formula is wrong
    rng = np.random.default_rng(3)
    Kc = K[1] * np.exp(1j * rng.uniform(0, 2 * np.pi, K[1].shape))
    _, rc = conj_symmetric_part(Kc)
    assert rc > 1e-3
    ref2 = orc.socs_tmpimg(mskf, np.array([K[0], Kc]), sc[:2].astype(np.float32), Nx, Nx)
    bad = orc.socs_tmpimg(mskf, np.array([np.sqrt(float(np.float32(sc[0]))) * K[0] + 1j *
                                        np.sqrt(float(np.float32(sc[1]))) * Kc]),
                        np.ones(1, np.float32), Nx, Nx)
    assert max_rel(bad, ref2) > 1e-3


def test_class_pairing_keeps_16bit_bank_error(orc, base_source):
    """Even kernels have real symmetric parts, odd ones imaginary; pairing within
    a class keeps each kernel in its own component of the entry, so a 16-bit
    bank with one scale per component quantises each kernel as it would alone
    (pairing across classes puts both in one component and does not)."""
    src, s = base_source
    N = 256
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    A, S = orc.pupil_matrix(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx)
    sc, K, _ = orc.calc_kernels_gram(A, S, 12, Nx, Nx)
    sc32 = sc.astype(np.float32)
    rng = np.random.default_rng(29)
    mskf = orc.mask_spectrum(rng.integers(0, 256, (N, N)).astype(np.float32) / np.float32(255))
    ref = orc.socs_tmpimg(mskf, K, sc32, Nx, Nx)
    plan, syms = pairing_plan(K, sc32)
    assert all(q is not None for _, q in plan)  # 12 kernels, 6 + 6 by class
    ent = np.array([np.sqrt(float(sc32[k])) * syms[k] + 1j * np.sqrt(float(sc32[q])) * syms[q]
                    for k, q in plan])
    one = np.ones(len(ent), np.float32)
    assert max_rel(orc.socs_tmpimg(mskf, ent, one, Nx, Nx), ref) < 1e-12
    e_alone = max_rel(orc.socs_tmpimg(mskf, q16(K), sc32, Nx, Nx), ref)
    e_class = max_rel(orc.socs_tmpimg(mskf, q16(ent), one, Nx, Nx), ref)
    assert abs(e_class - e_alone) <= 1e-3 * e_alone
    cross = np.array([np.sqrt(float(sc32[a])) * conj_symmetric_part(K[a])[0] +
                      1j * np.sqrt(float(sc32[a + 1])) * conj_symmetric_part(K[a + 1])[0]
                      for a in range(0, 12, 2)])
    assert max_rel(orc.socs_tmpimg(mskf, q16(cross), one, Nx, Nx), ref) > 3 * e_alone


def gpu_images(gpu, ctx, K, sc, Nx, masks):
    B, N = masks.shape[0], masks.shape[1]
    bank = ctx.bank(K, sc, Nx, Nx)
    try:
        info = ctx.bank_info(bank)
        dm = ctx.alloc(masks.nbytes)
        di = ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.socs_image_batch(bank, dm, B, N, di)
        img = ctx.d2h(np.zeros((B, N, N), np.float32), di)
        ctx.free(dm)
        ctx.free(di)
    finally:
        bank.close()
    return info, img


def oracle_images(orc, K, sc, Nx, masks):
    return [orc.calc_socs(orc.mask_spectrum(m), K.astype(np.complex128), sc.astype(np.float32), Nx,
                          Nx, nthreads=8) for m in masks]


@pytest.mark.gpu
def test_real_kernels_pair(gpu, orc, base_source):
    src, s = base_source
    N = 512
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 12)
    masks = np.stack([random_mask(N, 10 + i) for i in range(3)])
    print(f"[case] pairing: 12 zero-defocus kernels N={N} Nx={Nx}", flush=True)
    ctx = gpu.Context()
    try:
        (entries, pairs), img = gpu_images(gpu, ctx, K, sc, Nx, masks)
    finally:
        ctx.close()
    plan, _ = pairing_plan(K, sc.astype(np.float32))
    npair = sum(q is not None for _, q in plan)
    assert npair >= 5 and (entries, pairs) == (len(plan), npair)
    for g, r in zip(img, oracle_images(orc, K, sc, Nx, masks)):
        e = max_rel(g, r)
        print(f"[pairing {pairs} pairs] max rel err vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.gpu
def test_mixed_kernels_pair_where_they_can(gpu, orc, base_source):
    """Kernel 2 made non-symmetric and kernel 5 given a negative weight (both
    keep their own entries), kernel 7 turned by a phase (still pairs)."""
    src, s = base_source
    N = 512
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx, 12)
    rng = np.random.default_rng(11)
    K = K.copy()
    sc = sc.astype(np.float32).copy()
    K[2] = K[2] * np.exp(1j * rng.uniform(0, 2 * np.pi, K[2].shape))
    sc[5] = -0.5 * sc[5]
    K[7] = K[7] * np.exp(0.7j)
    masks = np.stack([random_mask(N, 20 + i) for i in range(2)])
    print("[case] pairing: mixed eligibility", flush=True)
    ctx = gpu.Context()
    try:
        (entries, pairs), img = gpu_images(gpu, ctx, K, sc, Nx, masks)
    finally:
        ctx.close()
    plan, _ = pairing_plan(K, sc)
    assert all(q != 2 and k != 2 and q != 5 and k != 5 for k, q in plan if q is not None)
    assert any(7 in (k, q) for k, q in plan if q is not None)
    assert (entries, pairs) == (len(plan), sum(q is not None for _, q in plan))
    for g, r in zip(img, oracle_images(orc, K, sc, Nx, masks)):
        e = max_rel(g, r)
        print(f"[pairing mixed] max rel err vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.gpu
def test_defocus_kernels_do_not_pair(gpu, orc, base_source):
    src, s = base_source
    N = 256
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.2, 4 * N, 4 * N, Nx, Nx, 8)
    masks = np.stack([random_mask(N, 30 + i) for i in range(2)])
    print("[case] pairing: defocus 0.2", flush=True)
    ctx = gpu.Context()
    try:
        (entries, pairs), img = gpu_images(gpu, ctx, K, sc, Nx, masks)
    finally:
        ctx.close()
    assert (entries, pairs) == (8, 0)
    for g, r in zip(img, oracle_images(orc, K, sc, Nx, masks)):
        assert max_rel(g, r) < 1e-5
