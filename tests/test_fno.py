"""Learned surrogate (FNO): the oracle's spectral convolution and gradients
(CPU), the GPU forward / backward kernels against the oracle, and a short
GPU training run on SOCS pairs that must beat the untrained network."""
import os
import sys

import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA, ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import fno_ref as F  # noqa: E402


def test_spectral_conv_matches_rfft2():
    """The explicit truncated DFTs used by the model equal the usual FNO
    spectral convolution written with numpy's rfft2 / irfft2."""
    rng = np.random.default_rng(1)
    N, B, Ci, Co = 32, 2, 3, 4
    x = rng.standard_normal((B, Ci, N, N))
    R = rng.standard_normal((Ci, Co, F.MY, F.MX)) + 1j * rng.standard_normal((Ci, Co, F.MY, F.MX))
    Fx, Fy, Hx = F.mats(N)
    X2 = np.einsum("qy,bcyk->bcqk", Fy, x @ Fx)
    Y2 = np.einsum("biqk,ioqk->boqk", X2, R)
    z = (np.einsum("qy,boqk->boyk", Fy.conj(), Y2) @ Hx).real
    assert np.allclose(z, F.spectral_conv_fft(x, R), atol=1e-12)


def test_oracle_gradient_against_finite_differences():
    rng = np.random.default_rng(2)
    N, B = 32, 2
    th = F.init_params(3)
    m = (rng.random((B, N, N)) < 0.4).astype(np.float64)
    t = rng.random((B, N, N)) * 0.5
    loss, g = F.loss_grad(th, m, t)
    eps = 1e-6
    for k in range(3):  # directional derivatives
        v = rng.standard_normal(th.size)
        v /= np.linalg.norm(v)
        fd = (F.loss_grad(th + eps * v, m, t)[0] - F.loss_grad(th - eps * v, m, t)[0]) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-6 * max(1e-12, abs(fd)) + 1e-12, (k, fd, g @ v)
    for name in ("pw", "rre0", "rim2", "w3", "b1", "q1", "q2b"):  # single coordinates
        o, n = F.LAYOUT[name]
        i = o + int(rng.integers(n))
        e = np.zeros_like(th)
        e[i] = 1e-5
        fd = (F.loss_grad(th + e, m, t)[0] - F.loss_grad(th - e, m, t)[0]) / 2e-5
        # 1e-11: the loss round-off over the step (~eps |L| / h)
        assert abs(fd - g[i]) <= 1e-5 * abs(fd) + 1e-11, (name, fd, g[i])


@pytest.mark.gpu
@pytest.mark.parametrize("N", [32, 64, 256])
def test_gpu_fno_matches_oracle(gpu, N):
    """Forward output, loss and every gradient tensor of the GPU kernels
    against the FP64 oracle at the same (bf16-rounded) weights, up to the
    surrogate's 256^2 clip; a second backward pass gives the same gradient
    bit for bit (the reductions add their partial sums in a fixed order)."""
    from litho_amd.fno import FNO
    print(f"[case] fno forward/backward N={N}", flush=True)
    rng = np.random.default_rng(N)
    B = 2
    th = F.round_bf16(F.init_params(N))
    m = (rng.random((B, N, N)) < 0.4).astype(np.float32)
    t = (rng.random((B, N, N)) * 0.5).astype(np.float32)
    ctx = gpu.Context()
    try:
        model = FNO(ctx, N, seed=1)
        assert model.nparam == F.NPARAM
        model.set_params(th.astype(np.float32))
        dm, dt, do = ctx.alloc(m.nbytes), ctx.alloc(t.nbytes), ctx.alloc(t.nbytes)
        ctx.h2d(dm, m)
        ctx.h2d(dt, t)
        model.forward_fp32(dm, B, do)
        out = ctx.d2h(np.zeros_like(t), do)
        ref_out = F.forward(th, m.astype(np.float64))
        assert np.abs(out - ref_out).max() <= 1e-5 * np.abs(ref_out).max()
        loss = model.loss_grad(dm, dt, B)
        g = model.params(grads=True).astype(np.float64)
        ref_loss, ref_g = F.loss_grad(th, m.astype(np.float64), t.astype(np.float64))
        assert abs(loss - ref_loss) <= 1e-5 * ref_loss
        worst = 0.0
        for name, (o, n) in F.LAYOUT.items():
            a, r = g[o:o + n], ref_g[o:o + n]
            e = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
            worst = max(worst, e)
            assert e < 1e-4, (name, e)
        print(f"[fno N={N}] loss {loss:.6e} vs {ref_loss:.6e}; worst gradient rel err {worst:.2e}",
              flush=True)
        loss2 = model.loss_grad(dm, dt, B)
        assert loss2 == loss
        assert np.array_equal(model.params(grads=True).astype(np.float64), g)
        for p in (dm, dt, do):
            ctx.free(p)
        model.close()
    finally:
        ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("N", [32, 64, 128, 256])
def test_gpu_fno_matrix_core_inference(gpu, N):
    """litho_fno_forward (fused matrix-core layer kernels, bf16 hi + lo
    operands, FP32 accumulate) against the FP64 oracle and the FP32
    vector-ALU forward at the same bf16-rounded weights, within 1e-4 of the
    output's range; binary, grey-level and empty masks; a second call with a
    smaller batch reuses the buffers of the first."""
    from litho_amd.fno import FNO
    print(f"[case] fno matrix-core inference N={N}", flush=True)
    rng = np.random.default_rng(100 + N)
    B = 3
    th = F.round_bf16(F.init_params(N + 7))
    m = np.stack([(rng.random((N, N)) < 0.4), rng.random((N, N)), np.zeros((N, N))])
    m = m.astype(np.float32)
    ctx = gpu.Context()
    try:
        model = FNO(ctx, N, seed=2)
        model.set_params(th.astype(np.float32))
        dm, do = ctx.alloc(m.nbytes), ctx.alloc(m.nbytes)
        ctx.h2d(dm, m)
        model.forward(dm, B, do)
        out = ctx.d2h(np.zeros_like(m), do)
        model.forward_fp32(dm, B, do)
        f32 = ctx.d2h(np.zeros_like(m), do)
        ref = F.forward(th, m.astype(np.float64))
        tol = 1e-4 * np.abs(ref).max()
        e_ref, e_f32 = np.abs(out - ref).max(), np.abs(out - f32).max()
        print(f"[fno mfma N={N}] max abs err vs oracle {e_ref:.3e}, vs FP32 kernels {e_f32:.3e}, "
              f"range {np.abs(ref).max():.3e}", flush=True)
        assert e_ref <= tol and e_f32 <= tol
        model.forward(dm, 1, do)
        one = ctx.d2h(np.zeros((1, N, N), np.float32), do)
        assert np.array_equal(one[0], out[0])
        ctx.free(dm)
        ctx.free(do)
        model.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_fno_verify_judge(gpu):
    """litho_fno_verify: per-clip max relative error against SOCS images on the
    device (numpy's on the host), the clips above tol (or NaN) overwritten
    with their SOCS images, every other clip untouched; a clip stride of 2
    checks clips 0, 2, 4 of the served batch."""
    import ctypes as C
    from litho_amd.fno import FNO
    print("[case] fno verify judge", flush=True)
    N, n = 32, 6
    rng = np.random.default_rng(7)
    ref = rng.random((n, N, N)).astype(np.float32) + 0.5
    img = ref.copy()
    scale = [0.0, 5e-3, 2e-2, 9.9e-3, 0.5, 0.0]
    for j, e in enumerate(scale):
        img[j, 3, 5] += e * ref[j].max()
    img[5, 7, 7] = np.nan
    ref[0] = 0.0
    img[0] = 0.0  # both all zero: error 0
    ctx = gpu.Context()
    try:
        model = FNO(ctx, N, seed=1)
        di, dr = ctx.alloc(img.nbytes), ctx.alloc(ref.nbytes)
        ctx.h2d(di, img)
        ctx.h2d(dr, ref)
        err = np.zeros(n, np.float32)
        fell = gpu.lib().litho_fno_verify(model.h, di, 1, dr, n, 1e-2, err)
        want = np.abs(img - ref).max(axis=(1, 2)) / np.where(ref.max(axis=(1, 2)) > 0,
                                                             np.abs(ref).max(axis=(1, 2)), 1)
        want[5] = np.inf
        assert fell == 3
        assert np.allclose(err[:5], want[:5], rtol=1e-6, atol=0) and np.isinf(err[5])
        out = ctx.d2h(np.zeros_like(img), di)
        for j in range(n):
            expect = ref[j] if j in (2, 4, 5) else img[j]
            assert np.array_equal(out[j], expect, equal_nan=True), j
        # stride 2: clips 0, 2, 4 of the batch against three SOCS images
        ctx.h2d(di, img)
        ref3 = np.ascontiguousarray(ref[[0, 2, 4]])
        ctx.h2d(dr, ref3)
        err3 = np.zeros(3, np.float32)
        assert gpu.lib().litho_fno_verify(model.h, di, 2, dr, 3, 1e-2, err3) == 2
        out = ctx.d2h(np.zeros_like(img), di)
        assert np.array_equal(out[2], ref[2]) and np.array_equal(out[4], ref[4])
        assert np.array_equal(out[1], img[1]) and np.array_equal(out[3], img[3])
        assert gpu.lib().litho_fno_verify(model.h, di, 0, dr, 3, 1e-2, err3) == -1
        ctx.free(di)
        ctx.free(dr)
        model.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_training_beats_untrained(gpu, orc, base_source):
    """Trained on SOCS pairs by its own kernels, the FNO's held-out error
    falls far below the untrained network's; the server verifies 1 in 16
    clips with SOCS and falls back above the tolerance."""
    from litho_amd import fno
    print("[case] fno training N=256", flush=True)
    src, s = base_source
    N = 256
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx,
                                       min(16, (2 * Nx + 1) ** 2))
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        ntr, nte = 512, 64
        tm, ti = fno.socs_pairs(ctx, bank, N, ntr, 10_000)
        vm, vi = fno.socs_pairs(ctx, bank, N, nte, 90_000)
        model = fno.FNO(ctx, N, seed=5)
        e0 = fno.max_rel_errors(ctx, model, vm, vi, nte)
        losses = fno.train(model, tm, ti, ntr, steps=400, batch=16, lr=2e-3)
        e1 = fno.max_rel_errors(ctx, model, vm, vi, nte)
        print(f"[fno training] loss {losses[0]:.3e} -> {losses[-1]:.3e}; held-out max rel err "
              f"median {np.median(e0):.3e} -> {np.median(e1):.3e}", flush=True)
        assert losses[-1] < losses[0] / 20
        assert np.median(e1) < np.median(e0) / 10
        d_out = ctx.alloc(nte * N * N * 4)
        ref = ctx.d2h(np.zeros((nte, N, N), np.float32), vi)
        # the serving contract, 1e-2 max rel err: every clip verified (default)
        srv = fno.SurrogateServer(ctx, model, bank, tol=1e-2)
        srv.serve(vm, nte, d_out)
        served = ctx.d2h(np.zeros((nte, N, N), np.float32), d_out)
        err = np.abs(served - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
        assert srv.verified == nte and srv.unverified == 0
        assert err.max() <= 1e-2 * (1 + 1e-6)
        assert srv.fallbacks == int((e1 > 1e-2).sum())
        # the sampled mode (1 in 16): verified clips meet 1e-2, the rest are
        # counted as unverified
        srv16 = fno.SurrogateServer(ctx, model, bank, tol=1e-2, verify_every=16)
        srv16.serve(vm, nte, d_out)
        assert srv16.verified == nte // 16 and srv16.unverified == nte - nte // 16
        served = ctx.d2h(np.zeros((nte, N, N), np.float32), d_out)
        for i in range(0, nte, 16):
            e = np.abs(served[i] - ref[i]).max() / np.abs(ref[i]).max()
            assert e <= 1e-2 * (1 + 1e-6)
        for p in (tm, ti, vm, vi, d_out):
            ctx.free(p)
        model.close()
        bank.close()
    finally:
        ctx.close()
