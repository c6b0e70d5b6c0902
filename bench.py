#!/usr/bin/env python3
"""SOCS aerial-image throughput on one MI355X.

Workload (the production clip of the reference's SOCS path): 1024 x 1024
binary masks at 4 nm/pixel (period 4096 nm), seeded Manhattan rectangles with
sides of 12..45 pixels (48..180 nm) to 40 % fill; lambda 193 nm, NA 0.85, zero
defocus, 41 x 41 annular source sigma 0.6-0.9 (572 points); Nx = 34 (69 x 69
window), K = 24 kernels extracted on the GPU; the 144 x 144 band conv grid
(the reference's nextPow2 grid is 256: FI makes the image independent of it).

A step is one pass of the hot path over one batch of B masks already resident
in HBM: mask spectrum window -> SOCS accumulate -> Fourier interpolation ->
B aerial images in HBM.  `value` = images per second over the timed steps.
The CPU baseline is the project's FP64 C restatement of the reference CPU path
(the oracle), timed on a bounded sample on the host cores, and doubles as the
error check of the GPU images against FP64.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fpga-litho-accel_amd"))

import litho_amd as L  # noqa: E402

HBM_PEAK_GBPS = 8000.0
N = 1024
PIX_NM = 4
LAMBDA, NA, DEFOCUS = 193.0, 0.85, 0.0
NK = 24
WMIN, WMAX, FILL = 12, 45, 0.40


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_alg_bytes(N, Nx, G, K, B, band=True):
    """Algorithmic HBM bytes per launch of each stage: its inputs and outputs
    (intermediates of this design counted once each way); K = kernel-bank
    entries the SOCS stage reads.  The band path keeps
    kx = 0..Nx of each mask row (a real mask's spectrum is Hermitian), the
    nextPow2 path the whole window."""
    W = 2 * Nx + 1
    WR = Nx + 1 if band else W  # row-DFT columns kept
    nb = 2 * Nx + 1
    nby = 4 * Nx + 1
    f2 = 8
    return {
        "mask_rows": B * (4 * N * N + N * WR * f2),
        "mask_cols": B * (N * WR * f2 + W * W * f2),
        "socs_accum": B * (W * W * f2 + 4 * G * G) + K * W * W * f2,
        "socs_band": B * (W * W * f2 + nby * nb * f2) + K * W * W * f2,
        "fi_rows": B * (4 * G * G + G * nb * f2),
        "fi_cols": B * (nby * nb * f2 + N * nb * f2),
        "image_rows": B * (N * nb * f2 + 4 * N * N),
    }


def cpu_baseline(masks_host, K, sc, Nx, threads, budget_s, gpu_image_sets):
    """Oracle (FP64 C, OpenMP over kernels like litho_thread.cpp L897-961) on
    a bounded sample: Module 3 (mask spectrum) + Module 7 (SOCS + FI) per image."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    done, t0 = 0, time.perf_counter()
    errs = [[] for _ in gpu_image_sets]
    Kd = K.astype(np.complex128)
    scf = sc.astype(np.float32)
    # cycle over the sample clips until the time budget is spent; the first
    # pass also checks the GPU images against FP64
    while True:
        i = done % len(masks_host)
        img = O.calc_socs(O.mask_spectrum(masks_host[i]), Kd, scf, Nx, Nx, nthreads=threads)
        if done < len(masks_host):
            for e, imgs in zip(errs, gpu_image_sets):
                e.append(float(np.abs(imgs[i].astype(np.float64) - img).max() / np.abs(img).max()))
        done += 1
        if done >= len(masks_host) and time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return done / dt, done, dt, [max(e) for e in errs]


def cpu_setup_baseline(src, sigma, Nx, nk, sc_gpu, K_gpu):
    """Modules 4 + 6 once on the host: the oracle's pupil matrix (calcTCC's
    float arithmetic, litho_thread.cpp L458-484) and its source-space Gram
    route (572^2 FP64 Gram, cyclic Jacobi, lift), the fastest exact CPU form of
    calcTCC + calcKernels here; also checks the GPU kernels against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    period = PIX_NM * N
    t0 = time.perf_counter()
    A, S = O.pupil_matrix(src, sigma, LAMBDA, NA, DEFOCUS, period, period, Nx, Nx)
    t1 = time.perf_counter()
    so, Ko, _ = O.calc_kernels_gram(A, S, nk, Nx, Nx)
    t2 = time.perf_counter()
    return {"pupil_s": round(t1 - t0, 3), "gram_eigen_lift_s": round(t2 - t1, 3),
            "threads": "oracle (OpenMP where the C code has it)",
            "gpu_eigenvalue_max_rel_err": float(np.abs(sc_gpu - so).max() / so[0]),
            "gpu_kernel_max_abs_err": float(np.abs(K_gpu - Ko).max())}


def mixed_size_bench(ctx, src, sigma, steps):
    """Batching form: 512 distinct clips with sizes drawn from {256, 512, 768,
    1024}^2 at 4 nm, one K=16 kernel set per size (four TCC builds), all clips
    in one graph-captured step over four buckets, in FP32 and in the compact
    form (1-bit masks, matrix-core band DFTs, fp16 storage and images);
    images/s and per-clip latency (a clip's latency is the device time from the
    step's start to the end of its bucket's last launch, all clips submitted
    together)."""
    rng = np.random.default_rng(2024)
    sizes = rng.choice([256, 512, 768, 1024], size=512)
    out = {"clips": 512, "K": 16, "counts": {}, "plans": {}}
    t0 = time.perf_counter()
    banks, bl, bl8, dev = [], [], [], []
    for N in (256, 512, 768, 1024):
        cnt = int((sizes == N).sum())
        out["counts"][str(N)] = cnt
        Nx = L.compute_nx(NA, PIX_NM * N, sigma, LAMBDA)
        sc, K, _ = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, PIX_NM * N, PIX_NM * N,
                                         Nx, Nx, 16)
        bank = ctx.bank(K, sc, Nx, Nx)
        banks.append(bank)
        masks = np.stack([L.manhattan_mask(N, 5000 + 10000 * N + i, WMIN, WMAX, FILL)
                          for i in range(cnt)])  # every clip distinct
        bits = np.packbits(masks.astype(bool), axis=-1, bitorder="little")
        dm, dm1, di = ctx.alloc(masks.nbytes), ctx.alloc(bits.nbytes), ctx.alloc(masks.nbytes)
        ctx.h2d(dm, masks)
        ctx.h2d(dm1, bits)
        dev += [dm, dm1, di]
        bl.append((bank, N, cnt, dm, di))
        bl8.append((bank, N, cnt, dm1, di))
        del masks
    out["tcc_builds_s"] = round(time.perf_counter() - t0, 3)
    compact = (L.LITHO_PREC_COMPACT & ~L.LITHO_PREC_MASK_U8) | L.LITHO_PREC_MASK_1BIT
    for name, prec, b in (("fp32", L.LITHO_PREC_FP32, bl), ("compact", compact, bl8)):
        t0 = time.perf_counter()
        ctx.autotune(b, prec, reps=3)
        tune_s = time.perf_counter() - t0
        ctx.socs_image_buckets(b, prec, False)
        ends = ctx.bucket_times(len(b))
        lat = np.concatenate([np.full(c, e) for (_, _, c, _, _), e in zip(b, ends)])
        for _ in range(3):
            ctx.socs_image_buckets(b, prec, True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.socs_image_buckets(b, prec, True)
        ctx.sync()
        rate = 512 * steps / (time.perf_counter() - t0)
        # each bucket alone through the same entry (its own graph): the sum of
        # these is the time of the set run one size after another, each
        # filling the chip by itself
        def bucket_ms(bk):
            for _ in range(3):
                ctx.socs_image_buckets([bk], prec, True)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.socs_image_buckets([bk], prec, True)
            ctx.sync()
            return (time.perf_counter() - t0) / steps * 1e3

        alone, marginal = [], []
        for bk in b:
            alone.append(bucket_ms(bk))
            # the same clips twice in one bucket: t(2c) - t(c) is the bucket's
            # cost without its fixed launch and fill overheads
            bank, N, cnt, dm, di = bk
            mb = cnt * N * N * 4 if prec == L.LITHO_PREC_FP32 else cnt * N * N // 8
            dm2, di2 = ctx.alloc(2 * mb), ctx.alloc(2 * cnt * N * N * 4)
            ctx.d2d(dm2, dm, mb)
            ctx.d2d(ctypes.c_void_p(dm2.value + mb), dm, mb)
            marginal.append(bucket_ms((bank, N, 2 * cnt, dm2, di2)) - alone[-1])
            ctx.free(dm2)
            ctx.free(di2)
        out[name] = {"value": round(rate, 1), "unit": "images/s",
                     "p50_latency_ms": round(float(np.percentile(lat, 50)), 4),
                     "p99_latency_ms": round(float(np.percentile(lat, 99)), 4),
                     "autotune_s": round(tune_s, 3),
                     "set_ms": round(512 / rate * 1e3, 4),
                     "bucket_alone_ms": {str(bk[1]): round(t, 4) for bk, t in zip(b, alone)},
                     "buckets_in_sequence_ms": round(sum(alone), 4),
                     "bucket_marginal_ms": {str(bk[1]): round(t, 4) for bk, t in zip(b, marginal)},
                     "work_conserving_ms": round(sum(marginal), 4),
                     "packing_efficiency": round(sum(marginal) / (512 / rate * 1e3), 3)}
        out["plans"][name] = {str(N): list(ctx.plan(N, L.compute_nx(NA, PIX_NM * N, sigma, LAMBDA),
                                                     16, prec)[:3]) for N in (256, 512, 768, 1024)}
        log(f"mixed sizes ({name}): {rate:.1f} images/s, per-clip latency p50 "
            f"{out[name]['p50_latency_ms']} ms p99 {out[name]['p99_latency_ms']} ms, "
            f"autotune {tune_s:.2f}s")
    for p_ in dev:
        ctx.free(p_)
    for bk in banks:
        bk.close()
    return out


def small_clip_bench(ctx, src, sigma, steps, with_oracle):
    """Config 1: 512^2 clips at 4 nm (period 2048 nm), Nx = 17, K = 10 from the
    GPU Gram route, plain FP32 on the 81^2 band grid; images/s over batches of
    1024 clips resident in HBM, and one clip against the FP64 oracle."""
    N1, B1, nk = 512, 1024, 10
    Nx = L.compute_nx(NA, PIX_NM * N1, sigma, LAMBDA)
    sc, K, flags = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, PIX_NM * N1,
                                         PIX_NM * N1, Nx, Nx, nk)
    bank = ctx.bank(K, sc, Nx, Nx)
    masks = np.stack([L.manhattan_mask(N1, 9000 + i, WMIN, WMAX, FILL) for i in range(B1)])
    dm, di = ctx.alloc(masks.nbytes), ctx.alloc(masks.nbytes)
    ctx.h2d(dm, masks)
    for _ in range(3):
        ctx.socs_image_batch(bank, dm, B1, N1, di)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        ctx.socs_image_batch(bank, dm, B1, N1, di)
    ctx.sync()
    rate = B1 * steps / (time.perf_counter() - t)
    out = {"N": N1, "Nx": Nx, "K": nk, "batch": B1, "conv_grid": 81, "flags": flags,
           "value": round(rate, 1), "unit": "images/s",
           "path_frac": round(8 * N1 * N1 * rate / 1e9 / HBM_PEAK_GBPS, 4)}
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        img = ctx.d2h(np.zeros((N1, N1), np.float32), di)
        ref = O.calc_socs(O.mask_spectrum(masks[0]), K.astype(np.complex128),
                          sc.astype(np.float32), Nx, Nx, nthreads=min(16, os.cpu_count() or 1))
        out["max_rel_err_vs_fp64"] = float(np.abs(img.astype(np.float64) - ref).max()
                                           / np.abs(ref).max())
    log(f"config 1 (512^2, K=10): {rate:.1f} images/s"
        + (f", {out['max_rel_err_vs_fp64']:.2e} vs FP64" if with_oracle else ""))
    ctx.free(dm)
    ctx.free(di)
    bank.close()
    return out


def v18_entry_bench(ctx, reps=50):
    """The HLS v18 accelerator entry at the reference's own operating point:
    its golden_1024 optics (1 nm pixels, 1024 x 1024 nm period, NA 0.8,
    defocus 0.2, 101^2 annulus 0.6-0.9, so nx = ny = 8), K = 10 kernels from
    the GPU, the fixed 128 x 128 tmpImgp (v18 .h L34-35), FI excluded.  Host
    entry: litho_socs_tmpimg on host buffers, as a v18 caller calls it (the
    1024^2 spectrum uploaded, the SOCS loop, both outputs read back); device
    entry: litho_socs_tmpimg_dev with the kernels resident in a bank and the
    spectrum in HBM, calls enqueued back to back.  The reference's figures for
    this call: the FPGA SOCS loop's C-synthesis estimate, 10.57 ms at 250 MHz
    (doc/论文/第五章.md L193-216), and the CPU Module 7, 35.6 ms on one thread
    (data/pc_cpp/run_log_10kernels.txt L112, FI included)."""
    src = L.createSource(101, "Annular", 0.6, 0.9)
    s = L.getOuterSigma(src)
    nx = L.compute_nx(0.8, 1024, s, LAMBDA)
    nk = 10
    sc, K, _ = L.kernels_from_source(src, s, LAMBDA, 0.8, 0.2, 1024, 1024, nx, nx, nk)
    mf = L.FT_r2c_padded(L.manhattan_mask(1024, 77))
    G = L.tmpimg_grid(nx)
    L.calc_socs_hls(mf, K, sc, nx, nx)  # first call: context and tables
    t0 = time.perf_counter()
    for _ in range(reps):
        L.calc_socs_hls(mf, K, sc, nx, nx)
    host_ms = (time.perf_counter() - t0) / reps * 1e3
    bank = ctx.bank(K, sc, nx, nx)
    mr, mi = L._planes(mf)
    dr, di = ctx.alloc(mr.nbytes), ctx.alloc(mi.nbytes)
    dt, do = ctx.alloc(G * G * 4), ctx.alloc(G * G * 4)
    try:
        ctx.h2d(dr, mr)
        ctx.h2d(di, mi)

        def call():
            L._check(L.lib().litho_socs_tmpimg_dev(ctx.h, bank, dr, di, 1024, 1024, G, G, dt, do),
                     "litho_socs_tmpimg_dev")

        call()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        ctx.sync()
        dev_us = (time.perf_counter() - t0) / reps * 1e6
    finally:
        for q in (dr, di, dt, do):
            ctx.free(q)
        bank.close()
    return {"nx": nx, "K": nk, "grid": G, "host_entry_ms": round(host_ms, 3),
            "device_entry_us": round(dev_us, 1),
            "reference": "FPGA SOCS loop 10.57 ms per call (C-synthesis estimate, 250 MHz, "
                         "FI excluded); CPU Module 7 35.6 ms on one thread (FI included)"}


def large_clip_bench(ctx, src, sigma, steps, with_oracle):
    """Config 5: 4096^2 clips at 4 nm (period 16384 nm), Nx = 137, K = 64 from
    the implicit TCC (source-space Gram route); the SOCS loop in FP32 on a
    576 = 24^2 conv grid (the intensity band 4 Nx + 1 = 549 fits; the
    reference's nextPow2 grid is 1024 and gives the same image), kernels stored
    as given (f32) and as bf16 with per-entry scales (config 5's storage),
    fp16, bf16 hi + lo.  Images/s over graph-replayed batches of 32 clips
    resident in HBM; error of one clip against the FP64 oracle fed the same
    kernels."""
    N, B, nk = 4096, 32, 64
    Nx = L.compute_nx(NA, PIX_NM * N, sigma, LAMBDA)
    t0 = time.perf_counter()
    sc, K, flags = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, PIX_NM * N, PIX_NM * N,
                                         Nx, Nx, nk)
    setup = {"wall_s": round(time.perf_counter() - t0, 3), **L.setup_timing()}
    bank = ctx.bank(K, sc, Nx, Nx)
    masks = np.stack([L.manhattan_mask(N, 7000 + i, WMIN, WMAX, FILL) for i in range(B)])
    dm, di = ctx.alloc(masks.nbytes), ctx.alloc(masks.nbytes)
    ctx.h2d(dm, masks)
    forms = {"fp32": L.LITHO_PREC_FP32, "bf16_bank": L.LITHO_PREC_BANK_BF16,
             "fp16_bank": L.LITHO_PREC_BANK_FP16, "bf16x2_bank": L.LITHO_PREC_BANK_BF16X2}
    out = {"N": N, "Nx": Nx, "K": nk, "batch": B, "conv_grid": 576 if 4 * Nx + 1 <= 576 else 1024,
           "entries": ctx.bank_info(bank)[0], "setup": setup, "forms": {}}
    imgs = {}
    for name, prec in forms.items():
        for _ in range(2):
            ctx.socs_image_batch_pipelined(bank, dm, B, N, di, prec, 1, True)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(steps):
            ctx.socs_image_batch_pipelined(bank, dm, B, N, di, prec, 1, True)
        ctx.sync()
        rate = B * steps / (time.perf_counter() - t)
        imgs[name] = ctx.d2h(np.zeros((N, N), np.float32), di)
        out["forms"][name] = {"value": round(rate, 2), "unit": "images/s",
                              "path_frac": round(8 * N * N * rate / 1e9 / HBM_PEAK_GBPS, 4)}
        log(f"config 5 (4096^2, K=64) {name}: {rate:.1f} images/s")
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t = time.perf_counter()
        ref = O.calc_socs(O.mask_spectrum(masks[0]), K.astype(np.complex128), sc.astype(np.float32),
                          Nx, Nx, nthreads=min(16, os.cpu_count() or 1))
        for name, img in imgs.items():
            out["forms"][name]["max_rel_err_vs_fp64"] = float(
                np.abs(img.astype(np.float64) - ref).max() / np.abs(ref).max())
        out["oracle_s"] = round(time.perf_counter() - t, 1)
        log("config 5 errors vs FP64: " + ", ".join(
            f"{n} {f['max_rel_err_vs_fp64']:.2e}" for n, f in out["forms"].items()))
    ctx.free(dm)
    ctx.free(di)
    bank.close()
    return out


def surrogate_bench(ctx, src, sigma, train_s):
    """Learned form: the FNO trained on 20k SOCS pairs (256^2 at 4 nm, K=16)
    by its own GPU kernels for `train_s` seconds; held-out max relative error
    against SOCS, inference and served images/s (1 in 16 clips re-run by SOCS,
    fallback above 1e-2), SOCS images/s at the same size."""
    from litho_amd import fno
    N, ntr, nte, batch = 256, 20000, 512, 32
    Nx = L.compute_nx(NA, PIX_NM * N, sigma, LAMBDA)
    sc, K, _ = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, PIX_NM * N, PIX_NM * N, Nx,
                                     Nx, 16)
    bank = ctx.bank(K, sc, Nx, Nx)
    t0 = time.perf_counter()
    tm, ti = fno.socs_pairs(ctx, bank, N, ntr, 10_000, 256, WMIN, WMAX, FILL)
    vm, vi = fno.socs_pairs(ctx, bank, N, nte, 900_000, 256, WMIN, WMAX, FILL)
    data_s = time.perf_counter() - t0
    model = fno.FNO(ctx, N, seed=5)
    steps = int(train_s / 0.014) + 1  # the cosine schedule spans the expected step count
    t0 = time.perf_counter()
    losses = fno.train(model, tm, ti, ntr, steps=steps, batch=batch, lr=2e-3, time_budget=train_s)
    tr_s = time.perf_counter() - t0
    e = fno.max_rel_errors(ctx, model, vm, vi, nte)
    d_o = ctx.alloc(max(256, nte) * N * N * 4)

    def rate(fn, B, reps):
        for _ in range(2):
            fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return B * reps / (time.perf_counter() - t)

    inf = rate(lambda: model.forward(vm, 256, d_o), 256, 10)
    inf32 = rate(lambda: model.forward_fp32(vm, 256, d_o), 256, 5)
    socs = rate(lambda: ctx.socs_image_batch(bank, vm, 256, N, d_o), 256, 20)
    srv = fno.SurrogateServer(ctx, model, bank, tol=1e-2, verify_every=16)
    served = rate(lambda: srv.serve(vm, nte, d_o), nte, 3)
    srv_all = fno.SurrogateServer(ctx, model, bank, tol=1e-2, verify_every=1)
    served_all = rate(lambda: srv_all.serve(vm, nte, d_o), nte, 3)
    out = {"form": "Fourier neural operator, width 32, 16x16 retained modes per ky sign, 4 layers, "
                   "GELU, bf16 weights; trained by its own HIP forward/backward kernels (Adam, MSE, "
                   "FP32 arithmetic) on SOCS images of seeded Manhattan masks; inference on the "
                   "matrix cores (one fused bf16-MFMA kernel per Fourier layer, activations and "
                   "tables split into bf16 hi + lo, FP32 accumulate)",
           "N": N, "K_socs": 16, "train_pairs": ntr, "data_s": round(data_s, 2),
           "train_s": round(tr_s, 1), "steps": int(len(losses)), "batch": batch,
           "samples_per_s": round(len(losses) * batch / tr_s, 1),
           "loss": [float(losses[0]), float(losses[-1])],
           "heldout_clips": nte,
           "max_rel_err": {"median": float(np.median(e)), "p90": float(np.percentile(e, 90)),
                           "max": float(e.max()), "frac_le_1e-2": float((e <= 1e-2).mean())},
           "inference_images_per_s": round(inf, 1),
           "inference_fp32_vector_images_per_s": round(inf32, 1),
           "served_images_per_s": round(served, 1),
           "served_1_in_16": srv.stats(),
           "served_all_verified_images_per_s": round(served_all, 1),
           "served_all_verified": srv_all.stats(),
           "socs_images_per_s_same_size": round(socs, 1),
           "inference_vs_socs": round(inf / socs, 5)}
    log(f"surrogate: {len(losses)} steps in {tr_s:.1f}s, held-out max rel err median "
        f"{np.median(e):.2e} p90 {np.percentile(e, 90):.2e} max {e.max():.2e}; inference "
        f"{inf:.0f} images/s (FP32 vector kernels {inf32:.0f}) vs SOCS {socs:.0f} images/s at "
        f"256^2")
    for p_ in (tm, ti, vm, vi, d_o):
        ctx.free(p_)
    model.close()
    bank.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--path", choices=["auto", "pow2"], default="auto",
                    help="conv grid: auto (band grid 144 at Nx=34) or the reference nextPow2 grid")
    ap.add_argument("--chunks", type=int, default=1,
                    help="sub-batches of the pipelined (chosen) form")
    ap.add_argument("--chosen-batch", type=int, default=512,
                    help="clips per step of the chosen form (its best batch; also timed at --batch)")
    ap.add_argument("--surrogate-s", type=float, default=100.0,
                    help="FNO training budget in seconds (0: skip the learned form)")
    ap.add_argument("--no-explicit", action="store_true",
                    help="skip the explicit n x n TCC + top-K eigensolve timing")
    ap.add_argument("--setup-only", action="store_true",
                    help="run the kernel extraction (Modules 4 + 6) timings only and print them")
    ap.add_argument("--setup-n", type=int, default=1024, choices=[512, 1024, 4096],
                    help="clip size of --setup-only (512: K=10, 4096: K=64, Gram route only)")
    ap.add_argument("--no-cpu-setup", action="store_true",
                    help="skip the CPU Modules 4 + 6 timing")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 4096^2 (config 5) measurement")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-size (batching form) measurement")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 512^2 (config 1) measurement")
    ap.add_argument("--no-v18", action="store_true",
                    help="skip the timing of the HLS v18 entry (litho_socs_tmpimg)")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 4096^2 (config 5) measurement and print its JSON")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        ctx = L.Context()
        a, b = ctx.alloc(args.copy_probe), ctx.alloc(args.copy_probe)
        for _ in range(3):
            ctx.d2d(b, a, args.copy_probe)
        log(f"copied {args.copy_probe} bytes device-to-device 3 times")
        ctx.close()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.large_only:
        src = L.createSource(41, "Annular", 0.6, 0.9)
        ctx = L.Context()
        out = large_clip_bench(ctx, src, L.getOuterSigma(src), 4, not args.no_cpu)
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    B = args.batch
    global N, NK
    if args.setup_only and args.setup_n != N:
        # Modules 4 + 6 at another clip of the same optics: 512^2 (Nx = 17,
        # K = 10) or 4096^2 (Nx = 137, K = 64, implicit TCC only)
        N, NK = args.setup_n, {512: 10, 4096: 64}.get(args.setup_n, NK)
        if (2 * L.compute_nx(NA, PIX_NM * N, 0.9, LAMBDA) + 1) ** 2 > 20000:
            args.no_explicit = True

    # ---- setup (one time): source, Nx, kernels on the GPU ----------------------
    src = L.createSource(41, "Annular", 0.6, 0.9)
    sigma = L.getOuterSigma(src)
    period = PIX_NM * N
    Nx = L.compute_nx(NA, period, sigma, LAMBDA)
    G = 1
    while G < 4 * Nx + 1:
        G *= 2
    log(f"N={N} Nx={Nx} K={NK} G={G} batch={B}")
    t0 = time.perf_counter()
    sc, K, flags = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, period, period, Nx, Nx, NK)
    setup_s = time.perf_counter() - t0
    setup_gram = L.setup_timing()
    log(f"kernels (Gram route, FP64 subspace eigensolve on GPU) {setup_s:.3f}s flags={flags} "
        f"{setup_gram}")
    # precision form of the setup: split-BF16 Gram contraction, FP64 eigensolve
    t0 = time.perf_counter()
    sc16, K16, flags16 = L.kernels_from_source(src, sigma, LAMBDA, NA, DEFOCUS, period, period, Nx,
                                               Nx, NK, tcc_prec=L.LITHO_TCC_SPLIT_BF16)
    setup16_s = time.perf_counter() - t0
    setup_gram16 = L.setup_timing()
    # the reference's own route (calcTCC L441-532 then calcKernels L597-663):
    # the explicit n x n TCC, n = 4761, then its top-24 eigenpairs
    explicit = None
    if not args.no_explicit:
        t0 = time.perf_counter()
        T = L.calcTCC(src, sigma, LAMBDA, NA, DEFOCUS, period, period, Nx, Nx)
        tcc_wall = time.perf_counter() - t0
        tcc_t = L.setup_timing()
        t0 = time.perf_counter()
        scx, Kx, _ = L.calcKernels(T, NK, Nx, Nx)
        eig_wall = time.perf_counter() - t0
        eig_t = L.setup_timing()
        del T
        explicit = {"n": (2 * Nx + 1) ** 2,
                    "pupil_ms": round(tcc_t["pupil_ms"], 3),
                    "tcc_contraction_ms": round(tcc_t["contraction_ms"], 3),
                    "tcc_call_s": round(tcc_wall, 3),
                    "eigensolve_ms": round(eig_t["eigensolve_ms"], 2),
                    "eigensolver": eig_t["solver"], "iterations": eig_t["iterations"],
                    "kernels_call_s": round(eig_wall, 3),
                    "tcc_contraction_gflop": round(8 * 572 * ((2 * Nx + 1) ** 2) ** 2 / 1e9, 1),
                    "eigenvalue_max_rel_err_vs_gram_route": float(np.abs(scx - sc).max() / sc[0]),
                    "kernel_max_abs_err_vs_gram_route": float(np.abs(Kx - K).max()),
                    "reference_48_threads": "TCC 1.386 s, zheevr 111.8 s at n = 4225 "
                                            "(output/Different_resolution_tests/TEST_REPORT.md "
                                            "L55, L57)",
                    "scope": "host wall clock inside the library calls; the call times add the "
                             "host staging of the n x n complex128 matrix (363 MB each way)"}
        log(f"explicit route n={explicit['n']}: TCC contraction {explicit['tcc_contraction_ms']} ms, "
            f"top-{NK} eigensolve {explicit['eigensolve_ms']} ms ({eig_t['solver']}, "
            f"{eig_t['iterations']} iterations)")
    eig_err = float(np.abs(sc16 - sc).max() / sc[0])
    krn_err = float(max(np.abs(K16[k] - K[k]).max() / np.abs(K[k]).max() for k in range(NK)))
    log(f"kernels (split-BF16 Gram contraction + FP64 Jacobi) {setup16_s:.3f}s; vs FP64 "
        f"eigenvalues {eig_err:.2e}, kernels {krn_err:.2e}")

    if args.setup_only:
        print(json.dumps({"N": N, "Nx": Nx, "K": NK, "gram_route_fp64": setup_gram,
                          "gram_route_split_bf16": setup_gram16,
                          "explicit_route_fp64": explicit}), flush=True)
        return
    ctx = L.Context()
    ctx.set_path(L.LITHO_PATH_POW2 if args.path == "pow2" else L.LITHO_PATH_AUTO)
    bank = ctx.bank(K, sc, Nx, Nx)
    bank16 = ctx.bank(K16, sc16, Nx, Nx)  # chosen form: kernels of the split-BF16 route
    masks = np.stack([L.manhattan_mask(N, 1000 + i, WMIN, WMAX, FILL) for i in range(B)])
    d_mask = ctx.alloc(masks.nbytes)
    d_img = ctx.alloc(masks.nbytes)
    ctx.h2d(d_mask, masks)

    # ---- warmup + timed steps ---------------------------------------------------
    # the timed steps record no per-stage events (the probe steps below do)
    ctx.set_timing(False)
    for _ in range(args.warmup):
        ctx.socs_image_batch(bank, d_mask, B, N, d_img)
    ctx.sync()
    log(f"warmup {args.warmup} steps done")
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.socs_image_batch(bank, d_mask, B, N, d_img)
    ctx.sync()
    wall = time.perf_counter() - t0
    # per-stage device time from HIP events on the library stream (extra steps)
    ctx.set_timing(True)
    nprobe = min(5, args.steps)
    stage_acc = {}
    for _ in range(nprobe):
        ctx.socs_image_batch(bank, d_mask, B, N, d_img)
        _, st = ctx.last_timing()
        for k, v in st.items():
            stage_acc[k] = stage_acc.get(k, 0.0) + v / nprobe
    ms_per_step = wall * 1e3 / args.steps
    images_per_s = B * args.steps / wall
    log(f"{images_per_s:.1f} images/s, {ms_per_step:.3f} ms/step, stages(ms) "
        + ", ".join(f"{n}={t:.3f}" for n, t in stage_acc.items()))
    Gimg = G if ("socs_accum" in stage_acc) else next(g for g in (36, 81, 144) if g >= 4 * Nx + 1)
    band = "socs_band" in stage_acc
    entries, pairs = ctx.bank_info(bank)  # kernels sharing one transform in the band path
    alg = stage_alg_bytes(N, Nx, Gimg, entries if band else NK, B, band=band)
    kernels = {n: {"ms": round(float(t), 4),
                   "alg_bytes": int(alg[n]),
                   "GBps": round(alg[n] / (t * 1e-3) / 1e9, 1)}
               for n, t in stage_acc.items()}
    dom = max(stage_acc, key=stage_acc.get)
    path_bytes = 8 * N * N  # mask read + image write per image
    step_dev_ms = float(sum(stage_acc.values()))
    achieved = path_bytes * B / (step_dev_ms * 1e-3) / 1e9
    dom_gbps = alg[dom] / (stage_acc[dom] * 1e-3) / 1e9

    # ---- chosen form: 1-bit (or 8-bit) masks, fp16 band DFTs on the matrix
    # cores, fp16 kernel bank, fp16 images; the step captured as one hipGraph ---
    def timed_form(precision, dm, bk=bank, graph=True, nb=B):
        for _ in range(max(2, args.warmup)):
            ctx.socs_image_batch_pipelined(bk, dm, nb, N, d_img2, precision, args.chunks, graph)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            ctx.socs_image_batch_pipelined(bk, dm, nb, N, d_img2, precision, args.chunks, graph)
        ctx.sync()
        return nb * args.steps / (time.perf_counter() - t)

    d_img2 = ctx.alloc(masks.nbytes)
    levels = np.rint(masks * 255).astype(np.uint8)  # binary masks: exact
    d_mask8 = ctx.alloc(levels.nbytes)
    ctx.h2d(d_mask8, levels)
    pipe_fp32 = timed_form(L.LITHO_PREC_FP32, d_mask)
    mixed = timed_form(L.LITHO_PREC_MIXED, d_mask)
    mixed_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float32), d_img2)
    bf16_bank_prec = L.LITHO_PREC_COMPACT & ~L.LITHO_PREC_BANK_FP16 | L.LITHO_PREC_BANK_BF16
    bf16_bank = timed_form(bf16_bank_prec, d_mask8, bank16)
    bf16_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float16), d_img2).astype(np.float32)
    bf16x2_prec = L.LITHO_PREC_COMPACT & ~L.LITHO_PREC_BANK_FP16 | L.LITHO_PREC_BANK_BF16X2
    bf16x2_bank = timed_form(bf16x2_prec, d_mask8, bank16)
    bf16x2_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float16), d_img2).astype(np.float32)
    compact_u8 = timed_form(L.LITHO_PREC_COMPACT, d_mask8, bank16)
    u8_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float16), d_img2).astype(np.float32)
    # chosen form: the same with the binary masks one bit per pixel (lossless
    # for 0/1 masks, the workload's masks), timed at its own batch size (512
    # distinct clips by default: its intermediates are a quarter of the FP32
    # path's, so the larger batch still runs from the last-level cache, where
    # the FP32 path is fastest at 256) and at the plain path's
    CB = max(B, args.chosen_batch)
    extra = [L.manhattan_mask(N, 1000 + i, WMIN, WMAX, FILL) for i in range(B, CB)]
    bits = np.packbits(np.concatenate([masks] + ([np.stack(extra)] if extra else [])).astype(bool),
                       axis=-1, bitorder="little")
    del extra
    d_mask1 = ctx.alloc(bits.nbytes)
    ctx.h2d(d_mask1, bits)
    if CB * N * N * 2 > masks.nbytes:
        ctx.free(d_img2)
        d_img2 = ctx.alloc(CB * N * N * 2)
    chosen_prec = (L.LITHO_PREC_COMPACT & ~L.LITHO_PREC_MASK_U8) | L.LITHO_PREC_MASK_1BIT
    chosen_at_b = timed_form(chosen_prec, d_mask1, bank16)
    chosen = timed_form(chosen_prec, d_mask1, bank16, nb=CB) if CB != B else chosen_at_b
    chosen_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float16), d_img2).astype(np.float32)
    log(f"graph FP32 {pipe_fp32:.1f} images/s; fp16 storage {mixed:.1f}; bf16 bank {bf16_bank:.1f}; "
        f"u8 masks {compact_u8:.1f}; chosen form (split-BF16 kernels, 1-bit masks, matrix-core "
        f"DFTs, fp16 window, bank and images) {chosen:.1f} images/s at {CB} clips per step "
        f"({chosen_at_b:.1f} at {B}) = {chosen / images_per_s:.2f}x plain")

    # measured HBM bytes (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated against a
    # device copy) for this pipeline, scaled to the batch
    traffic = path_traffic = None
    tfile = os.path.join(ROOT, "profiles", "socs_pipeline_1024_k24_plain_traffic.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        if set(tj["stages"]) == set(stage_acc) and tj["N"] == N and tj["Nx"] == Nx:
            path_traffic = int(tj["total_bytes_per_step"] * B / tj["batch"])
            traffic = int((tj["stages"][dom]["read"] + tj["stages"][dom]["write"]) * B / tj["batch"])

    out_imgs = ctx.d2h(np.zeros((min(B, 8), N, N), np.float32), d_img)
    mixed_sizes = None if args.no_mixed else mixed_size_bench(ctx, src, sigma,
                                                              max(3, args.steps // 4))
    small = None if args.no_small else small_clip_bench(ctx, src, sigma, 10, not args.no_cpu)
    large = None if args.no_large else large_clip_bench(ctx, src, sigma, 4, not args.no_cpu)
    v18 = None if args.no_v18 else v18_entry_bench(ctx)
    if v18:
        log(f"v18 entry (nx = ny = {v18['nx']}, K = {v18['K']}, {v18['grid']}^2 tmpImgp): host "
            f"{v18['host_entry_ms']} ms per call, device-resident {v18['device_entry_us']} us")
    surrogate = surrogate_bench(ctx, src, sigma, args.surrogate_s) if args.surrogate_s > 0 else None

    cpu = None
    max_err = None
    chosen_err = None
    mixed_err = None
    bf16_err = None
    bf16x2_err = None
    u8_err = None
    if not args.no_cpu:
        # the box's CPU share for one GPU is 16 threads (OMP_NUM_THREADS there)
        threads = min(16, os.cpu_count() or 1)
        rate, n_done, dt, errs = cpu_baseline(masks[:8], K, sc, Nx, threads, args.cpu_budget,
                                              [out_imgs, chosen_imgs, mixed_imgs, bf16_imgs,
                                               bf16x2_imgs, u8_imgs])
        max_err, chosen_err, mixed_err, bf16_err, bf16x2_err, u8_err = errs
        rate1, n1, dt1, _ = cpu_baseline(masks[:1], K, sc, Nx, 1, args.cpu_budget / 2, [])
        cpu = {"value": round(rate, 4), "unit": "images/s", "cores": threads, "kind": "port",
               "sample": f"{n_done} clip passes over 8 of the same 1024^2 clips, Module 3 + "
                         f"Module 7 (mask FFT, "
                         f"24-kernel SOCS on the 256^2 grid, FI) in FP64 with own FFT, "
                         f"{dt:.1f}s",
               "host_nproc": os.cpu_count(),
               "single_thread": {"value": round(rate1, 4), "unit": "images/s", "cores": 1,
                                 "sample": f"{n1} passes over one clip, {dt1:.1f}s"}}
        log(f"cpu baseline {rate:.3f} images/s on {threads} threads, {rate1:.3f} on 1; GPU vs "
            f"FP64 max rel err plain {max_err:.2e}, chosen {chosen_err:.2e}")
        if not args.no_cpu_setup:
            cpu["setup_modules_4_6"] = cpu_setup_baseline(src, sigma, Nx, NK, sc, K)
            log(f"cpu setup (Modules 4 + 6, Gram route): {cpu['setup_modules_4_6']}")

    line = {
        "metric": "1024^2 aerial images/s, plain vs chosen form at <=1e-3 max rel err; % HBM peak",
        "value": round(images_per_s, 2),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 SOCS, 1024^2 Manhattan clips at 4 nm, NA 0.85, "
                               "41^2 annulus 0.6-0.9, Nx=34, K=24",
                   "N": N, "Nx": Nx, "K": NK, "conv_grid": Gimg, "batch": B,
                   "socs_entries": entries if band else NK, "kernel_pairs": pairs if band else 0,
                   "form": "plain"},
        "roofline": {"bound": "hbm", "achieved": round(dom_gbps, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(dom_gbps / HBM_PEAK_GBPS, 4),
                     "traffic": traffic,
                     "kernel": dom,
                     "scope": f"the longest of the {len(stage_acc)} launches of a step: its "
                              "algorithmic bytes (inputs + outputs) per launch / its mean "
                              "HIP-event duration",
                     "path": {"achieved": round(achieved, 1),
                              "frac": round(achieved / HBM_PEAK_GBPS, 4),
                              "traffic": path_traffic,
                              "scope": "whole step; algorithmic bytes = 8 N^2 per image (mask "
                                       "read + image write) / summed launch durations"}},
        "path_roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                          "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": path_traffic,
                          "scope": "whole step of the plain path: 8 N^2 bytes per image (mask "
                                   "read + image write) / summed launch durations"},
        "kernels": kernels,
        "max_rel_err_vs_fp64": max_err,
        "chosen_form": {
            "form": "kernels from the split-BF16 Gram contraction (FP32 accumulate) and FP64 "
                    "eigensolve; batched mixed precision: binary masks one bit per pixel, "
                    "mask-row and image-row band DFTs as fp16 matrix products with FP32 "
                    "accumulate (image side split hi+lo below half the band), fp16 row-DFT "
                    "window, fp16 kernel bank with FP32 scales, FP32 column FFTs and SOCS, fp16 "
                    "images; the step captured as one hipGraph",
            "precision_flags": chosen_prec,
            "value": round(chosen, 2), "unit": "images/s", "batch": CB,
            f"value_at_batch_{B}": round(chosen_at_b, 2),
            "ratio_vs_plain": round(chosen / images_per_s, 3),
            "max_rel_err_vs_fp64": chosen_err,
            "roofline_frac": round((N * N / 8 + 2 * N * N) * chosen / 1e9 / HBM_PEAK_GBPS, 4),
            "roofline_scope": "algorithmic bytes of this form = N^2 / 8 + 2 N^2 per image "
                              "(1-bit mask read + fp16 image write)",
            "u8_masks_value": round(compact_u8, 2),
            "u8_masks_max_rel_err_vs_fp64": u8_err,
            "u8_masks_flags": L.LITHO_PREC_COMPACT,
            "graph_fp32_value": round(pipe_fp32, 2),
            "fp16_storage_value": round(mixed, 2),
            "fp16_storage_max_rel_err_vs_fp64": mixed_err,
            "tcc_split_bf16": {"setup_s": round(setup16_s, 3), "fp64_setup_s": round(setup_s, 3),
                               "eigenvalue_max_rel_err": eig_err,
                               "kernel_max_rel_err": krn_err},
            "bf16_bank_value": round(bf16_bank, 2),
            "bf16_bank_max_rel_err_vs_fp64": bf16_err,
            "bf16x2_bank_value": round(bf16x2_bank, 2),
            "bf16x2_bank_max_rel_err_vs_fp64": bf16x2_err,
            "bank_forms": "bf16: one bf16 per component x per-entry FP32 scale (8 significant "
                          "bits); bf16x2: bf16 hi + lo per component, same scales (16 bits); "
                          "fp16 (chosen): one fp16 per component (11 bits)"},
        "setup_s": round(setup_s, 3),
        "setup": {"gram_route_fp64": setup_gram, "gram_route_split_bf16": setup_gram16,
                  "explicit_route_fp64": explicit,
                  "scope": "host wall clock (ms) inside the library: pupil matrix, contraction, "
                           "eigensolve + kernel extraction"},
        "batching_form": mixed_sizes,
        "small_clip_form": small,
        "large_clip_form": large,
        "v18_entry": v18,
        "learned_form": surrogate,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    bank.close()
    bank16.close()
    ctx.free(d_mask)
    ctx.free(d_img)
    ctx.free(d_img2)
    ctx.free(d_mask8)
    ctx.free(d_mask1)
    ctx.close()


if __name__ == "__main__":
    main()
