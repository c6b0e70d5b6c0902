"""One context per host thread: two threads, each with its own context and
kernel bank, run batches (plain FP32 and the matrix-core form) at the same
time; every image equals the one the same calls give when run alone, bit for
bit.  The C-ABI keeps no global mutable state other than per-device launch
attributes, which are set under a lock."""
import threading

import numpy as np
import pytest

from conftest import BASE_LAMBDA, BASE_NA

pytestmark = pytest.mark.gpu

N, B, NK = 512, 6, 10


def run_batch(gpu, K, sc, Nx, masks, prec):
    ctx = gpu.Context()
    try:
        bank = ctx.bank(K, sc, Nx, Nx)
        data = masks if prec == 0 else (masks * 255).astype(np.uint8)
        dm = ctx.alloc(data.nbytes)
        di = ctx.alloc(masks.size * 4)
        ctx.h2d(dm, data)
        out = []
        for _ in range(3):
            if prec == 0:
                ctx.socs_image_batch(bank, dm, B, N, di)
                out.append(ctx.d2h(np.zeros((B, N, N), np.float32), di))
            else:
                ctx.socs_image_batch_pipelined(bank, dm, B, N, di, prec, 1, True)
                out.append(ctx.d2h(np.zeros((B, N, N), np.float16), di))
        ctx.free(dm)
        ctx.free(di)
        bank.close()
        return out
    finally:
        ctx.close()


def test_two_threads_match_sequential(gpu, orc, base_source):
    src, s = base_source
    Nx = orc.compute_nx(BASE_NA, 4 * N, s, BASE_LAMBDA)
    sc, K, _ = gpu.kernels_from_source(src, s, BASE_LAMBDA, BASE_NA, 0.0, 4 * N, 4 * N, Nx, Nx,
                                       NK)
    jobs = [(np.stack([gpu.manhattan_mask(N, 700 + i) for i in range(B)]), 0),
            (np.stack([gpu.manhattan_mask(N, 800 + i) for i in range(B)]),
             gpu.LITHO_PREC_COMPACT)]
    print("[case] sequential reference runs", flush=True)
    ref = [run_batch(gpu, K, sc, Nx, m, p) for m, p in jobs]
    got = [None, None]
    errs = []

    def worker(i):
        try:
            got[i] = run_batch(gpu, K, sc, Nx, *jobs[i])
        except Exception as e:  # reported on the main thread
            errs.append(e)

    print("[case] two threads at once", flush=True)
    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    assert all(not t.is_alive() for t in th)
    for i in range(2):
        for a, b in zip(ref[i], got[i]):
            assert np.array_equal(a, b), i
        assert np.array_equal(got[i][0], got[i][-1])  # repeated calls agree
