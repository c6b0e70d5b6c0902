"""Learned surrogate of the SOCS aerial image: a Fourier neural operator
trained on SOCS-generated pairs by the GPU's own forward / backward kernels
(litho_fno_* in the C-ABI), and a server in which SOCS checks and backs it up.

The reference has no learned model.  SOCS stays the ground truth: the
training targets are SOCS images of seeded Manhattan masks, and in serving
SOCS recomputes every clip (or one in `verify_every`); a verified clip whose
max relative error exceeds `tol` is served from SOCS instead.
"""
import ctypes as C
import time

import numpy as np

from . import _check, lib, manhattan_mask


class FNO:
    """Model handle on a Context (width 32, 16x16 modes, 4 layers)."""

    def __init__(self, ctx, N, seed=0):
        self.ctx, self.N = ctx, N
        h = C.c_void_p()
        _check(lib().litho_fno_create(ctx.h, N, seed, C.byref(h)), "litho_fno_create")
        self.h = h
        self.nparam = int(lib().litho_fno_param_count(self.h))

    def close(self):
        if self.h:
            lib().litho_fno_destroy(self.h)
            self.h = None

    def params(self, grads=False):
        p = np.zeros(self.nparam, np.float32)
        _check(lib().litho_fno_get_params(self.h, p, self.nparam, int(grads)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().litho_fno_set_params(self.h, p, self.nparam), "set_params")

    def forward(self, d_masks, B, d_out):
        """Inference on the matrix cores (fused bf16 hi + lo layer kernels)."""
        _check(lib().litho_fno_forward(self.h, d_masks, B, d_out), "litho_fno_forward")

    def forward_fp32(self, d_masks, B, d_out):
        """The training forward's FP32 vector-ALU kernels."""
        _check(lib().litho_fno_forward_fp32(self.h, d_masks, B, d_out), "litho_fno_forward_fp32")

    def loss_grad(self, d_masks, d_targets, B, sync=True):
        loss = C.c_float(0.0)
        _check(lib().litho_fno_loss_grad(self.h, d_masks, d_targets, B,
                                         C.byref(loss) if sync else None), "litho_fno_loss_grad")
        return loss.value

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        _check(lib().litho_fno_adam(self.h, lr, b1, b2, eps), "litho_fno_adam")

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path))


def socs_pairs(ctx, bank, N, n, seed0, batch=256, wmin=12, wmax=45, fill=0.40):
    """n seeded Manhattan masks and their SOCS images, resident on the device:
    returns (d_masks, d_images) of n x N x N f32 each."""
    d_m = ctx.alloc(n * N * N * 4)
    d_i = ctx.alloc(n * N * N * 4)
    for b0 in range(0, n, batch):
        nb = min(batch, n - b0)
        m = np.stack([manhattan_mask(N, seed0 + b0 + i, wmin, wmax, fill) for i in range(nb)])
        dm = C.c_void_p(d_m.value + b0 * N * N * 4)
        ctx.h2d(dm, m)
        ctx.socs_image_batch(bank, dm, nb, N, C.c_void_p(d_i.value + b0 * N * N * 4))
    ctx.sync()
    return d_m, d_i


def train(model, d_masks, d_images, n, steps, batch=16, lr=2e-3, time_budget=None, log=None):
    """Adam on minibatches drawn in order from the n resident pairs, with a
    cosine learning-rate decay; stops after `steps` or `time_budget` s.
    Returns the loss of every step."""
    N = model.N
    plane = N * N * 4
    losses, t0 = [], time.perf_counter()
    nbatch = max(1, n // batch)
    for s in range(steps):
        k = s % nbatch
        dm = C.c_void_p(d_masks.value + k * batch * plane)
        di = C.c_void_p(d_images.value + k * batch * plane)
        losses.append(model.loss_grad(dm, di, batch, sync=(s % 25 == 0 or s == steps - 1)))
        frac = s / max(1, steps - 1)
        model.adam(lr * 0.5 * (1 + np.cos(np.pi * frac)) + lr * 0.01)
        if log and s % 200 == 0:
            log(f"step {s} loss {losses[-1]:.3e}")
        if time_budget and time.perf_counter() - t0 > time_budget:
            break
    model.ctx.sync()
    return np.array(losses)


def max_rel_errors(ctx, model, d_masks, d_images, n, batch=64):
    """Per-clip max |FNO - SOCS| / max |SOCS| over n resident pairs."""
    N = model.N
    plane = N * N * 4
    d_o = ctx.alloc(batch * plane)
    errs = []
    for b0 in range(0, n, batch):
        nb = min(batch, n - b0)
        model.forward(C.c_void_p(d_masks.value + b0 * plane), nb, d_o)
        got = ctx.d2h(np.zeros((nb, N, N), np.float32), d_o)
        ref = ctx.d2h(np.zeros((nb, N, N), np.float32), C.c_void_p(d_images.value + b0 * plane))
        errs += list(np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2)))
    ctx.free(d_o)
    return np.array(errs)


class SurrogateServer:
    """Serve images from the FNO with SOCS as judge and fallback.

    Every `verify_every`-th clip of a batch is recomputed by SOCS and a
    verified clip above `tol` max relative error is served from SOCS.  With
    verify_every = 1 (the default) every served clip is within `tol` of SOCS;
    SOCS at the surrogate's sizes costs a fraction of a percent of the FNO
    forward pass, so the bound is nearly free.  With verify_every > 1 (one in
    16 clips in the sampled mode) the other clips are served unverified: their
    count is kept in `unverified`, and their error is only known
    statistically (held-out error distribution)."""

    def __init__(self, ctx, model, bank, tol=1e-2, verify_every=1):
        if verify_every < 1:
            raise ValueError("verify_every must be >= 1")
        self.ctx, self.model, self.bank = ctx, model, bank
        self.tol, self.every = tol, verify_every
        self.served = self.verified = self.fallbacks = 0
        self.worst = 0.0

    @property
    def unverified(self):
        return self.served - self.verified

    def stats(self):
        return {"served": self.served, "verified": self.verified,
                "unverified": self.unverified, "fallbacks": self.fallbacks,
                "unverified_fraction": self.unverified / max(1, self.served),
                "worst_verified_before_fallback": self.worst}

    def serve(self, d_masks, B, d_images):
        N = self.model.N
        plane = N * N * 4
        self.model.forward(d_masks, B, d_images)
        self.served += B
        idx = list(range(0, B, self.every))
        if not idx:
            return
        d_s = self.ctx.alloc(len(idx) * plane)
        if self.every == 1:
            self.ctx.socs_image_batch(self.bank, d_masks, B, N, d_s)
        else:
            d_m = self.ctx.alloc(len(idx) * plane)
            for j, i in enumerate(idx):
                self.ctx.d2d(C.c_void_p(d_m.value + j * plane),
                             C.c_void_p(d_masks.value + i * plane), plane)
            self.ctx.socs_image_batch(self.bank, d_m, len(idx), N, d_s)
            self.ctx.free(d_m)
        # the judge runs on the device: per-clip max relative error, and the
        # clips above tol are overwritten with their SOCS images there
        err = np.zeros(len(idx), np.float32)
        fell = lib().litho_fno_verify(self.model.h, d_images, self.every, d_s, len(idx),
                                      self.tol, err)
        _check(min(fell, 0), "litho_fno_verify")
        self.verified += len(idx)
        self.fallbacks += fell
        self.worst = max(self.worst, float(err.max()))
        self.ctx.free(d_s)
