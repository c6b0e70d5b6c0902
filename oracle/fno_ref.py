"""CPU restatement (numpy, FP64) of the Fourier-neural-operator surrogate --
test infrastructure only: the GPU kernels (fpga-litho-accel_amd/csrc/fno.hip)
are checked against it, and it is checked against finite differences and
against numpy's rfft2 / irfft2 definition of a mode-truncated spectral
convolution.  The reference repository has no learned model; this file
restates the library's own model, not reference code.

Model: width 32, 16 x-modes (kx = 0..15), 32 y-modes (ky = 0..15 and
-16..-1), 4 Fourier layers, GELU, projection 32 -> 128 -> 1; parameters in one
flat vector (layout below, the same as the GPU's); the GPU reads its weights
rounded to bf16, so `round_bf16` is applied before comparing.
"""
import numpy as np
from scipy.special import erf

CW, MX, MY, PJ, NL = 32, 16, 32, 128, 4
NM = MX * MY


def layout():
    o, L = 0, {}

    def take(name, n):
        nonlocal o
        L[name] = (o, n)
        o += n

    take("pw", CW)
    take("pb", CW)
    for l in range(NL):
        take(f"rre{l}", CW * CW * NM)
        take(f"rim{l}", CW * CW * NM)
        take(f"w{l}", CW * CW)
        take(f"b{l}", CW)
    take("q1", PJ * CW)
    take("q1b", PJ)
    take("q2", PJ)
    take("q2b", 1)
    return L, o


LAYOUT, NPARAM = layout()


def view(theta, name):
    o, n = LAYOUT[name]
    return theta[o:o + n]


def round_bf16(x):
    """FP32 -> bf16 (round to nearest even) -> float64, as the GPU copy."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(a):
    return 0.5 * a * (1.0 + erf(a / np.sqrt(2.0)))


def gelu_d(a):
    return 0.5 * (1.0 + erf(a / np.sqrt(2.0))) + a * np.exp(-0.5 * a * a) / np.sqrt(2 * np.pi)


def ky_values(N):
    return np.array([k if k < MY // 2 else k - MY for k in range(MY)])


def mats(N):
    x = np.arange(N)
    Fx = np.exp(-2j * np.pi * np.outer(x, np.arange(MX)) / N)           # N x MX
    Fy = np.exp(-2j * np.pi * np.outer(ky_values(N), x) / N)            # MY x N
    w = np.where(np.arange(MX) == 0, 1.0, 2.0)
    Hx = (w[:, None] * np.exp(2j * np.pi * np.outer(np.arange(MX), x) / N)) / (N * N)  # MX x N
    return Fx, Fy, Hx


def unpack(theta):
    P = {k: view(theta, k) for k in LAYOUT}
    R = [(P[f"rre{l}"] + 1j * P[f"rim{l}"]).reshape(CW, CW, MY, MX) for l in range(NL)]
    W = [P[f"w{l}"].reshape(CW, CW) for l in range(NL)]
    b = [P[f"b{l}"] for l in range(NL)]
    return P, R, W, b


def forward(theta, masks, keep=False):
    """masks [B][N][N] -> images [B][N][N] (and the activations if keep)."""
    B, N, _ = masks.shape
    Fx, Fy, Hx = mats(N)
    P, R, W, b = unpack(theta)
    x = P["pw"][None, :, None, None] * masks[:, None] + P["pb"][None, :, None, None]
    acts = {"x": [x], "a": [], "X2": []}
    for l in range(NL):
        X1 = x @ Fx                                          # [B,C,N,MX]
        X2 = np.einsum("qy,bcyk->bcqk", Fy, X1)              # [B,C,MY,MX]
        Y2 = np.einsum("biqk,ioqk->boqk", X2, R[l])
        Y1 = np.einsum("qy,boqk->boyk", Fy.conj(), Y2)       # [B,C,N,MX]
        z = (Y1 @ Hx).real
        a = z + np.einsum("oi,biyx->boyx", W[l], x) + b[l][None, :, None, None]
        x = gelu(a) if l < NL - 1 else a
        acts["a"].append(a)
        acts["X2"].append(X2)
        acts["x"].append(x)
    Q1 = P["q1"].reshape(PJ, CW)
    a1 = np.einsum("jc,bcyx->bjyx", Q1, x) + P["q1b"][None, :, None, None]
    out = np.einsum("j,bjyx->byx", P["q2"], gelu(a1)) + P["q2b"][0]
    if keep:
        acts["a1"] = a1
        return out, acts
    return out


def loss_grad(theta, masks, targets):
    """MSE over every pixel of the batch and its gradient (flat, as theta)."""
    B, N, _ = masks.shape
    Fx, Fy, Hx = mats(N)
    P, R, W, b = unpack(theta)
    out, A = forward(theta, masks, keep=True)
    n = out.size
    loss = float(np.mean((out - targets) ** 2))
    dout = 2.0 * (out - targets) / n
    g = np.zeros_like(theta)

    def acc(name, v):
        o, k = LAYOUT[name]
        g[o:o + k] += np.ravel(v)

    a1, h = A["a1"], gelu(A["a1"])
    Q1 = P["q1"].reshape(PJ, CW)
    acc("q2", np.einsum("byx,bjyx->j", dout, h))
    acc("q2b", [dout.sum()])
    da1 = P["q2"][None, :, None, None] * dout[:, None] * gelu_d(a1)
    acc("q1", np.einsum("bjyx,bcyx->jc", da1, A["x"][NL]))
    acc("q1b", da1.sum(axis=(0, 2, 3)))
    dv = np.einsum("jc,bjyx->bcyx", Q1, da1)
    for l in range(NL - 1, -1, -1):
        a, x = A["a"][l], A["x"][l]
        da = dv * gelu_d(a) if l < NL - 1 else dv
        acc(f"w{l}", np.einsum("boyx,biyx->oi", da, x))
        acc(f"b{l}", da.sum(axis=(0, 2, 3)))
        dx = np.einsum("oi,boyx->biyx", W[l], da)
        # spectral adjoints
        dY1 = da @ Hx.conj().T                                   # [B,C,N,MX]
        dY2 = np.einsum("qy,boyk->boqk", Fy, dY1)
        dR = np.einsum("biqk,boqk->ioqk", A["X2"][l].conj(), dY2)
        acc(f"rre{l}", dR.real)
        acc(f"rim{l}", dR.imag)
        dX2 = np.einsum("boqk,ioqk->biqk", dY2, R[l].conj())
        dX1 = np.einsum("qy,biqk->biyk", Fy.conj(), dX2)
        dx = dx + (dX1 @ Fx.conj().T).real
        dv = dx
    acc("pw", np.einsum("bcyx,byx->c", dv, masks))
    acc("pb", dv.sum(axis=(0, 2, 3)))
    return loss, g


def spectral_conv_fft(x, R):
    """The same mode-truncated spectral convolution through numpy's rfft2 /
    irfft2 (the usual FNO definition), for pinning the explicit-DFT form."""
    B, C, N, _ = x.shape
    xf = np.fft.rfft2(x)
    of = np.zeros((B, R.shape[1], N, N // 2 + 1), complex)
    ky = ky_values(N)
    rows = np.where(ky >= 0, ky, N + ky)
    of[:, :, rows, :MX] = np.einsum("biqk,ioqk->boqk", xf[:, :, rows, :MX], R)
    return np.fft.irfft2(of, s=(N, N))


def init_params(seed=0):
    """A random parameter vector of the GPU initialiser's distributions."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM)
    th[slice(*_span("pw"))] = rng.uniform(-1, 1, CW)
    for l in range(NL):
        th[slice(*_span(f"rre{l}"))] = rng.uniform(0, 1, CW * CW * NM) / (CW * CW)
        th[slice(*_span(f"rim{l}"))] = rng.uniform(0, 1, CW * CW * NM) / (CW * CW)
        th[slice(*_span(f"w{l}"))] = rng.uniform(-1, 1, CW * CW) / np.sqrt(CW)
        th[slice(*_span(f"b{l}"))] = rng.uniform(-0.1, 0.1, CW)
    th[slice(*_span("q1"))] = rng.uniform(-1, 1, PJ * CW) / np.sqrt(CW)
    th[slice(*_span("q1b"))] = rng.uniform(-0.1, 0.1, PJ)
    th[slice(*_span("q2"))] = rng.uniform(-1, 1, PJ) / np.sqrt(PJ)
    th[slice(*_span("pb"))] = rng.uniform(-0.1, 0.1, CW)
    return th


def _span(name):
    o, n = LAYOUT[name]
    return o, o + n
