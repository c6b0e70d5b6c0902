"""Mask inputs (PNG <-> BIN) and the reference's published 1024^2 run.

The reference's mask PNGs (ICCAD T1-T10, the resolution set) are decoded by
the package's own PNG codec; the oracle and the GPU path are then pinned to
the statistics litho.cpp printed for T1 (data/pc_cpp/run_log_*kernels.txt:
mask FFT samples, TCC, TCC-direct image, eigenvalues, SOCS images at
K = 10, 50 and 289 = full rank).  Printed values carry 6 decimals, so the
tolerance is 5e-7 of rounding plus FP32 image rounding.
"""
import json
import os

import numpy as np
import pytest

from conftest import golden

import litho_amd.maskio as mio

MASKS = sorted(f for f in os.listdir(golden("masks")) if f.endswith(".png"))
RUNLOG = json.load(open(golden("runlog_1024_nx8.json")))
TOL = 1.5e-6


def stats(a):
    a = np.asarray(a, np.float32)
    m = float(np.mean(a, dtype=np.float64))
    return {"mean": m, "std": float(np.sqrt(np.mean((a.astype(np.float64) - m) ** 2))),
            "max": float(a.max()), "min": float(a.min()),
            "samples": [float(v) for v in a.ravel()[:5]]}


def check_stats(got, want, what, tol=TOL):
    for k in ("mean", "std", "max", "min"):
        assert abs(got[k] - want[k]) <= tol, f"{what} {k}: {got[k]} vs {want[k]}"
    for a, b in zip(got["samples"], want.get("samples", [])):
        assert abs(a - b) <= tol, f"{what} samples: {got['samples']} vs {want['samples']}"


@pytest.mark.parametrize("name", MASKS)
def test_png_decoder_matches_pil(name):
    Image = pytest.importorskip("PIL.Image")
    path = golden(os.path.join("masks", name))
    assert np.array_equal(mio.read_png(path), np.array(Image.open(path)))


def test_png_roundtrip_every_filter(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    img[5:20, 10:30] = 255  # flat runs as in masks
    p = str(tmp_path / "x.png")
    mio.write_png(p, img, filters=[r % 5 for r in range(img.shape[0])])
    assert np.array_equal(mio.read_png(p), img)
    Image = pytest.importorskip("PIL.Image")
    assert np.array_equal(np.array(Image.open(p)), img)


def test_png_to_bin_convention(tmp_path):
    """png_to_bin divides by 255 (png_bin_converter.py L164-183); T1 is the
    run's input/mask/1024x1024.bin: mean and StdDev as printed (Module 2)."""
    b = str(tmp_path / "t1.bin")
    arr = mio.png_to_bin(golden("masks/T1.png"), b)
    assert np.array_equal(mio.read_bin(b), arr)
    assert np.array_equal(arr, mio.read_png(golden("masks/T1.png")).astype(np.float32) / 255)
    want = RUNLOG["runs"]["10"]["mask"]
    got = stats(arr)
    assert abs(got["mean"] - want["mean"]) <= TOL and abs(got["std"] - want["std"]) <= TOL
    assert np.array_equal(mio.mask_levels(arr), mio.read_png(golden("masks/T1.png")))
    # grey levels survive the 8-bit round trip as well
    g = mio.load_mask(golden("masks/res_256.png"))
    assert len(np.unique(g)) > 2
    assert np.array_equal(mio.mask_levels(g), mio.read_png(golden("masks/res_256.png")))
    u8 = mio.bin_to_png(b, str(tmp_path / "t1.png"))
    assert np.array_equal(u8, mio.read_png(golden("masks/T1.png")))


def runlog_setup(lib):
    cfg = RUNLOG["config"]
    mask = mio.load_mask(golden(cfg["mask"]))
    src = lib.source_annular(cfg["src_size"], cfg["inner"], cfg["outer"])
    sigma = lib.outer_sigma(src)
    Nx = lib.compute_nx(cfg["NA"], cfg["period"], sigma, cfg["wavelength"])
    return cfg, mask, src, sigma, Nx


def test_oracle_reproduces_reference_run(orc):
    """The oracle, end to end on T1, against every statistic the reference
    printed (FFT mask samples, TCC, TCC image, 10/50/289 eigenvalues, SOCS
    images; K=400 was clipped to n=289 by zheevr and equals the TCC image)."""
    cfg, mask, src, sigma, Nx = runlog_setup(orc)
    r10 = RUNLOG["runs"]["10"]
    assert Nx == r10["nx"]
    mskf = orc.mask_spectrum(mask)
    for y, x, re, im in r10["fft_samples"]:
        assert abs(mskf[y, x].real - re) <= TOL and abs(mskf[y, x].imag - im) <= TOL
    P = cfg["period"]
    tcc = orc.calc_tcc(src, sigma, cfg["wavelength"], cfg["NA"], cfg["defocus"], P, P, Nx, Nx)
    check_stats(stats(tcc.real.astype(np.float32)), r10["tcc_real"], "TCC real")
    itcc = orc.calc_image_tcc(mskf, tcc, Nx, Nx)
    check_stats(stats(itcc), r10["image_tcc"], "TCC image")
    n = tcc.shape[0]
    for key in ("10", "50", "400"):
        run = RUNLOG["runs"][key]
        nk = min(run["nk"], n)
        sc, K, split = orc.calc_kernels(tcc, nk, Nx, Nx)
        got = sc[:len(run["scales"])]
        assert np.all(np.abs(got - np.array(run["scales"])) <= 1e-6 * np.maximum(1, got)), key
        img = orc.calc_socs(mskf, K, sc.astype(np.float32), Nx, Nx, nthreads=8)
        if not split:  # a split cluster makes the image basis-dependent
            check_stats(stats(img), run["image_socs"], f"SOCS K={nk}")


@pytest.mark.gpu
def test_gpu_reproduces_reference_run(gpu, orc):
    """The GPU path (PNG mask -> spectrum, TCC, TCC image, kernels, SOCS) on
    T1 against the reference's printed statistics."""
    cfg, mask, src, sigma, Nx = runlog_setup(orc)
    P = cfg["period"]
    gsrc = gpu.createSource(cfg["src_size"], "Annular", cfg["inner"], cfg["outer"])
    assert np.array_equal(gsrc, src)
    r10 = RUNLOG["runs"]["10"]
    mskf = gpu.FT_r2c_padded(mask)
    for y, x, re, im in r10["fft_samples"]:
        assert abs(mskf[y, x].real - re) <= TOL and abs(mskf[y, x].imag - im) <= TOL
    tcc = gpu.calcTCC(gsrc, sigma, cfg["wavelength"], cfg["NA"], cfg["defocus"], P, P, Nx, Nx)
    check_stats(stats(tcc.real.astype(np.float32)), r10["tcc_real"], "GPU TCC real")
    itcc = gpu.calcImage(mask, tcc, Nx, Nx)
    check_stats(stats(itcc), r10["image_tcc"], "GPU TCC image")
    for key in ("10", "50"):
        run = RUNLOG["runs"][key]
        sc, K, flags = gpu.calcKernels(tcc, run["nk"], Nx, Nx)
        assert np.all(np.abs(sc - np.array(run["scales"])) <= 1e-6 * np.maximum(1, sc)), key
        if not flags & gpu.LITHO_FLAG_CLUSTER_SPLIT:
            img = gpu.socs_image(mask, K, sc, Nx, Nx)
            check_stats(stats(img), run["image_socs"], f"GPU SOCS K={run['nk']}")
