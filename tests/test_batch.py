"""Batch suites (litho_amd.batch, the reference's validation/batch_test_runner.py):
suite expansion and validation on the CPU; the ICCAD T1-T10 and resolution
suites through the GPU host CLI, pinned to the reference's eigenvalues
(output/Different_*_tests/*/kernels/kernel_info.txt) and to the oracle."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, golden, max_rel

from litho_amd.batch import BatchTestRunner

SUITES = os.path.join(ROOT, "tests", "configs", "batch")


def runner(name, out, **kw):
    return BatchTestRunner(os.path.join(SUITES, name), str(out), **kw)


def test_dry_run_expands_every_case(tmp_path):
    r = runner("iccad_t1_t10.json", tmp_path / "o", dry_run=True, verbose=0)
    s = r.run_all_tests()
    assert s["summary"] == {"total": 10, "success": 10, "failed": 0, "successRate": "100.0%"}
    assert all(x["status"] == "dry_run" for x in s["results"])
    suite = r.load_config()
    cfg = r.generate_test_config(suite["testCases"][2], {}, r.base_config(suite))
    assert cfg["optics"] == {"NA": 0.8, "wavelength": 193, "defocus": 0.2}
    assert cfg["mask"]["period"] == {"Lx": 2048, "Ly": 2048}
    assert cfg["kernel"]["count"] == 10
    # T3.bin is absent: the runner resolves the PNG beside it
    p, is_png = r.mask_source(cfg["mask"])
    assert is_png and p.endswith(os.path.join("masks", "T3.png"))
    assert not (tmp_path / "o").exists()


def test_validation_and_common_config(tmp_path):
    suite = {"testSuite": "s", "commonConfig": {"optics": {"NA": 0.7, "wavelength": 193,
                                                           "defocus": 0.0}},
             "testCases": [{"testName": "missing", "mask": {"type": "Import",
                                                            "inputFile": "nope.bin"}},
                           {"description": "no name"},
                           {"testName": "ls", "mask": {"type": "LineSpace"}}]}
    p = tmp_path / "suite.json"
    p.write_text(json.dumps(suite))
    r = BatchTestRunner(str(p), str(tmp_path / "o"), dry_run=True, verbose=0)
    s = r.run_all_tests()
    st = [x["status"] for x in s["results"]]
    assert st == ["validation_failed", "validation_failed", "dry_run"]
    assert "Mask file not found" in s["results"][0]["errors"][0]
    cfg = r.generate_test_config(suite["testCases"][2], suite["commonConfig"], {})
    assert cfg["optics"]["NA"] == 0.7 and cfg["mask"]["type"] == "LineSpace"


def eig_fixture(name, k):
    return np.loadtxt(golden(name))[:k]


@pytest.mark.gpu
def test_iccad_suite_on_gpu(tmp_path, orc):
    """T1-T10 (2048 nm period, Nx = 16): one kernel build, nine resumes; the
    eigenvalues equal the reference's T1 kernel_info.txt, the images the
    oracle's."""
    r = runner("iccad_t1_t10.json", tmp_path / "iccad", verbose=0)
    s = r.run_all_tests()
    assert s["summary"]["failed"] == 0, s["results"]
    assert sum("kernelsFrom" in x for x in s["results"]) == 9
    rep = json.load(open(tmp_path / "iccad" / "test_report.json"))
    assert rep["summary"]["total"] == 10
    sc = np.fromfile(tmp_path / "iccad" / "T1" / "scales.bin", np.float32)
    ref = eig_fixture("eig_resolution_2048x2048.txt", 10)  # = output/Different_mask_tests/T1
    assert np.all(np.abs(sc - ref) / ref < 6e-6)
    # T5 against the oracle: its mask centred in the 2048 field, the shared kernels
    import litho_amd.maskio as mio
    m = np.zeros((2048, 2048), np.float32)
    m[512:1536, 512:1536] = mio.load_mask(golden("masks/T5.png"))
    d1 = tmp_path / "iccad" / "T1"
    n = 33 * 33
    K = np.stack([np.fromfile(d1 / "kernels" / f"krn_{k}_r.bin", np.float32)
                  + 1j * np.fromfile(d1 / "kernels" / f"krn_{k}_i.bin", np.float32)
                  for k in range(10)]).astype(np.complex128)
    assert K.shape == (10, n)
    ref_img = orc.calc_socs(orc.mask_spectrum(m), K, sc, 16, 16, nthreads=8)
    img = np.fromfile(tmp_path / "iccad" / "T5" / "aerial_image_socs_kernel.bin",
                      np.float32).reshape(2048, 2048)
    assert max_rel(img, ref_img) < 1e-5


@pytest.mark.gpu
def test_resolution_suite_on_gpu(tmp_path):
    """256/512/1024 at 1 nm pixels (Nx 2/4/8): eigenvalues against the
    reference runs' kernel_info.txt; the TCC-direct image is written too."""
    r = runner("resolution_256_1024.json", tmp_path / "res", verbose=0)
    s = r.run_all_tests()
    assert s["summary"]["failed"] == 0, s["results"]
    for res in (256, 512, 1024):
        d = tmp_path / "res" / f"{res}x{res}"
        sc = np.fromfile(d / "scales.bin", np.float32)
        ref = eig_fixture(f"eig_resolution_{res}x{res}.txt", len(sc))
        assert np.all(np.abs(sc - ref) / ref < 6e-6), res
        assert (d / "aerial_image_tcc_direct.bin").exists()
