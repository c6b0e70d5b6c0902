"""Batch test suites: many host-flow runs from one JSON suite file.

Mirrors validation/batch_test_runner.py of the reference (BatchTestRunner,
L40-264): a suite names `testSuite`, an optional `baseConfig` (a full host
config) and/or `commonConfig` (source / optics / kernel sections), and a list
of `testCases`, each with a `testName` and its own `mask` section.  Every case
becomes a full config written to <output-base>/<testName>/config.json and is
run through the GPU host CLI (litho_gpu, the `litho <config> <outdir>`
replacement); a `test_report.json` sums the outcomes.

Differences, on purpose:
  * a mask `inputFile` that is a PNG -- or a missing .bin next to a .png of
    the same stem, as the reference's suites expect after running
    tool/png_bin_converter.py -- is converted to BIN in the case directory;
  * cases that share optics, source, kernel count and mask period reuse the
    first such case's kernels (the CLI's --load-kernels resume), so a suite of
    ten clips builds one TCC and runs one eigensolve.

    python -m litho_amd.batch --config suite.json [--output-base DIR] [--dry-run]
"""
import argparse
import copy
import json
import os
import subprocess
import sys
import time
from datetime import datetime

import numpy as np

from . import maskio

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_CLI = os.path.join(PKG, "litho_gpu")


def _resolve(path, *bases):
    if os.path.isabs(path):
        return path
    for b in bases:
        p = os.path.normpath(os.path.join(b, path))
        if os.path.exists(p):
            return p
    return os.path.normpath(os.path.join(bases[0], path)) if bases else path


class BatchTestRunner:
    def __init__(self, config_path, output_base, dry_run=False, verbose=1, executable=None,
                 reuse_kernels=True, search_dirs=()):
        self.config_path = os.path.abspath(config_path)
        self.output_base = os.path.abspath(output_base)
        self.dry_run = dry_run
        self.verbose = verbose
        self.executable = executable or DEFAULT_CLI
        self.reuse_kernels = reuse_kernels
        self.suite_dir = os.path.dirname(self.config_path)
        self.search = [self.suite_dir, os.getcwd(), *search_dirs]
        self.results = []
        self._kernel_dirs = {}

    def log(self, msg, level=1):
        if self.verbose >= level:
            print(msg, flush=True)

    def load_config(self):
        with open(self.config_path, encoding="utf-8") as f:
            return json.load(f)

    def base_config(self, suite):
        base = suite.get("baseConfig")
        if not base:
            return {}
        p = _resolve(base, *self.search)
        if not os.path.exists(p):
            self.log(f"[WARN] baseConfig {base} not found; using commonConfig only")
            return {}
        with open(p, encoding="utf-8") as f:
            return json.load(f)

    def generate_test_config(self, test_case, common_config, base):
        """Full host config of one case: the base config, then the suite's
        common sections, then the case's mask."""
        full = copy.deepcopy(base) if base else {}
        full["simulation"] = {"description": f"Test case: {test_case['testName']}",
                              "version": "1.0"}
        for sec in ("source", "optics", "kernel"):
            if sec in common_config:
                full[sec] = copy.deepcopy(common_config[sec])
        mask = copy.deepcopy(full.get("mask", {}))
        mask.update(copy.deepcopy(test_case.get("mask", {})))
        full["mask"] = mask
        return full

    def mask_source(self, mask_cfg):
        """(path, is_png) of a case's mask input, or (None, False)."""
        f = mask_cfg.get("inputFile")
        if not f:
            return None, False
        p = _resolve(f, *self.search)
        if os.path.exists(p):
            return p, p.lower().endswith(".png")
        stem = os.path.splitext(p)[0]
        for cand in (stem + ".png", _resolve(os.path.splitext(f)[0] + ".png", *self.search)):
            if os.path.exists(cand):
                return cand, True
        return None, False

    def validate_test_case(self, test_case):
        errors = []
        if "testName" not in test_case:
            errors.append("Missing 'testName' field")
        if "mask" not in test_case:
            errors.append("Missing 'mask' configuration")
        elif test_case["mask"].get("type", "Import") == "Import":
            if "inputFile" not in test_case["mask"]:
                errors.append("Missing 'mask.inputFile' field")
            elif self.mask_source(test_case["mask"])[0] is None:
                errors.append(f"Mask file not found: {test_case['mask']['inputFile']}")
        return errors

    def _kernel_key(self, cfg):
        m = cfg.get("mask", {})
        return json.dumps([cfg.get("source"), cfg.get("optics"), cfg.get("kernel"),
                           m.get("period"), m.get("size")], sort_keys=True)

    def run_single_test(self, test_case, common_config, base):
        name = test_case.get("testName", "unknown")
        t0 = time.time()
        result = {"testName": name, "description": test_case.get("description", ""),
                  "status": "pending", "startTime": datetime.now().isoformat(),
                  "duration": 0.0, "outputDir": "", "errors": []}
        errs = self.validate_test_case(test_case)
        if errs:
            result["status"] = "validation_failed"
            result["errors"] = errs
            return result
        cfg = self.generate_test_config(test_case, common_config, base)
        out = os.path.join(self.output_base, name)
        result["outputDir"] = out
        if self.dry_run:
            result["status"] = "dry_run"
            self.log(f"[DRY-RUN] {name} -> {out}\n{json.dumps(cfg, indent=2)}", 2)
            return result
        os.makedirs(out, exist_ok=True)
        try:
            if cfg["mask"].get("type", "Import") == "Import":
                src, is_png = self.mask_source(cfg["mask"])
                if is_png:
                    binp = os.path.join(out, "mask_input.bin")
                    maskio.png_to_bin(src, binp)
                    src = binp
                cfg["mask"]["inputFile"] = src
            cfg_path = os.path.join(out, "config.json")
            with open(cfg_path, "w", encoding="utf-8") as f:
                json.dump(cfg, f, indent=2)
            cmd = [self.executable, cfg_path, out, "--quiet" if self.verbose < 2 else "--verbose"]
            key = self._kernel_key(cfg)
            if self.reuse_kernels and key in self._kernel_dirs:
                cmd += ["--load-kernels", self._kernel_dirs[key]]
                result["kernelsFrom"] = self._kernel_dirs[key]
            p = subprocess.run(cmd, capture_output=True, text=True)
            result["returnCode"] = p.returncode
            if p.returncode != 0:
                result["status"] = "execution_failed"
                result["errors"].append(p.stderr.strip()[-2000:])
            else:
                ok, info = self.analyze_results(out)
                result.update(info)
                result["status"] = "success" if ok else "analysis_failed"
                if ok and key not in self._kernel_dirs:
                    self._kernel_dirs[key] = out
        except Exception as e:  # noqa: BLE001 -- reported per case, as the reference does
            result["status"] = "error"
            result["errors"].append(str(e))
        finally:
            result["duration"] = time.time() - t0
            result["endTime"] = datetime.now().isoformat()
        return result

    @staticmethod
    def analyze_results(out):
        """The aerial image exists, is finite and non-negative; its statistics."""
        p = os.path.join(out, "aerial_image_socs_kernel.bin")
        if not os.path.exists(p):
            return False, {"errors": ["aerial_image_socs_kernel.bin missing"]}
        img = np.fromfile(p, np.float32)
        ok = img.size > 0 and bool(np.isfinite(img).all()) and float(img.min()) >= 0.0
        return ok, {"image": {"mean": float(img.mean()), "max": float(img.max()),
                              "min": float(img.min()), "pixels": int(img.size)}}

    def run_all_tests(self):
        suite = self.load_config()
        name = suite.get("testSuite", "unknown")
        cases = suite.get("testCases", [])
        common = suite.get("commonConfig", {})
        base = self.base_config(suite)
        self.log(f"# Batch test suite {name}: {len(cases)} cases -> {self.output_base}"
                 f"{' (dry run)' if self.dry_run else ''}")
        if not self.dry_run:
            os.makedirs(self.output_base, exist_ok=True)
        for i, case in enumerate(cases, 1):
            r = self.run_single_test(case, common, base)
            self.results.append(r)
            self.log(f"[{i}/{len(cases)}] {r['testName']}: {r['status']} ({r['duration']:.2f}s)")
        return self.generate_summary(name)

    def generate_summary(self, suite_name):
        total = len(self.results)
        ok = sum(1 for r in self.results if r["status"] in ("success", "dry_run"))
        summary = {"testSuite": suite_name, "timestamp": datetime.now().isoformat(),
                   "summary": {"total": total, "success": ok, "failed": total - ok,
                               "successRate": f"{ok / total * 100:.1f}%" if total else "0%"},
                   "totalDuration": sum(r["duration"] for r in self.results),
                   "results": self.results}
        if not self.dry_run:
            with open(os.path.join(self.output_base, "test_report.json"), "w",
                      encoding="utf-8") as f:
                json.dump(summary, f, indent=2, ensure_ascii=False)
        return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description="Batch test suites through the GPU host CLI")
    ap.add_argument("--config", required=True, help="suite JSON")
    ap.add_argument("--output-base", default=None,
                    help="output directory (default: output/<suite file stem>)")
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--verbose", "-v", action="count", default=1)
    ap.add_argument("--executable", default=None, help="host CLI (default: the in-tree litho_gpu)")
    ap.add_argument("--no-reuse", action="store_true",
                    help="build kernels for every case even when optics repeat")
    ap.add_argument("--search", action="append", default=[],
                    help="extra directory to resolve baseConfig / mask paths against")
    a = ap.parse_args(argv)
    out = a.output_base or os.path.join("output",
                                        os.path.splitext(os.path.basename(a.config))[0])
    r = BatchTestRunner(a.config, out, a.dry_run, a.verbose, a.executable, not a.no_reuse,
                        a.search)
    s = r.run_all_tests()
    print(json.dumps(s["summary"]))
    return 0 if s["summary"]["failed"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
