"""Extract the published statistics of the reference's 1024^2 run logs
(data/pc_cpp/run_log_{10,50,400}kernels.txt: litho.cpp on
input/mask/1024x1024.bin, i.e. T1.png / 255, NA 0.8, defocus 0.2, 101^2
annulus 0.6-0.9, period 1024) into runlog_1024_nx8.json.

Run in a checkout that has the reference next to it:
    python tests/golden/make_runlog_fixture.py /root/reference
"""
import json
import os
import re
import sys


def parse(path):
    txt = open(path, encoding="utf-8").read()
    out = {}

    def stats(name):
        m = re.search(r"\[STATISTICS\] " + re.escape(name) + r":\s*\n"
                      r"\s*Mean:\s*(\S+)\s*\n\s*StdDev:\s*(\S+)\s*\n\s*Max:\s*(\S+)\s*\n"
                      r"\s*Min:\s*(\S+)\s*\n(?:\s*Samples:\s*([^\n]*)\n)?", txt)
        if not m:
            return None
        d = {"mean": float(m.group(1)), "std": float(m.group(2)), "max": float(m.group(3)),
             "min": float(m.group(4))}
        if m.group(5):
            d["samples"] = [float(v) for v in m.group(5).split()]
        return d

    out["nk"] = int(re.search(r"Number of Kernels:\s*(\d+)", txt).group(1))
    out["mask"] = stats("Mask")
    out["tcc_real"] = stats("TCC Real")
    out["image_tcc"] = stats("Aerial Image (TCC)")
    out["image_socs"] = stats("Aerial Image (SOCS)")
    out["fft_samples"] = [[int(a), int(b), float(c), float(d)] for a, b, c, d in re.findall(
        r"\[(\d+),(\d+)\] = \((\S+), (\S+)\)", txt)]
    out["scales"] = [float(v) for v in re.findall(r"Scale\[\d+\]:\s*(\S+)", txt)]
    out["nx"] = int(re.search(r"Nx:\s*(\d+)", txt).group(1))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    runs = {}
    for k in (10, 50, 400):
        runs[str(k)] = parse(os.path.join(ref, "data", "pc_cpp", f"run_log_{k}kernels.txt"))
    fixture = {
        "source": "data/pc_cpp/run_log_{10,50,400}kernels.txt (litho.cpp)",
        "config": {"N": 1024, "period": 1024, "NA": 0.8, "wavelength": 193.0, "defocus": 0.2,
                   "src_size": 101, "inner": 0.6, "outer": 0.9, "mask": "masks/T1.png"},
        "runs": runs,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runlog_1024_nx8.json")
    with open(dst, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
