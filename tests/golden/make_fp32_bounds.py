"""Measures the FP32 walker step's error per deterministic variation against
the FP64 oracle (tests/fp32_variation_error.py) and writes
fp32_variation_bounds.json beside this script: the measured statistics and
the bounds tests/test_gpu_stages.py::test_variation_fp32_walker enforces.

Bounds: p99 <= max(2 x measured p99, 1e-5); max <= max(4 x measured max,
1e-4); finite fraction >= measured - 0.005.  Needs the GPU library.

    python tests/golden/make_fp32_bounds.py [--markdown] [--out PATH]
"""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def up(x, floor):
    """x rounded up to two significant digits, at least floor."""
    x = max(x, floor)
    e = math.floor(math.log10(x)) - 1
    return math.ceil(x / 10 ** e) * 10 ** e


def main():
    import flam3_amd
    import pyoracle
    from flam3_amd import abi
    import fp32_variation_error as E

    flam3_amd.lib()
    pyoracle.lib()
    out = {}
    for vi, name in enumerate(abi.VARIATION_NAMES):
        if name in E.RANDOM_VARS:
            continue
        m = E.measure(flam3_amd, pyoracle, vi)
        m["bound_p99"] = up(2 * m["p99"], 1e-5)
        m["bound_max"] = up(4 * m["max"], 1e-4)
        m["bound_finite"] = round(max(0.0, m["finite"] - 0.005), 4)
        out[name] = m
        print("%-16s p50 %.2e p99 %.2e max %.2e finite %.4f excluded %.4f" % (
            name, m["p50"], m["p99"], m["max"], m["finite"], m["excluded"]), flush=True)
    path = sys.argv[sys.argv.index("--out") + 1] if "--out" in sys.argv else os.path.join(
        HERE, "fp32_variation_bounds.json")
    with open(path, "w") as fh:
        json.dump({"grid": E.GRID, "span": E.SPAN, "step": E.STEP, "gain": E.GAIN, "variations": out}, fh,
                  indent=1, sort_keys=True)
    if "--markdown" in sys.argv:
        print("| variation | p50 | p99 | max | finite | excluded |")
        print("|---|---|---|---|---|---|")
        for name, m in out.items():
            print("| %s | %.1e | %.1e | %.1e | %.3f | %.3f |" % (name, m["p50"], m["p99"], m["max"], m["finite"],
                                                              m["excluded"]))


if __name__ == "__main__":
    main()
