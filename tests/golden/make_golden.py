"""Writes the golden fixtures of tests/golden/ from the reference library
built in place (oracle/_ref, compiled from /root/reference by
oracle/Makefile).  The reference ships no test vectors of its own, so these
are generated here: ISAAC streams for string and integer seeds, and a few
tiny reference renders (raw image bytes) with the genome XML that made them.

    python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyref  # noqa: E402
from flam3_amd import abi  # noqa: E402
from flam3_amd import genome as G  # noqa: E402

RENDERS = {
    "small_q10": (lambda: G.synthetic_small(seed=11, width=48, height=36, quality=10.0), {}),
    "ss2_final_q8": (lambda: G.synthetic_1080p(seed=12, width=40, height=24, quality=8.0, supersample=2), {}),
    "b64_rgba16": (lambda: G.synthetic_small(seed=13, width=32, height=24, quality=12.0),
                   dict(bits=64, bpc=2, nchan=4, transp=1)),
}


def isaac_streams():
    L = pyref.ref("strict")
    out = {}
    for seed in ["flam3", "abc", "cfg1", "x" * 200]:
        rc = pyref.seeded_rc(seed)
        out[seed] = [int(L.ref_irand(ctypes.byref(rc))) for _ in range(100)]
    # time()-style seed: every randrsl word the same integer (flam3.c:2503-2506)
    rc = abi.randctx()
    for k in range(16):
        rc.randrsl[k] = 1234567890
    L.irandinit(ctypes.byref(rc), 1)
    out["int:1234567890"] = [int(L.ref_irand(ctypes.byref(rc))) for _ in range(100)]
    return out


def main():
    with open(os.path.join(HERE, "isaac_streams.json"), "w") as f:
        json.dump(isaac_streams(), f)
    for name, (mk, kw) in RENDERS.items():
        g = mk()
        xml = G.flame_document([g])
        img, st = pyref.render(xml, seed="golden", **kw)
        with open(os.path.join(HERE, name + ".flam3"), "w") as f:
            f.write(xml)
        np.save(os.path.join(HERE, name + ".npy"), img)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump({"settings": kw, "seed": "golden", "badvals": st["badvals"], "num_iters": st["num_iters"],
                       "shape": list(img.shape), "dtype": str(img.dtype)}, f)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
