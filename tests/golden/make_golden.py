"""Writes tests/golden/oracle_c1_c2.npz: the FP64 oracle's records for 256
seeded pixels of each of the two frame configurations (Schwarzschild
320x180 and Kerr a=0.9 1920x1080, the fixed-fraction step rule), with the
same-algorithm RK4 and with adaptive Tsit5 at tol 1e-10.

The reference ships no golden vectors for its tracer and cannot run here,
so these are the oracle's own outputs, frozen: a regression guard for the
oracle (tests/test_oracle_golden.py) and a fixed target for the GPU
(tests/test_gpu_golden.py).  Regenerate with `python tests/golden/make_golden.py`
after a deliberate change to the oracle."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

from spacetime_amd import scenes  # noqa: E402
from util import oracle_scene  # noqa: E402

N = 256


def pixels(sc, seed):
    rng = np.random.default_rng(seed)
    idx = rng.choice(sc.width * sc.height, N, replace=False)
    return np.stack([idx % sc.width, idx // sc.width], 1).astype(np.int32)


def main():
    out = {}
    for name, sc, seed in (("c1", scenes.schwarzschild_320x180(), 101), ("c2", scenes.kerr09_1080p(), 202)):
        o = oracle_scene(sc)
        ij = pixels(sc, seed)
        out[f"{name}_ij"] = ij
        out[f"{name}_rk4"] = o.trace_list(ij, "rk4", nthreads=8)
        out[f"{name}_tsit5"] = o.trace_list(ij, "tsit5", 1e-10, nthreads=8)
    np.savez_compressed(HERE / "oracle_c1_c2.npz", **out)
    print("wrote", HERE / "oracle_c1_c2.npz")


if __name__ == "__main__":
    main()
