"""Writes tests/golden/reference_curves.npz: resampling, dispersion-phase and
window curves generated by the reference's own, unmodified
src/core/processorconfiguration.cpp (compiled into oracle/_ref/ref_curves by
oracle/Makefile from the reference tree).  Run in the build container:

    make -C oracle ref && python tests/golden/make_curve_fixtures.py
"""

import itertools
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE.parents[1] / "oracle" / "_ref" / "ref_curves"

RESAMPLE = [(0.0, 1.0, 0.0, 0.0), (0.5, 2048.0, -100.0, 50.0), (3.25, 1500.5, 200.25, -30.125),
            (-20.0, 2100.0, 10.0, 0.0)]
DISPERSION = [(0.0, 0.0, 1.0, -3.0), (0.1, 2.5, 30.0, -12.5), (0.0, 0.0, 0.0, 0.0)]
WINDOWS = [(t, c, f) for t in range(6) for c, f in [(0.5, 0.95), (0.9, 0.85), (0.3, 1.2), (1.3, 0.5)]]
LENGTHS = [256, 1000, 2048, 4096]


def run(*args):
    out = subprocess.run([str(REF), *map(str, args)], capture_output=True, check=True).stdout
    return np.frombuffer(out, dtype=np.float32).copy()


def main():
    arrays = {}
    for n in LENGTHS:
        for i, c in enumerate(RESAMPLE):
            arrays[f"resample_n{n}_{i}"] = run("resample", n, *c)
        for i, d in enumerate(DISPERSION):
            arrays[f"dispersion_n{n}_{i}"] = run("dispersion", n, *d)
        for i, (t, c, f) in enumerate(WINDOWS):
            arrays[f"window_n{n}_{i}"] = run("window", n, t, c, f)
    arrays["params_resample"] = np.array(RESAMPLE, dtype=np.float32)
    arrays["params_dispersion"] = np.array(DISPERSION, dtype=np.float32)
    arrays["params_window"] = np.array(WINDOWS, dtype=np.float32)
    arrays["lengths"] = np.array(LENGTHS, dtype=np.int32)
    np.savez_compressed(HERE / "reference_curves.npz", **arrays)
    print(f"wrote {len(arrays)} arrays")


if __name__ == "__main__":
    main()
