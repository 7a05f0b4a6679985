"""Writes the golden vectors of tests/golden/ with the oracle (a C
restatement of the reference algorithm, pinned by tests/test_oracle_pins.py):

  brute3d.npz   the 3D brute force (old/BinPair.h:28-47 in FP64 and
                gpu/brute.cu:26-55 in FP32) on the reference's own sample
                files delta-nanoer / delta-nano / delta-micro / mock-10000,
                50 bins over [0, 100) Mpc/h
  c1mini.npz    a 100-quasar Lyman-alpha catalogue of the seeded generator
                (inputs stored) and its Cartesian, Polar and Observing
                histograms with every XiBin field, HEALPix order 4

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "turbo-octo-spice_amd"))
sys.path.insert(0, ROOT)

SAMPLES = ["delta-nanoer", "delta-nano", "delta-micro", "mock-10000"]
C0_RANGE = (0.0, 100.0, 50)
C1MINI = dict(n_qso=100, area_deg2=10.0, seed=2101)
GRIDS = {"cartesian": ("[0:200]*50", "[0:200]*50"), "polar": ("[0:200]*50", "[0:1]*20"),
         "observing": ("[0:0.02]*20", "[0:0.06]*6")}


def brute3d(ora):
    out = {}
    for name in SAMPLES:
        cols = np.loadtxt(os.path.join(HERE, name + ".dat"), usecols=(0, 1, 2, 3, 4)).T
        for mode, fn in (("f64", ora.brute3d_f64), ("f32", ora.brute3d_f32)):
            ds, ws, cn = fn(*cols, *C0_RANGE)
            out[f"{name}/{mode}/dsum"] = ds
            out[f"{name}/{mode}/wsum"] = ws
            out[f"{name}/{mode}/count"] = cn
    return out


def c1mini(ora):
    from tos_amd import GridSpec, synth
    cat = synth.lya_catalog(C1MINI["n_qso"], area_deg2=C1MINI["area_deg2"], seed=C1MINI["seed"])
    out = {"ra": cat.ra, "dec": cat.dec, "offsets": cat.offsets, "value": cat.value, "loglam": cat.loglam,
           "weight": cat.weight, "distance": cat.distance, "scale": np.float64(cat.scale)}
    for gname, (a1, a2) in GRIDS.items():
        grid = getattr(GridSpec, gname)(a1, a2)
        r = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=4)
        for f in r["raw"].dtype.names:
            out[f"{gname}/{f}"] = r["raw"][f]
        for k, v in r["stats"].items():
            out[f"{gname}/stats/{k}"] = np.int64(v)
    return out


def main():
    from oracle import oracle as ora
    ora.lib()
    np.savez_compressed(os.path.join(HERE, "brute3d.npz"), **brute3d(ora))
    np.savez_compressed(os.path.join(HERE, "c1mini.npz"), **c1mini(ora))
    print("wrote brute3d.npz and c1mini.npz")


if __name__ == "__main__":
    main()
