"""HDF5 delta files through the C++ tools: write a catalogue in the layout
HDF5Delta::loadForests reads (tos_mkdelta), read one back with the loader's
semantics (tos_h5dump), and run the h5healxi / gpuxi drivers."""
from __future__ import annotations

import os
import subprocess

import numpy as np

from ._lib import PKG_DIR
from .synth import Catalog

BIN = os.path.join(PKG_DIR, "_build", "bin")


def tool(name: str) -> str:
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
    return path


def write_dump(cat: Catalog, path: str, plates=None, value=None, weight=None):
    n, p = cat.n_los, cat.n_pix
    z = cat.zq if cat.zq is not None else np.zeros(n)
    plate = np.zeros(n, np.int32) if plates is None else np.asarray(plates, np.int32)
    with open(path, "wb") as f:
        np.array([n, p], np.int64).tofile(f)
        np.asarray(cat.offsets, np.int64).tofile(f)
        for a in (cat.ra, cat.dec, z):
            np.asarray(a, np.float64).tofile(f)
        plate.tofile(f)
        for a in (cat.value if value is None else value, cat.loglam,
                  cat.weight if weight is None else weight, cat.distance):
            np.asarray(a, np.float32).tofile(f)


def read_dump(path: str):
    with open(path, "rb") as f:
        n, p = np.fromfile(f, np.int64, 2)
        off = np.fromfile(f, np.int64, n + 1)
        ra, dec, z = (np.fromfile(f, np.float64, n) for _ in range(3))
        plate = np.fromfile(f, np.int32, n)
        v, l, w, d = (np.fromfile(f, np.float32, p) for _ in range(4))
    return Catalog(ra, dec, off, v, l, w, d), plate


def write_h5(cat: Catalog, path: str, plates=None, value=None, weight=None):
    dump = path + ".bin"
    write_dump(cat, dump, plates, value, weight)
    subprocess.run([tool("tos_mkdelta"), dump, path], check=True, capture_output=True)
    os.remove(dump)


def read_h5(path: str):
    dump = path + ".read.bin"
    subprocess.run([tool("tos_h5dump"), path, dump], check=True, capture_output=True)
    out = read_dump(dump)
    os.remove(dump)
    return out
