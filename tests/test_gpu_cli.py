"""End to end through the reference's command-line interface: an HDF5
delta file -> h5healxi (C++ XiEstimator, HDF5Delta, HealpixBins on the
C-ABI) -> .data / .details / subsample files, against the oracle on the same
sightlines; and gpuxi (bruteGPU, FP32 and FP64 binning) on the reference's
sample files."""
import os
import subprocess

import numpy as np
import pytest

import tos_amd
from tos_amd import GridSpec, h5io, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = (a != 0) | (b != 0)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a - b)[m] / np.maximum(np.abs(b[m]), 1e-300)))


@pytest.mark.parametrize("search", ["host", "device"])
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_h5healxi_cartesian(tmp_path, ora, precision, search):
    cat = synth.lya_catalog(150, area_deg2=15.0, seed=2024, omega_lambda=0.73, omega_matter=0.27)
    path = str(tmp_path / "delta.h5")
    h5io.write_h5(cat, path)
    out = str(tmp_path / "xi")
    cmd = [h5io.tool("h5healxi"), "-i", path, "-o", out, "--cart", "--axis2", "[0:200]*50", "--order", "4",
           "--nthreads", "4",
           "--omega-lambda", "0.73", "--omega-matter", "0.27", "--precision", precision, "--save-subsamples"]
    if search == "device":
        cmd.append("--device-search")
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    grid = GridSpec.cartesian()
    ref = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=4, per_skybin=True)
    raw = ref["raw"]
    data = np.loadtxt(out + ".data")
    det = np.loadtxt(out + ".details")
    assert data.shape == (grid.nbins_total, 2) and det.shape == (grid.nbins_total, 10)
    m = raw["wgt"] > 0
    xi_ref = np.where(m, raw["didj"] / np.where(m, raw["wgt"], 1), 0)
    assert rel_err(data[:, 1], xi_ref) <= 1e-9
    assert np.array_equal(det[:, 8].astype(np.int64), raw["num_pairs"])
    assert rel_err(det[:, 7], raw["wgt"]) <= 1e-9
    # bin centres of the 50 x 50 x 1 grid, axis 0 slowest
    b = np.arange(grid.nbins_total)
    assert np.allclose(det[:, 1], 4 * (b // 50) + 2) and np.allclose(det[:, 2], 4 * (b % 50) + 2)
    # one subsample file per occupied sky bin
    keys = np.asarray(ref["skybin_keys"])
    per = ref["per_skybin"]
    for k, key in enumerate(keys):
        sub = np.loadtxt(f"{out}-{int(key)}.data")
        assert np.array_equal(sub[:, 3].astype(np.int64), per["num_pairs"][k])
        assert rel_err(sub[:, 2], per["wgt"][k]) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_h5healxi_observing(tmp_path, ora, precision):
    """h5healxi with the Observing (QuasarGrid) grid, the CLI default when
    neither --cart nor --polar is given: Delta loglam x theta (radians)."""
    cat = synth.lya_catalog(150, area_deg2=15.0, seed=2026, omega_lambda=0.73, omega_matter=0.27)
    path = str(tmp_path / "delta.h5")
    h5io.write_h5(cat, path)
    out = str(tmp_path / "xi")
    cmd = [h5io.tool("h5healxi"), "-i", path, "-o", out, "--axis1", "[0:0.02]*20", "--axis2", "[0:0.06]*6",
           "--order", "4", "--nthreads", "4", "--omega-lambda", "0.73", "--omega-matter", "0.27",
           "--precision", precision]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    grid = GridSpec.observing("[0:0.02]*20", "[0:0.06]*6")
    raw = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=4)["raw"]
    assert raw["num_pairs"].sum() > 0
    data = np.loadtxt(out + ".data")
    det = np.loadtxt(out + ".details")
    assert data.shape == (grid.nbins_total, 2)
    m = raw["wgt"] > 0
    xi_ref = np.where(m, raw["didj"] / np.where(m, raw["wgt"], 1), 0)
    # xi relative to max(|xi|, 1e-4): a bin whose signed sum cancels is
    # compared against the cancellation floor
    assert float(np.max(np.abs(data[:, 1] - xi_ref) / np.maximum(np.abs(xi_ref), 1e-4))) <= 1e-9
    assert np.array_equal(det[:, 8].astype(np.int64), raw["num_pairs"])
    assert rel_err(det[:, 7], raw["wgt"]) <= 1e-9


def run_gpuxi(tmp_path, name, *flags):
    src = os.path.join(HERE, "golden", name)
    out = str(tmp_path / "xi.dat")
    r = subprocess.run([h5io.tool("gpuxi"), "-i", src, "-o", out, *flags],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return np.loadtxt(out), np.loadtxt(src, usecols=(0, 1, 2, 3, 4)).T


def test_gpuxi_fp64_nano_sample(tmp_path, ora):
    got, cols = run_gpuxi(tmp_path, "delta-nano.dat", "--axis1", "[0:200]*50", "--fp64")
    ds, ws, _ = ora.brute3d_f64(*cols, 0.0, 200.0, 50)
    xi_ref = np.where(ws > 0, ds / np.where(ws > 0, ws, 1), 0)
    assert got.shape == (50, 2)
    # "index xi" at the stream's default 6 significant digits (gpuxi.cc)
    assert rel_err(got[:, 1], xi_ref) <= 5e-6


@pytest.mark.parametrize("name,axis", [("mock-10000.dat", "[0:200]*50"), ("delta-nano.dat", "[0:100]*50")])
def test_gpuxi_reference_fp32_binning(tmp_path, ora, name, axis):
    """Default bruteGPU: the reference kernel's FP32 binning
    (gpu/brute.cu:26-55), xi from slots 1..nbins (brute.cu:184-212),
    against the oracle's restatement of that kernel."""
    got, cols = run_gpuxi(tmp_path, name, "--axis1", axis)
    lo, hi = (float(v) for v in axis[1:axis.index("]")].split(":"))
    ds, ws, _ = ora.brute3d_f32(*cols, lo, hi, 50)
    xi_ref = np.where(ws[1:51] > 0, ds[1:51] / np.where(ws[1:51] > 0, ws[1:51], 1), 0)
    assert got.shape == (50, 2)
    assert rel_err(got[:, 1], xi_ref) <= 5e-6


def test_h5healxi_covariance(tmp_path):
    """h5healxi --cov: the .cov file (upper triangle, row <= col) against the
    weighted covariance recomputed from the subsample files it writes."""
    cat = synth.lya_catalog(150, area_deg2=15.0, seed=2025, omega_lambda=0.73, omega_matter=0.27)
    path = str(tmp_path / "delta.h5")
    h5io.write_h5(cat, path)
    out = str(tmp_path / "xi")
    cmd = [h5io.tool("h5healxi"), "-i", path, "-o", out, "--polar", "--axis1", "[0:200]*10", "--axis2", "[0:1]*5",
           "--order", "4", "--nthreads", "4", "--omega-lambda", "0.73", "--omega-matter", "0.27",
           "--save-subsamples", "--cov"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import glob
    subs = sorted(glob.glob(out + "-*.data"))
    assert len(subs) > 1
    s = np.stack([np.loadtxt(f) for f in subs])          # [k, nbins, 4]: index xi wgt num_pairs
    x, g = s[:, :, 1], s[:, :, 2]
    w = np.where(g > 0, g, 0).sum(axis=1)
    W = w.sum()
    mu = (w[:, None] * x).sum(axis=0) / W
    ref = np.einsum("k,ka,kb->ab", w, x - mu, x - mu) / W
    c = np.loadtxt(out + ".cov")
    nb = x.shape[1]
    assert c.shape == (nb * (nb + 1) // 2, 3)
    rows, cols = c[:, 0].astype(int), c[:, 1].astype(int)
    assert (rows <= cols).all()
    scale = np.abs(ref).max()
    assert np.abs(c[:, 2] - ref[rows, cols]).max() <= 1e-9 * scale
    # the diagnostics the reference prints (XiEstimator.cc:82-87), for the
    # matrix the .cov file holds
    full = np.zeros((nb, nb))
    full[rows, cols] = c[:, 2]
    full[cols, rows] = c[:, 2]
    pd, ld = tos_amd.covariance_check(full)
    assert f"is pos def: {int(pd)}" in r.stdout
    if pd:
        printed = float(r.stdout.split("log(|C|): ")[1].split()[0])
        assert abs(printed - ld) <= 1e-9 * max(1.0, abs(ld))
    else:
        assert "log(|C|)" not in r.stdout


def test_h5healxi_plates_device_search_equals_host(tmp_path):
    """PlateBins through the C++ drop-in: the GPU candidate search
    (--device-search) gives the host enumeration's histograms and
    subsamples (counts equal, sums within 1e-10)."""
    cat = synth.lya_catalog(150, area_deg2=15.0, seed=2025, omega_lambda=0.73, omega_matter=0.27)
    pol, pid, pra, pdec = synth.plate_tiling(cat, spacing_deg=1.5)
    path = str(tmp_path / "delta.h5")
    h5io.write_h5(cat, path, plates=pol)
    plist = str(tmp_path / "plates.txt")
    with open(plist, "w") as f:
        for i, r, d in zip(pid, np.degrees(pra), np.degrees(pdec)):
            f.write(f"{int(i)} {r:.9g} {d:.9g}\n")
    outs = {}
    for search in ("host", "device"):
        out = str(tmp_path / f"xi_{search}")
        cmd = [h5io.tool("h5healxi"), "-i", path, "-o", out, "--cart", "--axis2", "[0:200]*50", "--order", "0",
               "--platelist-filename", plist, "--nthreads", "4", "--omega-lambda", "0.73", "--omega-matter", "0.27",
               "--save-subsamples"] + (["--device-search"] if search == "device" else [])
        print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs[search] = out
    a, b = np.loadtxt(outs["host"] + ".details"), np.loadtxt(outs["device"] + ".details")
    assert np.array_equal(a[:, 8], b[:, 8]) and a[:, 8].sum() > 0
    assert rel_err(b[:, 7], a[:, 7]) <= 1e-10
    assert rel_err(np.loadtxt(outs["device"] + ".data")[:, 1], np.loadtxt(outs["host"] + ".data")[:, 1]) <= 1e-10
    subs = {os.path.basename(f)[len("xi_host"):] for f in os.listdir(tmp_path) if f.startswith("xi_host-")}
    assert subs and subs == {os.path.basename(f)[len("xi_device"):] for f in os.listdir(tmp_path)
                             if f.startswith("xi_device-")}
    for suffix in subs:
        sa, sb = np.loadtxt(outs["host"] + suffix), np.loadtxt(outs["device"] + suffix)
        assert np.array_equal(sa[:, 3], sb[:, 3])
