"""The drop-in GPU models end to end (gs_cli -> adapter classes -> C ABI).

A run must leave what the reference's GPU model leaves
(CUDAKernel/src/OneDimensionBpModel.cu:149-200 and its siblings):
results/<MODEL>/<destination>/ with log.dat (one line per event in the
model's format), simulation_log.txt, and the Results pass's spectrum files;
plus gs_histogram.dat.  log.dat must hold exactly the events gs_run_plain
returns for the same parameter block, and the histogram its spectrum.
"""
import json
import math
import os
import subprocess

import pytest

from geliosphere_amd import abi
from geliosphere_amd.engine import default_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "geliosphere_amd", "gs_cli")
M0, Q, C = 1.67261e-27, 1.60219e-19, 2.99793e8
T0W = M0 * C * C


def cli(args, cwd):
    print("gs_cli", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def params_of(args):
    """The parameter block gs_cli runs with, rebuilt through the C ABI."""
    j = json.loads(subprocess.run([CLI, *args, "--print-params"], capture_output=True, text=True,
                                  check=True).stdout)
    p = default_params(j["model"])
    for k in ("injection", "flags", "K0", "V", "dt", "r_inj", "theta_inj", "tilt", "polarity", "ratio",
              "delta0", "T_min", "T_max", "grid_n", "n_particles", "seed", "n_bins", "n_weights"):
        setattr(p, k, j[k])
    for i, w in enumerate(j["weights"]):
        p.weights[i] = w
    return p


def p_si(T):
    tw = T * 1e9 * Q
    return math.sqrt(tw * (tw + 2 * T0W)) / Q * Q / C


def log_w(p):
    return (math.pow(M0 * M0 * C ** 4 + p * p * C * C, -1.85) / p) / 1e45


def close(a, b):
    return abs(a - b) <= 6e-6 * max(abs(a), abs(b)) + 1e-300


def check_histogram(path, spec):
    rows = [l.split() for l in open(path) if not l.startswith("#")]
    assert len(rows) == spec.n_bins
    for b, r in enumerate(rows):
        assert int(r[2]) == spec.count[b]
        for k in range(spec.n_weights):
            assert float(r[3 + 4 * k]) == spec.sum_w[k][b]


@pytest.mark.parametrize("flag,model,n,expect_lines", [
    ("-B", abi.MODEL_1D_BP, 20000, None),
    ("-T", abi.MODEL_2D_GELIOSPHERE, 1500, None),
    ("-E", abi.MODEL_2D_SOLARPROP, 1500, None),
    ("-F", abi.MODEL_1D_FP, 300000, None),
])
def test_cli_run_matches_engine(engine, tmp_path, flag, model, n, expect_lines):
    args = [flag, "--particles", str(n), "--seed", "11", "-p", "run1", "--keep-log"]
    if model in (abi.MODEL_1D_BP, abi.MODEL_1D_FP):
        args += ["-K", "5e22", "-V", "400", "-d", "500"]
    cli(args, tmp_path)
    d = tmp_path / "results" / {0: "1DFP", 1: "1DBP", 2: "SOLARPROPLike", 3: "Geliosphere2D"}[model] / "run1"
    for f in ("log.dat", "simulation_log.txt", "gs_histogram.dat"):
        assert (d / f).exists(), f
    outputs = ["Usoskin_2005.dat", "Weber.dat", "Burger.dat", "Ulysses.dat", "Ams.dat"] if model >= 2 else \
        ["output_1e2bin.dat", "output_1e3bin.dat", "output_4e2bin.dat", "output_logbin.dat"]
    for f in outputs:
        assert (d / f).exists(), f
    p = params_of(args)
    spec, recs, n_ev = engine.run_plain(p, records=True, rec_cap=max(16, 2 * n))
    check_histogram(d / "gs_histogram.dat", spec)
    lines = open(d / "log.dat").read().splitlines()
    assert len(lines) == n_ev == len(recs)
    for line, e in zip(lines, recs):
        v = [float(x) for x in line.split()]
        if model == abi.MODEL_1D_BP:
            want = [e.t_exit, e.t_inj, e.r, log_w(p_si(e.t_exit))]
        elif model == abi.MODEL_1D_FP:
            want = [p_si(e.t_inj), p_si(e.t_exit), e.r, log_w(p_si(e.t_inj)), e.sumac]
        else:
            want = [e.t_inj, e.t_exit, e.r, log_w(p_si(e.t_inj)), p.theta_inj, e.theta]
        assert len(v) == len(want) and all(close(a, b) for a, b in zip(v, want)), (line, want)
    report = (d / "simulation_log.txt").read_text()
    assert report.startswith("Model: %s\n" % abi.MODEL_NAMES[model]) and "Seed: 11" in report


def test_cli_log_removed_by_default(engine, tmp_path):
    cli(["-B", "--particles", "2000", "--seed", "2", "-p", "r"], tmp_path)
    d = tmp_path / "results" / "1DBP" / "r"
    assert not (d / "log.dat").exists() and (d / "output_1e2bin.dat").exists()


KEEP_LOG_SETTINGS = """[default_values]
K0 = 5e22
V = 400.0
dt = 50.0
r_injection = 1.0
uniform_energy_injection_maximum = 101.0

[2d_models_common_settings]
theta_injection = 90.0
use_uniform_injection = true

[SolarProp_like_model_settings]
SolarProp_ratio = 0.02

[Geliosphere_model_settings]
Geliosphere_ratio = 0.2
K0_ratio = 1.0
C_delta = 8.7e-5
default_tilt_angle = 0.1

[advanced_settings]
remove_log_files_after_simulation = false
"""

REFERENCE_FILES = {0: ["output_1e2bin.dat", "output_1e3bin.dat", "output_4e2bin.dat", "output_logbin.dat"],
                   1: ["output_1e2bin.dat", "output_1e3bin.dat", "output_4e2bin.dat", "output_logbin.dat"],
                   2: ["Usoskin_2005.dat", "Weber.dat", "Burger.dat", "Ulysses.dat", "Ams.dat"],
                   3: ["Usoskin_2005.dat", "Weber.dat", "Burger.dat", "Ulysses.dat", "Ams.dat"]}


def test_cli_batch_rows(engine, tmp_path):
    """gs_cli -b runs every row in one gs_run_batch_records call and leaves in
    each row's directory what the row's own run leaves (the reference's
    BatchRun runs each row's Algorithm, Algorithm/src/BatchRun.cpp:50-78):
    log.dat (kept here by the row's settings file), the Results pass's
    spectrum files, simulation_log.txt and gs_histogram.dat.  log.dat holds
    exactly the records gs_run_plain returns for the row's parameter block,
    and the histogram its spectrum."""
    keep = tmp_path / "keep.toml"
    keep.write_text(KEEP_LOG_SETTINGS)
    f = tmp_path / "batch.csv"
    f.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,1,1.0,90,%s,rowA,1D Bp\n"
                 "-1,-1,1e22,600,100,1,5.0,90,%s,rowB,1D Fp\n"
                 "-1,-1,5e22,,50,1,1.0,60,%s,rowC,2D Geliosphere\n"
                 "-1,-1,5e22,,50,1,2.0,120,,rowD,2D SolarProp-like\n" % (keep, keep, keep))
    args = ["-b", str(f), "--particles", "3000", "--seed", "40"]
    prepared = [json.loads(l) for l in subprocess.run([CLI, *args, "--print-params"], capture_output=True,
                                                      text=True, check=True).stdout.splitlines()]
    assert [r["destination"] for r in prepared] == ["rowA", "rowB", "rowC", "rowD"]
    cli(args, tmp_path)
    for r in prepared:
        model = r["model"]
        d = tmp_path / "results" / {0: "1DFP", 1: "1DBP", 2: "SOLARPROPLike", 3: "Geliosphere2D"}[model] / \
            r["destination"]
        for name in REFERENCE_FILES[model] + ["gs_histogram.dat", "simulation_log.txt"]:
            assert (d / name).exists(), (r["destination"], name)
        assert (d / "log.dat").exists() == (r["destination"] != "rowD")   # rowD: batch defaults remove it
        assert "batched" in (d / "simulation_log.txt").read_text()
        p = default_params(model)
        for k in ("injection", "flags", "K0", "V", "dt", "r_inj", "theta_inj", "tilt", "polarity", "ratio",
                  "delta0", "T_min", "T_max", "grid_n", "n_particles", "seed", "n_bins", "n_weights"):
            setattr(p, k, r[k])
        for i, w in enumerate(r["weights"]):
            p.weights[i] = w
        spec, recs, n_ev = engine.run_plain(p, records=True, rec_cap=max(16, 2 * p.n_particles))
        check_histogram(d / "gs_histogram.dat", spec)
        if r["destination"] != "rowD":
            lines = open(d / "log.dat").read().splitlines()
            assert len(lines) == n_ev
            for line, e in zip(lines, recs):
                v = [float(x) for x in line.split()]
                t_col = {1: 1, 0: None, 2: 0, 3: 0}[model]
                if t_col is not None:
                    assert close(v[t_col], e.t_inj), (line, e.t_inj)
