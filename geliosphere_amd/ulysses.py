"""The reference's Ulysses comparison (Visualization/create_ulysses_Geliosphere_flux.py
and the scripts it calls) run through the drop-in CLI: for every Carrington
rotation of 1994-1998, a Geliosphere (-T) run at Ulysses' position on the
rotation's start day, with that rotation's K0, tilt and polarity taken from
the measured-value table by -m/-y, dt = 1000 s and uniform injection up to
3 GeV (batch_run_geliosphere.py); the model flux is the Ulysses spectrum's
bin from 0.25 GeV (speAvg / 1e7, create_plot.py), to set beside the Ulysses
KET P190 proton flux averaged over the rotation.

    python3 -m geliosphere_amd.ulysses [--millions 1] [--out DIR]
"""
import argparse
import csv
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "gs_cli")
# Per rotation: Ulysses' position and the KET P190 flux, made from the
# reference's data files by data/make_ulysses_table.py.
ROTATIONS = os.path.join(HERE, "data", "ulysses_1994_1998.csv")

SETTINGS = """[default_values]
K0 = 5e22
V = 400.0
dt = 50.0
r_injection = {r}
uniform_energy_injection_maximum = 3.0

[2d_models_common_settings]
theta_injection = {theta}
use_uniform_injection = true

[SolarProp_like_model_settings]
SolarProp_ratio = 0.02

[Geliosphere_model_settings]
Geliosphere_ratio = 0.2
K0_ratio = 1.0
C_delta = 8.7e-5
default_tilt_angle = 0.1

[advanced_settings]
remove_log_files_after_simulation = true
"""


def load_rotations(path=ROTATIONS):
    """(table header, rows) with row = dict(table_row, year, month, day, r, theta, ket_p116, ket_p190)."""
    with open(path) as f:
        header = f.readline()[2:].strip()
        f.readline()
        rows = []
        for line in f:
            table_row, r, theta, p116, p190, _ = line.strip().split(";")
            c = table_row.split(",")
            rows.append(dict(table_row=table_row, cr=int(c[0]), year=int(c[5]), month=int(c[6]), day=int(c[7]),
                             r=float(r), theta=float(theta), ket_p116=float(p116), ket_p190=float(p190)))
    return header, rows


def model_flux(results_dir):
    """speAvg / 1e7 of the Ulysses spectrum's 0.25 GeV row (create_plot.py)."""
    with open(os.path.join(results_dir, "Ulysses.csv")) as f:
        for row in csv.DictReader(f):
            if abs(float(row["Tkin"]) - 0.25) < 1e-9:
                return float(row["speAvg"]) / 1e7
    raise ValueError("no 0.25 GeV row in " + results_dir)


def run(workdir, millions=1, seed=1, cli=CLI, rows=None):
    """Runs every rotation in `workdir` (which gets the table and a link to
    gs_cli, as the reference keeps its tables beside the executable) and
    returns [(row, model flux)]."""
    header, all_rows = load_rotations()
    rows = all_rows if rows is None else rows
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "Geliosphere_K0_phi_table.csv"), "w") as f:
        f.write(header + "\n")
        for r in all_rows:
            f.write(r["table_row"] + "\n")
    app = os.path.join(workdir, "gs_cli")
    if not os.path.exists(app):
        os.symlink(os.path.abspath(cli), app)
    out = []
    for i, r in enumerate(rows):
        st = os.path.join(workdir, "Settings_batch.toml")
        with open(st, "w") as f:
            f.write(SETTINGS.format(r=r["r"], theta=r["theta"]))
        dest = "%d_%d_%d" % (r["year"], r["month"], r["day"])
        res = subprocess.run([app, "-T", "-c", "-s", st, "-d", "1000", "-m", str(r["month"]), "-y", str(r["year"]),
                              "-p", dest, "-N", str(millions), "--seed", str(seed + i)],
                             cwd=workdir, capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError("gs_cli failed for %s: %s" % (dest, res.stderr))
        out.append((r, model_flux(os.path.join(workdir, "results", "Geliosphere2D", dest))))
        print("CR %d %04d-%02d-%02d r %.2f theta %.1f  model %.4e  KET P190 %.4e" %
              (r["cr"], r["year"], r["month"], r["day"], r["r"], r["theta"], out[-1][1], r["ket_p190"]),
              flush=True)
    return out


def summary(out):
    import numpy as np
    m = np.array([f for _, f in out])
    k = np.array([r["ket_p190"] for r, _ in out])
    ratio = m / k
    rm = np.argsort(np.argsort(m)); rk = np.argsort(np.argsort(k))
    spearman = float(np.corrcoef(rm, rk)[0, 1])
    return {"rotations": len(out), "median_ratio_model_to_ket": float(np.median(ratio)),
            "ratio_p10": float(np.percentile(ratio, 10)), "ratio_p90": float(np.percentile(ratio, 90)),
            "spearman": spearman}


def main(argv):
    ap = argparse.ArgumentParser()
    ap.add_argument("--millions", type=int, default=1)
    ap.add_argument("--out", default="ulysses_runs")
    a = ap.parse_args(argv)
    out = run(os.path.abspath(a.out), a.millions)
    print(json.dumps(summary(out)))


if __name__ == "__main__":
    main(sys.argv[1:])
