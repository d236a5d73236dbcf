"""Host side of the drop-in GPU models, on the CPU (no device calls).

* ParamsCarrier (standalone build) behaves like the reference's
  (Input/include/ParamsCarrier.hpp:158-300);
* gs_cli maps the reference's flags and carrier keys onto the parameter
  block the way the reference sets its constants
  (CUDAKernel/src/CosmicConstants.cu:43-111, Input/src/InputValidation.cpp:57-118):
  float arithmetic, K0 used for 1D models only when entered by the user,
  Geliosphere K0 times K0_ratio, V sign for B-p, particle counts rounded up
  to whole 4096 x 256 launches;
* the Results passes (--evaluation) bin and weight log.dat lines and write
  the spectrum files in the reference's formats
  (Algorithm/src/*Results.cpp, Utils/src/ResultsUtils.cpp).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "geliosphere_amd", "gs_cli")

M0, Q, C = 1.67261e-27, 1.60219e-19, 2.99793e8
T0W = M0 * C * C
T0 = M0 * C * C / (Q * 1e9)
f32 = np.float32


def cli(*args, cwd=None, check=True):
    if not os.path.exists(CLI):
        pytest.fail("gs_cli is not built; run __graft_entry__.build()")
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError("gs_cli %s failed (%d): %s" % (" ".join(args), r.returncode, r.stderr))
    return r


def params(*args):
    return json.loads(cli(*args, "--print-params").stdout)


def test_params_carrier_semantics(tmp_path):
    exe = tmp_path / "pc_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "geliosphere_amd", "host", "compat"),
                    os.path.join(ROOT, "tests", "cpp", "params_carrier_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr


def conv_k0(k_cm2s):
    return float(f32(float(f32(k_cm2s)) * 4.4683705e-27))


def conv_v(v_kms):
    return float(f32(float(f32(v_kms)) * 6.68458712e-9))


def deg_f32(x):
    return float((f32(x) * f32(3.1415926535)) / f32(180.0))


def test_cli_maps_1d_backward_like_the_reference():
    p = params("-B", "-K", "5e22", "-V", "400", "-d", "500", "--seed", "7")
    assert p["model"] == 1 and p["seed"] == 7
    assert p["K0"] == conv_k0(5e22)
    assert p["V"] == conv_v(400)           # magnitude; the engine applies -V for B-p
    assert p["dt"] == 500.0
    # 1D injection: the 1e-4 GeV grid up to 101 GeV (CosmicUtils.cu:17-34)
    assert p["injection"] == 1 and p["grid_n"] == 1010000 and p["T_min"] == 1e-4
    assert p["n_particles"] == 1 << 20     # 1 million rounded up to one launch
    assert p["weights"][:2] == [1, 0] and p["n_weights"] == 2 and p["n_bins"] == 30


SETTINGS = """# the reference's schema (Input/src/TomlSettings.cpp), its shipped values
[default_values]
K0 = 5e22
V = 400.0
dt = 50.0
r_injection = 1.0
uniform_energy_injection_maximum = 101.0

[2d_models_common_settings]
theta_injection = 90.0
use_uniform_injection = false

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


def settings_file(tmp_path, **changes):
    text = SETTINGS
    for k, v in changes.items():
        lines = [("%s = %s" % (k, v)) if l.split("=")[0].strip() == k else l for l in text.splitlines()]
        text = "\n".join(lines) + "\n"
    f = tmp_path / "settings.toml"
    f.write_text(text)
    return str(f)


def test_cli_1d_k0_only_when_entered_by_user(tmp_path):
    # Without -K the 1D GPU models keep their built-in K0 (CosmicConstants.cu:16,69-73)
    assert params("-B", "--seed", "1")["K0"] == float(f32(0.000222))
    assert params("-F", "--seed", "1")["K0"] == float(f32(0.000222))
    # no settings file: the device defaults, dt 5 s and V 2.66667e-6 AU/s (CosmicConstants.cu:15,74-83)
    p = params("-F", "--seed", "1")
    assert p["dt"] == 5.0 and p["V"] == float(f32(2.66667e-6))
    # the shipped settings: dt 50 s, V 400 km/s
    p = params("-F", "--seed", "1", "-s", settings_file(tmp_path))
    assert p["dt"] == 50.0 and p["V"] == conv_v(400)


def test_settings_file_values_and_errors(tmp_path):
    f = settings_file(tmp_path, dt="500.0", theta_injection="45.0", use_uniform_injection="true",
                      Geliosphere_ratio="0.25", K0_ratio="2.0", C_delta="1e-4", default_tilt_angle="20.0")
    g = params("-T", "--seed", "1", "-s", f)
    assert g["dt"] == 500.0 and g["theta_inj"] == deg_f32(45.0)
    assert g["injection"] == 1                                  # uniform grid injection
    assert g["ratio"] == float(f32(0.25)) and g["delta0"] == float(f32(1e-4)) and g["tilt"] == deg_f32(20.0)
    assert g["K0"] == float(f32(f32(conv_k0(5e22)) * f32(2.0)))
    # a missing key is an error; a missing file only a warning
    bad = tmp_path / "bad.toml"
    bad.write_text(SETTINGS.replace("dt = 50.0\n", ""))
    r = cli("-B", "-s", str(bad), "--print-params", check=False)
    assert r.returncode != 0 and "default_values.dt" in r.stderr
    r = cli("-B", "-s", str(tmp_path / "none.toml"), "--print-params")
    assert "No settings file exists on entered path" in r.stderr


def test_cli_2d_models(tmp_path):
    # no settings: K0 stays the device default (no K0_default to scale)
    assert params("-T", "--seed", "3")["K0"] == float(f32(0.000222))
    g = params("-T", "--seed", "3", "--theta", "45", "--tilt", "10", "--polarity", "-1", "-s", settings_file(tmp_path))
    assert g["model"] == 3 and g["injection"] == 2          # SolarProp energies
    assert g["K0"] == float(f32(f32(conv_k0(5e22)) * f32(1.0)))   # settings K0 x K0_ratio
    assert g["theta_inj"] == deg_f32(45) and g["tilt"] == deg_f32(10) and g["polarity"] == -1.0
    assert g["ratio"] == float(f32(0.2)) and g["delta0"] == float(f32(8.7e-5))
    assert g["weights"][:3] == [2, 3, 4]
    g_user = params("-T", "--seed", "3", "-K", "1e22")
    assert g_user["K0"] == conv_k0(1e22)                      # user K0 is not scaled
    e = params("-E", "--seed", "3", "--tilt", "30")
    assert e["tilt"] == deg_f32(30) and e["flags"] == 0 and e["ratio"] == float(f32(0.02))
    eg = params("-E", "--seed", "3", "--tilt", "30", "--solarprop-gpu-compat")
    assert eg["tilt"] == deg_f32(5.75) and eg["flags"] == 1


def test_cli_particle_counts_and_range_checks():
    assert params("-B", "-N", "3", "--seed", "1")["n_particles"] == 3 * (1 << 20)
    assert params("-B", "--particles", "777", "--seed", "1")["n_particles"] == 777
    assert cli("-B", "-V", "50", check=False).returncode != 0         # V range 100-1500 km/s
    assert cli("-B", "-d", "1", check=False).returncode != 0          # dt range 3-5000 s
    assert cli("-B", "-N", "0", check=False).returncode != 0
    assert cli(check=False).returncode != 0                           # no model


def test_batch_csv_rows(tmp_path):
    f = tmp_path / "batch.csv"
    f.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,runA,1D Bp\n"
                 "2010,5,,,,2,30,45,,\"run B\",2D Geliosphere\n")
    rows = [json.loads(l) for l in cli("-b", str(f), "--seed", "7", "--print-params").stdout.splitlines()]
    assert [r["row"] for r in rows] == [1, 2] and [r["destination"] for r in rows] == ["runA", "run B"]
    assert rows[0]["model"] == 1 and rows[0]["dt"] == 500 and rows[0]["K0"] == pytest.approx(conv_k0(5e22))
    assert rows[0]["n_particles"] == 1 << 20 and rows[1]["n_particles"] == 2 << 20
    assert rows[1]["model"] == 3 and rows[1]["r_inj"] == 30 and rows[1]["theta_inj"] == pytest.approx(deg_f32(45))
    assert [r["seed"] for r in rows] == [7, 8]          # one particle stream per row
    assert all(r["injection"] == 1 for r in rows)       # the batch defaults: uniform injection
    f.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,0,1.0,90,,bad,1D Bp\n-1,-1,5e22,400,500,1,1.0,90,,ok,1D Fp\n")
    r = cli("-b", str(f), "--seed", "7", "--print-params")
    assert [json.loads(l)["destination"] for l in r.stdout.splitlines()] == ["ok"] and "skipping" in r.stderr


def test_model_flags_exclude_each_other():
    """-F, -B, -E, -T, --custom-model and -b exclude each other
    (Input/src/ParseParams.cpp:54-59); the run does not start."""
    for a, b in [("-B", "-F"), ("-E", "-T"), ("-F", "-T"), ("-B", "--custom-model")]:
        args = [a, b] if b != "--custom-model" else [a, b, "1D Bp"]
        r = cli(*args, "--print-params", check=False)
        assert r.returncode != 0 and "excludes" in r.stderr and r.stdout == ""
    r = cli("-b", "x.csv", "-B", "--print-params", check=False)
    assert r.returncode != 0 and "excludes" in r.stderr


def test_custom_model_resolves_through_the_factory():
    """--custom-model NAME runs what the reference's factory returns for NAME
    (Factory/src/CosmicFactory.cpp:8-31): a model name selects that model;
    an unknown name is an error (main.cpp:34-39)."""
    assert params("--custom-model", "2D Geliosphere", "--seed", "1")["model"] == 3
    assert params("--custom-model", "1D Fp", "--seed", "1")["model"] == 0
    r = cli("--custom-model", "My model", "--print-params", check=False)
    assert r.returncode != 0 and "not defined in factory" in r.stderr


def test_cpu_only_is_rejected_with_a_message():
    """--cpu-only selects the reference's CPU models (ParseParams.cpp:36,
    136-139), which the GPU drop-in does not provide."""
    r = cli("-B", "--cpu-only", "--print-params", check=False)
    assert r.returncode != 0 and "--cpu-only" in r.stderr and r.stdout == ""


def test_batch_duplicate_names_reject_the_batch(tmp_path):
    """Non-blank names must be unique (BatchRun.cpp:344-365); blank and empty
    names may repeat.  A row with model "batch run" is skipped."""
    f = tmp_path / "dup.csv"
    f.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,same,1D Bp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,,1D Bp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,same,1D Fp\n")
    r = cli("-b", str(f), "--print-params", check=False)
    assert r.returncode != 0 and "duplicated values" in r.stderr and r.stdout == ""
    f.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,a,1D Bp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,  ,1D Bp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,  ,1D Bp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,,1D Fp\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,,1D Fp\n")
    assert len(cli("-b", str(f), "--print-params").stdout.splitlines()) == 5


# ---------------------------------------------------------------- Results passes

def p_si(T):
    tw = T * 1e9 * Q
    return math.sqrt(tw * (tw + 2 * T0W)) / Q * Q / C


def yamada(p):
    return math.exp(-1.85 * math.log(M0 * M0 * C ** 4 + p * p * C * C)) / p


def read_rows(path, skip_header=False):
    lines = open(path).read().splitlines()
    if skip_header:
        lines = lines[1:]
    return [[float(v) for v in l.replace(",", " ").split()] for l in lines]


def test_results_1d_bp(tmp_path):
    # boundary energy, 1 AU energy, r, (ignored) weight; the last line is not read
    recs = [(2.0, 1.5, 100.01, 0.0), (0.7, 0.31, 100.2, 0.0), (3.0, 1.52, 100.1, 0.0), (60.0, 50.2, 100.0, 0.0)]
    (tmp_path / "in.dat").write_text("".join("%g %g %g %g\n" % r for r in recs))
    cli("-B", "--evaluation", "in.dat", cwd=tmp_path)
    e2 = read_rows(tmp_path / "output_1e2bin.dat")
    e3 = read_rows(tmp_path / "output_1e3bin.dat")
    e4 = read_rows(tmp_path / "output_4e2bin.dat")
    lg = read_rows(tmp_path / "output_logbin.dat")
    assert (len(e2), len(e3), len(e4), len(lg)) == (100, 1000, 400, 24)
    assert all(len(r) == 3 for r in e2) and all(len(r) == 2 for r in lg)
    w = [yamada(p_si(tb)) * p_si(t1) ** 2 for tb, t1, _, _ in recs]
    # trunc(T) bins: 1.5 and 1.52 -> bin 1 (row 0), 0.31 -> none
    assert e2[0][0] == 1.0 and e2[0][1] == pytest.approx(w[0] + w[2], rel=1e-12)
    assert e2[0][2] == pytest.approx((w[0] + w[2]) / 2, rel=1e-12)
    assert e2[49][1] == 0.0                       # the dropped last line (50.2 GeV)
    # trunc((T + 0.05) * 10): 1.5 -> 15, 1.52 -> 15, 0.31 -> 3
    assert e3[14][1] == pytest.approx(w[0] + w[2], rel=1e-12) and e3[2][1] == pytest.approx(w[1], rel=1e-12)
    # trunc((T + 0.125) * 4): 1.5 -> 6, 0.31 -> 1
    assert e4[5][1] == pytest.approx(w[0] + w[2], rel=1e-12) and e4[0][1] == pytest.approx(w[1], rel=1e-12)
    cli("-B", "--evaluation", "in.dat", "-c", cwd=tmp_path)
    head = open(tmp_path / "output_1e2bin.csv").readline().strip()
    assert head == "Tkin,spe,speAvg,speCount"
    assert open(tmp_path / "output_logbin.csv").readline().strip() == "binc,spelog/binw"


def test_results_1d_fp_zero_weight_inside_reflection_radius(tmp_path):
    pinj = p_si(2.0)
    recs = [(pinj, p_si(1.2), 1.0, 0.0, 0.01), (pinj, p_si(1.3), 0.2, 0.0, 0.01), (pinj, p_si(9.0), 1.0, 0.0, 0.0)]
    (tmp_path / "in.dat").write_text("".join(" %g  %g  %g  %g %g \n" % r for r in recs))
    cli("-F", "--evaluation", "in.dat", cwd=tmp_path)
    e2 = read_rows(tmp_path / "output_1e2bin.dat")
    # first two records bin at trunc(T) = 1; the one inside r < 0.3 carries 0
    p100 = float("%g" % pinj)
    p1 = float("%g" % p_si(1.2))
    w = yamada(p100) / 1e45 * p1 * p1 * math.exp(0.01)
    assert e2[0][1] == pytest.approx(w, rel=1e-9)
    assert e2[0][2] == pytest.approx(w / 2, rel=1e-9)    # two events, one with zero weight


def test_results_2d_files(tmp_path):
    sp = [0.01 * 1.5 ** k for k in range(30)]
    recs = [(sp[5], 3.0, 100.0, 0.0, 1.5708, 1.0), (sp[5], 4.0, 100.0, 0.0, 1.5708, 2.0),
            (0.2, 1.0, 100.0, 0.0, 1.5708, 1.0), (1.0, 2.0, 100.0, 0.0, 1.5708, 1.0),
            (sp[9], 5.0, 100.0, 0.0, 1.5708, 1.0)]
    (tmp_path / "in.dat").write_text("".join("%g %g %g %g %g %g\n" % r for r in recs))
    cli("-T", "--evaluation", "in.dat", cwd=tmp_path)
    for name, n in (("Usoskin_2005", 29), ("Weber", 29), ("Burger", 29), ("Ulysses", 2), ("Ams", 70)):
        rows = read_rows(tmp_path / (name + ".dat"))
        assert len(rows) == n and all(len(r) == 4 for r in rows), name
    weber = read_rows(tmp_path / "Weber.dat")
    # SolarProp energy index 6 (0.0759375 GeV) holds the first two records
    assert weber[5][3] == 2.0 and weber[5][0] == pytest.approx(0.0759375, abs=1e-4)
    assert sum(r[3] for r in weber) == 2.0        # the last line (index 10) is not read
    uly = read_rows(tmp_path / "Ulysses.dat")
    assert uly[0][3] == 1.0 and uly[1][3] == 1.0   # 0.2 GeV in (0.125, 0.25); 1.0 in (0.25, 2)


# ---- measured-value tables (-m / -y)

TABLES = os.path.join(ROOT, "tests", "golden", "k0_tables")   # rows excerpted from the reference's CSV files


def carrington(month, year):
    """The reference's rotation number of the 15th of a month
    (MeasureValuesTransformation.cpp:80-115), restated with datetime."""
    import datetime
    s = (datetime.datetime(year, month, 15) - datetime.datetime(1976, 5, 27)).total_seconds()
    r = s / 2356560.0
    frac = math.modf(r)[0]
    return int(math.trunc(r) if frac <= 0.5499 else math.ceil(r)) + 1642


def app_with_tables(tmp_path):
    """gs_cli started from a directory that holds the tables: the reference
    reads them next to its executable (ParseParams.cpp getApplicationPath)."""
    for f in os.listdir(TABLES):
        (tmp_path / f).write_bytes(open(os.path.join(TABLES, f), "rb").read())
    link = tmp_path / "gs_cli"
    if not os.path.exists(CLI):
        pytest.fail("gs_cli is not built; run __graft_entry__.build()")
    os.symlink(CLI, link)
    return str(link)


def run_app(app, *args, check=True):
    r = subprocess.run([app, *args], capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError("gs_cli %s failed (%d): %s" % (" ".join(args), r.returncode, r.stderr))
    return r


def test_carrington_rotation_rule():
    # the table rows: rotation 1642 begins on 1976-05-27, 1643 on 06-23, 2000 on 2003-02-20
    assert carrington(5, 1976) == 1642 and carrington(6, 1976) == 1643 and carrington(8, 1976) == 1645
    assert carrington(2, 2003) == 2000


def test_month_year_1d_takes_v_from_the_table(tmp_path):
    app = app_with_tables(tmp_path)
    p = json.loads(run_app(app, "-B", "-m", "6", "-y", "1990", "--seed", "1", "--print-params").stdout)
    assert p["V"] == conv_v(435)          # row 1990.45: V = 435 km/s
    # the 1D GPU models keep their built-in K0 unless it was entered (CosmicConstants.cu:69-73)
    assert p["K0"] == float(f32(0.000222))
    p = json.loads(run_app(app, "-F", "-m", "12", "-y", "2008", "--seed", "1", "--print-params").stdout)
    assert p["V"] == conv_v(380)          # row 2008.95


def test_month_year_2d_takes_k0_tilt_polarity(tmp_path):
    app = app_with_tables(tmp_path)
    g = json.loads(run_app(app, "-T", "-m", "6", "-y", "1976", "--seed", "1", "--print-params").stdout)
    # rotation 1643, x K0_ratio (5 when no settings file sets it, CosmicConstants.cu:108)
    assert g["K0"] == float(f32(f32(5.47252112676056e-05) * f32(5.0)))
    assert g["tilt"] == deg_f32(6.2) and g["polarity"] == 1.0
    e = json.loads(run_app(app, "-E", "-m", "2", "-y", "2003", "--seed", "1", "--print-params").stdout)
    assert e["K0"] == float(f32(6.94405829596413e-05))                   # SolarProp table, not scaled
    assert e["tilt"] == deg_f32(43.8) and e["polarity"] == -1.0
    g2 = json.loads(run_app(app, "-T", "-m", "2", "-y", "2003", "--seed", "1", "--print-params").stdout)
    assert g2["tilt"] == deg_f32(43.8) and g2["polarity"] == -1.0       # rotation 2000, A < 0


def test_month_year_missing_row_and_option_rules(tmp_path):
    app = app_with_tables(tmp_path)
    r = run_app(app, "-T", "-m", "1", "-y", "1960", "--seed", "1", "--print-params")
    assert "not found in input table" in r.stderr      # logged; the run keeps its defaults
    d = json.loads(r.stdout)
    assert d["polarity"] == 1.0
    assert run_app(app, "-B", "-m", "6", "--print-params", check=False).returncode != 0          # -m needs -y
    assert run_app(app, "-B", "-K", "5e22", "-m", "6", "-y", "1990", "--print-params", check=False).returncode != 0
    assert run_app(app, "-B", "-V", "400", "-y", "1990", "--print-params", check=False).returncode != 0


def test_ulysses_fixture_and_settings():
    """The Ulysses comparison's inputs: 67 rotations of 1994-1998 with
    Ulysses' range and colatitude, the rotation's table row and the KET
    fluxes; the settings written for each run parse with the CLI's reader."""
    from geliosphere_amd import ulysses
    header, rows = ulysses.load_rotations()
    assert header.startswith("carrington_rotation,") and len(rows) == 67
    assert all(1.0 < r["r"] < 6.0 and 0.0 < r["theta"] < 180.0 for r in rows)
    assert all(r["ket_p190"] > 0 and r["ket_p116"] > 0 for r in rows)
    assert [r["cr"] for r in rows] == list(range(rows[0]["cr"], rows[0]["cr"] + 67))


REF_INPUT = "/root/reference/Input"


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="the reference tree is not present on this machine")
def test_dropin_builds_against_the_reference_params_carrier(tmp_path):
    """The drop-in host code compiles and links against the reference's own
    ParamsCarrier (Input/include/ParamsCarrier.hpp and Input/src/
    ParamsCarrier.cpp, read where they lie; compat/ is not on the include
    path), and gs_cli built that way maps flags, settings and batch rows to
    the same parameter blocks as the standalone build."""
    host = os.path.join(ROOT, "geliosphere_amd", "host")
    exe = tmp_path / "gs_cli_ref"
    srcs = [os.path.join(host, f) for f in ("gs_cli.cpp", "gs_dropin.cpp", "gs_results.cpp", "gs_batch.cpp",
                                            "gs_tables.cpp", "gs_settings.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(REF_INPUT, "include"), "-I", host, "-o", str(exe), *srcs,
           os.path.join(REF_INPUT, "src", "ParamsCarrier.cpp"), "-L", os.path.join(ROOT, "geliosphere_amd"),
           "-lgeliosphere_amd", "-Wl,-rpath," + os.path.join(ROOT, "geliosphere_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    st = settings_file(tmp_path, r_injection=5.0)
    b = tmp_path / "b.csv"
    b.write_text("year,month,K0,V,dt,N,r,theta,pathToCustomSettingsFile,name,model\n"
                 "-1,-1,5e22,400,500,1,1.0,90,,runA,1D Bp\n-1,-1,,,,2,30,45,,,2D Geliosphere\n")
    for args in (["-B", "--seed", "5"], ["-F", "-K", "3e22", "-V", "600", "-d", "100", "--seed", "1"],
                 ["-T", "-K", "1e22", "-V", "500", "--tilt", "40", "--polarity", "-1", "--seed", "2"],
                 ["-E", "-s", str(st), "--seed", "3"], ["--custom-model", "2D SolarProp-like", "--seed", "4"],
                 ["-b", str(b), "--seed", "9"]):
        ref = subprocess.run([str(exe), *args, "--print-params"], capture_output=True, text=True, timeout=60)
        own = cli(*args, "--print-params")
        assert ref.returncode == 0 and ref.stdout == own.stdout, args
