"""Benchmark of the plain trajectory engine on one MI355X.

Workload (one "step"): the 1D backward-in-time (-B) and forward-in-time (-F)
models, N = 1e8 particles each, K0 = 5e22 cm^2/s, V = 400 km/s, dt = 500 s,
injection log-uniform 0.01-100 GeV, 30 log bins, FP64.  Inputs are a
parameter block; nothing is transferred per step except the 30-bin result.

Prints one JSON line: particle-steps/s (whole job), the roofline of the
dominant kernel (1D B-p trajectory kernel, FP64-VALU bound; algorithmic FLOPs
per step from the oracle's op count) and the CPU baseline (the oracle on the
host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6          # MI355X dense FP64 vector
# Algorithmic FLOPs per particle-step of the reference formulation, counted
# by the oracle's op-counting scalar (tests/test_oracle.py::test_flop_count_per_step).
ALG_FLOPS_PER_STEP = {0: 42.0, 1: 40.0, 2: 164.0, 3: 258.6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=float, default=1e8, help="particles per model per step")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--model", choices=["bp+fp", "bp", "fp", "sp", "ge"], default="bp+fp",
                    help="models run per step (the headline workload is bp+fp)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib):
    """One device-to-device hipMemcpy (the runtime's copy kernel) of known
    size, to scale FETCH_SIZE / WRITE_SIZE readings into bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib * (1 << 20))
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(n)):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3):   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n}), flush=True)


def cpu_baseline(budget_s):
    """The oracle (FP64 restatement of the reference CPU path) on all host
    cores, on a bounded sample of the same 1D B-p workload."""
    from geliosphere_amd import abi
    from oracle import oracle_ffi
    threads = oracle_ffi.default_threads()
    n = 2000
    steps = 0
    t_total = 0.0
    sample_n = 0
    seed = 99
    while True:
        p = abi.make_params(abi.MODEL_1D_BP, N=n, seed=seed, first=sample_n)
        t0 = time.perf_counter()
        rc, s, _, _ = oracle_ffi.run(p, n_threads=threads)
        dt = time.perf_counter() - t0
        steps += s.total_steps
        t_total += dt
        sample_n += n
        if t_total > budget_s * 0.5 or sample_n > 2_000_000:
            break
        n = int(n * max(1.5, min(4.0, (budget_s * 0.5) / max(dt, 1e-3))))
    return {"value": steps / t_total, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "1D B-p, %d particles (log-uniform 0.01-100 GeV, dt=500 s), %d steps in %.1f s"
                      % (sample_n, steps, t_total)}


def main():
    a = parse()
    if a.calibrate_copy > 0:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path in this build "
              "(particles shard by id range; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from geliosphere_amd import abi
    from geliosphere_amd.engine import Engine

    n = int(a.n)
    eng = Engine(0)
    models = {"bp+fp": (abi.MODEL_1D_BP, abi.MODEL_1D_FP), "bp": (abi.MODEL_1D_BP,),
              "fp": (abi.MODEL_1D_FP,), "sp": (abi.MODEL_2D_SOLARPROP,),
              "ge": (abi.MODEL_2D_GELIOSPHERE,)}[a.model]
    plist = [abi.make_params(m, N=n, seed=a.seed + m) for m in models]
    dom = models[0]

    def step(i):
        res = []
        for p in plist:
            q = abi.clone_params(p, seed=p.seed + 1000 * i)
            s, _, _ = eng.run_plain(q)
            st = eng.last_stats()
            res.append((p.model, s.total_steps, st.kernel_ms, s))
        return res

    for i in range(a.warmup):
        step(-1 - i)
    t0 = time.perf_counter()
    runs = []
    for i in range(a.steps):
        runs.extend(step(i))
    wall = time.perf_counter() - t0

    total_steps = sum(r[1] for r in runs)
    value = total_steps / wall
    bp_runs = [r for r in runs if r[0] == dom]
    bp_steps = sum(r[1] for r in bp_runs) / len(bp_runs)
    bp_ms = sum(r[2] for r in bp_runs) / len(bp_runs)
    achieved = ALG_FLOPS_PER_STEP[dom] * bp_steps / (bp_ms * 1e-3) / 1e12
    kname = {0: "Fp", 1: "Bp", 2: "Sp", 3: "Ge"}[dom]
    out = {
        "metric": "particle-steps/s (plain MC, %s)" % a.model,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-generated particles, log-uniform injection)",
        "config": {"workload": "1D B-p and 1D F-p plain MC, N=1e8 particles each, K0=5e22 cm^2/s, "
                               "V=400 km/s, dt=500 s, T_inj log-uniform 0.01-100 GeV, 30 log bins",
                   "particles_per_model": n},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "plain_kernel<%s>" % kname, "kernel_ms": bp_ms,
                     "alg_flops_per_step": ALG_FLOPS_PER_STEP[dom],
                     "steps_per_launch": bp_steps},
        "detail": {
            "dominant_steps_per_s_kernel": bp_steps / (bp_ms * 1e-3),
            "steps_per_launch": {abi.MODEL_NAMES[m]: sum(r[1] for r in runs if r[0] == m) / a.steps
                                 for m in models},
            "spectra_per_s": len(models) * a.steps / wall,
        },
    }
    eng.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
