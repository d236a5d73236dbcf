"""Benchmark of the trajectory engine on one MI355X.

Headline (value): plain Monte Carlo, one "step" = the 1D backward-in-time
(-B) and forward-in-time (-F) models, N = 1e8 particles each, K0 = 5e22
cm^2/s, V = 400 km/s, dt = 500 s, injection log-uniform 0.01-100 GeV, 30 log
bins, FP64.  Inputs are a parameter block; nothing is transferred per step
except the 30-bin result.  Reported as particle-steps/s (whole job), with
the roofline of the dominant kernel (1D B-p trajectory kernel, FP64-VALU
bound; algorithmic FLOPs per step from the oracle's op count) and the CPU
baseline (the oracle on the host cores, bounded sample).

Also measured in the same run (detail):
  batched  - spectra/s of gs_run_batch vs sequential gs_run_plain on the
             same heterogeneous simulations (batch sampler, particle counts
             scaled by --batch-scale), with a bit-identity check;
  surrogate - the trained ensemble's inference rate (65,536 queries), its
             median per-bin error over 0.1-10 GeV on the held-out Monte Carlo
             spectra, and the plain Monte Carlo spectra/s it replaces.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6          # MI355X dense FP64 vector
# HBM bytes per particle of the 1D B-p trajectory kernel: (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0) KiB per launch, scaled by a 256 MiB copy, over 1e8
# particles (profiles/plain_bp_hbm_traffic.md).  Algorithmic: one 64-byte
# event record per particle.
BP_TRAFFIC_BYTES_PER_PARTICLE = 83.4
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
    ap.add_argument("--batch-sims", type=int, default=256,
                    help="simulations in the batched leg (0: skip); 256 as in the batch benchmark")
    ap.add_argument("--batch-scale", type=float, default=0.01,
                    help="particle counts of the batched leg relative to the batch sampler's 1e5-1e7 "
                         "(default 1 %%: the full-scale comparison is --batch-only --batch-scale 1)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched comparison (the batch benchmark at --batch-scale) and print it")
    ap.add_argument("--batch-mode", choices=["both", "batched", "sequential"], default="both",
                    help="with --batch-only: time gs_run_batch, sequential gs_run_plain, or both")
    ap.add_argument("--batch-slice", default="",
                    help="with --batch-only: simulations START:END of the sampled batch")
    ap.add_argument("--batch-parts", type=int, default=1,
                    help="with --batch-only: run the batched leg as this many consecutive gs_run_batch calls")
    ap.add_argument("--hash-out", default="", help="with --batch-only: write per-simulation spectrum hashes here")
    ap.add_argument("--autotune-cache", default=os.path.join(ROOT, "geliosphere_amd", "data", "autotune.txt"),
                    help="launch shapes from the autotuner's cache (geliosphere_amd.autotune); '' for defaults")
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


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s):
    """The oracle (FP64 restatement of the reference CPU path) on all host
    cores, on a bounded sample of the same 1D B-p workload, in two modes:
    the reference's own random source and thread schedule (std::mt19937 +
    std::normal_distribution<float>, threads spawned per iteration of 101 x
    250 particles each; one iteration), reported as the baseline, and the
    oracle's Philox streams with dynamic scheduling (the parity mode)."""
    from geliosphere_amd import abi
    from oracle import oracle_ffi
    threads = oracle_ffi.default_threads()
    # reference-faithful: one iteration of threads x 25,250 particles
    n_f = threads * 101 * 250
    p = abi.make_params(abi.MODEL_1D_BP, N=n_f, seed=98, flags=oracle_ffi.FLAG_FAITHFUL)
    t0 = time.perf_counter()
    rc, s, _, _ = oracle_ffi.run(p, n_threads=threads)
    t_f = time.perf_counter() - t0
    steps_f = s.total_steps
    # parity mode: a growing sample until half the budget is spent
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
    return {"value": steps_f / t_f, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "mode": "reference-faithful: FP64 step, std::mt19937 + std::normal_distribution<float>, "
                    "threads spawned per iteration of 101 x 250 particles",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "1D B-p, %d particles (one iteration of %d threads; log-uniform 0.01-100 GeV, dt=500 s), "
                      "%d steps in %.1f s" % (n_f, threads, steps_f, t_f),
            "wall_s": t_f, "total_steps": steps_f,
            "philox_mode": {"value": steps / t_total, "wall_s": t_total, "total_steps": steps,
                            "sample": "1D B-p, %d particles, Philox streams, dynamic scheduling" % sample_n}}


def progress(msg):
    """One line on stderr when a leg of the run ends (stdout carries only the
    result line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def batch_leg(eng, n_sims, scale, seed=2024):
    """gs_run_batch vs one gs_run_plain per simulation, same simulations."""
    from geliosphere_amd import batch
    sims = batch.sample_batch(n_sims, seed=seed, n_scale=scale)
    eng.run_batch(sims)                       # builds the graph of this layout
    t0 = time.perf_counter()
    outs = eng.run_batch(sims)
    t_batch = time.perf_counter() - t0
    st = eng.last_stats()
    t0 = time.perf_counter()
    same = True
    for p, o in zip(sims, outs):
        s, _, _ = eng.run_plain(p)
        same &= bytes(s) == bytes(o)
    t_seq = time.perf_counter() - t0
    steps = sum(o.total_steps for o in outs)
    return {"sims": n_sims, "particles": sum(p.n_particles for p in sims),
            "scale": scale,
            "note": "the batch benchmark at %g %% of its particle counts" % (100 * scale),
            "particle_count_range": [int(1e5 * scale), int(1e7 * scale)],
            "batched_spectra_per_s": n_sims / t_batch, "sequential_plain_spectra_per_s": n_sims / t_seq,
            "speedup": t_seq / t_batch, "batched_particle_steps_per_s": steps / t_batch,
            "lane_utilisation": st.active_cycles / max(st.lane_cycles, 1), "bit_identical_to_plain": same}


def spectrum_hash(s):
    import hashlib
    return hashlib.sha256(bytes(s)).hexdigest()[:16]


def batch_only(eng, a):
    """The batch benchmark at --batch-scale (scale 1: the benchmark as
    specified, 256 simulations at N 1e5-1e7): gs_run_batch over the sampled
    simulations (in --batch-parts consecutive calls) and/or one gs_run_plain
    per simulation, each leg timed; per-simulation spectrum hashes let
    sessions that time different legs or slices check bit-identity."""
    from geliosphere_amd import batch
    sims = batch.sample_batch(a.batch_sims, seed=2024, n_scale=a.batch_scale)
    lo, hi = 0, len(sims)
    if a.batch_slice:
        lo, hi = (int(v) for v in a.batch_slice.split(":"))
    idx = list(range(lo, hi))
    part = [sims[i] for i in idx]
    res = {"sims": len(part), "slice": [lo, hi], "scale": a.batch_scale,
           "particle_count_range": [int(1e5 * a.batch_scale), int(1e7 * a.batch_scale)],
           "particles": sum(p.n_particles for p in part)}
    hashes = {}
    if a.batch_mode in ("both", "batched"):
        k = max(1, a.batch_parts)
        cuts = [len(part) * j // k for j in range(k + 1)]
        t_batch, steps, act, lanes = 0.0, 0, 0, 0
        for j in range(k):
            chunk = part[cuts[j]:cuts[j + 1]]
            t0 = time.perf_counter()
            outs = eng.run_batch(chunk)
            dt = time.perf_counter() - t0
            st = eng.last_stats()
            t_batch += dt
            steps += sum(o.total_steps for o in outs)
            act += st.active_cycles
            lanes += st.lane_cycles
            for i, o in zip(idx[cuts[j]:cuts[j + 1]], outs):
                hashes.setdefault(i, {})["batched"] = spectrum_hash(o)
            progress("batched part %d/%d: %d simulations in %.1f s" % (j + 1, k, len(chunk), dt))
        res.update({"batched_seconds": t_batch, "batched_spectra_per_s": len(part) / t_batch,
                    "batched_particle_steps_per_s": steps / t_batch, "batched_calls": k,
                    "lane_utilisation": act / max(lanes, 1), "particle_steps": steps})
    if a.batch_mode in ("both", "sequential"):
        t_seq = 0.0
        for n_done, (i, p) in enumerate(zip(idx, part)):
            t0 = time.perf_counter()
            s, _, _ = eng.run_plain(p)
            dt = time.perf_counter() - t0
            t_seq += dt
            hashes.setdefault(i, {})["plain"] = spectrum_hash(s)
            progress("plain %d/%d (simulation %d, %s, N=%d): %.2f s" % (n_done + 1, len(part), i,
                                                                       abi_name(p.model), p.n_particles, dt))
        res.update({"sequential_seconds": t_seq, "sequential_plain_spectra_per_s": len(part) / t_seq})
    if "batched_seconds" in res and "sequential_seconds" in res:
        res["speedup"] = res["sequential_seconds"] / res["batched_seconds"]
        res["bit_identical_to_plain"] = all(h["batched"] == h["plain"] for h in hashes.values())
    if a.hash_out:
        with open(a.hash_out, "w") as f:
            json.dump({str(i): h for i, h in sorted(hashes.items())}, f, indent=0)
    return res


def surrogate_leg(eng):
    """Inference rate and held-out accuracy of the committed ensemble, and
    what answering every query costs: the ensemble serves the queries it
    trusts, batched Monte Carlo (gs_run_batch, at the training particle
    counts) the rest, and every 1D F-p query (the ensemble does not serve
    that model).  Effective spectra/s = 1 / (1 / inference rate + fallback
    share / Monte Carlo rate), per model and over the served models."""
    import numpy as np
    from geliosphere_amd import surrogate, surrogate_eval, surrogate_data
    data_dir = os.path.join(ROOT, "geliosphere_amd", "data")
    wpath = os.path.join(data_dir, "surrogate_weights.npz")
    dpath = os.path.join(data_dir, "surrogate_dataset.npz")
    if not (os.path.exists(wpath) and os.path.exists(dpath)):
        return {"status": "no trained ensemble in geliosphere_amd/data"}
    sg = surrogate.Surrogate(0)
    sg.load(wpath)
    d = surrogate_eval.load([dpath])
    acc = surrogate_eval.evaluate(sg, d)
    rate_host, dt_host = surrogate_eval.inference_rate(sg)
    rate, dt = surrogate_eval.inference_rate_device(sg)
    n_train = dict(zip(surrogate_data.MODELS, np.load(dpath)["n_train"].tolist()))
    n_eval = dict(zip(surrogate_data.MODELS, np.load(dpath)["n_eval"].tolist()))
    mc = {}
    for m in surrogate_data.MODELS:
        n = int(n_train[m]) or int(surrogate_data.DEFAULT_N_TRAIN[m])
        sims = surrogate_data.sims_for(m, 16, n, 99, 777)
        eng.run_batch(sims[:4])
        t0 = time.perf_counter()
        eng.run_batch(sims)
        mc[abi_name(m)] = {"particles": n, "batched_spectra_per_s": len(sims) / (time.perf_counter() - t0)}
    eff = {}
    for m in surrogate_data.MODELS:
        name = abi_name(m)
        served = m in surrogate.SURROGATE_MODELS and name in acc
        f = acc[name]["fallback_share"] if served else 1.0
        eff[name] = 1.0 / (1.0 / rate + f / mc[name]["batched_spectra_per_s"])
    served_names = [abi_name(m) for m in surrogate.SURROGATE_MODELS if abi_name(m) in acc]
    f_all = float(np.mean([acc[k]["fallback_share"] for k in served_names])) if served_names else 1.0
    mc_all = len(served_names) / sum(1.0 / mc[k]["batched_spectra_per_s"] for k in served_names)
    return {"inference_spectra_per_s": rate, "inference_batch": 65536, "inference_batch_ms": dt * 1e3,
            "inference_buffers": "device-resident (gs_surrogate_predict_device)",
            "inference_spectra_per_s_host_buffers": rate_host, "inference_batch_ms_host_buffers": dt_host * 1e3,
            "served_models": served_names, "mc_only_models": [abi_name(m) for m in surrogate_data.MODELS
                                                              if m not in surrogate.SURROGATE_MODELS],
            "training_particles": {abi_name(m): int(n_train[m]) for m in surrogate_data.MODELS},
            "evaluation_particles": {abi_name(m): int(n_eval[m]) for m in surrogate_data.MODELS},
            "mc_fallback": mc,
            "fallback_share": {k: v["fallback_share"] for k, v in acc.items()},
            "effective_spectra_per_s": eff,
            "effective_spectra_per_s_served": 1.0 / (1.0 / rate + f_all / mc_all),
            "effective_speedup_vs_batched_mc_served": (1.0 / (1.0 / rate + f_all / mc_all)) / mc_all,
            "surrogate_only_speedup_vs_batched_mc_served": rate / mc_all,
            "median_rel_err_0p1_10GeV": {k: v["median_rel_err_surrogate"] for k, v in acc.items()},
            "median_rel_err_trusted_0p1_10GeV": {k: v["median_rel_err_trusted"] for k, v in acc.items()},
            "median_rel_err_nearest_neighbour": {k: v["median_rel_err_nearest_neighbour"] for k, v in acc.items()},
            "mc_reference_median_rel_se": {k: v["median_rel_se_mc_reference"] for k, v in acc.items()}}


def abi_name(m):
    from geliosphere_amd import abi
    return abi.MODEL_NAMES[m]


def main():
    a = parse()
    if a.calibrate_copy > 0:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: a run on several GPUs would shard by particle id "
              "range, one range per GPU, and add the parts with gs_spectrum_merge; no multi-GPU "
              "launcher is part of this build", file=sys.stderr)
        sys.exit(2)
    from geliosphere_amd import abi
    from geliosphere_amd.engine import Engine

    n = int(a.n)
    eng = Engine(0)
    tuned = eng.autotune_load(a.autotune_cache) if a.autotune_cache and os.path.exists(a.autotune_cache) else 0
    shapes = {abi.MODEL_NAMES[m]: {"plain": list(eng.launch_shape(m, abi.KERNEL_PLAIN)),
                                   "batch": list(eng.launch_shape(m, abi.KERNEL_BATCH))} for m in range(4)}
    shapes_info = {"cache": os.path.relpath(a.autotune_cache, ROOT) if tuned else None, "entries_applied": tuned,
                   "shape": "[block_threads, blocks_per_cu, ids per queue grab]", "per_model": shapes}
    if a.batch_only:
        res = batch_only(eng, a)
        eng.close()
        print(json.dumps({"batch": res, "launch_shapes": shapes_info}), flush=True)
        return
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
            util = st.active_cycles / st.lane_cycles if st.lane_cycles else 0.0
            res.append((p.model, s.total_steps, st.kernel_ms, s, util, st.max_particle_steps))
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
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": BP_TRAFFIC_BYTES_PER_PARTICLE * n if dom == 1 else None,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                     "algorithmic_bytes": 64 * n,
                     "kernel": "plain_kernel<%s>" % kname, "kernel_ms": bp_ms,
                     "alg_flops_per_step": ALG_FLOPS_PER_STEP[dom],
                     "steps_per_launch": bp_steps},
        "detail": {
            "dominant_steps_per_s_kernel": bp_steps / (bp_ms * 1e-3),
            "steps_per_launch": {abi.MODEL_NAMES[m]: sum(r[1] for r in runs if r[0] == m) / a.steps
                                 for m in models},
            "spectra_per_s": len(models) * a.steps / wall,
            "lane_utilisation": {abi.MODEL_NAMES[m]: sum(r[4] for r in runs if r[0] == m) / a.steps
                                 for m in models},
            "longest_particle_steps": {abi.MODEL_NAMES[m]: max(r[5] for r in runs if r[0] == m) for m in models},
            "launch_shapes": shapes_info,
        },
    }
    progress("plain MC legs done: %.4g particle-steps/s" % out["value"])
    if a.batch_sims > 0:
        out["detail"]["batched"] = batch_leg(eng, a.batch_sims, a.batch_scale)
        progress("batched leg done: %.1fx sequential plain" % out["detail"]["batched"]["speedup"])
    if not a.no_surrogate:
        out["detail"]["surrogate"] = surrogate_leg(eng)
        progress("surrogate leg done")
    eng.close()
    if not a.no_cpu:
        cb = cpu_baseline(a.cpu_seconds)
        # spectra/s of each side for the step's models at the step's N: the
        # steps one spectrum takes (measured above) over each side's rate
        steps_per_step = sum(r[1] for r in runs) / a.steps
        cb["spectra_per_s_at_step_size"] = len(models) * cb["value"] / steps_per_step
        cb["gpu_spectra_per_s_at_step_size"] = len(models) * a.steps / wall
        out["cpu_baseline"] = cb
        progress("CPU baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
