"""Benchmark: RK3 steps/s of the pseudospectral jet step at 512^3 on one MI355X.

Workload (the metric's configuration): 512^3 temporal jet, L = 2 pi, Re = 5000, theta = 0.02,
one scalar Sc = 1.0, divergence-free noise of rms 0.05 on 1 <= |m| <= 32, dt = 2.5e-4, seeded.
Synthetic: the state is the seeded initial condition built by the engine (no dataset).

One "step" = one full low-storage RK3 step (3 substages: 13 3D transforms each, products,
projection, integrating-factor update).  `value` = FP32-form steps/s; the FP64 plain form runs
the same seed and its steps/s, the FP32/FP64 speedup and the FP32 error against FP64 are
reported beside it.  The standalone 3D FFT (R2C + C2R) bandwidth is reported as fft3d.

The K timed steps run with nothing but the step's kernels on the stream; K further steps with a HIP
event pair around every kernel give the per-kernel tables (kernels_f32, kernels_f64).
roofline: the dominant kernel of those FP32 steps, timed with HIP events on the engine's
stream; achieved = algorithmic bytes per launch / mean launch time; peak = the vendor 8 TB/s,
with the rate of a plain device copy measured on the same GPU reported beside it (fft3d.*.copy_gbs,
roofline.copy_gbs_measured).  cpu_baseline: the FP64
oracle (test infrastructure, C + FFTW3 API) timed for a bounded sample of the same workload.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--config jet512]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gnss-1_amd"))
import gnss  # noqa: E402

HBM_PEAK_GBS = 8000.0
ORACLE_LIB = os.path.join(ROOT, "oracle", "libgnss_oracle.so")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_gpu(cfg, steps, warmup):
    """W untimed warm-up steps, K steps timed on the host clock with nothing but the step's kernels
    on the stream (the headline rate), then K more steps with a HIP event pair around every kernel
    (the per-kernel table and the dominant kernel's launch time; the events add a little per
    launch, so these steps are not the ones the rate is taken from)."""
    s = gnss.Solver(cfg).setup()
    s.step(warmup)
    t0 = time.perf_counter()
    s.step(steps)  # returns after the stream has drained (NaN watchdog reads back a flag)
    t1 = time.perf_counter()
    s.timing(True)
    s.step(steps)
    rep = s.timing_report()
    s.timing(False)
    return s, (t1 - t0), rep


def fft_bandwidth(cfg, reps=5):
    s = gnss.Solver(cfg).setup()
    ms_f, ms_i = s.bench_fft(reps)
    ms_c, bytes_c = s.bench_copy(reps)
    n = cfg.n
    b = 8 if cfg.precision == gnss.FP64 else 4
    modes = n[0] * n[1] * (n[2] // 2 + 1)
    b_fft = 3 * 2 * modes * 2 * b  # 3 passes x (read + write) x complex field
    out = {
        "precision": "f64" if cfg.precision == gnss.FP64 else "f32",
        "algorithmic_bytes": b_fft,
        "ms_r2c": ms_f, "ms_c2r": ms_i,
        "gbs_r2c": b_fft / (ms_f * 1e-3) / 1e9, "gbs_c2r": b_fft / (ms_i * 1e-3) / 1e9,
    }
    out["frac_peak"] = 0.5 * (out["gbs_r2c"] + out["gbs_c2r"]) / HBM_PEAK_GBS
    # a plain device copy of one spectral field on the same stream: the practical HBM rate
    out["copy_gbs"] = bytes_c / (ms_c * 1e-3) / 1e9
    out["frac_of_copy"] = 0.5 * (out["gbs_r2c"] + out["gbs_c2r"]) / out["copy_gbs"]
    s.close()
    return out


def dominant_kernel(rep):
    name = max(rep, key=lambda k: rep[k]["ms"])
    r = rep[name]
    per_launch_ms = r["ms"] / r["launches"]
    per_launch_bytes = r["bytes"] / r["launches"]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        t = json.load(open(TRAFFIC_FILE))
        traffic = t.get(name, {}).get("bytes_per_launch")
    return name, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                  "kernel": name, "bytes_per_launch": per_launch_bytes,
                  "ms_per_launch": per_launch_ms}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(cfg, threads, steps=3, one_core_steps=1):
    """Bounded sample: `steps` FP64 RK3 steps of the same workload on the oracle (FFTW3 API) with
    `threads` threads, then `one_core_steps` steps on one thread (the single-core figure)."""
    lib = gnss.load(ORACLE_LIB)
    lib.gnss_oracle_set_threads.argtypes = [gnss.ctypes.c_int]
    lib.gnss_oracle_set_backend.argtypes = [gnss.ctypes.c_int]
    lib.gnss_oracle_backend_name.restype = gnss.ctypes.c_char_p
    lib.gnss_oracle_max_threads.restype = gnss.ctypes.c_int
    backend = "fftw3-api" if lib.gnss_oracle_set_backend(1) == 0 else "builtin"
    name = lib.gnss_oracle_backend_name().decode()
    c64 = gnss.Config(**{**cfg.__dict__, "precision": gnss.FP64})
    n = cfg.n

    def timed_steps(nthreads, k):
        lib.gnss_oracle_set_threads(nthreads)
        o = gnss.Solver(c64, lib).setup()
        t0 = time.perf_counter()
        o.step(k)
        dt = time.perf_counter() - t0
        o.close()
        return dt

    dt = timed_steps(threads, steps)
    out = {"value": round(steps / dt, 5), "unit": "steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} FP64 RK3 steps of the same {n[0]}^3 workload on the oracle ({backend}: {name}; "
                     f"{threads} threads of {os.cpu_count()} on {cpu_model()}); {dt:.1f} s"}
    if one_core_steps > 0:
        d1 = timed_steps(1, one_core_steps)
        out["one_core"] = {"value": round(one_core_steps / d1, 5), "unit": "steps/s", "cores": 1,
                           "sample": f"{one_core_steps} FP64 RK3 step(s) of the same workload, 1 thread; {d1:.1f} s"}
    lib.gnss_oracle_set_threads(threads)
    return out


def step_traffic(rep, steps, ms_per_step, cfg):
    """Algorithmic HBM bytes of one timed RK3 step (the sum of the kernels' per-launch byte counts:
    the pruned, fused passes the step runs), the rate they imply, and the unfused, unpruned
    per-step figure B_step = 3 [(9 + 4S) B_fft + 5 (3 + S) C] for comparison."""
    b = 8 if cfg.precision == gnss.FP64 else 4
    n, S = cfg.n, len(cfg.sc)
    modes = n[0] * n[1] * (n[2] // 2 + 1)
    b_fft = 3 * 2 * modes * 2 * b
    c = modes * 2 * b
    b_step = 3 * ((9 + 4 * S) * b_fft + 5 * (3 + S) * c)
    step_bytes = sum(v["bytes"] for v in rep.values()) / steps
    gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
    return {"step_bytes": step_bytes, "step_gbs": round(gbs, 1), "step_frac": round(gbs / HBM_PEAK_GBS, 4),
            "b_step_unfused": b_step,
            "note": "step_bytes counts what the fused, 2/3-pruned passes move (C2R + products + R2C in one "
                    "z pass, known-zero modes neither loaded nor stored); b_step_unfused is the same step as "
                    "separate full transforms and pointwise passes"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="jet512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the step shards over x/y slabs with an all-to-all per transform, but no multi-GPU "
              "path is built)", file=sys.stderr)
        sys.exit(2)

    cfg32 = gnss.config(args.config, precision=gnss.FP32)
    cfg64 = gnss.config(args.config, precision=gnss.FP64)
    n = cfg32.n
    npts = n[0] * n[1] * n[2]
    print(f"[bench] {args.config} FP32 warmup {args.warmup} steps {args.steps}", file=sys.stderr, flush=True)
    s32, t32, rep = run_gpu(cfg32, args.steps, args.warmup)
    print(f"[bench] FP32 {args.steps / t32:.3f} steps/s", file=sys.stderr, flush=True)
    s64, t64, rep64 = run_gpu(cfg64, args.steps, args.warmup)
    print(f"[bench] FP64 {args.steps / t64:.3f} steps/s", file=sys.stderr, flush=True)
    errs = {}
    names = ["u", "v", "w"] + [f"phi{i}" for i in range(len(cfg32.sc))]
    for f, nm in enumerate(names):
        errs[nm] = rel_l2(s32.get_spectral(f), s64.get_spectral(f))
    st32, _, _ = s32.stats(False, False)
    st64, _, _ = s64.stats(False, False)
    stat_err = {k: abs(st32[k] - st64[k]) / abs(st64[k]) for k in ["ke", "epsilon", "enstrophy", "cfl"]}
    stat_err["chi0"] = abs(st32["chi"][0] - st64["chi"][0]) / abs(st64["chi"][0])
    div32 = st32["divergence"]
    s32.close()
    s64.close()
    print("[bench] fft3d", file=sys.stderr, flush=True)
    fft32 = fft_bandwidth(cfg32)
    fft64 = fft_bandwidth(cfg64)
    kname, roof = dominant_kernel(rep)
    roof["copy_gbs_measured"] = round(fft32["copy_gbs"], 1)
    roof["frac_of_copy"] = round(roof["achieved"] / fft32["copy_gbs"], 4)

    sps32, sps64 = args.steps / t32, args.steps / t64
    line = {
        "metric": "RK3 steps/s at 512^3, FP32 vs FP64 plain, 1 GPU; 3D-FFT achieved HBM GB/s %",
        "value": round(sps32, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * t32 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded tanh jet + divergence-free Gaussian noise, built on the host in FP64)",
        "config": {"workload": f"{args.config}: {n[0]}^3 temporal jet, Re={cfg32.re:g}, theta={cfg32.theta:g}, "
                               f"Sc={list(cfg32.sc)}, noise 0.05 on |m|<={cfg32.noise_kmax:g}, dt={cfg32.dt:g}",
                   "grid": list(n), "scalars": len(cfg32.sc), "parallelism": "1 GPU"},
        "fp64_steps_per_s": round(sps64, 4),
        "fp64_ms_per_step": round(1000.0 * t64 / args.steps, 3),
        "fp32_over_fp64": round(sps32 / sps64, 3),
        "grid_point_updates_per_s": {"f32": sps32 * npts, "f64": sps64 * npts},
        "fp32_error_vs_fp64": {"rel_l2": errs, "stats_rel": stat_err, "divergence_fp32": div32,
                               "after_steps": args.warmup + 2 * args.steps},
        "fft3d": {"f32": fft32, "f64": fft64},
        "kernels_f32": rep,
        "kernels_f64": rep64,
        "roofline": roof,
        "step_traffic": {"f32": step_traffic(rep, args.steps, 1000.0 * t32 / args.steps, cfg32),
                         "f64": step_traffic(rep64, args.steps, 1000.0 * t64 / args.steps, cfg64)},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(cfg64, threads)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
