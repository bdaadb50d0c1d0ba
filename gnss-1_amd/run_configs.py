"""Runs the engine on the measured workloads beyond the headline bench line and prints JSON.

  jet256    FP64, 256^3, two scalars (Sc 0.7, 2.0), 200 steps, statistics every 20 steps
  ensemble  FP32 batched ensemble 32 x 64^3 + 16 x 96^3 + 8 x 128^3, one hipGraph per RK3 step,
            200 steps; per-member error against the FP64 form of sampled members
  jet1024   FP32 at 1024^3 (50 steps) with the VALU radix stages and with the split-FP16 MFMA
            radix-16 stage; errors against each other and against FP64 at 512^3 (same physical
            parameters) compared on the 512^3 modes
  colloid512  the 512^3 jet plus the colloid scalar, FP32 and FP64, 100 steps
  jet512    the headline 512^3 jet, FP32 and FP64 forms on the same seed, 100 steps: the FP32
            form's errors against FP64 (fields, integral statistics, profiles, spectrum) and its
            divergence, against the precision form's bounds (1e-5, 1e-4, 1e-6)

usage: python gnss-1_amd/run_configs.py {jet256,ensemble,jet1024,colloid512,jet512} [--steps K]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gnss  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_jet256(steps):
    cfg = gnss.config("jet256")
    s = gnss.Solver(cfg).setup()
    out = {"workload": "jet256 fp64", "stats": []}
    st, yp, sp = s.stats()
    out["stats"].append(st)
    t_step = 0.0
    for c in range(20, steps + 1, 20):
        t0 = time.perf_counter()
        s.step(20)
        t_step += time.perf_counter() - t0
        st, yp, sp = s.stats()
        out["stats"].append(st)
        log(f"[jet256] step {st['step']} KE {st['ke']:.8e} eps {st['epsilon']:.6e} CFL {st['cfl']:.4f}")
    out["steps_per_s"] = steps / t_step
    out["grid_point_updates_per_s"] = steps / t_step * 256 ** 3
    out["final_profiles_rows"] = int(yp.shape[0])
    out["final_spectrum_peak_shell"] = int(np.argmax(sp[0]))
    return out


def run_ensemble(steps):
    cfgs = gnss.ensemble_configs()
    t0 = time.perf_counter()
    ens = gnss.Ensemble(cfgs)
    ens.init()
    t_init = time.perf_counter() - t0
    ens.step(2)  # warm-up replays
    t0 = time.perf_counter()
    ens.step(steps)
    dt = time.perf_counter() - t0
    pts = sum(c.n[0] * c.n[1] * c.n[2] for c in cfgs)
    out = {"workload": "ensemble 32x64^3 + 16x96^3 + 8x128^3 fp32", "members": len(cfgs),
           "init_s": t_init, "steps": steps, "ensemble_steps_per_s": steps / dt,
           "member_steps_per_s": len(cfgs) * steps / dt, "grid_point_updates_per_s": pts * steps / dt}
    # errors of sampled members against their FP64 form (same seed, same number of steps)
    errs = {}
    for i in (0, 31, 32, 47, 48, 55):
        c64 = gnss.Config(**{**cfgs[i].__dict__, "precision": gnss.FP64})
        s = gnss.Solver(c64).setup()
        s.step(steps + 2)
        e = [rel(ens.get_spectral(i, f), s.get_spectral(f)) for f in range(4)]
        st32, st64 = ens.stats(i), s.stats(False, False)[0]
        errs[i] = {"n": cfgs[i].n[0], "re": cfgs[i].re, "theta": cfgs[i].theta, "rel_l2_max": max(e),
                   "ke_rel": abs(st32["ke"] - st64["ke"]) / st64["ke"]}
        s.close()
    out["fp32_vs_fp64"] = errs
    return out


def downsample(a, n_small):
    """Spectral field [ky][kx][kz] of a large grid -> the modes of an n_small grid."""
    ny, nx, nzh = a.shape
    h = n_small // 2
    iy = np.r_[0:h, ny - h:ny]
    ix = np.r_[0:h, nx - h:nx]
    return a[np.ix_(iy, ix, np.arange(h + 1))]


def run_jet1024(steps):
    out = {"workload": "jet1024 fp32 (VALU and MFMA split-FP16 radix-16)", "steps": steps}
    fields = {}
    for name, flags in (("mfma", gnss.FLAG_MFMA_R16), ("valu", 0)):
        cfg = gnss.config("jet1024", precision=gnss.FP32, flags=flags)
        log(f"[jet1024] {name}: init")
        s = gnss.Solver(cfg).setup()
        s.step(1)
        t0 = time.perf_counter()
        s.step(steps - 1)
        dt = time.perf_counter() - t0
        out[f"{name}_steps_per_s"] = (steps - 1) / dt
        out[f"{name}_ms_per_step"] = 1000 * dt / (steps - 1)
        st = s.stats(False, False)[0]
        out[f"{name}_stats"] = {k: st[k] for k in ("ke", "epsilon", "cfl", "divergence")}
        fields[name] = [downsample(s.get_spectral(f), 512) for f in range(4)]
        log(f"[jet1024] {name}: {out[f'{name}_ms_per_step']:.1f} ms/step")
        s.close()
    out["mfma_vs_valu_rel_l2"] = [rel(fields["mfma"][f], fields["valu"][f]) for f in range(4)]
    cfg = gnss.config("jet1024", n=(512,) * 3, precision=gnss.FP64)
    s = gnss.Solver(cfg).setup()
    s.step(steps)
    ref = [s.get_spectral(f) * 8.0 for f in range(4)]  # unnormalised: 1024^3 / 512^3 = 8
    mask = s.mask().astype(bool)
    out["fp32_1024_vs_fp64_512_rel_l2"] = {
        k: [rel(fields[k][f][mask], ref[f][mask]) for f in range(4)] for k in fields}
    return out


def run_colloid512(steps):
    """The 512^3 jet with the reference's colloid added as a second scalar (alpha 0.0033,
    Sc_c 0.7): FP32 and FP64 rates and the FP32 error against FP64 after `steps` steps."""
    out = {"workload": "jet512 + colloid (alpha 0.0033, Sc_c 0.7)", "steps": steps}
    fields = {}
    for name, prec in (("fp32", gnss.FP32), ("fp64", gnss.FP64)):
        cfg = gnss.with_colloid(gnss.config("jet512", precision=prec))
        s = gnss.Solver(cfg).setup()
        s.step(1)
        t0 = time.perf_counter()
        s.step(steps - 1)
        dt = time.perf_counter() - t0
        out[f"{name}_steps_per_s"] = (steps - 1) / dt
        st = s.stats(False, False)[0]
        out[f"{name}_stats"] = {k: st[k] for k in ("ke", "epsilon", "divergence")}
        out[f"{name}_chi"] = list(st["chi"]) if "chi" in st else None
        fields[name] = [s.get_spectral(f) for f in range(5)]
        log(f"[colloid512] {name}: {out[f'{name}_steps_per_s']:.2f} steps/s")
        s.close()
    out["fp32_vs_fp64_rel_l2"] = [rel(fields["fp32"][f], fields["fp64"][f]) for f in range(5)]
    return out


def run_jet512(steps):
    """FP32 form against the FP64 plain form, 512^3 jet, same seed, `steps` RK3 steps.  Reductions
    (statistics, profiles, spectrum) are FP64 in both forms; errors are relative (L2 for fields
    and profiles, absolute value over the FP64 value for scalar statistics)."""
    out = {"workload": "jet512 fp32 vs fp64", "steps": steps}
    res = {}
    for name, prec in (("fp32", gnss.FP32), ("fp64", gnss.FP64)):
        s = gnss.Solver(gnss.config("jet512", precision=prec)).setup()
        t0 = time.perf_counter()
        s.step(steps)
        dt = time.perf_counter() - t0
        out[f"{name}_steps_per_s"] = steps / dt
        st, yp, sp = s.stats()
        res[name] = (st, yp, sp, [s.get_spectral(f) for f in range(4)])
        log(f"[jet512] {name}: {steps / dt:.2f} steps/s")
        s.close()
    (s32, y32, p32, f32), (s64, y64, p64, f64) = res["fp32"], res["fp64"]
    out["rel_l2"] = {k: rel(f32[i], f64[i]) for i, k in enumerate(("u", "v", "w", "phi0"))}
    keys = ("ke", "vrms", "epsilon", "enstrophy", "cfl", "umax", "lint", "lambda", "eta")
    out["stats_rel"] = {k: abs(s32[k] - s64[k]) / abs(s64[k]) for k in keys if k in s64 and s64[k] != 0}
    out["stats_rel"]["chi0"] = abs(s32["chi"][0] - s64["chi"][0]) / abs(s64["chi"][0])
    out["stats_rel"]["scalar_var0"] = abs(s32["scalar_var"][0] - s64["scalar_var"][0]) / abs(s64["scalar_var"][0])
    # profile rows that vanish in exact arithmetic (the mean of v: d<v>/dy = 0 for a divergence-free
    # field, and its volume mean is 0) hold only roundoff in FP64: they are reported by their
    # largest absolute FP32 value instead
    big = [r for r in range(y64.shape[0]) if np.max(np.abs(y64[r])) > 1e-8]
    out["profiles_rel_l2_max"] = max(rel(y32[r], y64[r]) for r in big)
    out["profiles_zero_rows_absmax"] = max((float(np.max(np.abs(y32[r]))) for r in range(y64.shape[0])
                                            if r not in big), default=0.0)
    out["spectrum_rel_l2"] = rel(p32, p64)
    out["divergence_fp32"] = s32["divergence"]
    out["divergence_fp64"] = s64["divergence"]
    out["within_bounds"] = bool(max(out["rel_l2"].values()) <= 1e-5 and max(out["stats_rel"].values()) <= 1e-4
                                and out["divergence_fp32"] <= 1e-6)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["jet256", "ensemble", "jet1024", "colloid512", "jet512"])
    ap.add_argument("--steps", type=int, default=0)
    a = ap.parse_args()
    steps = a.steps or {"jet256": 200, "ensemble": 200, "jet1024": 50, "colloid512": 100, "jet512": 100}[a.which]
    fn = {"jet256": run_jet256, "ensemble": run_ensemble, "jet1024": run_jet1024,
          "colloid512": run_colloid512, "jet512": run_jet512}[a.which]
    print(json.dumps(fn(steps)))


if __name__ == "__main__":
    main()
