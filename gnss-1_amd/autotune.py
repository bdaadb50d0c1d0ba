"""Autotuner driver: times the engine's compiled pass variants for each workload on the GPU and
writes the winning choices to tuning/gnss_tuning.json, which gnss_create reads.

Tuned per (precision, grid, scalars, colloid): the column tile widths of the strided x and y
passes, persistent or one-block-per-tile column and z passes, the fused z pass variant, and the
x-pass mode and its tile (`gnss_autotune` inside one context).  With --ensemble the batched
engines of the uneven ensemble (32 x 64^3 + 16 x 96^3 + 8 x 128^3, FP32) are tuned as they run,
all members per launch (`gnss_ensemble_autotune`), giving the "<key>b" entries.  The row padding
kp keeps the value the table already holds (0 by default).

The table is updated in place: keys this run tunes are replaced, every other entry is kept.  The
raw candidate timings go to --report (default autotune_report.json in the working directory, an
untracked file), not beside the table.

usage: python gnss-1_amd/autotune.py [--reps 5] [--out gnss-1_amd/tuning/gnss_tuning.json]
                                     [--big] [--only KEY ...] [--ensemble | --ensemble-only]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gnss  # noqa: E402


def cases(big: bool):
    out = []
    for n in (64, 96, 128):
        out.append(gnss.Config(n=(n,) * 3, sc=(0.7,), precision=gnss.FP32))
    out.append(gnss.config("jet64"))
    out.append(gnss.config("jet64", precision=gnss.FP32))
    out.append(gnss.config("jet256"))
    out.append(gnss.config("jet256", precision=gnss.FP32))
    out.append(gnss.config("jet512", precision=gnss.FP32))
    out.append(gnss.config("jet512"))
    out.append(gnss.with_colloid(gnss.config("jet512", precision=gnss.FP32)))
    out.append(gnss.with_colloid(gnss.config("jet512")))
    if big:
        out.append(gnss.config("jet1024", precision=gnss.FP32))
    return out


def read_table(path):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def write_table(path, table):
    with open(path, "w") as f:
        f.write("{\n")
        items = sorted(table.items())
        for i, (k, v) in enumerate(items):
            f.write(f'"{k}": {json.dumps(v)}{"," if i + 1 < len(items) else ""}\n')
        f.write("}\n")


def choice_of(rep, old):
    """Table entry of one tuning report; kp (not tuned) keeps its old value."""
    ch = {k: int(v) for k, v in (kv.split("=") for kv in rep["choice"].split())}
    ch["kp"] = int(old.get("kp", ch.get("kp", 0)))
    return ch


def key_of(cfg):
    """The engine's tuning key of a workload: precision nx ny nz S[c]."""
    return (f"{'f32' if cfg.precision == gnss.FP32 else 'f64'} {cfg.n[0]} {cfg.n[1]} {cfg.n[2]} {len(cfg.sc)}"
            f"{'c' if cfg.flags & gnss.FLAG_COLLOID else ''}")


def tune_single(reps, big, table, reports, only=None):
    for cfg in cases(big):
        if only and key_of(cfg) not in only:
            continue
        t0 = time.time()
        s = gnss.Solver(cfg).setup()
        rep = s.autotune(reps)
        s.close()
        table[rep["key"]] = choice_of(rep, table.get(rep["key"], {}))
        reports.append(rep)
        print(f"[autotune] {rep['key']}: {rep['choice']}  ({time.time() - t0:.1f} s)", flush=True)


def tune_ensemble(reps, table, reports):
    t0 = time.time()
    ens = gnss.Ensemble(gnss.ensemble_configs())
    ens.init()
    for rep in ens.autotune(reps):
        table[rep["key"]] = choice_of(rep, table.get(rep["key"], {}))
        reports.append(rep)
        print(f"[autotune] {rep['key']} ({rep['members']} members): {rep['choice']}", flush=True)
    ens.close()
    print(f"[autotune] ensemble tuned in {time.time() - t0:.1f} s", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "gnss_tuning.json"))
    ap.add_argument("--report", default="autotune_report.json")
    ap.add_argument("--big", action="store_true", help="also tune the 1024^3 FP32 workload")
    ap.add_argument("--ensemble", action="store_true", help="also tune the batched ensemble engines")
    ap.add_argument("--ensemble-only", action="store_true", help="tune only the batched ensemble engines")
    ap.add_argument("--only", action="append", default=[],
                    help="tune only this workload key (e.g. 'f32 512 512 512 1'); repeatable")
    args = ap.parse_args()
    table, reports = read_table(args.out), []
    if not args.ensemble_only:
        tune_single(args.reps, args.big, table, reports, set(args.only))
    if args.ensemble or args.ensemble_only:
        # the batched engines read the table's single-context entries first: write them before
        write_table(args.out, table)
        os.environ["GNSS_TUNING"] = os.path.abspath(args.out)
        tune_ensemble(args.reps, table, reports)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    write_table(args.out, table)
    if args.report:
        os.makedirs(os.path.dirname(os.path.abspath(args.report)), exist_ok=True)
        with open(args.report, "w") as f:
            json.dump(reports, f, indent=1)


if __name__ == "__main__":
    main()
