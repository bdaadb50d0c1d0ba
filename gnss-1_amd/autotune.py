"""Autotuner driver: times the engine's compiled pass variants for each workload on the GPU and
writes the winning choices to tuning/gnss_tuning.json, which gnss_create reads.

Tuned per (precision, grid, scalars, colloid): the column tile widths of the strided x and y
passes, persistent or one-block-per-tile column and z passes, the fused z pass variant, and the
x-pass mode and its tile (`gnss_autotune` inside one context).  The row padding kp keeps the value
the table already holds (0 by default).

usage: python gnss-1_amd/autotune.py [--reps 5] [--out gnss-1_amd/tuning/gnss_tuning.json] [--big]
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


def write_table(path, table):
    with open(path, "w") as f:
        f.write("{\n")
        items = sorted(table.items())
        for i, (k, v) in enumerate(items):
            f.write(f'"{k}": {json.dumps(v)}{"," if i + 1 < len(items) else ""}\n')
        f.write("}\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "gnss_tuning.json"))
    ap.add_argument("--big", action="store_true", help="also tune the 1024^3 FP32 workload")
    args = ap.parse_args()
    table, reports = {}, []
    for cfg in cases(args.big):
        t0 = time.time()
        s = gnss.Solver(cfg).setup()
        rep = s.autotune(args.reps)
        s.close()
        choice = dict(kv.split("=") for kv in rep["choice"].split())
        table[rep["key"]] = {k: int(v) for k, v in choice.items()}
        reports.append(rep)
        print(f"[autotune] {rep['key']}: {rep['choice']}  ({time.time() - t0:.1f} s)", flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    write_table(args.out, table)
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump(reports, f, indent=1)


if __name__ == "__main__":
    main()
