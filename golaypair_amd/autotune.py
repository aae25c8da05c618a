"""Project-written autotuner: measures the library's launch and batch parameters
on the device and writes the fastest ones into golaypair_amd/tuning.tsv, the
table every context loads (gp_create; keys in golaypair_amd/csrc/tuning.hip).

    python -m golaypair_amd.autotune [--out PATH] [--workloads v50 v45 v90] [--reps R]

Per workload (order, compression pair, PAF constant):
  gen_ranks_per_lane/<n>/<k>   ranks per lane of the gen kernel at word length
                               n over k letters: candidates 16 .. 256, timed by
                               gp_generate at the coarse compression;
  unc_waves/<N>                waves per SIMD of the uncompression filter at
                               finer length N (capped by each operand class's
                               occupancy): candidates 1 .. 5;
  unc_leaf_budget/<N>          leaves per uncompression device batch:
                               candidates 2^27 .. 2^31;
  unc_counter_rounds/<N>       rounds of filter work items handed out by the
                               work counters, the rest by wave index:
                               candidates 0 .. 4 (default: every item);
  unc_counter_streams/<N>      work counters per filter launch: 1 .. 4
                               (default 8, one per XCD);
the uncompression candidates timed on a fixed slice of the level's matched
pairs.  Each candidate runs `reps` times (minimum wall time kept) and its
output must equal the default setting's, byte for byte; the winner replaces
the default only if it is at least 3 % faster (run-to-run noise is about 2 %).
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

import golaypair_amd as g

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(PKG, "tuning.tsv")

# name -> (order, coarse compression, finer compression, paf, matched pairs [lo, hi) timed)
WORKLOADS = {
    "v50": (50, 5, 1, 0, (0, 400)),
    "v45": (45, 3, 1, -2, (0, 2500)),
    "v90": (90, 10, 5, 0, (0, 156)),
}
GEN_RANKS = [16, 32, 64, 128, 256]
UNC_WAVES = [1, 2, 3, 4, 5]
UNC_BUDGET = [1 << 27, 1 << 28, 1 << 29, 1 << 30, 1 << 31]
UNC_COUNTER_ROUNDS = [0, 1, 2, 4]
UNC_COUNTER_STREAMS = [1, 2, 4]


def _digest(parts):
    h = hashlib.sha256()
    for p in parts:
        h.update(np.ascontiguousarray(p).tobytes())
    return h.hexdigest()


def _set(ctx, key, value):
    rc = ctx.lib.gp_set_tuning(ctx.h, key.encode(), -1 if value is None else int(value))
    if rc != g.GP_OK:
        raise g.GolayError("gp_set_tuning %s failed" % key)


def _time(fn, reps):
    best, out = None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, out


def _sweep(ctx, key, candidates, run, reps, log):
    """Times run() with key unset (the default) and at every candidate; returns
    (chosen value or None for the default, {value: seconds})."""
    _set(ctx, key, None)
    t_def, ref = _time(run, reps)
    times = {"default": t_def}
    best_v, best_t = None, t_def
    for v in candidates:
        _set(ctx, key, v)
        t, out = _time(run, reps)
        if out != ref:
            raise RuntimeError("%s=%d changed the output" % (key, v))
        times[v] = t
        if t < best_t * 0.97:
            best_v, best_t = v, t
    _set(ctx, key, best_v)
    log("%-28s default %.4f s; %s -> %s" % (key, t_def, ", ".join("%s: %.4f" % (k, t) for k, t in times.items()
                                                                  if k != "default"),
                                           "default" if best_v is None else best_v))
    return best_v, times


def tune(workloads, reps, log):
    ctx = g.Context(0)
    chosen, notes = {}, []
    try:
        for name in workloads:
            order, c0, c1, paf, (q0, q1) = WORKLOADS[name]
            n0, N = order // c0, order // c1
            k = len(range(-c0, c0 + 1, 2))

            def gen():
                a, b, _ = ctx.generate(order, c0, paf)
                return _digest([a.words, a.keys, b.words, b.keys])

            key = "gen_ranks_per_lane/%d/%d" % (n0, k)
            v, t = _sweep(ctx, key, GEN_RANKS, gen, reps, log)
            if v is not None:
                chosen[key] = v
            notes.append("%s %s" % (key, t))
            a, b, _ = ctx.generate(order, c0, paf)
            rc, pa, pb, _ = ctx.match(order, c0, paf, ctx.sort_candidates(a), ctx.sort_candidates(b))
            pa, pb = pa[q0:q1], pb[q0:q1]

            def unc():
                oa, ob, _, _, _ = ctx.uncompress(order, c0, c1, paf, pa, pb)
                return _digest([oa, ob])

            for key, cand in (("unc_waves/%d" % N, UNC_WAVES), ("unc_leaf_budget/%d" % N, UNC_BUDGET),
                              ("unc_counter_rounds/%d" % N, UNC_COUNTER_ROUNDS),
                              ("unc_counter_streams/%d" % N, UNC_COUNTER_STREAMS)):
                v, t = _sweep(ctx, key, cand, unc, reps, log)
                if v is not None:
                    chosen[key] = v
                notes.append("%s %s" % (key, t))
    finally:
        ctx.close()
    return chosen, notes


def write(path, chosen, notes, workloads, reps):
    with open(path, "w") as f:
        f.write("# Tuning table of golaypair_amd, written by python -m golaypair_amd.autotune\n")
        f.write("# workloads %s, %d timed runs per candidate (minimum kept); keys absent here use the\n"
                "# built-in defaults.  Measured wall times in seconds per candidate:\n" % (" ".join(workloads), reps))
        for n in notes:
            f.write("#   %s\n" % n)
        for key in sorted(chosen):
            f.write("%s %d\n" % (key, chosen[key]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    os.environ["GOLAY_TUNING"] = os.devnull  # start from the built-in defaults
    chosen, notes = tune(args.workloads, args.reps, lambda s: print(s, file=sys.stderr, flush=True))
    write(args.out, chosen, notes, args.workloads, args.reps)
    print(args.out)


if __name__ == "__main__":
    main()
