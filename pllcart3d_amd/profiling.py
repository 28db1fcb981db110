"""Profiling workload and rocprofv3 summaries for the pllcart3d time step.

    python pllcart3d_amd/profiling.py run [--config front_256x128x128] [--steps 2] [--precision plain|mixed]
        set-up, one UPDATE_OLD (12 whole-field copies of known size: the
        calibration of the HBM counters) and `steps` fused steps; run it
        under rocprofv3 --kernel-trace --stats or --pmc FETCH_SIZE / WRITE_SIZE.
        The steps run on one stream (PLLC_FORK=0) unless --fork is given, so
        that each kernel's duration is its own.

    python pllcart3d_amd/profiling.py run --closure [--steps 20]
        the 2x-coarser mixed-precision grid of the configuration with the
        learned closure's network (seeded initial weights, the correction
        scaled to zero so the state is the plain one) making a correction
        every step: the network's kernels beside the step's.

    python pllcart3d_amd/profiling.py run --ensemble [--steps 4]
        the 64-member mixed-precision ensemble (three batches of equal grids,
        their step pairs the branches of one graph) after two warm-up steps;
        a separate handle's UPDATE_OLD (12 copy kernels) marks where the
        summarised dispatches start.

    python pllcart3d_amd/profiling.py summary --trace T.db [--fetch F.db --write W.db]
        a markdown summary: per-kernel durations from the kernel trace and the
        HBM bytes per dispatch from the two counter passes, scaled by what the
        same counters read for the copy kernel (one field read, one written).
"""
import argparse
import collections
import os
import re
import sqlite3
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _short(name):
    s = name.replace("(anonymous namespace)::", "")
    s = re.sub(r"\(.*", "", s).replace("void ", "").replace("pllc::", "")
    return s.strip()


def run(args):
    sys.path.insert(0, ROOT)
    if not args.fork:
        # one stream per step (no side stream for the momentum explicit
        # parts), so each kernel's duration is its own, as the bench's
        # per-kernel timing measures it
        os.environ["PLLC_FORK"] = "0"
    import pllcart3d_amd as P
    from pllcart3d_amd.configs import CONFIGS, config_params
    prm = config_params(args.config)
    cfg = CONFIGS[args.config]
    c0 = P.ic_fill(prm, cfg["ic"], cfg["seed"], cfg["amp"])
    if args.closure:
        import numpy as np
        from pllcart3d_amd import closure as CL
        cp = CL.coarse_params(prm)
        net = CL.Network(cp.nx, cp.ny, cp.nz)
        net.set_params(CL.flat_from_dict(CL.init_params(1)))
        net.set_norm(np.zeros(CL.NIN), np.ones(CL.NIN), 0.0)
        sim = P.Sim(cp, "mixed")
        sim.init(CL.coarse_c0(prm, c0))
        CL.set_closure(sim, net, 1)
        sim.run(args.steps)
        sim.synchronize()
        print("profile run: closure on the %dx%dx%d grid, %d steps" % (cp.nx, cp.ny, cp.nz, args.steps), flush=True)
        sim.close()
        net.close()
        return
    if args.ensemble:
        from pllcart3d_amd.configs import ensemble_members
        mem = ensemble_members(64)
        ens = P.Ensemble([m[0] for m in mem], "mixed")
        for (p, rec, seed, amp), s in zip(mem, ens.members):
            s.init(P.ic_fill(p, rec, seed, amp))
        ens.run(2)
        mk = P.Sim(mem[0][0], "mixed")
        mk.init(P.ic_fill(*mem[0]))
        mk.op("UPDATE_OLD")
        mk.synchronize()
        mk.close()
        ens.run(args.steps)
        print("profile run: ensemble of %d members in %d batches, %d steps" % (len(mem), ens.batches(), args.steps),
              flush=True)
        ens.close()
        return
    if args.tune:
        # the sweeps' block shapes tuned on a throwaway handle first, as
        # bench.py does; later handles of this grid and precision use them
        t = P.Sim(prm, precision=args.precision)
        t.init(c0)
        t.autotune(3)
        t.close()
    sim = P.Sim(prm, precision=args.precision)
    sim.init(c0)
    # the 12 copies of UPDATE_OLD calibrate the counters and mark where the
    # summarised dispatches start
    sim.op("UPDATE_OLD")
    sim.run(args.steps)
    sim.synchronize()
    print("profile run: %s %s, %d steps, %d points per field" % (args.config, args.precision, args.steps, sim.n),
          flush=True)


def _counters(db, cname, last):
    """Per kernel, the counter's values of its last `last[kernel]` dispatches
    (the profiled handle's, which come last in the process)."""
    con = sqlite3.connect(db)
    rows = list(con.execute("select kernel_name, dispatch_id, value from counters_collection where counter_name=? "
                            "order by dispatch_id", (cname,)))
    d = collections.defaultdict(list)
    for name, di, v in rows:
        d[_short(name)].append(v)
    return {k: v[-last[k]:] for k, v in d.items() if last.get(k)}


def _trace_rows(db):
    """(name, calls, total us, average us, %) per kernel over the profiled
    handle's dispatches (the stream of the last dispatch; a tuning handle
    runs on its own stream) from its first copy kernel on."""
    con = sqlite3.connect(db)
    rows = list(con.execute("select name, dispatch_id, stream_id, start, end from kernels order by dispatch_id"))
    st = rows[-1][2]
    rows = [r for r in rows if r[2] == st]
    first = min((di for n, di, s_, a, b in rows if _short(n).startswith("k_copy")), default=0)
    agg = collections.defaultdict(lambda: [0, 0.0])
    for name, di, s_, a, b in rows:
        if di >= first:
            agg[name][0] += 1
            agg[name][1] += (b - a) / 1e3
    tot = sum(v[1] for v in agg.values()) or 1.0
    return [(n, c, t, t / c, 100.0 * t / tot) for n, (c, t) in agg.items()]


def ensemble_summary(args):
    """Dispatches after the last copy kernel (the marker), every stream."""
    con = sqlite3.connect(args.trace)
    rows = list(con.execute("select name, dispatch_id, start, end, grid_x, grid_y, grid_z from kernels "
                            "order by dispatch_id"))
    last = max(di for n, di, a, b, gx, gy, gz in rows if _short(n).startswith("k_copy"))
    rows = [r for r in rows if r[1] > last]
    span = (max(r[3] for r in rows) - min(r[2] for r in rows)) / 1e6
    agg = collections.defaultdict(lambda: [0, 0.0])
    for name, di, a, b, gx, gy, gz in rows:
        k = (_short(name), gx, gy, gz)
        agg[k][0] += 1
        agg[k][1] += (b - a) / 1e3
    out = ["%d dispatches over %.2f ms: %.2f ms per step (%d steps; the batches' kernels overlap, so "
           "durations below add up to more than the span)" % (len(rows), span, span / args.steps, args.steps), "",
           "| kernel | grid (work-items) | calls | total ms | avg us |", "|---|---|---:|---:|---:|"]
    for (n, gx, gy, gz), (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append("| `%s` | %d x %d x %d | %d | %.3f | %.2f |" % (n[:60], gx, gy, gz, c, t / 1e3, t / c))
    sys.stdout.write("\n".join(out) + "\n")


def issue_summary(args):
    """Vector-issue table from a --pmc SQ_INSTS_VALU ... SQ_WAVES run: per
    kernel its average duration, VALU instructions and waves per dispatch,
    and the share of the device's vector-issue slots its instructions fill
    (SQ_INSTS_VALU x 4 cycles / (CUs x 4 SIMDs x duration x clock))."""
    con = sqlite3.connect(args.trace)
    dur = collections.defaultdict(list)
    for name, a, b in con.execute("select name, start, end from kernels"):
        dur[_short(name)].append((b - a) / 1e3)
    cnt = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, c, v in con.execute("select kernel_name, counter_name, value from counters_collection"):
        cnt[_short(name)][c].append(v)
    simds = 4 * args.cus
    out = ["| kernel | avg us | VALU instructions per launch | waves | per wave | VALU issue share |",
           "|---|---:|---:|---:|---:|---:|"]
    rows = []
    for k, cs in cnt.items():
        if not k.startswith("k_") or "SQ_INSTS_VALU" not in cs or "SQ_WAVES" not in cs or k not in dur:
            continue
        t = sum(dur[k]) / len(dur[k])
        ins = sum(cs["SQ_INSTS_VALU"]) / len(cs["SQ_INSTS_VALU"])
        wav = sum(cs["SQ_WAVES"]) / len(cs["SQ_WAVES"])
        share = ins * 4 / (simds * t * 1e-6 * args.clock_ghz * 1e9)
        rows.append((t * len(dur[k]), "| `%s` | %.1f | %.3g | %d | %.0f | %.0f%% |" % (k[:60], t, ins, wav, ins / wav,
                                                                                   100 * share)))
    out += [r for _, r in sorted(rows, key=lambda r: -r[0])]
    sys.stdout.write("\n".join(out) + "\n")


def summary(args):
    if getattr(args, "issue", False):
        return issue_summary(args)
    if getattr(args, "ensemble", False):
        return ensemble_summary(args)
    out = []
    rows = _trace_rows(args.trace)
    out.append("| kernel | calls | total ms | avg us | % |")
    out.append("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in sorted(rows, key=lambda r: -r[2]):
        out.append("| `%s` | %d | %.3f | %.2f | %.1f |" % (_short(name)[:60], calls, tot / 1e3, avg, pct))
    if args.fetch and args.write:
        last = {_short(n): c for n, c, t, a, p in rows}
        f = _counters(args.fetch, "FETCH_SIZE", last)
        w = _counters(args.write, "WRITE_SIZE", last)
        copy = [k for k in f if k.startswith("k_copy")]
        if not copy:
            raise SystemExit("no copy kernel in the counter passes")
        ck = copy[0]
        field_kib = args.points * args.bytes / 1024.0
        cf = sum(f[ck]) / len(f[ck])
        cw = sum(w[ck]) / len(w[ck])
        sf, sw = field_kib / cf, field_kib / cw
        out.append("")
        out.append("Calibration: `%s` copies one field of %d points x %d B = %.0f KiB; per dispatch "
                   "FETCH_SIZE reads %.0f KiB and WRITE_SIZE %.0f KiB, so the factors are %.3f (fetch) "
                   "and %.3f (write)." % (ck, args.points, args.bytes, field_kib, cf, cw, sf, sw))
        out.append("")
        out.append("| kernel | dispatches | HBM read MB | HBM write MB | read+write fields |")
        out.append("|---|---:|---:|---:|---:|")
        for k in sorted(f, key=lambda k: -sum(f[k]) / len(f[k])):
            rf = sum(f[k]) / len(f[k]) * sf
            wf = (sum(w[k]) / len(w[k]) * sw) if k in w else float("nan")
            out.append("| `%s` | %d | %.1f | %.1f | %.2f |" % (k[:60], len(f[k]), rf * 1024 / 1e6, wf * 1024 / 1e6,
                                                               (rf + wf) / field_kib))
    text = "\n".join(out) + "\n"
    sys.stdout.write(text)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--config", default="front_256x128x128")
    r.add_argument("--steps", type=int, default=2)
    r.add_argument("--precision", default="plain", choices=["plain", "mixed"])
    r.add_argument("--tune", action="store_true", help="autotune the sweeps' block shapes first (as bench.py)")
    r.add_argument("--closure", action="store_true", help="the coarse grid with the learned closure instead")
    r.add_argument("--ensemble", action="store_true", help="the 64-member mixed-precision ensemble instead")
    r.add_argument("--fork", action="store_true",
                   help="keep the side stream of the momentum explicit parts (default: one stream)")
    s = sub.add_parser("summary")
    s.add_argument("--trace", required=True)
    s.add_argument("--fetch")
    s.add_argument("--write")
    s.add_argument("--points", type=int, default=262 * 130 * 130, help="points of one field (with ghosts)")
    s.add_argument("--bytes", type=int, default=8, help="bytes per value")
    s.add_argument("--ensemble", action="store_true", help="an ensemble run's trace (every stream, per grid)")
    s.add_argument("--steps", type=int, default=4, help="the ensemble run's profiled steps")
    s.add_argument("--issue", action="store_true", help="a --pmc SQ_INSTS_VALU ... SQ_WAVES run: vector-issue table")
    s.add_argument("--cus", type=int, default=256, help="compute units of the device (issue table)")
    s.add_argument("--clock-ghz", type=float, default=2.4, help="engine clock (issue table)")
    a = ap.parse_args()
    run(a) if a.cmd == "run" else summary(a)


if __name__ == "__main__":
    main()
