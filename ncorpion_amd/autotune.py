"""Autotuner of the tiled gravity launch shapes.

For each body count N (a seeded disk of that size), every candidate shape --
bodies per lane, j-splits per i-block, tile source (scalar loads or LDS) and,
for the mixed form, the near/far threshold theta -- is timed with HIP events
around the pair kernel.  A theta is admissible only if the mixed form's max
relative error on the mutual acceleration, against the FP64 form, stays below
`max_err`.  The fastest admissible shape per (form, N) is written to the tuning
table that libncg.so reads at context creation (ncg_tuning.txt beside it).

    python -m ncorpion_amd.autotune --n 4096 16384 65536 262144 \
        --out ncorpion_amd/ncg_tuning.txt
"""
import argparse
import itertools
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "ncg_tuning.txt")


def _workload(n, seed=0):
    from . import ic
    m, R = ic.disk_c3_masses(n, seed)
    return ic.disk(n, seed=seed, masses=m, radii=R)


def _time(ctx, bt, acc, prec, reps, soft):
    ctx.gravity_allpairs_dev(bt, acc, None, None, 1.0, soft, prec)
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_read("pairs_" + prec)
    for _ in range(reps):
        ctx.gravity_allpairs_dev(bt, acc, None, None, 1.0, soft, prec)
    ctx.synchronize()
    ms, cnt = ctx.profile_read("pairs_" + prec)
    ctx.profile(False)
    return ms / max(cnt, 1)


def tune(ctx, n, reps=3, max_err=5e-7, soft=1e-6, thetas=(0.25, 0.35, 0.5), log=print):
    """Returns {form: (ipl, nsplit, theta, src, ms)} for N = n."""
    import torch
    b = _workload(n)
    bt = torch.from_numpy(b).cuda()
    acc = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    ref = torch.zeros_like(acc)
    ctx.set_tiling(1, 0, 0.35, 0)
    ctx.gravity_allpairs_dev(bt, ref, None, None, 1.0, soft, "fast64")
    ctx.synchronize()
    rnorm = torch.linalg.norm(ref, dim=1)
    cus = 256
    nib1 = max(1, -(-n // 256))
    splits = sorted({s for s in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512) if s * nib1 <= 512 * cus})
    out = {}
    for prec in ("mixed", "fast64"):
        best = None
        for theta in (thetas if prec == "mixed" else (0.35,)):
            if prec == "mixed":
                ctx.set_tiling(1, 0, theta, 0)
                ctx.gravity_allpairs_dev(bt, acc, None, None, 1.0, soft, "mixed")
                ctx.synchronize()
                err = float(torch.max(torch.linalg.norm(acc - ref, dim=1) / rnorm))
                log("  n=%d theta=%.2f max rel err %.2e" % (n, theta, err))
                if err > max_err:
                    continue
            for ipl, src, ns in itertools.product((1, 2), (0, 1), splits):
                ctx.set_tiling(ipl, ns, theta, src)
                ms = _time(ctx, bt, acc, prec, reps, soft)
                if best is None or ms < best[4]:
                    best = (ipl, ns, theta, src, ms)
        out[prec] = best
        log("  n=%d %s best ipl=%d nsplit=%d theta=%.2f src=%d %.3f ms" % ((n, prec) + best))
    return out


def write_table(rows, path):
    with open(path, "w") as f:
        f.write("# Launch shapes of the tiled gravity forms, read by libncg.so at ncg_create.\n")
        f.write("# Written by ncorpion_amd/autotune.py on an MI355X; one row per (form, N bucket):\n")
        f.write("# the row with the largest n_min <= N applies.\n")
        f.write("# form   n_min   ipl  nsplit  theta  src   (measured ms per evaluation at n_min)\n")
        for prec, nmin, (ipl, ns, theta, src, ms) in rows:
            f.write("%-7s  %-7d %-4d %-7d %-6.2f %-4d  # %.3f ms\n" % (prec, nmin, ipl, ns, theta, src, ms))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, nargs="+", default=[4096, 16384, 65536, 262144])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--max-err", type=float, default=5e-7)
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    from . import Context
    ctx = Context(0)
    rows, res = [], {}
    ns = sorted(args.n)
    for i, n in enumerate(ns):
        r = tune(ctx, n, max_err=args.max_err, log=lambda s: print(s, flush=True))
        res[n] = r
        # A row applies from the geometric midpoint to the previous size.
        nmin = 1 if i == 0 else int((ns[i - 1] * n) ** 0.5)
        for prec in ("mixed", "fast64"):
            if r[prec] is not None:
                rows.append((prec, nmin, r[prec]))
    write_table(rows, args.out)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({str(k): v for k, v in res.items()}, f)
    print("wrote", args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
