"""Autotuner of the gravity launch shapes.

Single systems: for each body count N (a seeded disk of that size), every
candidate shape -- bodies per lane, j-splits per i-block, tile source (scalar
loads or LDS) and, for the mixed form, the near/far threshold theta -- is
timed with HIP events around the pair kernel.  A theta is admissible only if
the mixed form's max relative error on the mutual acceleration, against the
FP64 form, stays below `max_err`.

Ensembles: on 512 seeded disks of log-uniform N in [256, 8192], bodies per
lane of the FP64-fast ensemble pass are timed the same way, and the mixed
form's theta is chosen under the same error bound.

The fastest admissible shape per (form, N) is written to the tuning table
that libncg.so reads at context creation (ncg_tuning.txt beside it); rows of
forms that were not tuned are kept.

    python -m ncorpion_amd.autotune --n 8192 16384 65536 262144 \
        --out ncorpion_amd/ncg_tuning.txt

Sizes up to 4096 are skipped (the direct form serves them); the first tuned
size's row applies from n_min 1.
    python -m ncorpion_amd.autotune --forms ensemble
"""
import argparse
import itertools
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "ncg_tuning.txt")
# Up to this many bodies the fast forms run the direct kernel (k_small), which
# has no tiled launch shape: every candidate would time the same.
DIRECT_MAX_N = 4096


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


def tune(ctx, n, reps=3, max_err=5e-7, soft=1e-6, thetas=(0.3, 0.35, 0.5), log=print):
    """Returns {form: (ipl, nsplit, theta, src, ms)} for N = n.  When no
    theta keeps the mixed form within max_err, the most accurate one is
    used (and said so).  The thetas start at 0.3: a far pair's FP32
    separation error is at most u (1 + 2/theta) relative (u = 2^-24), 4.6e-7
    there, and smaller thetas, admissible on one tuning disk, were not on
    others."""
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
        cands = (0.35,)
        if prec == "mixed":
            errs = {}
            for theta in thetas:
                ctx.set_tiling(1, 0, theta, 0)
                ctx.gravity_allpairs_dev(bt, acc, None, None, 1.0, soft, "mixed")
                ctx.synchronize()
                errs[theta] = float(torch.max(torch.linalg.norm(acc - ref, dim=1) / rnorm))
                log("  n=%d theta=%.2f max rel err %.2e" % (n, theta, errs[theta]))
            cands = tuple(t for t in thetas if errs[t] <= max_err)
            if not cands:
                cands = (min(errs, key=errs.get),)
                log("  n=%d: no theta within %.1e; using the most accurate, %.2f" %
                    (n, max_err, cands[0]))
        timed = []
        for theta in cands:
            for ipl, src, ns in itertools.product((1, 2), (0, 1), splits):
                ctx.set_tiling(ipl, ns, theta, src)
                timed.append((ipl, ns, theta, src, _time(ctx, bt, acc, prec, reps, soft)))
        # Of the shapes within 1 % of the fastest, the one with the fewest
        # j-splits: each split writes (and the finish reads) 24 B per body.
        fastest = min(t[4] for t in timed)
        best = min((t for t in timed if t[4] <= 1.01 * fastest), key=lambda t: (t[1], t[4]))
        out[prec] = best
        log("  n=%d %s best ipl=%d nsplit=%d theta=%.2f src=%d %.3f ms" % ((n, prec) + best))
    return out


def tune_ensemble(ctx, reps=3, max_err=5e-7, thetas=(0.3, 0.35, 0.5), log=print):
    """Returns {form: (ipl, 0, theta, src, ms)} for the 512-disk ensemble; the
    table row applies to every ensemble (n_min 1).  The mixed
    form runs one body per lane; its theta is admissible when the max
    relative error of the mutual accelerations of every body, against the
    FP64-fast ensemble form, stays below max_err (else the most accurate
    theta is used)."""
    import numpy as np
    import torch
    from . import ic
    sizes = ic.seeded_ensemble_sizes(512)
    systems = [ic.disk(int(nn), seed=s, mass=1e-8, radius=2e-4) for s, nn in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    bt = torch.from_numpy(np.concatenate(systems)).cuda()
    cen = np.zeros((len(sizes), 8))
    cen[:, 6] = 1.0
    cen[:, 7] = 1.0
    ct = torch.from_numpy(cen).cuda()
    c0 = torch.zeros_like(ct)  # central bodies absent: the mutual part only
    acc = torch.zeros((bt.shape[0], 3), dtype=torch.float64, device="cuda")
    ref = torch.zeros_like(acc)
    cm = torch.zeros((len(sizes), 3), dtype=torch.float64, device="cuda")
    ctx.set_tiling(1, 0, 0.35, 0)
    ctx.ensemble_gravity_dev(bt, off, c0, ref, cm, 1.0, 0.0, "fast64")
    ctx.synchronize()
    rnorm = torch.linalg.norm(ref, dim=1)
    out = {}
    for prec in ("mixed", "fast64"):
        cands = (0.35,)
        if prec == "mixed":
            errs = {}
            for theta in thetas:
                ctx.set_tiling(1, 0, theta, 0)
                ctx.ensemble_gravity_dev(bt, off, c0, acc, cm, 1.0, 0.0, "mixed")
                ctx.synchronize()
                errs[theta] = float(torch.max(torch.linalg.norm(acc - ref, dim=1) / rnorm))
                log("  ensemble theta=%.2f max rel err %.2e" % (theta, errs[theta]))
            cands = tuple(t for t in thetas if errs[t] <= max_err)
            if not cands:
                cands = (min(errs, key=errs.get),)
                log("  ensemble: no theta within %.1e; using the most accurate, %.2f" %
                    (max_err, cands[0]))
        best = None
        ipls = (1,) if prec == "mixed" else (1, 2)
        for theta, ipl in itertools.product(cands, ipls):
            ctx.set_tiling(ipl, 0, theta, 0)
            ctx.ensemble_gravity_dev(bt, off, ct, acc, cm, 1.0, 0.0, prec)
            ctx.synchronize()
            ctx.profile(True)
            ctx.profile_read("ens_" + prec)
            for _ in range(reps):
                ctx.ensemble_gravity_dev(bt, off, ct, acc, cm, 1.0, 0.0, prec)
            ctx.synchronize()
            ms, cnt = ctx.profile_read("ens_" + prec)
            ctx.profile(False)
            ms /= max(cnt, 1)
            log("  ensemble %s ipl=%d theta=%.2f %.3f ms" % (prec, ipl, theta, ms))
            if best is None or ms < best[4]:
                best = (ipl, 0, theta, 0, ms)
        out[prec] = best
    return out


def read_table(path):
    """Rows (form, n_min, (ipl, nsplit, theta, src, ms)) of an existing table."""
    rows = []
    if not os.path.exists(path):
        return rows
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split("#")[0].split()
        ms = float(line.split("#")[1].split()[0]) if "#" in line else 0.0
        rows.append((f[0], int(f[1]), (int(f[2]), int(f[3]), float(f[4]), int(f[5]), ms)))
    return rows


def write_table(rows, path):
    with open(path, "w") as f:
        f.write("# Launch shapes of the tiled gravity forms, read by libncg.so at ncg_create.\n")
        f.write("# Written by ncorpion_amd/autotune.py on an MI355X; one row per (form, N bucket):\n")
        f.write("# the row with the largest n_min <= N applies.\n")
        f.write("# form   n_min   ipl  nsplit  theta  src   (measured ms per evaluation at n_min)\n")
        f.write("# (ens_* forms: nsplit and src unused; theta used by ens_mixed)\n")
        for prec, nmin, (ipl, ns, theta, src, ms) in rows:
            f.write("%-7s  %-7d %-4d %-7d %-6.2f %-4d  # %.3f ms\n" % (prec, nmin, ipl, ns, theta, src, ms))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, nargs="+", default=[8192, 16384, 65536, 262144])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--max-err", type=float, default=5e-7)
    ap.add_argument("--json", default=None)
    ap.add_argument("--forms", nargs="+", default=["single", "ensemble"],
                    choices=["single", "ensemble"])
    args = ap.parse_args(argv)
    from . import Context
    ctx = Context(0)
    log = lambda s: print(s, flush=True)  # noqa: E731
    tuned = set()
    rows, res = [], {}
    if "single" in args.forms:
        tuned |= {"mixed", "fast64"}
        skipped = sorted(n for n in args.n if n <= DIRECT_MAX_N)
        if skipped:
            log("skipping n = %s: the direct form serves n <= %d" % (skipped, DIRECT_MAX_N))
        ns = sorted(n for n in args.n if n > DIRECT_MAX_N)
        if not ns:
            ap.error("no size above %d to tune" % DIRECT_MAX_N)
        for i, n in enumerate(ns):
            r = tune(ctx, n, max_err=args.max_err, log=log)
            res[n] = r
            # A row applies from the geometric midpoint to the previous size.
            nmin = 1 if i == 0 else int((ns[i - 1] * n) ** 0.5)
            for prec in ("mixed", "fast64"):
                if r[prec] is not None:
                    rows.append((prec, nmin, r[prec]))
    if "ensemble" in args.forms:
        tuned |= {"ens_mixed", "ens_fast64"}
        r = tune_ensemble(ctx, max_err=args.max_err, log=log)
        res["ensemble"] = r
        for prec in ("mixed", "fast64"):
            rows.append(("ens_" + prec, 1, r[prec]))
    kept = [row for row in read_table(args.out) if row[0] not in tuned]
    write_table(kept + rows, args.out)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({str(k): v for k, v in res.items()}, f)
    print("wrote", args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
