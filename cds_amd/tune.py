"""Autotune the engine's launch shapes on seeded synthetic graphs and write
the cache file the engine loads (cds_amd/tune_cache.txt by default).

  python -m cds_amd.tune [--out PATH] [--quick]

Each (k, edge class) entry comes from cds_autotune on one graph of that
class: Chung-Lu graphs (exponent 2.3, mean degree 12, one planted
G(40, 0.9)) for the small classes, R-MAT (a=.57 b=.19 c=.19, edge factor
16) for the large ones.  Every candidate shape must reproduce the answer of
the first solve, so the cache only changes speed.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CACHE = os.path.join(HERE, "tune_cache.txt")

# (label, k, generator): one graph per (k, edge class)
PLAN = [
    ("chung_lu n=2000", 3, ("chung_lu", 2000)),
    ("chung_lu n=20000", 3, ("chung_lu", 20000)),
    ("rmat scale 14", 3, ("rmat", 14)),
    ("rmat scale 16", 3, ("rmat", 16)),
    ("rmat scale 18", 3, ("rmat", 18)),
    ("rmat scale 20", 3, ("rmat", 20)),
    ("chung_lu n=2000", 4, ("chung_lu", 2000)),
    ("chung_lu n=20000", 4, ("chung_lu", 20000)),
    ("rmat scale 14", 4, ("rmat", 14)),
    ("chung_lu n=20000", 2, ("chung_lu", 20000)),
    ("rmat scale 16", 2, ("rmat", 16)),
    ("rmat scale 18", 2, ("rmat", 18)),
]
QUICK = {("chung_lu n=20000", 3), ("rmat scale 16", 3)}


def make(gen, seed):
    import cds_amd

    kind, size = gen
    if kind == "chung_lu":
        return cds_amd.Graph.chung_lu(size, 2.3, 12.0, 40, 0.9, seed=seed)
    return cds_amd.Graph.rmat(size, 16, 0.57, 0.19, 0.19, seed=seed)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--quick", action="store_true", help="two graphs only")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--skip-existing", action="store_true",
                    help="leave (k, class) entries already in --out as they are")
    args = ap.parse_args(argv)
    import cds_amd

    have = set()
    if args.skip_existing and os.path.exists(args.out):
        for line in open(args.out):
            f = line.split()
            if f and f[0] == "cds-tune":
                have.add((int(f[4]), int(f[5])))
    eng = cds_amd.Engine(device=args.device, tune_cache="")
    for i, (label, k, gen) in enumerate(PLAN):
        if args.quick and (label, k) not in QUICK:
            continue
        g = make(gen, 100 + i)
        cls = max(int(g.m).bit_length() - 1, 0)
        if (k, cls) in have:
            print(f"[tune] {label} k={k}: class {cls} already tuned", flush=True)
            continue
        dg = eng.upload(g)
        t0 = time.time()
        best, best_ms, first_ms = eng.autotune(dg, k, args.out)
        print(f"[tune] {label} k={k} m={g.m}: {json.dumps(best)} {best_ms:.2f} ms "
              f"(untuned {first_ms:.2f} ms, {time.time() - t0:.1f} s)", flush=True)
        del dg
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
