"""Autotuner of the generated kernels' per-class register budget.

Each canonical class is compiled with a register budget (doubles per target
group, codegen/gen_eri.plan_class): a larger budget means fewer groups and
less recomputation, a smaller one fewer spills and more waves.  The best value
differs between classes and is measured, not modelled:

  python -m libint_amd.tune build   --dir D [--budgets 128,144,160] [--classes 2222,2122]
        (CPU) one library per budget, holding only the tuned classes (ERI
        kernels), under D/b<budget>/
  python -m libint_amd.tune measure --dir D
        (GPU) times every class's ERI kernel in every library on the same
        seeded quartets of the workload-2 recipe (60 atoms), each library in
        a child process; writes D/timings.json
  python -m libint_amd.tune select  --dir D [--out libint_amd/tuned.cfg]
        writes the fastest budget of each class as class=budget:0 entries,
        which the Makefile passes to the generator (CLASSCFG)
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_BUDGETS = (128, 144, 160)


def heavy_classes():
    sys.path.insert(0, os.path.join(HERE, "codegen"))
    import gen_eri
    return ["".join(map(str, c)) for c in gen_eri.canonical_classes(2) if sum(c) >= 5]


def build(args):
    budgets = [int(b) for b in args.budgets.split(",")]
    classes = args.classes or ",".join(heavy_classes())
    procs = []
    for b in budgets:
        d = os.path.join(os.path.abspath(args.dir), f"b{b}")
        os.makedirs(d, exist_ok=True)
        cmd = ["make", "-s", "-C", HERE, f"GEN={d}/gen", f"OBJ={d}/obj", f"LIB={d}/libl2g.so",
               f"ONLY={classes}", "MODES=eri", f"BUDGET={b}", "-j2"]
        procs.append((b, subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)))
    for b, p in procs:
        _, err = p.communicate()
        if p.returncode:
            raise SystemExit(f"budget {b}: build failed\n{err.decode()[-2000:]}")
        print(f"budget {b}: built", flush=True)
    with open(os.path.join(args.dir, "classes.txt"), "w") as f:
        f.write(classes)


def time_library(args):
    """Child process: time the ERI kernel of each class with the library
    named by LIBINT_AMD_LIB; prints one JSON object."""
    import numpy as np

    import libint_amd as la
    from libint_amd import molecule
    classes = args.classes.split(",")
    mol = molecule.make_molecule(2)
    B = mol.basis
    ctx = la.Context(0)
    b = la.Basis(ctx, B)
    out = {}
    rng = np.random.default_rng(11)
    for name in classes:
        la_, lb_, lc_, ld_ = (int(x) for x in name)
        bra = [(i, j) for i in range(B.nshell) for j in range(B.nshell)
               if B.l[i] == la_ and B.l[j] == lb_ and (la_ != lb_ or i >= j)]
        ket = [(i, j) for i in range(B.nshell) for j in range(B.nshell)
               if B.l[i] == lc_ and B.l[j] == ld_ and (lc_ != ld_ or i >= j)]
        bq = np.array(bra, np.int32)[rng.choice(len(bra), min(len(bra), 4000), replace=False)]
        kq = np.array(ket, np.int32)[rng.choice(len(ket), min(len(ket), 4000), replace=False)]
        q = np.concatenate([np.repeat(bq, len(kq), 0), np.tile(kq, (len(bq), 1))], 1)
        q = q[rng.choice(len(q), min(len(q), 400000), replace=False)]
        plan = la.EriPlan(b, q)
        d = plan.device_output()
        plan.execute(d)
        ctx.synchronize()
        ms = []
        for _ in range(3):
            plan.execute(d)
            ctx.synchronize()
            ms.append(sum(s["ms"] for s in plan.stats()))
        out[name] = float(np.median(ms))
        del plan
    del b
    ctx.close()
    print(json.dumps(out), flush=True)


def measure(args):
    classes = open(os.path.join(args.dir, "classes.txt")).read().strip()
    res = {}
    for sub in sorted(os.listdir(args.dir)):
        lib = os.path.join(os.path.abspath(args.dir), sub, "libl2g.so")
        if not (sub.startswith("b") and os.path.exists(lib)):
            continue
        env = dict(os.environ, LIBINT_AMD_LIB=lib)
        print(f"budget {sub[1:]}: timing {classes}", flush=True)
        r = subprocess.run([sys.executable, "-m", "libint_amd.tune", "_time", "--classes", classes],
                           env=env, cwd=os.path.dirname(HERE), capture_output=True, text=True, timeout=900)
        if r.returncode:
            raise SystemExit(f"budget {sub[1:]}: timing failed ({r.returncode})\n{r.stderr[-2000:]}")
        res[sub[1:]] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"  {res[sub[1:]]}", flush=True)
    with open(os.path.join(args.dir, "timings.json"), "w") as f:
        json.dump(res, f, indent=1)


def select(args):
    res = json.load(open(os.path.join(args.dir, "timings.json")))
    classes = sorted({c for r in res.values() for c in r})
    entries = []
    for c in classes:
        best = min((r[c], int(b)) for b, r in res.items() if c in r)
        entries.append(f"{c}={best[1]}:0")
        print(f"{c}: budget {best[1]} ({best[0]:.2f} ms); " +
              ", ".join(f"{b}: {r[c]:.2f}" for b, r in sorted(res.items(), key=lambda x: int(x[0])) if c in r))
    with open(args.out, "w") as f:
        f.write(",".join(entries) + "\n")
    print(f"wrote {args.out}")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("step", choices=["build", "measure", "select", "_time"])
    ap.add_argument("--dir", default="tune_build")
    ap.add_argument("--budgets", default=",".join(map(str, DEFAULT_BUDGETS)))
    ap.add_argument("--classes", default="")
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.cfg"))
    args = ap.parse_args()
    {"build": build, "measure": measure, "select": select, "_time": time_library}[args.step](args)


if __name__ == "__main__":
    main()
