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
        seeded quartets (s-d classes: the 60-atom s4p3d2 recipe; f classes:
        the 120-atom s4p3d2f1 recipe), each library in a child process;
        writes D/timings.json
  python -m libint_amd.tune select  --dir D [--out libint_amd/tuned.cfg]
        writes the fastest budget of each measured class as a
        class=budget:0:fp32 entry, merged into the file's other entries; the
        Makefile passes them to the generator (CLASSCFG)
  python -m libint_amd.tune fp32 [--classes ...] [--out libint_amd/tuned.cfg]
        (GPU) times each class's FP64 kernel and its FP32 variant in the
        built library on the quartets a mixed batch would run in FP32
        (K K < 1e-6, S S < 16e-6) and sets the fp32 field (1: the FP32
        variant is at least 5 % faster there, so mixed batches route to it
        by default); rebuild the library afterwards
  python -m libint_amd.tune fewgen [--out libint_amd/tuned.cfg]
        (GPU) for each class whose quartets of the f-recipe batch run a wave
        per quartet, times them on its register kernel and on the generic
        kernel (quartets split over blocks) and sets the fewgen field (1: the
        generic kernel is at least 10 % faster, so such quartets go there);
        rebuild the library afterwards
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_BUDGETS = (128, 144, 160)
DEFAULT_BUDGET = 160  # the Makefile's BUDGET


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
               f"ONLY={classes}", "MODES=eri,fp32", f"BUDGET={b}", "CLASSCFG=", "-j2"]
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
    ctx = la.Context(0)
    bases = {}
    out = {}
    rng = np.random.default_rng(11)
    for name in classes:
        la_, lb_, lc_, ld_ = (int(x) for x in name)
        # s-d classes on the 60-atom recipe, f classes on the 120-atom one
        w = 3 if max(la_, lb_, lc_, ld_) == 3 else 2
        if w not in bases:
            m = molecule.make_molecule(w)
            bases[w] = (m.basis, la.Basis(ctx, m.basis))
        B, b = bases[w]
        bra = [(i, j) for i in range(B.nshell) for j in range(B.nshell)
               if B.l[i] == la_ and B.l[j] == lb_ and (la_ != lb_ or i >= j)]
        ket = [(i, j) for i in range(B.nshell) for j in range(B.nshell)
               if B.l[i] == lc_ and B.l[j] == ld_ and (lc_ != ld_ or i >= j)]
        # significant pairs (Schwarz factor >= 1e-9), quartets with
        # K_ab K_cd >= 1e-12, as the workloads screen them
        bra = np.array(bra, np.int32)
        ket = np.array(ket, np.int32)
        kb, sb = la.precision_bounds(b, bra)
        kk, sk = (kb, sb) if (la_, lb_) == (lc_, ld_) else la.precision_bounds(b, ket)
        bra, kb, sb = bra[kb >= 1e-9], kb[kb >= 1e-9], sb[kb >= 1e-9]
        ket, kk, sk = ket[kk >= 1e-9], kk[kk >= 1e-9], sk[kk >= 1e-9]
        i1 = rng.integers(0, len(bra), 1600000)
        i2 = rng.integers(0, len(ket), 1600000)
        sig = kb[i1] * kk[i2] >= 1e-12
        # the quartets a mixed batch at threshold 1e-6 runs in FP32
        # (K K < 1e-6 and S S < 16e-6): where the FP32 variant is compared
        elig = sig & (kb[i1] * kk[i2] < 1e-6) & (sb[i1] * sk[i2] < la.FP32_HRR_FACTOR * 1e-6)
        qe = np.concatenate([bra[i1[elig][:400000]], ket[i2[elig][:400000]]], 1)
        i1, i2 = i1[sig][:400000], i2[sig][:400000]
        q = np.concatenate([bra[i1], ket[i2]], 1)
        variants = [("", 0.0, q)]
        if args.fp32 and max(la_, lb_, lc_, ld_) <= la.fp32_max_l() and len(qe):
            # the eligible quartets in FP64 and, every one below the bounds, in FP32
            variants += [("/e64", 0.0, qe), ("/fp32", 1e300, qe)]
        for suffix, thr, qq in variants:
            ctx.set_fp32_policy(la.FP32_ALWAYS)
            plan = la.EriPlan(b, qq, fp32_threshold=thr)
            ctx.set_fp32_policy(la.FP32_MEASURED)
            if thr and not all(s["fp32"] for s in plan.stats()):
                raise SystemExit(f"{name}: the FP32 plan is not all FP32")
            d = plan.device_output()
            plan.execute(d)
            ctx.synchronize()
            ms = []
            for _ in range(3):
                plan.execute(d)
                ctx.synchronize()
                ms.append(sum(s["ms"] for s in plan.stats()))
            out[name + suffix] = float(np.median(ms))
            del plan
    bases.clear()
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
        r = subprocess.run([sys.executable, "-m", "libint_amd.tune", "_time", "--fp32", "--classes", classes],
                           env=env, cwd=os.path.dirname(HERE), capture_output=True, text=True, timeout=900)
        if r.returncode:
            raise SystemExit(f"budget {sub[1:]}: timing failed ({r.returncode})\n{r.stderr[-2000:]}")
        res[sub[1:]] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"  {res[sub[1:]]}", flush=True)
    with open(os.path.join(args.dir, "timings.json"), "w") as f:
        json.dump(res, f, indent=1)


def read_cfg(path):
    """class -> [budget, waves, fp32, fewgen] of a tuned.cfg file."""
    out = {}
    try:
        text = open(path).read().strip()
    except OSError:
        return out
    for item in filter(None, text.split(",")):
        k, v = item.split("=")
        f = [int(x) for x in v.split(":")]
        out[k] = f + [0] * (4 - len(f))
    return out


def write_cfg(path, cfg):
    """Entries that differ from the defaults (budget 160, waves 0, fp32 0,
    fewgen 0); the fewgen field only where set."""
    keep = {c: v for c, v in cfg.items() if list(v) != [DEFAULT_BUDGET, 0, 0, 0]}
    with open(path, "w") as f:
        f.write(",".join(f"{c}={v[0]}:{v[1]}:{v[2]}" + (f":{v[3]}" if v[3] else "")
                         for c, v in sorted(keep.items())) + "\n")
    print(f"wrote {path}")


def fp32_pays(t64, t32):
    return 1 if t32 is not None and t32 < 0.95 * t64 else 0


def select(args):
    res = json.load(open(os.path.join(args.dir, "timings.json")))
    classes = sorted({c for r in res.values() for c in r if "/" not in c})
    cfg = read_cfg(args.out)
    for c in classes:
        best = min((r[c], int(b)) for b, r in res.items() if c in r)
        r = res[str(best[1])]
        cfg[c] = [best[1], 0, fp32_pays(r.get(c + "/e64", r[c]), r.get(c + "/fp32")), cfg.get(c, [0] * 4)[3]]
        print(f"{c}: budget {best[1]} ({best[0]:.2f} ms), fp32 {cfg[c][2]}; " +
              ", ".join(f"{b}: {r_[c]:.2f}" for b, r_ in sorted(res.items(), key=lambda x: int(x[0])) if c in r_))
    write_cfg(args.out, cfg)


def fp32(args):
    """FP64 kernel against FP32 variant of each class, with the built library."""
    sys.path.insert(0, os.path.join(HERE, "codegen"))
    import gen_eri
    classes = args.classes or ",".join("".join(map(str, c)) for c in gen_eri.canonical_classes(2))
    r = subprocess.run([sys.executable, "-m", "libint_amd.tune", "_time", "--fp32", "--classes", classes],
                       cwd=os.path.dirname(HERE), capture_output=True, text=True, timeout=1200)
    if r.returncode:
        raise SystemExit(f"timing failed ({r.returncode})\n{r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    cfg = read_cfg(args.out)
    for c in classes.split(","):
        if c not in res:
            continue
        v = cfg.setdefault(c, [DEFAULT_BUDGET, 0, 0, 0])
        t64, t32 = res.get(c + "/e64"), res.get(c + "/fp32")
        v[2] = fp32_pays(t64, t32) if t64 is not None else 0
        print(f"{c}: all significant quartets fp64 {res[c]:.2f} ms; FP32-eligible quartets fp64 "
              f"{t64 if t64 is None else round(t64, 2)} ms, fp32 {t32 if t32 is None else round(t32, 2)} ms "
              f"-> fp32 {v[2]}", flush=True)
    write_cfg(args.out, cfg)


def canonical_class(ls):
    """Class name of a quartet's four angular momenta as the plans name
    it: each pair high-l first, bra and ket swapped when the bra's total
    angular momentum is the larger."""
    p1 = tuple(sorted(ls[:2], reverse=True))
    p2 = tuple(sorted(ls[2:], reverse=True))
    if sum(p1) > sum(p2):
        p1, p2 = p2, p1
    return "%d%d%d%d" % (p1 + p2)


def fewgen(args):
    """Few deep quartets: each class's register kernel (a wave per quartet)
    against the generic kernel (quartets split over blocks) on that class's
    quartets of the f-recipe batch (400k quartets of the 120-atom molecule,
    seed 21, as bench.py's eri_f leg), each class alone, median of 3."""
    import numpy as np

    import libint_amd as la
    from libint_amd import molecule
    ctx = la.Context(0)
    mol = molecule.make_molecule(3)
    B = mol.basis
    b = la.Basis(ctx, B)
    i, j = np.tril_indices(B.nshell)
    K = la.schwarz(b, np.stack([i, j], 1))
    keep = K >= 1e-9
    i, j, K = i[keep], j[keep], K[keep]
    rng = np.random.default_rng(21)
    n, picks, got = 400000, [], 0
    while got < n:
        p1 = rng.integers(0, len(K), 4 * n)
        p2 = rng.integers(0, len(K), 4 * n)
        ok = K[p1] * K[p2] >= 1e-12
        picks.append((p1[ok], p2[ok]))
        got += int(ok.sum())
    p1 = np.concatenate([p[0] for p in picks])[:n]
    p2 = np.concatenate([p[1] for p in picks])[:n]
    q = np.stack([i[p1], j[p1], i[p2], j[p2]], 1).astype(np.int32)
    names = np.array([canonical_class(tuple(int(B.l[s]) for s in row)) for row in q])
    cfg = read_cfg(args.out)

    def timed(qc, mode):
        plan = la.EriPlan(b, qc)
        plan.set_wave_split(mode)
        d = plan.device_output()
        plan.execute(d)
        ms = []
        for _ in range(3):
            plan.execute(d)
            ctx.synchronize()
            ms.append(plan.elapsed_ms())
        kinds = {("generic" if c["generic"] else "wave" if c["wave"] else "lane") for c in plan.stats()}
        return float(np.median(ms)), kinds

    for c in sorted(set(names)):
        qc = q[names == c]
        tw, kw = timed(qc, 2)
        if kw != {"wave"}:
            continue  # not a class of few deep quartets here
        ctx.set_generic_kernels(True)
        try:
            tg, _ = timed(qc, 2)
        finally:
            ctx.set_generic_kernels(False)
        v = cfg.setdefault(c, [DEFAULT_BUDGET, 0, 0, 0])
        v[3] = 1 if tg < 0.9 * tw else 0
        print(f"{c}: {len(qc)} quartets, wave {tw:.2f} ms, generic {tg:.2f} ms -> fewgen {v[3]}", flush=True)
    write_cfg(args.out, cfg)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("step", choices=["build", "measure", "select", "fp32", "fewgen", "_time"])
    ap.add_argument("--fp32", action="store_true", help="_time: also the FP32 variants")
    ap.add_argument("--dir", default="tune_build")
    ap.add_argument("--budgets", default=",".join(map(str, DEFAULT_BUDGETS)))
    ap.add_argument("--classes", default="")
    ap.add_argument("--out", default=os.path.join(HERE, "tuned.cfg"))
    args = ap.parse_args()
    {"build": build, "measure": measure, "select": select, "fp32": fp32, "fewgen": fewgen,
     "_time": time_library}[args.step](args)


if __name__ == "__main__":
    main()
