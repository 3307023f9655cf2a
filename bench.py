#!/usr/bin/env python3
"""Throughput of the batched FP64 (dd|dd) ERI path on one MI355X.

Workload (the configuration the headline metric is quoted on): 60 atoms at
seeded random positions in a 14 Angstrom cube, per atom s x4, p x3, d x2 with
K ~ U{1..6} primitives and exponents log-uniform in [0.05, 1e4]; every
permutationally-unique (dd|dd) quartet whose Schwarz bound K12*K34 is at
least 1e-12 (K = sqrt(max|(ab|ab)|), computed on the GPU with no primitive
screening).  One step = one pass of the (dd|dd) kernel over that whole
quartet list, inputs (pair data, work list) resident in HBM, output
(625 doubles per quartet, libint layout) written to HBM.

Prints one JSON line.  `roofline` is the (dd|dd) kernel's model FLOP rate
(libint_amd/flop_model.py) over its HIP-event duration on the library's
stream, against the MI355X FP64 peak; `cpu_baseline` times the CPU oracle
(a restatement of libint's Engine::compute2) on a seeded sample of the same
quartets with every host thread.  More legs follow the ERI step (skip with
--no-fock): `config1`, the plain-method 10-atom workload batched, per call
and on the CPU; `next_rows` (attenuated Coulomb, derivatives, 3-centre,
one-body); `eri_f`, a batch of the f recipe; `fock` and `fock_f`, J/K builds
(compute_2body_fock on the GPU) in FP64 and in mixed precision with their
model FLOP rate; `serve` and `serve_f`, Poisson streams of Fock-build
requests through the request server with their p50/p99 latency
(`p99_fock_build_ms`).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
METRIC = "ERI quartets/s + % FP64 peak, (dd|dd) class, 1 GPU; p99 Fock-build latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Model name and logical CPU count of the host (lscpu), for the record."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(B, quartets, seed=5, target_s=4.0, reps=3):
    """The CPU restatement of Engine::compute2 (oracle/), compiled here for
    this host's ISA, one Engine per OpenMP thread on every host thread, on a
    seeded sample of the same quartet list sized for about target_s seconds
    per repetition: a warm-up pass, then `reps` timed passes (steady clock),
    median reported."""
    import oracle
    try:
        isa = oracle.use_native_build(tempfile.mkdtemp(prefix="oracle_native_"))
    except (OSError, subprocess.SubprocessError) as e:  # keep the portable build
        isa = f"portable build (-march=x86-64-v3; native build failed: {e})"
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    rng = np.random.default_rng(seed)
    cal = quartets[rng.choice(len(quartets), min(2048, len(quartets)), replace=False)]
    oracle.eri_quartets(B, cal[:64], nthreads=nthreads)  # table build + warm-up
    t0 = time.perf_counter()
    oracle.eri_quartets(B, cal, nthreads=nthreads)
    rate = len(cal) / max(time.perf_counter() - t0, 1e-6)
    n = int(min(len(quartets), max(2048, rate * target_s)))
    sample = quartets[rng.choice(len(quartets), n, replace=False)]
    oracle.eri_quartets(B, sample, nthreads=nthreads)  # warm-up over the sample
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.eri_quartets(B, sample, nthreads=nthreads)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return dict(value=n / dt, unit="quartets/s", cores=nthreads, kind="port",
                sample=f"{n} (dd|dd) quartets drawn (seed {seed}) from the same screened list; warm-up, then "
                       f"{reps} timed passes of {', '.join(f'{t:.2f}' for t in ts)} s, median; {nthreads} OpenMP "
                       f"threads, one Engine restatement per thread; built {isa}",
                host=host_cpu())


def measured_traffic(nquartets):
    """HBM bytes per launch of the (dd|dd) kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/eri_2222_workload2_traffic.md,
    calibrated against a 1 GiB copy); None when the profile was taken on a
    different quartet list."""
    path = os.path.join(HERE, "profiles", "eri_2222_workload2_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 6 and f[0] == "traffic_bytes_per_launch" and f[3] == "quartets":
                return float(f[2]) if int(f[5]) == nquartets else None
    except OSError:
        pass
    return None


def density(nbf, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(0.0, 0.1, (nbf, max(1, nbf // 5)))
    return C @ C.T


def fock_leg(la, molecule, flop_model, ctx, natoms, reps, workload=2, same_density=False):
    """J/K build (compute_2body_fock on the GPU) for a molecule of a workload
    recipe, FP64 and mixed (FP32 where K12 K34 < 1e-6 and S12 S34 < 16e-6):
    median wall time per build, kernel time, and the largest difference of
    the mixed G from the FP64 G."""
    mol = molecule.make_molecule(workload, natoms=natoms, same_density=same_density)
    B = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(B)
    D = density(mol.nbf, 12)
    spec = molecule.WORKLOADS[workload]
    box = molecule.box_edge(workload, natoms) if same_density else spec["box"]
    shells = "s4p3d2f1, K~U{1..8}, exponents [0.03,2e4]" if workload >= 3 else "s4p3d2, K~U{1..6}"
    out = {"workload": f"{natoms} atoms in a {box:.2f} A cube ({shells}), nbf {mol.nbf}, D = C C^T, "
                       f"C ~ N(0, 0.1), precision 1e-12, SchwarzInf"}
    G = {}
    for name, thr in (("fp64", 0.0), ("mixed", 1e-6)):
        fb.set_fp32_threshold(thr)
        G[name], _, _ = fb.build(D, 1e-12)  # warm-up
        wall, kern = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            G[name], _, _ = fb.build(D, 1e-12)
            wall.append((time.perf_counter() - t0) * 1e3)
            kern.append(fb.stats()["kernel_ms"])
        st = fb.stats()
        cs = fb.class_stats()
        cls = sorted(cs, key=lambda c: -c["ms"])
        # model FLOPs of the build: ERI work per class (flop_model) plus the
        # J/K digest, over the kernels' event time
        flops = sum(flop_model.class_flops(*(int(x) for x in c["cls"]), c["nquartets"], c["nprim"]) +
                    c["nquartets"] * flop_model.fock_digest_flops(*(int(x) for x in c["cls"])) for c in cs)
        kms = float(np.median(kern))
        out[name] = {"build_ms": float(np.median(wall)), "kernel_ms": kms,
                     "quartets": st["nquartets"], "quartets_fp32": st["nquartets_fp32"],
                     "primitive_quartets": st["nprim"],
                     "roofline": {"bound": "fp64_valu", "achieved": flops / (kms * 1e-3) / 1e12,
                                  "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": flops / (kms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                                  "model_flops": flops,
                                  "note": "HGP model FLOPs per class + 12 FLOP of J/K digest per integral, "
                                          "over the sum of the Fock kernels' event times"},
                     "top_classes": [f"{c['cls']}{'/fp32' if c['fp32'] else ''}: {c['ms']:.1f} ms, "
                                     f"{c['nquartets']} quartets, {c['nprim']} primitive quartets"
                                     for c in cls[:6]]}
        log(f"fock {natoms} atoms (recipe {workload}) {name}: build {out[name]['build_ms']:.1f} ms, kernels {out[name]['kernel_ms']:.1f} ms")
    out["mixed_speedup_kernel"] = out["fp64"]["kernel_ms"] / out["mixed"]["kernel_ms"]
    out["mixed_speedup_build"] = out["fp64"]["build_ms"] / out["mixed"]["build_ms"]
    out["mixed_max_abs_dG"] = float(np.abs(G["mixed"] - G["fp64"]).max())
    del fb, B
    return out


def config4_leg(la, molecule, mode, rows_out=None, natoms=200):
    """One J/K build of the 200-atom f recipe (nbf 6000, D = C C^T with C ~
    N(0, 0.1) of nbf/5 columns, seed 12, precision 1e-12, SchwarzInf), FP64
    or mixed (every class's FP32 variant where K_ab K_cd < 1e-6): build and
    kernel time, the classes that take the most time, and (rows_out) the
    rows of G, J, K of molecule.sampled_row_shells plus every 50th row of G
    and its diagonal, for the check against the oracle's exact rows
    (tests/test_config4_rows.py)."""
    mol = molecule.make_molecule(4, natoms=natoms, same_density=natoms != 200)
    shells = molecule.sampled_row_shells(mol)
    sizes = mol.basis.sizes()
    first = np.concatenate([[0], np.cumsum(sizes)])
    rows = np.concatenate([np.arange(first[s], first[s + 1]) for s in shells])
    every = np.arange(0, mol.nbf, 50)
    D = density(mol.nbf, 12)
    ctx = la.Context(0)
    if mode == "mixed":
        ctx.set_fp32_policy(la.FP32_ALWAYS)
    b = la.Basis(ctx, mol.basis)
    fb = la.FockBuilder(b)
    fb.set_fp32_threshold(1e-6 if mode == "mixed" else 0.0)
    state = {"t0": time.perf_counter(), "last": time.perf_counter()}

    def progress(done, total):
        now = time.perf_counter()
        if now - state["last"] > 20 or done == total:
            log(f"  {done:.4e} / {total:.4e} quartets, {now - state['t0']:.0f} s")
            state["last"] = now

    fb.set_progress(progress)
    log(f"config 4: {natoms} atoms, nbf {mol.nbf}, {mode}")
    state["t0"] = time.perf_counter()
    G, J, K = fb.build(D, 1e-12)
    dt = time.perf_counter() - state["t0"]
    st = fb.stats()
    top = sorted(fb.class_stats(), key=lambda c: -c["ms"])[:8]
    out = {"workload": f"{natoms} atoms of the f recipe (s4p3d2f1, K~U{{1..8}}, exponents [0.03,2e4]), nbf {mol.nbf}, "
                       "D = C C^T, C ~ N(0, 0.1) (seed 12), precision 1e-12, SchwarzInf",
           "mode": mode, "build_s": dt, "kernel_s": st["kernel_ms"] / 1e3, "quartets": st["nquartets"],
           "quartets_fp32": st["nquartets_fp32"], "quartets_per_s": st["nquartets"] / dt,
           "top_classes": [f"{c['cls']}{'/fp32' if c['fp32'] else ''}{'/generic' if c['generic'] else ''}: "
                           f"{c['ms'] / 1e3:.1f} s, {c['nquartets']:.3e} quartets, {c['nprim']:.3e} primitive quartets"
                           for c in top],
           "symmetric": bool(np.allclose(G, G.T, rtol=0, atol=1e-10))}
    if rows_out:
        np.savez(rows_out, shells=np.array(shells), rows=rows, G=G[rows], J=J[rows], K=K[rows], every=every,
                 Ge=G[every], diag=np.diag(G), build_s=dt, mode=mode)
    return out


def eri_f_leg(la, molecule, flop_model, ctx, nquartets, reps=3, seed=21):
    """Batched ERIs of the high-angular-momentum recipe (120 atoms in a 20 A
    cube, s4p3d2f1, K~U{1..8}): a seeded sample of quartets with
    K_ab K_cd >= 1e-12 over every class up to (ff|ff), FP64 and mixed
    (FP32 where K_ab K_cd < 1e-6 and S_ab S_cd < 16e-6, l <= the FP32 kernels'
    l), and FP64 with the wave split off (a lane per quartet, or the generic
    kernel for few deep high-l classes); quartets/s and model TFLOP/s of the
    whole batch, and the classes that take the most time."""
    mol = molecule.make_molecule(3)
    B = mol.basis
    b = la.Basis(ctx, B)
    i, j = np.tril_indices(B.nshell)
    K = la.schwarz(b, np.stack([i, j], 1))
    keep = K >= 1e-9
    i, j, K = i[keep], j[keep], K[keep]
    rng = np.random.default_rng(seed)
    picks = []
    got = 0
    while got < nquartets:
        p1 = rng.integers(0, len(K), 4 * nquartets)
        p2 = rng.integers(0, len(K), 4 * nquartets)
        ok = K[p1] * K[p2] >= 1e-12
        picks.append((p1[ok], p2[ok]))
        got += int(ok.sum())
    p1 = np.concatenate([p[0] for p in picks])[:nquartets]
    p2 = np.concatenate([p[1] for p in picks])[:nquartets]
    q = np.stack([i[p1], j[p1], i[p2], j[p2]], 1).astype(np.int32)
    out = {"workload": f"120 atoms, 20 A cube, s4p3d2f1/atom, K~U{{1..8}}, exponents [0.03,2e4]; "
                       f"{nquartets} quartets drawn (seed {seed}) with K_ab K_cd >= 1e-12, all classes to (ff|ff)"}
    for name, thr, wave in (("fp64", 0.0, True), ("fp64_no_wave_split", 0.0, False), ("mixed", 1e-6, True)):
        plan = la.EriPlan(b, q, fp32_threshold=thr)
        if not wave:
            plan.set_wave_split(False)
        d = plan.device_output()
        plan.execute(d)
        ctx.synchronize()
        ms = []
        for _ in range(reps):
            plan.execute(d)
            ctx.synchronize()
            ms.append(plan.elapsed_ms())
        st = plan.stats()
        t = float(np.median(ms))
        fl = sum(flop_model.class_flops(*(int(x) for x in c["cls"]), c["nquartets"], c["nprim"]) for c in st)
        top = sorted(st, key=lambda c: -c["ms"])[:6]
        out[name] = {"ms": t, "quartets_per_s": nquartets / (t * 1e-3), "model_tflops": fl / (t * 1e-3) / 1e12,
                     "quartets_fp32": sum(c["nquartets"] for c in st if c["fp32"]),
                     "top_classes": [f"{c['cls']}{'/fp32' if c['fp32'] else ''}{'/generic' if c['generic'] else ''}"
                                     f"{'/wave' if c['wave'] else ''}: "
                                     f"{c['ms']:.1f} ms, {c['nquartets']} quartets, "
                                     f"{flop_model.class_flops(*(int(x) for x in c['cls']), c['nquartets'], c['nprim']) / (c['ms'] * 1e9):.2f} TFLOP/s"
                                     for c in top]}
        log(f"eri f-recipe {name}: {t:.1f} ms, {out[name]['quartets_per_s']:.3e} quartets/s, "
            f"{out[name]['model_tflops']:.2f} TFLOP/s model")
        del plan
    out["mixed_speedup"] = out["fp64"]["ms"] / out["mixed"]["ms"]
    del b
    return out


def config1_leg(la, molecule, ctx, reps=3, cpu_target_s=3.0, ncall=3000, seed=11):
    """The plain-method workload (10 atoms in a 6 A cube; per atom s x3 with
    K 6, 1, 1, p x2 with K 3, 1, d x1; exponents [0.1, 5e3]; spherical d):
    every unique quartet with K_ab K_cd >= 1e-12, all classes to (dd|dd).
    GPU: one batched plan over the whole list (median of `reps` executions,
    HIP events), and the drop-in per-call path (the resident single-quartet
    evaluator, one call per quartet as libint's Engine::compute loop makes
    them) on the first `ncall` quartets.  CPU: the restatement of
    Engine::compute2 on every host thread over the whole list, or a seeded
    sample sized for about cpu_target_s per pass when the whole list would
    take longer (warm-up, then `reps` passes, median), with the GPU values of
    the same quartets checked against it at the per-integral bar."""
    import oracle
    mol = molecule.make_molecule(1)
    B = mol.basis
    b = la.Basis(ctx, B)
    i, j = np.tril_indices(B.nshell)
    pairs = np.stack([i, j], 1).astype(np.int32)
    K = la.schwarz(b, pairs, precision=0.0)
    p1, p2 = np.triu_indices(len(pairs))
    keep = K[p1] * K[p2] >= 1e-12
    q = np.stack([i[p1[keep]], j[p1[keep]], i[p2[keep]], j[p2[keep]]], 1).astype(np.int32)
    plan = la.EriPlan(b, q)
    offs = plan.offsets()
    d = plan.device_output()
    plan.execute(d)
    ctx.synchronize()
    ms = []
    for _ in range(reps):
        plan.execute(d)
        ctx.synchronize()
        ms.append(plan.elapsed_ms())
    t_gpu = float(np.median(ms)) * 1e-3
    g, _ = plan.download()
    # the drop-in per-call path
    qe = la.QuartetEngine(ctx, int(B.nprim.max()), int(B.l.max()))
    off = np.concatenate([[0], np.cumsum(B.nprim)])
    shells = [la.Shell(int(B.l[s]), list(zip(B.alpha[off[s]:off[s + 1]], B.coeff[off[s]:off[s + 1]])),
                       B.origin[3 * s:3 * s + 3], pure=bool(B.pure[s]), renorm=False) for s in range(B.nshell)]
    calls = q[:ncall]
    qe.compute([shells[x] for x in calls[0]])
    t0 = time.perf_counter()
    for quad in calls:
        qe.compute([shells[x] for x in quad])
    t_call = (time.perf_counter() - t0) / len(calls)
    qe._release()
    # CPU restatement
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    rng = np.random.default_rng(seed)
    cal = q[rng.choice(len(q), min(4096, len(q)), replace=False)]
    oracle.eri_quartets(B, cal[:64], nthreads=nthreads)
    t0 = time.perf_counter()
    oracle.eri_quartets(B, cal, nthreads=nthreads)
    rate = len(cal) / max(time.perf_counter() - t0, 1e-6)
    if rate * cpu_target_s >= len(q):
        idx = np.arange(len(q))
    else:
        idx = np.sort(rng.choice(len(q), int(rate * cpu_target_s), replace=False))
    sample = q[idx]
    o, ooff, _ = oracle.eri_quartets(B, sample, nthreads=nthreads)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.eri_quartets(B, sample, nthreads=nthreads)
        ts.append(time.perf_counter() - t0)
    t_cpu = float(np.median(ts))
    ends = np.append(offs[1:], g.size)
    gs = np.concatenate([g[offs[k]:ends[k]] for k in idx])
    dd = np.abs(gs - o)
    nfail = int(((dd > 1e-12) & (dd > 1e-10 * np.abs(o))).sum())
    out = {"workload": f"10 atoms, 6 A cube, s3p2d1/atom (K 6,1,1 / 3,1 / 1), exponents [0.1,5e3], spherical d; "
                       f"{len(q)} of {len(p1)} unique quartets with K_ab K_cd >= 1e-12, all classes",
           "quartets": int(len(q)),
           "gpu_batched": {"ms": t_gpu * 1e3, "quartets_per_s": len(q) / t_gpu},
           "gpu_per_call": {"calls": int(len(calls)), "us_per_call": t_call * 1e6,
                            "quartets_per_s": 1.0 / t_call,
                            "note": "Python ctypes call per quartet into the resident single-quartet evaluator "
                                    "(one copy in, one launch, one copy out, a wait)"},
           "cpu": {"quartets": int(len(sample)), "full_list": bool(len(sample) == len(q)), "s": t_cpu,
                   "quartets_per_s": len(sample) / t_cpu, "cores": nthreads,
                   "timing": f"warm-up, then {reps} passes of {', '.join(f'{x:.2f}' for x in ts)} s, median"},
           "gpu_batched_over_cpu": (len(q) / t_gpu) / (len(sample) / t_cpu),
           "checked_integrals": int(gs.size), "max_abs_vs_cpu": float(dd.max(initial=0.0)),
           "values_beyond_1e-12_abs_and_1e-10_rel": nfail}
    log(f"config 1: {len(q)} quartets, GPU batched {t_gpu * 1e3:.1f} ms ({len(q) / t_gpu:.3e} q/s), "
        f"per call {t_call * 1e6:.0f} us, CPU {len(sample) / t_cpu:.3e} q/s on {nthreads} threads, "
        f"max |gpu - cpu| {dd.max(initial=0.0):.1e}")
    del plan, b
    return out


def next_rows_leg(la, mol, B, q, ctx, nq=400000, nderiv=50000, seed=7):
    """Throughput of the paths beside the headline one, on the same 60-atom
    molecule: erf/erfc-attenuated Coulomb (omega 0.4) on the first nq
    headline quartets next to plain Coulomb on them; first derivatives of
    nderiv of them; 3-centre (X|ab) for the headline's d pairs; the one-body
    matrices with the 60 nuclei as charges.  GPU times only (ERI plans by
    their events, the rest by wall clock around a synchronised call)."""
    out = {}
    qs = q[:nq]
    for name, oper in (("coulomb", la.COULOMB), ("erf_coulomb", la.ERF_COULOMB), ("erfc_coulomb", la.ERFC_COULOMB)):
        plan = la.EriPlan(B, qs)
        if oper != la.COULOMB:
            plan.set_operator(oper, 0.4)
        d = plan.device_output()
        plan.execute(d)
        ctx.synchronize()
        ms = []
        for _ in range(3):
            plan.execute(d)
            ctx.synchronize()
            ms.append(plan.elapsed_ms())
        t = float(np.median(ms))
        out[name] = {"quartets": int(len(qs)), "ms": t, "quartets_per_s": len(qs) / (t * 1e-3)}
        log(f"{name} (dd|dd) x{len(qs)}: {t:.1f} ms, {len(qs) / (t * 1e-3):.3e} quartets/s")
        del plan
    qd = q[:nderiv]
    dp = la.Deriv1Plan(B, qd)
    dp.execute()
    ctx.synchronize()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        dp.execute()
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out["deriv1"] = {"quartets": int(len(qd)), "ms": t * 1e3, "quartets_per_s": len(qd) / t,
                     "derivative_integrals_per_s": 12 * dp.size / t}
    log(f"first derivatives (dd|dd) x{len(qd)}: {t * 1e3:.1f} ms, {len(qd) / t:.3e} quartets/s "
        f"({12 * dp.size / t:.3e} derivative integrals/s)")
    del dp
    # 3-centre (X|ab): X over the shells of the first 20 atoms, ab the headline's d pairs
    a = mol.basis
    n = a.nshell
    unit = la.BasisArrays(np.append(a.l, 0), np.append(a.pure, 0), np.append(a.nprim, 1), np.append(a.alpha, 0.0),
                          np.append(a.coeff, 1.0), np.append(a.origin, [0.0, 0.0, 0.0]))
    bu = la.Basis(ctx, unit)
    X = np.nonzero(mol.atom_of_shell < 20)[0]
    pairs = np.unique(np.sort(q[:, :2], axis=1)[:, ::-1], axis=0)[:2000]
    q3 = np.stack([np.repeat(X, len(pairs)), np.full(len(X) * len(pairs), n),
                   np.tile(pairs[:, 0], len(X)), np.tile(pairs[:, 1], len(X))], 1).astype(np.int32)
    plan = la.EriPlan(bu, q3)
    d = plan.device_output()
    plan.execute(d)
    ctx.synchronize()
    ms = []
    for _ in range(3):
        plan.execute(d)
        ctx.synchronize()
        ms.append(plan.elapsed_ms())
    t = float(np.median(ms))
    out["three_centre"] = {"triples": int(len(q3)), "ms": t, "triples_per_s": len(q3) / (t * 1e-3)}
    log(f"3-centre (X|ab) x{len(q3)}: {t:.1f} ms, {len(q3) / (t * 1e-3):.3e} triples/s")
    del plan, bu
    charges = [(8.0, c) for c in mol.coords]
    la.one_body(B, charges)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        la.one_body(B, charges)
        ts.append(time.perf_counter() - t0)
    out["one_body"] = {"nbf": mol.nbf, "charges": len(charges), "ms": float(np.median(ts)) * 1e3}
    log(f"one-body S, T, V nbf {mol.nbf}, {len(charges)} charges: {np.median(ts) * 1e3:.1f} ms (with copies)")
    return out


def serve_leg(la, molecule, nreq, rate_hz, seed=5000, atoms=(2, 12), f_recipe=False, deterministic=False):
    """Fock-build requests arriving as a Poisson stream, served with
    continuous batching; latency = submit to result.  Molecules of atoms[0]
    to atoms[1] atoms (log-uniform): the two s-d recipes alternating, or
    (f_recipe) the s4p3d2f1 recipe at the 200-atom molecule's density."""
    rng = np.random.default_rng(seed)
    mols = []
    for k in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1])))))
        if f_recipe:
            m = molecule.make_molecule(4, natoms=n, seed=seed + k, same_density=True)
        else:
            m = molecule.make_molecule(1 + k % 2, natoms=n)
        mols.append((m, density(m.nbf, seed + k)))
    gaps = rng.exponential(1.0 / rate_hz, nreq)
    srv = la.Server(0, deterministic=deterministic)
    try:
        srv.result(srv.submit(mols[0][0].basis, mols[0][1], 1e-12))  # warm-up
        st0 = srv.stats()
        t0 = time.perf_counter()
        ids = []
        for (m, D), g in zip(mols, gaps):
            time.sleep(g)
            ids.append(srv.submit(m.basis, D, 1e-12))
        lat = np.array([srv.result(i)[3][1] for i in ids])
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    recipe = "the f recipe at the 200-atom density" if f_recipe else "the two s-d recipes"
    out = {"workload": f"{nreq} requests, Poisson {rate_hz:g}/s, {atoms[0]}-{atoms[1]} atoms of {recipe} (nbf "
                       f"{min(m.nbf for m, _ in mols)}-{max(m.nbf for m, _ in mols)}), precision 1e-12"
                       f"{', deterministic digest' if deterministic else ''}",
           "work_items_per_s": (st["work_items"] - st0["work_items"]) / wall,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "requests_per_s": nreq / wall, "rounds": st["rounds"],
           "max_requests_per_round": st["max_requests_per_round"]}
    log(f"serve: p50 {out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms, {out['requests_per_s']:.1f} req/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--max-quartets", type=int, default=0, help="cap the quartet list (0 = all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fock", action="store_true", help="skip the J/K build and serving legs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (the known-size kernel that calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--fock-atoms", type=int, default=20, help="atoms of the J/K build leg")
    ap.add_argument("--fock-f-atoms", type=int, default=10,
                    help="atoms of the f-shell J/K build leg (0: skip), at the 200-atom recipe's density")
    ap.add_argument("--eri-f-quartets", type=int, default=400000,
                    help="quartets of the high-angular-momentum ERI leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=150, help="requests of the serving leg")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--fock-f-only", action="store_true",
                    help="only the f-recipe J/K build leg (for profiling its kernels) and exit")
    ap.add_argument("--eri-f-only", action="store_true",
                    help="only the f-recipe ERI batch leg (for profiling its kernels) and exit")
    ap.add_argument("--config4", choices=["fp64", "mixed"],
                    help="only the 200-atom f-recipe J/K build (about 16 min on one GPU) and exit")
    ap.add_argument("--rows-out", help="with --config4: .npz of the sampled rows for tests/test_config4_rows.py")
    ap.add_argument("--serve-f-requests", type=int, default=24,
                    help="requests of the f-recipe serving leg (10-20 atoms at the 200-atom density; 0: skip)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.empty(2 ** 27, dtype=torch.float64, device="cuda")
        x.fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)  # 1 GiB read + 1 GiB written
        torch.cuda.synchronize()
        print("copied 1 GiB", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import libint_amd as la
    from libint_amd import flop_model, molecule

    if args.config4:
        print(json.dumps(config4_leg(la, molecule, args.config4, args.rows_out)), flush=True)
        return

    if args.eri_f_only:
        ctx = la.Context(0)
        print(json.dumps(eri_f_leg(la, molecule, flop_model, ctx, args.eri_f_quartets)), flush=True)
        return

    if args.fock_f_only:
        ctx = la.Context(0)
        print(json.dumps(fock_leg(la, molecule, flop_model, ctx, args.fock_f_atoms or 10, reps=2, workload=4,
                                  same_density=True)), flush=True)
        return

    t_setup = time.perf_counter()
    mol = molecule.make_molecule(2)
    ctx = la.Context(0)
    B = la.Basis(ctx, mol.basis)
    q, n_unique = molecule.headline_quartets(B, 1e-12)
    if args.max_quartets and len(q) > args.max_quartets:
        q = q[: args.max_quartets]
    log(f"workload: {len(q)} screened (dd|dd) quartets of {n_unique} unique")
    plan = la.EriPlan(B, q)
    d_out = plan.device_output()
    t_setup = time.perf_counter() - t_setup
    log(f"setup {t_setup:.1f}s; output {plan.size * 8 / 1e9:.2f} GB")

    for _ in range(args.warmup):
        plan.execute(d_out)
    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(d_out)
        ctx.synchronize()
        kernel_ms.append(sum(s["ms"] for s in plan.stats()))
    dt = time.perf_counter() - t0
    st = plan.stats()
    assert len(st) == 1 and st[0]["cls"] == "2222", st
    nprim = st[0]["nprim"]
    ms_step = dt / args.steps * 1e3
    value = len(q) / (dt / args.steps)
    # per-unit work: the measurement plan's (dd|dd) count per primitive
    # quartet (24,116 with prefactors and Boys) and the HGP + solid-harmonics
    # count per quartet; flop_model's own derivation of the primitive part is
    # reported beside it
    flops = flop_model.class_flops(2, 2, 2, 2, len(q), nprim, model="survey")
    flops_derived = flop_model.class_flops(2, 2, 2, 2, len(q), nprim, model="derived")
    kms = float(np.mean(kernel_ms))
    achieved = flops / (kms * 1e-3) / 1e12
    log(f"step {ms_step:.1f} ms, kernel {kms:.1f} ms, {value:.3e} quartets/s, {achieved:.2f} TFLOP/s model")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "60 atoms, 14 A cube, s4p3d2/atom, K~U{1..6}, exponents [0.05,1e4]; "
                        "(dd|dd) class, Schwarz >= 1e-12, precision eps, Original screening",
            "quartets": int(len(q)),
            "unique_dddd_quartets": int(n_unique),
            "primitive_quartets": int(nprim),
            "outputs_per_quartet": 625,
        },
        "pct_fp64_peak": 100.0 * achieved / FP64_PEAK_TFLOPS,
        "roofline": {
            "bound": "fp64_valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(len(q)),
            "kernel": "l2g_eri_2222",
            "kernel_ms": kms,
            "model_flops_per_launch": flops,
            "model_flops_per_launch_derived": flops_derived,
            "note": "FP64 vector (VALU) roof, 78.6 TFLOP/s; the kernel issues no MFMA. FLOPs from the canonical "
                    "HGP model (24,116 per primitive quartet + 21,992 per quartet), not from instruction counts",
        },
        "setup_s": t_setup,
    }
    # the same list in mixed precision (FP32 where K12 K34 < 1e-6 and
    # S12 S34 < 16e-6, in the classes where FP32 was measured faster)
    g64 = None
    if not args.no_fock:
        mplan = la.EriPlan(B, q, fp32_threshold=1e-6)
        dm = mplan.device_output()
        mplan.execute(dm)
        ctx.synchronize()
        mms = []
        for _ in range(args.steps):
            mplan.execute(dm)
            ctx.synchronize()
            mms.append(sum(s["ms"] for s in mplan.stats()))
        n32 = sum(s["nquartets"] for s in mplan.stats() if s["fp32"])
        gm, _ = mplan.download()
        del mplan, dm
        plan.execute(d_out)
        g64, _ = plan.download()
        t_m = float(np.mean(mms))
        out["mixed"] = {"kernel_ms": t_m, "quartets_fp32": int(n32), "quartets_per_s": len(q) / (t_m * 1e-3),
                        "speedup_vs_fp64": kms / t_m, "max_abs_vs_fp64": float(np.abs(gm - g64).max())}
        log(f"mixed: {t_m:.1f} ms ({n32} of {len(q)} quartets in FP32), {kms / t_m:.2f}x the FP64 kernel, "
            f"max |mixed - fp64| {out['mixed']['max_abs_vs_fp64']:.1e}")
        del gm, g64
    del plan
    if not args.no_fock:
        out["config1"] = config1_leg(la, molecule, ctx)
        out["next_rows"] = next_rows_leg(la, mol, B, q, ctx)
    if not args.no_fock and args.eri_f_quartets:
        out["eri_f"] = eri_f_leg(la, molecule, flop_model, ctx, args.eri_f_quartets)
    if not args.no_fock:
        out["fock"] = fock_leg(la, molecule, flop_model, ctx, args.fock_atoms, reps=2)
        if args.fock_f_atoms:
            out["fock_f"] = fock_leg(la, molecule, flop_model, ctx, args.fock_f_atoms, reps=2, workload=4,
                                     same_density=True)
        out["serve"] = serve_leg(la, molecule, args.serve_requests, rate_hz=args.serve_rate)
        if args.serve_f_requests:
            out["serve_f"] = serve_leg(la, molecule, args.serve_f_requests, rate_hz=args.serve_rate, seed=7000,
                                       atoms=(10, 20), f_recipe=True, deterministic=True)
        out["p99_fock_build_ms"] = out["serve"]["p99_ms"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mol.basis, q)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
