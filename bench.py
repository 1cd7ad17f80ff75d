#!/usr/bin/env python3
"""Benchmark: time-to-1e-8 of the 823,875-DOF tet elasticity solve on one
MI355X, FP32-inner + FP64-refinement vs plain FP64 ILU(0)-PCG.

Workload: 64^3 hexes x 6 tets (274,625 nodes, 1,572,864 tets), E_e =
exp(0.5 z_e) from the seeded portable RNG, nu = 0.3, z=0 clamped, body
force (0,0,-1e-3), 4-colour node order.  One step = CopyValue + Decompose
+ the full solve to relative residual 1e-8 with the matrix, factors' pattern
and right-hand side resident in HBM.  The JSON line reports the
FP32+refinement time (value) beside the FP64 time, DOF*iterations/s, the
HBM roofline of the dominant kernel and the CPU reference (the reference's
own code, compiled) solving the same cube in full, in its natural node order
and in the GPU's.  The same line then carries, under other_workloads, one
run of each other workload (--workload cube128, batch512, newmark64,
stiffwarp64, adjoint96, order64: the 128^3 mixed solve to 1e-10, the
512-mesh batch, the two time steps, the compliance gradient and the cube
from its natural node order, as given and multicolour-relabelled), each
measured by a child process after the headline's timed steps (--others none
skips them).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def spmv_bytes(nb: int, nnzb: int, s: int, read_y: bool, sv: int = 0) -> int:
    """Algorithmic bytes of one 3x3 BSR MatVec: off-diagonal blocks and their
    column indices, diagonal blocks, row starts, x read once, y written
    (+ read once when y is scaled by beta); s = bytes per stored value, sv =
    bytes per vector entry (default s)."""
    sv = sv or s
    ndof = 3 * nb
    b = nnzb * (9 * s + 4) + nb * 9 * s + (nb + 1) * 4 + 2 * ndof * sv
    return b + (ndof * sv if read_y else 0)


def kernel_bytes(nb: int, nnzb: int, s: int, n0: int = 0, nu0: int = 0, sv: int = 0) -> dict:
    """Algorithmic bytes of one launch class of the fused PCG iteration (s =
    bytes per scalar; column indices 4 B; nnzb off-diagonal blocks, nL =
    nnzb/2 lower = upper blocks; n0 rows without lower blocks = forward level 0, computed by the
    update; nu0 rows without upper blocks = backward level 0, no work):
      spmv   off-diagonal blocks + indices, diagonal blocks, row starts,
             p read, Ap written
      update r, Ap, x, p read; r, x written; D^-1 of the n0 level-0 rows
             read and their Pr written; two 4 B masks per 21-row slice
      fwd    (levels >= 1) lower blocks + indices, D^-1 blocks, 8 B per row
             of schedule, r read, Pr written
      bwd    (levels >= 1) upper blocks + indices, 8 B per row, Pr read and
             written, r read for (r, Pr)
      pupd   p, Pr read; p written
    s: bytes per stored matrix / factor value; sv: per vector entry (default s)"""
    sv = sv or s
    ndof = 3 * nb
    nl = nnzb // 2
    nf, nbk = nb - n0, nb - nu0
    return {
        "spmv": spmv_bytes(nb, nnzb, s, False, sv),
        "update": 6 * ndof * sv + n0 * (9 * s + 3 * sv) + 8 * ((nb + 20) // 21),
        "fwd_sweep": nl * (9 * s + 4) + nf * 9 * s + nf * 8 + 2 * 3 * nf * sv,
        "bwd_sweep": nl * (9 * s + 4) + nbk * 8 + 3 * 3 * nbk * sv,
        "pupdate": 3 * ndof * sv,
    }


def host_cpu() -> dict:
    """CPU model (the kernel's model name, with family / model / stepping,
    clock and cache size as /proc/cpuinfo reports them) and logical core
    count of the machine the CPU baseline ran on."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # the first processor's block
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        pass
    model = info.get("model name") or platform.processor() or platform.machine()
    detail = ", ".join(f"{k} {info[k]}" for k in ("cpu family", "model", "stepping", "cpu MHz",
                                                  "cache size") if k in info)
    return {"cpu_model": f"{model} ({detail})" if detail else model, "nproc": os.cpu_count()}


def stream_copy_gbs(mib: int = 1024, reps: int = 10) -> float:
    """Measured device-to-device copy rate (read + write bytes per second, GB/s)
    of the library's plain 16-byte-vector copy of `mib` MiB, the achievable
    counterpart of the 8 TB/s specification peak."""
    from delfem2_amd import solver as S

    return S.measure_copy_gbs(mib << 20, reps)


def level0_rows(psup_ind, psup, nb: int):
    """(rows with no lower neighbour, rows with no upper neighbour) of the
    one-ring pattern: the sizes of forward / backward level 0."""
    import numpy as np
    ind = np.asarray(psup_ind, np.int64)
    ps = np.asarray(psup, np.int64)
    cnt = np.diff(ind)
    rows = np.repeat(np.arange(nb), cnt)
    has_lo = np.zeros(nb, bool)
    has_up = np.zeros(nb, bool)
    has_lo[rows[ps < rows]] = True
    has_up[rows[ps > rows]] = True
    return int((~has_lo).sum()), int((~has_up).sum())


def pcg_iter_bytes(nb: int, nnzb: int, s: int, n0: int = 0, nu0: int = 0, sv: int = 0) -> int:
    return sum(kernel_bytes(nb, nnzb, s, n0, nu0, sv).values())


def cpu_reference_run(prob, tol: float = 1e-8, max_it: int = 20000, keep_x: bool = False) -> dict:
    """The CPU reference (oracle/_ref: the reference's own code compiled) or,
    without it, the oracle restatement, on the whole workload, single
    thread: pattern + assembly + BC, Decompose, and Solve_PCG to `tol`, each
    timed.  Returns the phase times, the iteration count and the final
    relative residual."""
    import numpy as np

    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    t0 = time.perf_counter()
    ind, ps = oracle.psup(prob.tet, prob.np_)
    Sys = oracle.RefSystem(ind, ps) if kind == "reference" else oracle.System(ind, ps)
    b = np.zeros(prob.ndof)
    Sys.assemble(b, prob.myu, prob.lam, prob.rho, prob.g, prob.xyz, prob.tet)
    if prob.f_ext is not None:
        b += prob.f_ext
    Sys.set_fixed_bc(prob.flag)
    oracle.set_rhs_zero(b, prob.flag)
    t1 = time.perf_counter()
    Sys.factor()
    t2 = time.perf_counter()
    x, hist = Sys.solve_pcg(b.copy(), tol, max_it)
    t3 = time.perf_counter()
    # a run stopped by the cap ends with a repeated entry (lsitrsol.h:146)
    out = {"kind": kind, "setup_s": t1 - t0, "decompose_s": t2 - t1, "pcg_s": t3 - t2,
           "iterations": int(min(len(hist) - 1, max_it)),
           "rel_residual": float(hist[-1] / hist[0]) if hist[0] > 0 else 0.0}
    if keep_x:
        out["x"] = x
    return out


def cpu_worker(spec: str) -> None:
    """Child-process entry (--cpu-worker n:sigma:ordering): the CPU reference
    run of one cube workload, as one JSON line on stdout."""
    from delfem2_amd import workload as W
    n, sigma, ordering = spec.split(":")
    prob = W.cube(int(n), sigma=float(sigma), seed=20261203, ordering=ordering)
    out = cpu_reference_run(prob)
    out["ndof"] = prob.ndof
    print(json.dumps(out), flush=True)


def cpu_baseline(n: int, sigma: float, ordering: str) -> dict:
    """The CPU reference solved in full to 1e-8, single thread, in the
    reference's own (natural) node order -- how its callers run it -- and in
    the GPU workload's order, as two child processes side by side (one core
    each).  value = the natural-order time-to-1e-8 (Decompose + PCG)."""

    orders = ["natural"] + ([ordering] if ordering != "natural" else [])
    procs = {o: subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                                  f"{n}:{sigma}:{o}"], stdout=subprocess.PIPE, text=True)
             for o in orders}
    runs = {}
    for o, p in procs.items():
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"CPU reference run ({o} order) failed")
        runs[o] = json.loads(out.strip().splitlines()[-1])
    nat = runs["natural"]

    def solve_s(r):
        return r["decompose_s"] + r["pcg_s"]

    res = {
        "value": round(solve_s(nat), 3),
        "unit": "s (time-to-1e-8 = Decompose + FP64 PCG, measured in full)",
        "cores": 1,
        "kind": nat["kind"],
        "measured_iterations": nat["iterations"],
        "sample": (f"the whole workload ({nat['ndof']} DOF) in the reference's natural node "
                   f"order: pattern+assembly+BC {nat['setup_s']:.2f} s, Decompose "
                   f"{nat['decompose_s']:.3f} s, {nat['iterations']} PCG iterations "
                   f"{nat['pcg_s']:.2f} s ({nat['pcg_s'] / max(nat['iterations'], 1) * 1e3:.1f} "
                   f"ms/iter), final relative residual {nat['rel_residual']:.2e}; the "
                   f"{ordering} order beside it, in a second process on another core"),
        "dof_iter_per_s": round(nat["ndof"] * nat["iterations"] / nat["pcg_s"], 1),
        **host_cpu(),
    }
    for o in orders[1:]:
        r = runs[o]
        res[o] = {"value": round(solve_s(r), 3), "measured_iterations": r["iterations"],
                  "setup_s": round(r["setup_s"], 3), "decompose_s": round(r["decompose_s"], 3),
                  "pcg_s": round(r["pcg_s"], 3), "rel_residual": r["rel_residual"],
                  "dof_iter_per_s": round(r["ndof"] * r["iterations"] / r["pcg_s"], 1)}
    return res


def cpu_baseline_sample(prob, iters_fp64: int, max_cpu_iters: int) -> dict:
    """For workloads whose full CPU solve takes many minutes (128^3): set-up,
    Decompose and `max_cpu_iters` PCG iterations of the CPU reference, the
    time-to-tolerance extrapolated with the GPU's FP64 iteration count."""
    r = cpu_reference_run(prob, 1e-30, max_cpu_iters)
    n_it = max(r["iterations"], 1)
    t_it = r["pcg_s"] / n_it
    return {
        "value": round(r["decompose_s"] + t_it * iters_fp64, 3),
        "unit": "s (Decompose + FP64 PCG, iterations x measured s/iteration)",
        "cores": 1,
        "kind": r["kind"],
        "measured_iterations": n_it,
        "sample": (f"{prob.name} ({prob.ndof} DOF): pattern+assembly+BC {r['setup_s']:.2f} s, "
                   f"Decompose {r['decompose_s']:.3f} s, {n_it} PCG iterations "
                   f"{r['pcg_s']:.2f} s ({t_it * 1e3:.1f} ms/iter) extrapolated to "
                   f"{iters_fp64} iterations"),
        "dof_iter_per_s": round(prob.ndof * n_it / r["pcg_s"], 1),
        **host_cpu(),
    }


def cpu_baseline_batch(bat, it64, budget_s: float = 15.0):
    """The CPU reference (or the oracle) solving meshes of the batch one after
    another (pattern + assembly + Decompose + FP64 PCG to 1e-8, single
    thread) until `budget_s` has passed; the whole batch extrapolated by the
    per-mesh work weight ndof_m x iterations_m (FP64 GPU counts)."""
    import numpy as np

    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    ndof = np.diff(bat.node_offset.astype(np.int64)) * 3
    w = ndof * np.maximum(it64.astype(np.int64), 1)
    t0 = time.perf_counter()
    done = 0
    for m in range(bat.nmesh):
        p = bat.mesh(m)
        ind, ps = oracle.psup(p.tet, p.np_)
        Sys = oracle.RefSystem(ind, ps) if kind == "reference" else oracle.System(ind, ps)
        b = np.zeros(p.ndof)
        Sys.assemble(b, p.myu, p.lam, p.rho, p.g, p.xyz, p.tet)
        Sys.set_fixed_bc(p.flag)
        oracle.set_rhs_zero(b, p.flag)
        Sys.factor()
        Sys.solve_pcg(b, 1e-8, 20000)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t = time.perf_counter() - t0
    est = t * w.sum() / w[:done].sum()
    return {"value": round(est, 3),
            "unit": "s (all meshes: set-up + Decompose + FP64 PCG to 1e-8, extrapolated)",
            "cores": 1, "kind": kind,
            "sample": f"first {done} of {bat.nmesh} meshes solved in {t:.2f} s; scaled by "
                      f"sum(ndof x iterations) of all meshes / of the sample",
            **host_cpu()}


def bench_batch(args):
    """512 independent cube meshes (4..24 hexes per axis, lognormal E, random
    body forces), one batched solve to 1e-8 per step: CopyValue + Decompose of
    the union, then the batch solver (FP32 + FP64 reliable updates; FP64
    alongside)."""
    import numpy as np
    import torch

    from delfem2_amd import workload as W
    from delfem2_amd.problem import BatchSystem

    print("[bench] building the 512-mesh batch", file=sys.stderr, flush=True)
    bat = W.batch()
    G = BatchSystem(bat, device=0)
    assert G.assemble()
    assert G.factor()
    ndof = np.diff(bat.node_offset.astype(np.int64)) * 3
    tune = {}
    for fp32 in (True, False):
        best, ms = G.solver.tune(fp32)
        tune["fp32" if fp32 else "fp64"] = {"best": G.solver.variant(fp32)[1],
                                            "ms_per_iter": [round(v, 4) for v in ms]}

    def step(fp32):
        G.factor()
        return G.solve_batch(1e-8, fp32=fp32, delta=args.delta)

    for _ in range(args.warmup):
        step(True)
        step(False)
    res = {}
    for name, fp32 in (("mixed", True), ("fp64", False)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            x, it, rr, st = step(fp32)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        assert np.all(st == 1)
        res[name] = {"time_s": round(t, 5), "iters_min": int(it.min()),
                     "iters_max": int(it.max()), "iters_mean": round(float(it.mean()), 2),
                     "dof_iter_per_s": round(float((ndof * it).sum()) / t, 1),
                     "_it": it}
        print(f"[bench] batch {name} {t * 1e3:.1f} ms, iterations {it.min()}..{it.max()}",
              file=sys.stderr, flush=True)
    it64 = res["fp64"].pop("_it")
    res["mixed"].pop("_it")
    res["mixed"]["speedup_vs_fp64"] = round(res["fp64"]["time_s"] / res["mixed"]["time_s"], 3)
    cpu = cpu_baseline_batch(bat, it64) if args.cpu_iters > 0 else None
    out = {
        "metric": "time-to-1e-8 of 512 independent tet meshes, batched FP32+refine (s)",
        "value": res["mixed"]["time_s"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(res["mixed"]["time_s"] * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-inner/f64-refine", "data": "synthetic (seeded voxel tet cubes)",
        "config": {"workload": f"512 cube meshes, 4..24 hexes per axis x6 tets, "
                               f"{int(ndof.sum())} DOF total, E~lognormal(0,0.5), nu=0.3, "
                               f"g~U[-1e-3,1e-3]^3, color4 order, ILU(0)-PCG to 1e-8 per mesh",
                   "ndof": int(ndof.sum()), "nmesh": bat.nmesh,
                   "ndof_min": int(ndof.min()), "ndof_max": int(ndof.max())},
        "mixed": res["mixed"], "fp64": res["fp64"], "autotune": tune,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """SIMP compliance gradient of the 96x48x48 cantilever (1,327,104 tets,
    698,691 DOF, rho_e ~ U(0.2, 1), E_e = rho_e^3, nu = 0.3, x=0 clamped, unit
    load on the central 3x3 nodes of the x=max face, colour-4 order): one
    step = E from rho, assembly, fixed BCs, ILU(0) CopyValue + Decompose,
    FP64 PCG to 1e-10 for u, the adjoint (K lam = f: the forward system, so
    lam = u) and the per-element contraction dC/drho_e = -3 rho_e^2 lam_e^T
    K_e^1 u_e.  Value: seconds per gradient (all 1.33M elements)."""
    import numpy as np
    import torch

    from delfem2_amd import solver as S
    from delfem2_amd import workload as W
    from delfem2_amd.adjoint import ElasticSystem, compliance_and_grad, unit_lame

    tol, p = 1e-10, 3.0
    print("[bench] building the 96x48x48 cantilever", file=sys.stderr, flush=True)
    prob = W.cantilever(96, 48, 48, seed=20261204, ordering="color4")
    rho_np = prob.extra["density"]
    rho = torch.from_numpy(rho_np).cuda()
    G = ElasticSystem(prob)
    for _ in range(args.warmup):
        compliance_and_grad(G, rho, p, tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Cv, grad, u = compliance_and_grad(G, rho, p, tol)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # one more gradient, phase by phase (synchronised: for the breakdown only)
    ph = {}

    def timed(name, fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ph[name] = round((time.perf_counter() - t) * 1e3, 3)
        return r

    timed("assembly_bc_ms", lambda: (G.set_modulus(rho ** p), G.assemble()))
    timed("ilu_copy_decompose_ms", G.factor)
    uu, hist, _ = timed("forward_pcg_ms", lambda: G.solve(tol, 20000))
    ph["forward_iterations"] = len(hist) - 1
    # dC/du = f: the adjoint system is the forward one, so the backward pass
    # takes lam = u (what a second, deterministic solve would return)
    r = G.b.clone()
    S.setRHS_Zero(r, G.flag, 0)
    assert torch.equal(r, G.b)
    lam = uu
    ph["adjoint_pcg_ms"] = 0.0
    ph["adjoint_iterations"] = 0
    timed("contraction_ms", lambda: G.adjoint_grad(uu, lam))
    nb, nnzb = prob.np_, G.A.ncrs
    # the contraction reads coordinates (4 nodes), tets, u and lam per element
    # and writes one value: algorithmic bytes per element
    con_bytes = prob.ne * (16 + 4 * 24 + 2 * 4 * 24 + 8)
    ach = con_bytes / (ph["contraction_ms"] * 1e-3) / 1e9
    cpu = None
    if args.cpu_iters > 0:
        import oracle
        print("[bench] CPU baseline: the reference's forward solve in full",
              file=sys.stderr, flush=True)
        m1, l1 = unit_lame(prob.extra["nu"])
        E = rho_np ** p
        cprob = W.cantilever(96, 48, 48, seed=20261204, ordering="color4")
        cprob.myu, cprob.lam = E * m1, E * l1
        r = cpu_reference_run(cprob, tol, 20000, keep_x=True)
        uc = r.pop("x")
        # the contraction has no reference counterpart: the oracle's, timed
        t_d = time.perf_counter()
        oracle.adjoint_grad(m1, l1, prob.xyz, prob.tet, uc, uc)
        t_con = time.perf_counter() - t_d
        value = r["setup_s"] + r["decompose_s"] + r["pcg_s"] + t_con
        cpu = {"value": round(value, 3), "unit": "s per gradient (measured)", "cores": 1,
               "kind": r["kind"], "measured_iterations": r["iterations"],
               "sample": f"the whole gradient, single thread: the reference's pattern + "
                         f"assembly + BC {r['setup_s']:.2f} s, Decompose "
                         f"{r['decompose_s']:.3f} s, forward PCG to {tol} in "
                         f"{r['iterations']} iterations {r['pcg_s']:.2f} s (the adjoint reuses "
                         f"it: lam = u), and the oracle's element contraction {t_con:.3f} s "
                         f"(no reference counterpart)",
               **host_cpu()}
    out = {
        "metric": "SIMP compliance gradient dC/drho of the 96x48x48 tet cantilever (s)",
        "value": round(t_step, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SIMP densities)",
        "config": {"workload": f"adjoint compliance gradient, cantilever 96x48x48 hex x6 tets, "
                               f"{prob.ndof} DOF, {prob.ne} elements, rho~U(0.2,1), E=rho^3, "
                               f"nu=0.3, ILU(0)-PCG to {tol} (adjoint = forward for the "
                               f"compliance), colour-4",
                   "ndof": prob.ndof, "elements": prob.ne},
        "compliance": float(Cv),
        "phases": ph,
        "roofline": {"bound": "hbm", "kernel": "k_adjoint_grad_tet", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": int(con_bytes)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_newmark(args):
    """Implicit Newmark-beta time stepping of the 64^3 cube (823,875 DOF, the
    headline cube's mesh, material and z=0 clamp; gravity (0, 0, -3)): one
    step = assembly of M + beta dt^2 K and the Newmark residual, fixed BCs,
    ILU(0) CopyValue + Decompose, FP64 Solve_PCG to 1e-4, and the u / v / a
    updates -- the step of 618_Fem3Tet/main.cpp:361-416 with its gamma = 1,
    beta = 0.5, dt = 0.03 and tolerance.  Value: seconds per step (steps
    from rest; the state evolves over the timed steps)."""
    import numpy as np
    import torch

    from delfem2_amd import workload as W
    from delfem2_amd.problem import DynamicSystem

    dt, gamma, beta, tol = 0.03, 1.0, 0.5, 1e-4
    print(f"[bench] building the {args.n}^3 Newmark problem", file=sys.stderr, flush=True)
    prob = W.cube(args.n, sigma=args.sigma, seed=20261203, ordering=args.ordering)
    prob.g = np.array([0.0, 0.0, -3.0])
    G = DynamicSystem(prob, "newmark", dt=dt, gamma_newmark=gamma, beta_newmark=beta)
    for _ in range(args.warmup):
        G.step(tol, 1000)
    torch.cuda.synchronize()
    its = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its.append(G.step(tol, 1000))
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # one more step, phase by phase (synchronised: for the breakdown only)
    ph = {}
    import delfem2_amd.solver as S

    def timed(name, fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ph[name] = round((time.perf_counter() - t) * 1e3, 3)
        return r

    timed("assembly_bc_ms", G.assemble)
    timed("ilu_copy_decompose_ms", G.factor)
    x = torch.zeros_like(G.b)
    hist = timed("pcg_ms", lambda: S.Solve_PCG(G.b, x, tol, 1000, G.A, G.ilu))
    ph["pcg_iterations"] = len(hist) - 1

    def upd():
        S.XPlusAYBZCW(G.u, G.flag, dt, G.v, 0.5 * dt * dt, G.a, dt * dt * beta, x)
        S.XPlusAYBZ(G.v, G.flag, dt * gamma, x, dt, G.a)
        S.XPlusAY(G.a, G.flag, 1.0, x)
    timed("update_ms", upd)
    nb, nnzb = prob.np_, G.A.ncrs
    # dominant phase against HBM: its algorithmic bytes over its time
    phase_bytes = {
        # A and b written once, coordinates / tets / Lame pairs / u, v, a read once
        "assembly_bc_ms": (nnzb + nb) * 72 + nb * 24 + prob.ne * 32 + nb * 24 * 4,
        # A values read, factors written (CopyValue + Decompose)
        "ilu_copy_decompose_ms": 2 * (nnzb + nb) * 72,
        # FP64 PCG iterations
        "pcg_ms": pcg_iter_bytes(nb, nnzb, 8) * max(ph["pcg_iterations"], 1),
    }
    dom = max(phase_bytes, key=lambda k: ph[k])
    ach = phase_bytes[dom] / (ph[dom] * 1e-3) / 1e9
    cpu = None
    if args.cpu_iters > 0:
        import oracle
        print("[bench] CPU baseline sample (one Newmark step)", file=sys.stderr, flush=True)
        kind = "reference" if oracle.ref_available() else "port"
        run = oracle.DynamicRun(prob, "newmark", dt=dt, gamma_newmark=gamma, beta_newmark=beta,
                                use_ref=kind == "reference")
        tc = time.perf_counter()
        it_c = run.step(tol, 1000)
        tc = time.perf_counter() - tc
        cpu = {"value": round(tc, 3), "unit": "s/step", "cores": 1, "kind": kind,
               "measured_iterations": it_c,
               "sample": f"one Newmark step from rest of the same {args.n}^3 problem "
                         f"({it_c} PCG iterations to {tol})", **host_cpu()}
    out = {
        "metric": "Newmark-beta implicit time step of the 823k-DOF tet cube (s/step)",
        "value": round(t_step, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded voxel tet cube, lognormal E)",
        "config": {"workload": f"Newmark-beta step, tet cube {args.n}^3 hex x6 tets, "
                               f"{prob.ndof} DOF, E~lognormal(0,{args.sigma}), nu=0.3, rho=1, "
                               f"g=(0,0,-3), dt={dt}, gamma={gamma}, beta={beta}, "
                               f"ILU(0)-PCG to {tol}, {args.ordering} order",
                   "ndof": prob.ndof},
        "pcg_iterations_per_step": its,
        "phases": ph,
        "roofline": {"bound": "hbm", "phase": dom, "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_phase": int(phase_bytes[dom])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_stiffwarp(args):
    """Stiffness-warped backward-Euler time stepping of the 64^3 cube
    (823,875 DOF, the headline cube's mesh, material and z=0 clamp; gravity
    (0, 0, -3)): one step = node rotations (polar decomposition over each
    one-ring), assembly of R K R^T + lumped M / dt^2 and its residual, fixed
    BCs, ILU(0) CopyValue + Decompose, device BiCGStab to 1e-4 (1000
    iterations at most) and the u / v updates -- the step of
    603_Fem3DStiffwarp/main.cpp:180-227 with its dt and tolerance; E = 1,
    nu = 0.3 everywhere (the reference's driver takes one Lame pair).  Value:
    seconds per step (steps from rest; the state evolves)."""
    import numpy as np
    import torch

    import delfem2_amd.solver as S
    from delfem2_amd import workload as W
    from delfem2_amd.problem import DynamicSystem

    dt, tol = 0.03, 1e-4
    print(f"[bench] building the {args.n}^3 stiffness-warped problem", file=sys.stderr,
          flush=True)
    # one Lame pair for the whole mesh: the reference's driver takes scalars
    prob = W.cube(args.n, sigma=0.0, seed=20261203, ordering=args.ordering)
    prob.g = np.array([0.0, 0.0, -3.0])
    G = DynamicSystem(prob, "stiffwarp", dt=dt)
    for _ in range(args.warmup):
        G.step(tol, 1000)
    torch.cuda.synchronize()
    its = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        its.append(G.step(tol, 1000))
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    ph = {}

    def timed(name, fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ph[name] = round((time.perf_counter() - t) * 1e3, 3)
        return r

    timed("rotations_assembly_bc_ms", G.assemble)
    timed("ilu_copy_decompose_ms", G.factor)
    x = torch.zeros_like(G.b)
    r = G.b.clone()
    hist = timed("bicgstab_ms", lambda: S.Solve_PBiCGStab(r, x, tol, 1000, G.A, G.ilu))
    ph["bicgstab_iterations"] = len(hist)
    nb, nnzb = prob.np_, G.A.ncrs
    kb = kernel_bytes(nb, nnzb, 8)
    # per iteration: two MatVecs, two full ILU solves (level 0 included), and
    # the vector updates (s: 3 passes, x / r: 7, p: 4 of n doubles)
    it_bytes = (2 * spmv_bytes(nb, nnzb, 8, True) + 2 * (kb["fwd_sweep"] + kb["bwd_sweep"])
                + 14 * 8 * 3 * nb)
    ach = it_bytes * max(len(hist), 1) / (ph["bicgstab_ms"] * 1e-3) / 1e9
    cpu = None
    if args.cpu_iters > 0:
        import oracle
        print("[bench] CPU baseline (one stiffness-warped step)", file=sys.stderr, flush=True)
        kind = "reference" if oracle.ref_available() else "port"
        run = oracle.DynamicRun(prob, "stiffwarp", dt=dt, use_ref=kind == "reference")
        tc = time.perf_counter()
        it_c = run.step(tol, 1000)
        tc = time.perf_counter() - tc
        cpu = {"value": round(tc, 3), "unit": "s/step", "cores": 1, "kind": kind,
               "measured_iterations": it_c,
               "sample": f"one stiffness-warped step from rest of the same {args.n}^3 problem "
                         f"({it_c} BiCGStab iterations to {tol})", **host_cpu()}
    out = {
        "metric": "stiffness-warped backward-Euler time step of the 823k-DOF tet cube (s/step)",
        "value": round(t_step, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded voxel tet cube, lognormal E)",
        "config": {"workload": f"stiffness-warped step, tet cube {args.n}^3 hex x6 tets, "
                               f"{prob.ndof} DOF, E=1, nu=0.3, rho=1, "
                               f"g=(0,0,-3), dt={dt}, ILU(0)-BiCGStab to {tol}, "
                               f"{args.ordering} order",
                   "ndof": prob.ndof},
        "bicgstab_iterations_per_step": its,
        "phases": ph,
        "roofline": {"bound": "hbm", "phase": "bicgstab_ms", "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_iteration": int(it_bytes)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def kernel_table(G, S, nb, nnzb, n0, nu0, step_mixed, step_fp64, args) -> dict:
    """Per-kernel times of the fused iterations (HIP events on the solver
    stream, right after a solve of the same precision: same buffers, same
    kernels), with each kernel class's algorithmic bytes and GB/s."""
    nlev_f, nlev_b = G.ilu.levels()
    nlev = {"fwd_sweep": max(nlev_f - 1, 1), "bwd_sweep": max(nlev_b - 1, 1)}
    kern = {}
    for prec, fp32, s in (("fp32", True, 4), ("fp64", False, 8)):
        if fp32:
            step_mixed()
        else:
            step_fp64()
        ms = S.pcg_kernel_times(G.A, G.ilu, fp32, args.spmv_reps)
        sv = 8 if (not fp32 or args.mixed_vectors == 64) else 4
        kb = kernel_bytes(nb, nnzb, s, n0, nu0, sv)
        kern[prec] = {}
        for k, b in kb.items():
            launches = nlev.get(k, 1)
            kern[prec][k] = {"ms_per_call": round(ms[k], 5), "launches_per_call": launches,
                             "bytes_per_call": b,
                             "gbs": round(b / (ms[k] * 1e-3) / 1e9, 1)}
        for k in ("alpha_step", "spmv_back_to_back"):
            kern[prec][k + "_ms"] = round(ms[k], 5)
        kern[prec]["beta_step_folded"] = ms["folded"]
    return kern


def bench_cube128(args):
    """128^3 hexes x 6 tets (6,440,067 DOF), E lognormal(0, 1.0), nu ~ U[0.25,
    0.35], 4-colour order: FP32 inner + FP64 reliable updates to relative
    residual 1e-10 (value) beside plain FP64 ILU(0)-PCG to 1e-10; one step =
    CopyValue + Decompose + solve.  Error of the mixed solution against an
    FP64 solve to 1e-12 is reported."""
    import torch

    from delfem2_amd import solver as S
    from delfem2_amd import workload as W
    from delfem2_amd.problem import DeviceSystem

    print("[bench] building the 128^3 problem", file=sys.stderr, flush=True)
    prob = W.cube(128, sigma=1.0, nu_range=(0.25, 0.35), ordering="color4")
    G = DeviceSystem(prob, device=0)
    G.ilu.set_mixed_options(args.mixed_vectors == 64, args.update_period)
    assert G.assemble()
    nb, nnzb = prob.np_, G.A.ncrs

    def step(fp32):
        G.factor()
        if fp32:
            return G.solve_mixed(1e-10, delta=args.delta, max_outer=400, max_inner=40000)
        return G.solve(1e-10, 40000)

    for _ in range(args.warmup):
        step(True)
        step(False)
    res = {}
    for name, fp32 in (("mixed", True), ("fp64", False)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = step(fp32)
        torch.cuda.synchronize()
        res[name] = ((time.perf_counter() - t0) / args.steps, out)
    t_m, (xm, hm, n_inner) = res["mixed"]
    t_64, (x64, h64, _) = res["fp64"]
    it64 = len(h64) - 1
    xr, _, _ = G.solve(1e-12, 40000)
    err = (torch.linalg.norm(xm - xr) / torch.linalg.norm(xr)).item()
    err64 = (torch.linalg.norm(x64 - xr) / torch.linalg.norm(xr)).item()
    rr = G.b.clone()
    G.A.MatVec(rr, -1.0, xm, 1.0)
    relres = (torch.linalg.norm(rr) / torch.linalg.norm(G.b)).item()
    print(f"[bench] 128^3 mixed {t_m * 1e3:.1f} ms ({n_inner} inner), fp64 {t_64 * 1e3:.1f} ms "
          f"({it64} it), err {err:.2e}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_iters > 0:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline_sample(prob, it64, min(args.cpu_iters, 20))
        cpu["unit"] = "s (time-to-1e-10 = Decompose + FP64 PCG, iterations x measured s/iteration)"
    n0, nu0 = level0_rows(G.psup_ind, G.psup, nb)
    b32 = pcg_iter_bytes(nb, nnzb, 4, n0, nu0, 4 if args.mixed_vectors == 32 else 8)
    kern = kernel_table(G, S, nb, nnzb, n0, nu0, lambda: step(True), lambda: step(False), args)
    out = {
        "metric": "time-to-1e-10 PCG at 6.44M DOF: FP32+refine vs FP64 (s)",
        "value": round(t_m, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_m * 1e3, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-inner/f64-refine" if args.mixed_vectors == 32 else
                 "f32-matrix+factors/f64-vectors+refine",
        "data": "synthetic (seeded voxel tet cube, lognormal E, uniform nu)",
        "config": {"workload": f"tet elasticity cube 128^3 hex x6 tets, {prob.ndof} DOF, "
                               f"E~lognormal(0,1.0), nu~U[0.25,0.35], ILU(0)-PCG to 1e-10, "
                               f"color4 order", "ndof": prob.ndof, "nnzb": nnzb},
        "mixed": {"time_s": round(t_m, 5), "inner_iterations": n_inner,
                  "outer_iterations": len(hm) - 1, "true_rel_residual": relres,
                  "x_rel_err_vs_fp64_1e-12": err,
                  "dof_iter_per_s": round(prob.ndof * n_inner / t_m, 1),
                  "hbm_frac_per_iter": round(b32 * n_inner / t_m / 1e9 / HBM_PEAK_GBS, 4),
                  "speedup_vs_fp64": round(t_64 / t_m, 3)},
        "x_reference": "the device FP64 solve of the same system to 1e-12 (its FP64 path is "
                       "checked against the CPU reference on the smaller configurations; a CPU "
                       "solve of this one to 1e-12 takes hours)",
        "fp64": {"time_s": round(t_64, 5), "iterations": it64, "x_rel_err_vs_fp64_1e-12": err64,
                 "dof_iter_per_s": round(prob.ndof * it64 / t_64, 1)},
        "kernels": kern,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_order64(args):
    """The 823,875-DOF cube (E lognormal(0, 0.5), nu = 0.3) as a caller's
    mesh gives it, in its natural node order (193 dependency levels per
    triangular sweep), against the same mesh relabelled by
    dfm2h_multicolor_order (4 levels): FP64 ILU(0)-PCG to 1e-8 on both, the
    FP32+refine solve on the relabelled one; one step = CopyValue + Decompose
    + solve.  value = the relabelled FP64 step.  The solutions agree once the
    relabelled one is mapped back (reported)."""
    import numpy as np
    import torch

    from delfem2_amd import workload as W
    from delfem2_amd.problem import DeviceSystem

    print("[bench] building the 64^3 problem in natural order", file=sys.stderr, flush=True)
    prob = W.cube(64, sigma=args.sigma, seed=20261203, ordering="natural")
    t0 = time.perf_counter()
    new2old, ncolor = W.multicolor_permutation(prob.tet, prob.np_)
    t_order = time.perf_counter() - t0
    xyz, tet, _ = W.permute_mesh(prob.xyz, prob.tet, new2old)
    flag = np.ascontiguousarray(prob.flag.reshape(-1, 3)[new2old].reshape(-1))
    relabelled = W.Problem("cube64_multicolor", xyz, tet, prob.myu, prob.lam, flag, rho=prob.rho,
                           g=prob.g, ordering="multicolor")
    res, xs = {}, {}
    for name, P, kinds in (("natural", prob, ("fp64",)),
                           ("multicolor", relabelled, ("fp64", "mixed"))):
        G = DeviceSystem(P, device=0)
        G.ilu.set_mixed_options(args.mixed_vectors == 64, args.update_period)
        assert G.assemble()
        res[name] = {"levels_per_sweep": int(G.ilu.levels()[0])}
        for kind in kinds:
            def step():
                assert G.factor()
                if kind == "mixed":
                    x, h, n = G.solve_mixed(1e-8, delta=args.delta)
                    return x, n
                x, h, _ = G.solve(1e-8, 20000)
                return x, len(h) - 1
            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                x, it = step()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / args.steps
            res[name][kind] = {"time_s": round(t, 5), "iterations": it}
            xs[(name, kind)] = x.cpu().numpy().reshape(-1, 3)
        print(f"[bench] {name}: {res[name]}", file=sys.stderr, flush=True)
    x_nat = xs[("natural", "fp64")]
    err = {}
    for kind in ("fp64", "mixed"):
        xb = np.empty_like(x_nat)
        xb[new2old] = xs[("multicolor", kind)]
        err[kind] = float(np.linalg.norm(xb - x_nat) / np.linalg.norm(x_nat))
    t_mc = res["multicolor"]["fp64"]["time_s"]
    out = {
        "metric": "time-to-1e-8 FP64 PCG at 823k DOF from the caller's natural node order: "
                  "multicolour-relabelled vs as given (s)",
        "value": t_mc, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mc * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded voxel tet cube, lognormal E)",
        "config": {"workload": f"tet elasticity cube 64^3 hex x6 tets, {prob.ndof} DOF, "
                               f"E~lognormal(0,{args.sigma}), nu=0.3, ILU(0)-PCG to 1e-8, the "
                               f"natural node order and its multicolour relabelling",
                   "ndof": prob.ndof, "colours": ncolor},
        "ordering_s": round(t_order, 4),
        "natural": res["natural"], "multicolor": res["multicolor"],
        "speedup_vs_natural": round(res["natural"]["fp64"]["time_s"] / t_mc, 2),
        "x_rel_err_vs_natural": err,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cube",
                    choices=["cube", "cube128", "batch512", "newmark64", "adjoint96",
                             "stiffwarp64", "order64"],
                    help="cube: the 823,875-DOF single solve to 1e-8 (headline); cube128: "
                         "the 6,440,067-DOF mixed-precision solve to 1e-10; batch512: 512 "
                         "independent meshes in one batched solve; newmark64: implicit "
                         "Newmark-beta time steps of the 823,875-DOF cube; adjoint96: the "
                         "compliance gradient of the 96x48x48 cantilever; stiffwarp64: "
                         "stiffness-warped time steps (BiCGStab) of the 823,875-DOF cube; "
                         "order64: the 823,875-DOF cube from its natural node order, as given "
                         "and relabelled by the multicolour ordering")
    ap.add_argument("--n", type=int, default=64, help="hexes per axis (64 = 823,875 DOF)")
    ap.add_argument("--sigma", type=float, default=0.5)
    ap.add_argument("--ordering", default="color4",
                    choices=["color8", "color4", "multicolor", "natural"])
    ap.add_argument("--delta", type=float, default=1e-1,
                    help="mixed: refine in FP64 whenever the FP32 residual fell by this factor")
    ap.add_argument("--update-period", type=int, default=100,
                    help="mixed: force a reliable update after this many inner iterations "
                         "(0: only the delta rule)")
    ap.add_argument("--mixed-vectors", type=int, default=32, choices=[32, 64],
                    help="Krylov vector precision of the mixed solver (matrix and factors "
                         "are FP32 either way)")
    ap.add_argument("--cpu-iters", type=int, default=200,
                    help="CPU baseline: 0 = skip; cube: any other value runs the reference's "
                         "full solves; other workloads: PCG iterations of the bounded sample")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--spmv-reps", type=int, default=50,
                    help="launches per kernel class when timing the fused kernels")
    ap.add_argument("--others", default=",".join(OTHER_WORKLOADS),
                    help="cube only: after the headline, also run these workloads (comma "
                         "list, or 'none'), each as a child process with --steps 1 --warmup 1, "
                         "and fold their results into the same JSON line")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MB (the known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return

    if args.copy_mb:
        import torch
        src = torch.ones(args.copy_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 8, "launches": 5}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of "
                                   "it (see DESIGN.md, sharding)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "batch512":
        bench_batch(args)
        return
    if args.workload == "cube128":
        bench_cube128(args)
        return
    if args.workload == "newmark64":
        bench_newmark(args)
        return
    if args.workload == "adjoint96":
        bench_adjoint(args)
        return
    if args.workload == "stiffwarp64":
        bench_stiffwarp(args)
        return
    if args.workload == "order64":
        bench_order64(args)
        return

    import torch

    from delfem2_amd import solver as S
    from delfem2_amd import workload as W
    from delfem2_amd.problem import DeviceSystem

    print(f"[bench] building {args.n}^3 problem ({args.ordering})", file=sys.stderr, flush=True)
    prob = W.cube(args.n, sigma=args.sigma, seed=20261203, ordering=args.ordering)
    G = DeviceSystem(prob, device=0)
    G.ilu.set_mixed_options(args.mixed_vectors == 64, args.update_period)
    torch.cuda.synchronize()
    t = time.perf_counter()
    assert G.assemble()
    torch.cuda.synchronize()
    t_asm_first = time.perf_counter() - t  # includes building the element lists
    t_warm = []
    for _ in range(5):
        t = time.perf_counter()
        assert G.assemble()
        torch.cuda.synchronize()
        t_warm.append(time.perf_counter() - t)
    t_asm = sorted(t_warm)[2]
    nb, nnzb = prob.np_, G.A.ncrs
    print(f"[bench] ndof={prob.ndof} nnzb={nnzb} setup={G.t_setup:.2f}s asm={t_asm * 1e3:.1f}ms",
          file=sys.stderr, flush=True)

    def step_fp64():
        G.factor()
        return G.solve(1e-8, 20000)

    def step_mixed():
        G.factor()
        return G.solve_mixed(1e-8, delta=args.delta, max_outer=200, max_inner=20000)

    for _ in range(args.warmup):
        step_fp64()
        step_mixed()
    torch.cuda.synchronize()

    # CopyValue + Decompose alone (part of every step below; T_pcg = step - this)
    t0 = time.perf_counter()
    for _ in range(max(args.steps, 3)):
        G.factor()
    torch.cuda.synchronize()
    t_factor = (time.perf_counter() - t0) / max(args.steps, 3)

    # ---- timed: FP32 + FP64 refinement (the headline) ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        xm, hm, n_inner = step_mixed()
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps

    # ---- timed: plain FP64 ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x64, h64, _ = step_fp64()
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / args.steps
    it64 = len(h64) - 1
    print(f"[bench] fp64 {t_fp64 * 1e3:.1f} ms ({it64} it), mixed {t_mixed * 1e3:.1f} ms "
          f"({n_inner} inner, {len(hm) - 1} outer)", file=sys.stderr, flush=True)

    err_mixed = (torch.linalg.norm(xm - x64) / torch.linalg.norm(x64)).item()
    rr = G.b.clone()
    G.A.MatVec(rr, -1.0, xm, 1.0)
    relres_mixed = (torch.linalg.norm(rr) / torch.linalg.norm(G.b)).item()

    # ---- per-kernel times of the fused iterations: HIP events on the solver
    # stream, right after the timed solves (same buffers, same kernels) ----
    n0, nu0 = level0_rows(G.psup_ind, G.psup, nb)
    kern = kernel_table(G, S, nb, nnzb, n0, nu0, step_mixed, step_fp64, args)
    # dominant kernel of the headline (FP32 + refinement) solve: largest time per iteration
    dom = max(("spmv", "fwd_sweep", "bwd_sweep", "update", "pupdate"),
              key=lambda k: kern["fp32"][k]["ms_per_call"])
    dk = kern["fp32"][dom]
    launch_ms = dk["ms_per_call"] / dk["launches_per_call"]
    launch_bytes = dk["bytes_per_call"] / dk["launches_per_call"]
    in_sequence_ms = launch_ms
    if dom == "spmv":  # its launches back to back: no event packets between the kernels
        launch_ms = kern["fp32"]["spmv_back_to_back_ms"]
    dom_gbs = launch_bytes / (launch_ms * 1e-3) / 1e9
    vt = "float, float" if args.mixed_vectors == 32 else "double, float"
    kname = {"spmv": f"k_pcg_spmv<{vt}>", "fwd_sweep": f"k_pcg_fwd<{vt}>",
             "bwd_sweep": f"k_pcg_bwd<{vt}>", "update": f"k_pcg_update<{vt}>",
             "pupdate": f"k_pcg_pupdate<{vt}>"}[dom]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if kname in t:
            traffic = t[kname]

    shapes = {}
    for prec, fp32 in (("fp32", True), ("fp64", False)):
        v, name, tms = G.ilu.sweep_shape(fp32)
        shapes[prec] = {"shape": name, "tune_ms_per_sweep_pair": [round(t, 5) for t in tms]}

    copy_gbs = stream_copy_gbs()
    b_iter64 = pcg_iter_bytes(nb, nnzb, 8, n0, nu0)
    t_it64 = t_fp64 / max(it64, 1)

    # one-off packing of A's sliced FP32 / FP64 copies (redone only when A's
    # values change, so outside the timed steps): a step right after A is
    # marked changed, minus a plain step
    t_pack = []
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        step_mixed()
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t
        G.A.values_changed()
        torch.cuda.synchronize()
        t = time.perf_counter()
        step_mixed()
        torch.cuda.synchronize()
        t_pack.append(time.perf_counter() - t - t_plain)
    pack_ms = sorted(t_pack)[1] * 1e3

    cpu = None
    if args.cpu_iters > 0:
        print("[bench] CPU reference (natural and workload order, full solves)",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.n, args.sigma, args.ordering)

    out = {
        "metric": "time-to-1e-8 PCG at 823k DOF: FP32+refine vs FP64 (s, DOF·iter/s, %HBM)",
        "value": round(t_mixed, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refine" if args.mixed_vectors == 32 else
                 "f32-matrix+factors/f64-vectors+refine",
        "data": "synthetic (seeded voxel tet cube, lognormal E)",
        "config": {"workload": f"tet elasticity cube {args.n}^3 hex x6 tets, {prob.ndof} DOF, "
                               f"E~lognormal(0,{args.sigma}), nu=0.3, ILU(0)-PCG to 1e-8, "
                               f"{args.ordering} order",
                   "ndof": prob.ndof, "nnzb": nnzb, "ordering": args.ordering,
                   "delta": args.delta},
        "fp64": {"time_s": round(t_fp64, 5), "iterations": it64,
                 "dof_iter_per_s": round(prob.ndof * it64 / t_fp64, 1),
                 "hbm_frac_per_iter": round(b_iter64 / t_it64 / 1e9 / HBM_PEAK_GBS, 4)},
        "mixed": {"time_s": round(t_mixed, 5), "inner_iterations": n_inner,
                  "outer_iterations": len(hm) - 1,
                  "dof_iter_per_s": round(prob.ndof * n_inner / t_mixed, 1),
                  "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
                  "x_rel_err_vs_fp64": err_mixed, "rel_residual": relres_mixed},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": int(launch_bytes), "avg_launch_ms": round(launch_ms, 5),
                     "in_sequence_ms": round(in_sequence_ms, 5),
                     "stream_copy_gbs": round(copy_gbs, 1),
                     "frac_of_stream_copy": round(dom_gbs / copy_gbs, 4)},
        "factor_ms": round(t_factor * 1e3, 3),
        "pcg_s": {"mixed": round(t_mixed - t_factor, 5), "fp64": round(t_fp64 - t_factor, 5)},
        "pack_ms": round(pack_ms, 3),  # A's sliced copies, packed once, outside the steps
        "kernels": kern,
        "sweep_shape": shapes,
        "cpu_baseline": cpu,
        "setup_s": round(G.t_setup, 3),
        "assembly_ms": round(t_asm * 1e3, 3),  # warm median; the first call also builds lists
        "assembly_first_ms": round(t_asm_first * 1e3, 3),
    }
    if args.others != "none":
        out["other_workloads"] = other_workloads(args)
    print(json.dumps(out), flush=True)


# the BASELINE-style workloads besides the headline, run after it by default
OTHER_WORKLOADS = ("cube128", "batch512", "newmark64", "stiffwarp64", "adjoint96", "order64")
# keys of a child's JSON line left out of the folded copy (per-kernel tables
# and autotuner timings; the child prints them when run on its own)
OTHER_DROP = ("kernels", "autotune", "sweep_shape", "n_gpus", "scaling", "vs_baseline",
              "higher_is_better", "steps", "warmup")


def other_workloads(args) -> dict:
    """Runs each workload of --others as `bench.py --workload W --steps 1
    --warmup 1` in a child process (one at a time, after the headline's
    timed steps) and returns its JSON line without the bulky tables; a child
    that fails or times out is recorded with its error."""
    res = {}
    for w in [x for x in args.others.split(",") if x]:
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", w, "--steps", "1",
               "--warmup", "1", "--cpu-iters", str(args.cpu_iters), "--others", "none"]
        print(f"[bench] workload {w}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=600)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if r.returncode != 0 or not lines:
                res[w] = {"error": f"exit status {r.returncode}"}
                continue
            d = json.loads(lines[-1])
            res[w] = {k: v for k, v in d.items() if k not in OTHER_DROP}
            res[w]["wall_s"] = round(time.perf_counter() - t0, 1)
        except subprocess.TimeoutExpired:
            res[w] = {"error": "timed out after 600 s"}
    return res


if __name__ == "__main__":
    main()
