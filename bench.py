"""bench.py -- FMM matvecs/s and time to a 1e-6 GMRES residual, mixed vs FP64,
on one MI355X.

A step is one FMM-BEM matvec (Burton-Miller operator: near field + P2M, M2M,
M2L, L2L, L2P) over all right-hand sides of the workload; value = right-hand-
side matvecs per second of the mixed (FP32 pass) path with the inputs already
in HBM.  The FP64 path is timed the same way, then both solve to a 1e-6 true
(FP64) residual with GMRES(30).  A third solve reaches the same FP64 residual
of the same operator by iterative refinement around GMRES on an inner plan of
the same tree with a looser truncation (fmb_gmres_refined, --inner-eps).

Workloads (synthetic geometry generated in-process):
  sphere-l7-10k : icosphere level 7 (327,680 triangles), r = 0.5 m, 10 kHz,
                  depth 7, truncNum orders, plane wave + 8 extra directions
  sphere-l6-5k  : icosphere level 6 (81,920 triangles), r = 0.3 m, 5 kHz,
                  depth 6, truncNum orders, one seeded plane-wave direction
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C_AIR = 343.0
PEAK = {"fp64": 78.6, "fp32": 157.3}  # TFLOP/s, MI355X vector / matrix dense

WORKLOADS = {
    "sphere-l7-10k": dict(level=7, radius=0.5, freq=10000.0, depth=7, nrhs=9, cpu_stride=8),
    "sphere-l6-5k": dict(level=6, radius=0.3, freq=5000.0, depth=6, nrhs=1, cpu_stride=1),
}


def fibonacci_dirs(n, seed=0):
    if n == 1:
        v = np.random.default_rng(seed).standard_normal(3)
        return (v / np.linalg.norm(v))[None, :]
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    d[0] = [0.0, 0.0, 1.0]
    return d


def m2l_flops(tree, plan, nrhs):
    """RCR algorithmic flops per level: P_l * 8 (10 p^3 - p) / 3 per RHS."""
    p = plan.orders()
    _, L, _, _ = tree.info()
    out = {}
    for l in range(2, L + 1):
        P = len(tree.pairs(l)[0])
        out[l] = P * 8.0 * (10.0 * p[l] ** 3 - p[l]) / 3.0 * nrhs
    return out


def m2m_l2l_flops(tree, plan, nrhs):
    """RCR algorithmic flops of the upward and downward passes per matvec:
    sum over l of 2 B_l * 8 (10 p_{l-1}^3 - p_{l-1}) / 3 per RHS (one M2M and
    one L2L per box, counted at the parent's order)."""
    p = plan.orders()
    _, L, _, _ = tree.info()
    tot = 0.0
    for l in range(3, L + 1):
        B = len(tree.level(l))
        tot += 2 * B * 8.0 * (10.0 * p[l - 1] ** 3 - p[l - 1]) / 3.0 * nrhs
    return tot


def timed_matvecs(torch, plan, x, y, prec, steps, warmup, flops):
    import fmbem_gpu_c_amd as F
    for _ in range(warmup):
        plan.matvec(x, y, prec=prec)
    torch.cuda.synchronize()
    plan.set_timing(True)
    m2l_ms = {l: 0.0 for l in flops}
    # matvec phases (HIP events on the plan's stream): gather + P2M, M2M,
    # M2L + L2L, L2P + near field + scatter
    names = {1: "gather_p2m", 2: "m2m", 3: "m2l_l2l", 5: "l2p_near"}
    phase_ms = {v: 0.0 for v in names.values()}
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.matvec(x, y, prec=prec)
        ph, m2l = plan.timing()
        for l in flops:
            m2l_ms[l] += m2l[l]
        for i, nm in names.items():
            phase_ms[nm] += ph[i]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.set_timing(False)
    return (dt, {l: v / steps for l, v in m2l_ms.items()},
            {nm: v / steps for nm, v in phase_ms.items()})


def cpu_baseline(V, F_, k, depth, stride):
    """The FP64 oracle FMM (C + OpenMP) on the host cores for the same mesh,
    tree and orders: P2M and M2M in full, M2L / L2L / L2P / near field on the
    first 1/stride of the leaf boxes and their ancestors (levels too small to
    sample run whole), each level's time scaled to all its boxes.  Returns
    (matvecs/s for one right-hand side, threads, setup s, sample s, est s)."""
    import oracle as O
    ob = O.Bem(V, F_, k)
    t0 = time.perf_counter()
    ob.fmm_setup(s=5, depth=depth)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    est, _ = ob.fmm_time_sample(stride)
    sample = time.perf_counter() - t0
    return 1.0 / est, O.lib().fo_num_threads(), setup, sample, est


def traffic_of(workload, lbig):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 counter summary of this workload (FETCH_SIZE, WRITE_SIZE,
    each scaled by what it reads for an elementwise kernel of known size),
    or None when there is none for this level."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    lv = d.get("levels", {}).get(str(lbig))
    if not lv:
        return None
    return lv["fetch_kb"] * 1024 * d["fetch_scale"] + lv["write_kb"] * 1024 * d["write_scale"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sphere-l7-10k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fixed tile-shape rule instead of timing the shapes at plan creation")
    ap.add_argument("--split-bf16", action="store_true",
                    help="split-BF16 rotations in the FP32 pass's LDS-mode translations")
    ap.add_argument("--inner-eps", type=float, default=0.5,
                    help="truncation eps of the refined solve's inner plan (0: no refined solve)")
    ap.add_argument("--inner-tol", type=float, default=1e-3,
                    help="relative tolerance of each inner solve of the refined solve")
    ap.add_argument("--serial-levels", action="store_true",
                    help="all translations on one stream (per-level times without overlap)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the FMM-BEM path here is single-GPU "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    import fmbem_gpu_c_amd as F

    w = WORKLOADS[a.workload]
    log = (lambda *s: print(*s, file=sys.stderr, flush=True)) if a.verbose else (lambda *s: None)
    k = 2 * np.pi * w["freq"] / C_AIR
    t0 = time.perf_counter()
    mesh = F.Mesh.icosphere(w["level"], w["radius"])
    V, Fc = mesh.arrays()
    tree = F.Tree(mesh, depth=w["depth"])
    plan = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=w["nrhs"],
                  autotune=0 if a.no_autotune else 1,
                  serial_levels=1 if a.serial_levels else 0,
                  split_bf16=1 if a.split_bf16 else 0)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log("setup %.2fs (autotune %s) orders %s mem %.2f GB"
        % (setup_s, "off" if a.no_autotune else "on", plan.orders()[2:w["depth"] + 1],
           plan.memory() / 1e9))
    n = Fc.shape[0]
    nrhs = w["nrhs"]
    # one elementwise kernel of known traffic (reads 256 MiB, writes 256 MiB):
    # hardware byte counters of the FMM kernels are scaled by what they read
    # for this launch in the same run
    ca = torch.zeros(64 << 20, dtype=torch.float32, device="cuda")
    torch.add(ca, 1.0, out=torch.empty_like(ca))
    del ca
    dirs = fibonacci_dirs(nrhs, seed=5)
    b = plan.rhs_plane_wave(dirs)
    x = torch.randn(nrhs, n, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    flops = m2l_flops(tree, plan, nrhs)
    flops_mv = sum(flops.values()) + m2m_l2l_flops(tree, plan, nrhs)

    dt32, m2l32, ph32 = timed_matvecs(torch, plan, x, y, F.FMB_MIXED, a.steps, a.warmup, flops)
    dt64, m2l64, ph64 = timed_matvecs(torch, plan, x, y, F.FMB_FP64, a.steps, a.warmup, flops)
    mv32 = a.steps * nrhs / dt32
    mv64 = a.steps * nrhs / dt64
    log("mixed %.3f ms/step, fp64 %.3f ms/step" % (1e3 * dt32 / a.steps, 1e3 * dt64 / a.steps))
    # accuracy of one mixed matvec against the FP64 one on the same input
    y64 = plan.matvec(x, prec=F.FMB_FP64)
    y32 = plan.matvec(x, prec=F.FMB_MIXED)
    err_mv = float(torch.linalg.norm(y32 - y64) / torch.linalg.norm(y64))
    del y64, y32

    # dominant kernel: the M2L launch of the level with the largest time
    lbig = max(m2l32, key=lambda l: m2l32[l])
    ach32 = flops[lbig] / (m2l32[lbig] * 1e-3) / 1e12
    ach64 = flops[lbig] / (m2l64[lbig] * 1e-3) / 1e12
    traffic = traffic_of(a.workload, lbig)

    solve = {}
    if not a.no_solve:
        for name, prec in (("fp64", F.FMB_FP64), ("mixed", F.FMB_MIXED)):
            plan.set_work_precision(prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs, info = plan.gmres(b, restart=30, tol=1e-6, maxit=3000, check=False)
            torch.cuda.synchronize()
            solve[name] = dict(seconds=time.perf_counter() - t0, iters=info.iters,
                               fp64_matvecs=info.matvecs_fp64, work_matvecs=info.matvecs_work,
                               relres=max(info.relres[r] for r in range(nrhs)), x=xs)
            log("solve %s %.2fs iters %d relres %.2e" % (name, solve[name]["seconds"],
                                                        info.iters, solve[name]["relres"]))
        u64, u32 = solve["fp64"]["x"].cpu().numpy(), solve["mixed"]["x"].cpu().numpy()
        err_mixed = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
        err_ref = None
        if a.inner_eps > 0:
            t0 = time.perf_counter()
            inner = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=nrhs, eps=a.inner_eps,
                           autotune=0 if a.no_autotune else 1,
                           serial_levels=1 if a.serial_levels else 0)
            inner.set_work_precision(F.FMB_MIXED)
            torch.cuda.synchronize()
            isetup = time.perf_counter() - t0
            t0 = time.perf_counter()
            xs, info = plan.gmres_refined(inner, b, restart=30, tol=1e-6, inner_tol=a.inner_tol,
                                          maxit=3000, check=False)
            torch.cuda.synchronize()
            solve["refined"] = dict(
                seconds=time.perf_counter() - t0, inner_eps=a.inner_eps, inner_tol=a.inner_tol,
                inner_orders=[int(q) for q in inner.orders()[2:w["depth"] + 1]],
                inner_setup_s=isetup, cycles=info.restarts, inner_iters=info.iters,
                inner_matvecs=info.matvecs_work, fp64_matvecs=info.matvecs_fp64,
                relres=max(info.relres[r] for r in range(nrhs)))
            log("solve refined %.2fs cycles %d relres %.2e" % (
                solve["refined"]["seconds"], info.restarts, solve["refined"]["relres"]))
            err_ref = float(np.linalg.norm(xs.cpu().numpy() - u64) / np.linalg.norm(u64))
            del inner
        import oracle as O
        tri = V[Fc]
        c = tri.mean(1)
        d0 = dirs[0]
        ct = (c @ d0) / np.linalg.norm(c, axis=1)
        mie = O.mie_rigid(k * w["radius"], ct, nterms=int(k * w["radius"]) + 40)
        err_mie = float(np.linalg.norm(u64[0] - mie) / np.linalg.norm(mie))
        for s in solve.values():
            s.pop("x", None)

    cpu = None
    if not a.no_cpu:
        st = w["cpu_stride"]
        val, cores, csetup, csample, cest = cpu_baseline(V, Fc, k, w["depth"], st)
        what = ("one FP64 FMM matvec (one right-hand side) of the same %s mesh, tree and "
                "orders by the C/OpenMP oracle" % a.workload)
        if st > 1:
            what += (", timed on a sample: P2M/M2M whole, M2L/L2L/L2P/near on the first 1/%d "
                     "of the leaf boxes and their ancestors, each level scaled to all its "
                     "boxes (sample %.1fs, estimated full matvec %.1fs" % (st, csample, cest))
        else:
            what += " (matvec %.1fs" % cest
        what += "; oracle setup %.1fs excluded)" % csetup
        cpu = {"value": val, "unit": "matvecs/s", "cores": cores, "kind": "port",
               "sample": what}

    out = {
        "metric": "FMM matvecs/s + time-to-1e-6 GMRES solve, mixed vs FP64, 1 GPU",
        "value": mv32,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt32 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 (mixed: split-BF16 rotations with FP32 accumulation in the LDS-mode "
                  "translations, FP32 elsewhere; FP64 recurrences and Arnoldi)" if a.split_bf16 else
                  "fp32 (mixed: FP32 translations/near field, FP64 recurrences and Arnoldi)"),
        "data": "synthetic icosphere and plane waves generated in-process",
        "config": {"workload": a.workload, "triangles": int(n), "freq_hz": w["freq"],
                   "radius_m": w["radius"], "depth": w["depth"], "nrhs": nrhs,
                   "orders": [int(q) for q in plan.orders()[2:w["depth"] + 1]],
                   "parallelism": "1 GPU",
                   "translations": "split-bf16 rotations" if a.split_bf16 else "fp32"},
        "fp64_matvecs_per_s": mv64,
        "speedup_mixed_vs_fp64_matvec": mv32 / mv64,
        "matvec_rel_err_mixed_vs_fp64": err_mv,
        "setup_s": setup_s,
        "roofline": {"bound": "mfma", "achieved": ach32, "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": ach32 / PEAK["fp32"],
                     "traffic": traffic,
                     # HBM rate of that launch: its counted bytes over its time
                     "hbm_GBps": (traffic / (m2l32[lbig] * 1e-3) / 1e9) if traffic else None,
                     "kernel": "M2L level %d, FP32 matrix-core kernel" % lbig,
                     "fp64_achieved": ach64, "fp64_frac": ach64 / PEAK["fp64"],
                     "m2l_ms_per_level_fp32": {str(l): v for l, v in m2l32.items()},
                     "m2l_ms_per_level_fp64": {str(l): v for l, v in m2l64.items()}},
        "phases_ms": {"fp32": ph32, "fp64": ph64},
        # whole matvec: counted RCR flops (M2L, M2M, L2L) over the step time
        "matvec_rcr_flops": flops_mv,
        "matvec_achieved_tflops": {"fp32": flops_mv / dt32 * a.steps / 1e12,
                                   "fp64": flops_mv / dt64 * a.steps / 1e12},
        "cpu_baseline": cpu,
        "tile_shapes": {
            name: {str(l): list(plan.tile_shape_bf(0, l, prec)[:5]) for l in range(2, w["depth"] + 1)}
            for name, prec in (("m2l_fp32", F.FMB_MIXED), ("m2l_fp64", F.FMB_FP64))},
    }
    if solve:
        out["solve"] = {kk: {q: v for q, v in s.items()} for kk, s in solve.items()}
        out["solve"]["speedup_mixed_vs_fp64"] = solve["fp64"]["seconds"] / solve["mixed"]["seconds"]
        out["solve"]["surface_pressure_rel_err_mixed_vs_fp64"] = err_mixed
        if "refined" in solve:
            out["solve"]["speedup_refined_vs_fp64"] = (solve["fp64"]["seconds"] /
                                                       solve["refined"]["seconds"])
            out["solve"]["surface_pressure_rel_err_refined_vs_fp64"] = err_ref
        out["solve"]["surface_pressure_rel_err_fp64_vs_mie"] = err_mie
    print(json.dumps(out))


if __name__ == "__main__":
    main()
