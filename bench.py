"""bench.py -- FMM matvecs/s and time to a 1e-6 GMRES residual, mixed vs FP64,
on one MI355X.

A step is one FMM-BEM matvec (Burton-Miller operator: near field + P2M, M2M,
M2L, L2L, L2P) over all right-hand sides of the workload; value = right-hand-
side matvecs per second of the mixed (FP32 pass) path with the inputs already
in HBM.  The FP64 path is timed the same way, then both solve to a 1e-6 true
(FP64) residual with GMRES(30).  A third solve reaches the same FP64 residual
of the same operator by iterative refinement around GMRES on an inner plan of
the same tree with a looser truncation (fmb_gmres_refined, --inner-eps).
Every solve is reported with and without the setup of the plans it needs.

Workloads (synthetic geometry generated in-process):
  sphere-l7-10k       : icosphere level 7 (327,680 triangles), r = 0.5 m,
                        10 kHz, depth 7, truncNum orders, plane wave + 8 extra
                        directions (the default)
  sphere-l6-5k        : icosphere level 6 (81,920 triangles), r = 0.3 m,
                        5 kHz, depth 6, truncNum orders, one seeded direction
  ellipsoid-60k-batch : perturbed ellipsoid (60,500 triangles), 16 frequencies
                        1-16 kHz (depth 5 below 4 kHz, 6 above), 64 plane waves
                        each, all frequencies in one graph-captured batch
  configs-table       : GPU FP64 / mixed matvecs/s and the CPU oracle (median
                        of 5 after one warm-up) on the 1,280-, 20,480- and
                        81,920-triangle spheres (not a metric line)

The CPU baseline is the FP64 oracle (C + OpenMP restatement of the reference
algorithm, the reference having no runnable CPU FMM) on the host cores:
complete matvecs of one right-hand side on the same mesh, tree and orders,
nothing sampled or extrapolated.
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
PEAK = {"fp64": 78.6, "fp32": 157.3}  # TFLOP/s, MI355X dense matrix

WORKLOADS = {
    "sphere-l7-10k": dict(kind="sphere", level=7, radius=0.5, freq=10000.0, depth=7, nrhs=9),
    "sphere-l6-5k": dict(kind="sphere", level=6, radius=0.3, freq=5000.0, depth=6, nrhs=1),
    "ellipsoid-60k-batch": dict(kind="batch", nu=55, axes=(0.08, 0.07, 0.10), seed=7,
                                nfreq=16, fmin=1000.0, fmax=16000.0, nrhs=64),
    "configs-table": dict(kind="table"),
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def timed_matvecs(torch, plan, x, y, prec, steps, warmup, levels):
    for _ in range(warmup):
        plan.matvec(x, y, prec=prec)
    torch.cuda.synchronize()
    plan.set_timing(True)
    m2l_ms = {l: 0.0 for l in levels}
    # matvec phases (HIP events on the plan's stream): gather + P2M, M2M,
    # M2L + L2L, L2P + near field + scatter
    names = {1: "gather_p2m", 2: "m2m", 3: "m2l_l2l", 5: "l2p_near"}
    phase_ms = {v: 0.0 for v in names.values()}
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.matvec(x, y, prec=prec)
        ph, m2l = plan.timing()
        for l in levels:
            m2l_ms[l] += m2l[l]
        for i, nm in names.items():
            phase_ms[nm] += ph[i]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.set_timing(False)
    return (dt, {l: v / steps for l, v in m2l_ms.items()},
            {nm: v / steps for nm, v in phase_ms.items()})


def oracle_matvec_seconds(V, F_, k, depth, orders, reps=1, warmup=0):
    """Seconds of complete FP64 oracle FMM matvecs (one right-hand side) on
    the host cores: same mesh, tree depth and per-level orders as the device
    plan.  Returns (list of seconds, threads, oracle setup seconds)."""
    import oracle as O
    ob = O.Bem(V, F_, k)
    t0 = time.perf_counter()
    ob.fmm_setup(depth=depth, plev={l: int(orders[l]) for l in range(2, depth + 1)})
    setup = time.perf_counter() - t0
    x = np.random.default_rng(3).standard_normal(F_.shape[0]) + 0j
    for _ in range(warmup):
        ob.fmm_matvec(x)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ob.fmm_matvec(x)
        ts.append(time.perf_counter() - t0)
    return ts, O.lib().fo_num_threads(), setup


def traffic_of(workload, lbig, shape):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary of this workload (FETCH_SIZE, WRITE_SIZE, each
    scaled by what it reads for an elementwise kernel of known size), or None
    when there is none for this level or it was taken on another tile shape
    (a stale record)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    lv = d.get("levels", {}).get(str(lbig))
    if not lv or lv.get("tile_shape") != list(shape):
        return None
    return lv["fetch_kb"] * 1024 * d["fetch_scale"] + lv["write_kb"] * 1024 * d["write_scale"]


def mfma_busy_of(workload, lbig, shape):
    """Matrix-core busy fraction of the dominant kernel from the same
    committed counter record (SQ_VALU_MFMA_BUSY_CYCLES over the launch's
    cycles on all SIMDs), or None without a record for this level and tile
    shape."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    lv = d.get("levels", {}).get(str(lbig))
    if not lv or lv.get("tile_shape") != list(shape):
        return None
    return lv.get("mfma_busy")


def plan_kwargs(a):
    return dict(autotune=0 if a.no_autotune else 1, serial_levels=1 if a.serial_levels else 0,
                split_bf16=1 if a.split_bf16 else 0, ldso=1 if a.ldso else 0,
                m2l_tile=a.m2l_tile)


def run_sphere(a, w, torch, F, log):
    k = 2 * np.pi * w["freq"] / C_AIR
    mesh = F.Mesh.icosphere(w["level"], w["radius"])
    V, Fc = mesh.arrays()
    tree = F.Tree(mesh, depth=w["depth"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=w["nrhs"], **plan_kwargs(a))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    # the same plan again: tile shapes now come from the autotuner's cache
    t0 = time.perf_counter()
    again = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=w["nrhs"], **plan_kwargs(a))
    torch.cuda.synchronize()
    setup_warm_s = time.perf_counter() - t0
    del again
    L = w["depth"]
    log("setup %.2fs (warm %.2fs, autotune %s) orders %s mem %.2f GB"
        % (setup_s, setup_warm_s, "off" if a.no_autotune else "on", plan.orders()[2:L + 1],
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
    y64 = plan.matvec(x, prec=F.FMB_FP64)
    y32 = plan.matvec(x, prec=F.FMB_MIXED)
    err_mv = float(torch.linalg.norm(y32 - y64) / torch.linalg.norm(y64))
    del y64, y32

    # the plain FP64 path: an FP64 plan, one right-hand side per pass (neither
    # of the method's precision and batching changes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p64 = F.Plan(tree, k, precision=F.FMB_FP64, nrhs_max=1, **plan_kwargs(a))
    torch.cuda.synchronize()
    plain_setup = time.perf_counter() - t0
    x1, y1 = x[:1].contiguous(), torch.empty_like(x[:1])
    dt1, _, _ = timed_matvecs(torch, p64, x1, y1, F.FMB_FP64, a.steps, a.warmup, flops)
    plain = {"matvecs_per_s": a.steps / dt1, "ms_per_matvec": 1e3 * dt1 / a.steps,
             "setup_s": plain_setup,
             "path": "FP64 plan, one right-hand side per pass, same tree and orders"}
    del x1, y1
    log("plain fp64 %.3f ms/matvec" % plain["ms_per_matvec"])

    # dominant kernel: the M2L launch of the level with the largest time
    lbig = max(m2l32, key=lambda l: m2l32[l])
    ach32 = flops[lbig] / (m2l32[lbig] * 1e-3) / 1e12
    ach64 = flops[lbig] / (m2l64[lbig] * 1e-3) / 1e12
    shape32 = plan.tile_shape_bf(0, lbig, F.FMB_MIXED)[:5]
    traffic = traffic_of(a.workload, lbig, shape32)
    mfma_busy = mfma_busy_of(a.workload, lbig, shape32)

    solve = {}
    field = None
    err_mie = err_mixed = err_ref = None
    if not a.no_solve:
        for name, prec in (("fp64", F.FMB_FP64), ("mixed", F.FMB_MIXED)):
            plan.set_work_precision(prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs, info = plan.gmres(b, restart=30, tol=1e-6, maxit=3000, check=False)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
            solve[name] = dict(seconds=sec, seconds_incl_setup=sec + setup_s,
                               seconds_incl_warm_setup=sec + setup_warm_s, iters=info.iters,
                               fp64_matvecs=info.matvecs_fp64, work_matvecs=info.matvecs_work,
                               relres=max(info.relres[r] for r in range(nrhs)), x=xs)
            log("solve %s %.2fs iters %d relres %.2e" % (name, sec, info.iters,
                                                        solve[name]["relres"]))
        # the plain FP64 path solves the right-hand sides one after another
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters, rel = [], 0.0
        for r in range(nrhs):
            xr, info = p64.gmres(b[r:r + 1].contiguous(), restart=30, tol=1e-6, maxit=3000,
                                 check=False)
            iters.append(info.iters)
            rel = max(rel, info.relres[0])
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        solve["plain_fp64"] = dict(seconds=sec, seconds_incl_setup=sec + plain_setup,
                                   iters_per_rhs=iters, relres=rel)
        log("solve plain fp64 %.2fs iters %s" % (sec, iters))
        u64, u32 = solve["fp64"]["x"].cpu().numpy(), solve["mixed"]["x"].cpu().numpy()
        err_mixed = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
        if a.inner_eps > 0:
            ikw = dict(plan_kwargs(a), split_bf16=0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            inner = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=nrhs, eps=a.inner_eps, **ikw)
            torch.cuda.synchronize()
            isetup = time.perf_counter() - t0
            del inner
            # the inner plan again, its tile shapes from the autotuner's cache
            t0 = time.perf_counter()
            inner = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=nrhs, eps=a.inner_eps, **ikw)
            torch.cuda.synchronize()
            isetup_warm = time.perf_counter() - t0
            inner.set_work_precision(F.FMB_MIXED)
            t0 = time.perf_counter()
            xs, info = plan.gmres_refined(inner, b, restart=30, tol=1e-6, inner_tol=a.inner_tol,
                                          maxit=3000, check=False)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
            solve["refined"] = dict(
                seconds=sec, seconds_incl_setup=sec + setup_s + isetup,
                seconds_incl_warm_setup=sec + setup_warm_s + isetup_warm,
                inner_eps=a.inner_eps, inner_tol=a.inner_tol,
                inner_orders=[int(q) for q in inner.orders()[2:L + 1]],
                inner_setup_s=isetup, inner_setup_warm_s=isetup_warm, cycles=info.restarts,
                inner_iters=info.iters, inner_matvecs=info.matvecs_work,
                fp64_matvecs=info.matvecs_fp64, relres=max(info.relres[r] for r in range(nrhs)))
            log("solve refined %.2fs (inner setup %.2fs, warm %.2fs) cycles %d relres %.2e" % (
                sec, isetup, isetup_warm, info.restarts, solve["refined"]["relres"]))
            err_ref = float(np.linalg.norm(xs.cpu().numpy() - u64) / np.linalg.norm(u64))
            del inner
        import oracle as O
        c = V[Fc].mean(1)
        ct = (c @ dirs[0]) / np.linalg.norm(c, axis=1)
        mie = O.mie_rigid(k * w["radius"], ct, nterms=int(k * w["radius"]) + 40)
        err_mie = float(np.linalg.norm(u64[0] - mie) / np.linalg.norm(mie))

    if a.field_points > 0:
        # total pressure on spheres of 2 and 1.2 radii (the second within a
        # few leaf boxes of the surface), through the tree and by the direct
        # sum: from the first right-hand side's FP64 solution, each against
        # the exterior Mie series, or (no solves) from a seeded random
        # surface vector
        if solve:
            x0 = solve["fp64"]["x"][0:1].contiguous()
        else:
            rng = np.random.default_rng(11)
            x0 = torch.tensor(rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n)),
                              device="cuda")
        field = {"points": int(a.field_points),
                 "source": "FP64 solution" if solve else "random surface vector"}
        for rf in (2.0, 1.2):
            fpts = rf * w["radius"] * fibonacci_dirs(a.field_points)
            fres = {}
            for meth in ("direct", "fmm"):
                plan.field(x0, fpts[:64], dirs[0:1], method=meth)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                pf = plan.field(x0, fpts, dirs[0:1], method=meth)
                torch.cuda.synchronize()
                fres[meth] = (time.perf_counter() - t0, pf.cpu().numpy()[0])
            fr = {"radius_m": rf * w["radius"], "direct_s": fres["direct"][0],
                  "fmm_s": fres["fmm"][0],
                  "speedup_fmm_vs_direct": fres["direct"][0] / fres["fmm"][0],
                  "rel_diff_fmm_vs_direct": float(
                      np.linalg.norm(fres["fmm"][1] - fres["direct"][1]) /
                      np.linalg.norm(fres["direct"][1]))}
            if solve:
                import oracle as O
                mief = O.mie_rigid_field(k, w["radius"], fpts, dirs[0])
                for meth in ("direct", "fmm"):
                    fr["rel_err_%s_vs_mie" % meth] = float(
                        np.linalg.norm(fres[meth][1] - mief) / np.linalg.norm(mief))
            field["r%.1fa" % rf] = fr
            log("field %d points at %.2f m: direct %.3fs, through the tree %.3fs" % (
                a.field_points, fr["radius_m"], fr["direct_s"], fr["fmm_s"]))
    for sv in solve.values():
        sv.pop("x", None)
    del p64
    cpu = None
    if not a.no_cpu:
        ts, cores, csetup = oracle_matvec_seconds(V, Fc, k, L, plan.orders())
        cpu = {"value": 1.0 / ts[0], "unit": "matvecs/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "sample": ("one complete FP64 FMM matvec of one of the %d right-hand sides "
                          "(every right-hand side costs the same) on the same %s mesh, tree and "
                          "orders by the C/OpenMP oracle: %.1fs, not sampled or extrapolated "
                          "(oracle setup %.1fs excluded)" % (nrhs, a.workload, ts[0], csetup))}
        log("cpu oracle matvec %.1fs on %d threads" % (ts[0], cores))

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
                   "radius_m": w["radius"], "depth": L, "nrhs": nrhs,
                   "orders": [int(q) for q in plan.orders()[2:L + 1]],
                   "parallelism": "1 GPU",
                   "translations": "split-bf16 rotations" if a.split_bf16 else "fp32"},
        "fp64_matvecs_per_s": mv64,
        "speedup_mixed_vs_fp64_matvec": mv32 / mv64,
        # the same-batch ratio above isolates precision; against the plain
        # FP64 path (one right-hand side per pass) batching counts too
        "plain_fp64": plain,
        "speedup_mixed_batched_vs_plain_fp64_matvec": mv32 / plain["matvecs_per_s"],
        "matvec_rel_err_mixed_vs_fp64": err_mv,
        "setup_s": setup_s,
        "setup_warm_s": setup_warm_s,
        "roofline": {"bound": "mfma", "achieved": ach32, "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": ach32 / PEAK["fp32"],
                     "traffic": traffic,
                     # HBM rate of that launch: its counted bytes over its time
                     "hbm_GBps": (traffic / (m2l32[lbig] * 1e-3) / 1e9) if traffic else None,
                     # matrix-core busy fraction of that launch (rocprofv3 SQ
                     # counters, committed beside the traffic record)
                     "mfma_busy": mfma_busy,
                     "kernel": "M2L level %d, FP32 matrix-core kernel" % lbig,
                     "tile_shape": list(shape32),
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
            name: {str(l): list(plan.tile_shape_bf(0, l, prec)[:5]) for l in range(2, L + 1)}
            for name, prec in (("m2l_fp32", F.FMB_MIXED), ("m2l_fp64", F.FMB_FP64))},
    }
    if solve:
        s = out["solve"] = {kk: dict(v) for kk, v in solve.items()}
        f64 = solve["fp64"]
        s["speedup_mixed_vs_fp64"] = f64["seconds"] / solve["mixed"]["seconds"]
        s["surface_pressure_rel_err_mixed_vs_fp64"] = err_mixed
        if "refined" in solve:
            r = solve["refined"]
            # the refined solve needs the inner plan besides the plan both
            # other solves use: with its setup counted, cold (tile shapes
            # timed) and warm (from the autotuner's cache)
            s["speedup_refined_vs_fp64"] = f64["seconds"] / r["seconds"]
            s["speedup_refined_vs_fp64_incl_inner_setup"] = (
                f64["seconds"] / (r["seconds"] + r["inner_setup_s"]))
            s["speedup_refined_vs_fp64_incl_warm_inner_setup"] = (
                f64["seconds"] / (r["seconds"] + r["inner_setup_warm_s"]))
            s["surface_pressure_rel_err_refined_vs_fp64"] = err_ref
        # against the plain FP64 path (one right-hand side at a time)
        pl = solve["plain_fp64"]["seconds"]
        s["speedup_mixed_batched_vs_plain_fp64"] = pl / solve["mixed"]["seconds"]
        if "refined" in solve:
            s["speedup_refined_batched_vs_plain_fp64"] = pl / solve["refined"]["seconds"]
            s["speedup_refined_batched_vs_plain_fp64_incl_warm_inner_setup"] = (
                pl / (solve["refined"]["seconds"] + solve["refined"]["inner_setup_warm_s"]))
        s["surface_pressure_rel_err_fp64_vs_mie"] = err_mie
    if field:
        out["field_points"] = field
    return out


def run_batch(a, w, torch, F, log):
    """The batched-frequency workload: 16 plans (one per frequency) on one
    mesh, 64 plane waves each; a step is one graph-captured batch matvec of
    all plans (fmb_matvec_batch replaying one hipGraph)."""
    mesh = F.Mesh.ellipsoid(w["nu"], *w["axes"], 2, 8, 0.05, seed=w["seed"])
    V, Fc = mesh.arrays()
    n = Fc.shape[0]
    freqs = [w["fmin"] * (w["fmax"] / w["fmin"]) ** (i / (w["nfreq"] - 1.0))
             for i in range(w["nfreq"])]
    depth = [5 if f < 4000.0 else 6 for f in freqs]
    nrhs = w["nrhs"]
    trees = {d: F.Tree(mesh, depth=d) for d in sorted(set(depth))}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plans = [F.Plan(trees[d], 2 * np.pi * f / C_AIR, precision=F.FMB_MIXED, nrhs_max=nrhs,
                    **plan_kwargs(a)) for f, d in zip(freqs, depth)]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log("setup %.1fs, %.1f GB" % (setup_s, sum(p.memory() for p in plans) / 1e9))
    nr = [nrhs] * len(plans)
    tot = nrhs * len(plans)
    x = torch.randn(tot, n, dtype=torch.complex128, device="cuda")
    # per-plan M2L launch times (eager passes with events) -> dominant launch
    best = (0.0, None)
    flops_all = 0.0
    for i, (f, pl) in enumerate(zip(freqs, plans)):
        fl = m2l_flops(pl.tree, pl, nrhs)
        flops_all += sum(fl.values()) + m2m_l2l_flops(pl.tree, pl, nrhs)
        xs = x[i * nrhs:(i + 1) * nrhs].contiguous()
        pl.matvec(xs, prec=F.FMB_MIXED)
        pl.set_timing(True)
        pl.matvec(xs, prec=F.FMB_MIXED)
        _, m2l = pl.timing()
        pl.set_timing(False)
        for l in fl:
            if m2l[l] > best[0]:
                best = (m2l[l], (f, l, fl[l], pl.tile_shape_bf(0, l, F.FMB_MIXED)[:5]))
    rates = {}
    for name, prec in (("mixed", F.FMB_MIXED), ("fp64", F.FMB_FP64)):
        F.matvec_batch(plans, nr, x, graph=True, prec=prec, reps=a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F.matvec_batch(plans, nr, x, graph=True, prec=prec, reps=a.steps)
        torch.cuda.synchronize()
        rates[name] = time.perf_counter() - t0
        log("%s batch step %.1f ms" % (name, 1e3 * rates[name] / a.steps))
    mv32 = a.steps * tot / rates["mixed"]
    mv64 = a.steps * tot / rates["fp64"]
    solve = {}
    if not a.no_solve:
        dirs = fibonacci_dirs(nrhs, seed=5)
        for p in plans:
            p.set_work_precision(F.FMB_MIXED)
        b = torch.cat([p.rhs_plane_wave(dirs) for p in plans], 0).contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xb, infos = F.gmres_batch(plans, nr, b, graph=True, tol=1e-6, maxit=3000, check=False)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        log("batch solve %.1fs" % tb)
        t0 = time.perf_counter()
        iters = []
        worst = 0.0
        xseq = torch.empty_like(xb)
        for i, p in enumerate(plans):
            xs, info = p.gmres(b[i * nrhs:(i + 1) * nrhs].contiguous(), tol=1e-6, maxit=3000,
                               check=False)
            xseq[i * nrhs:(i + 1) * nrhs] = xs
            iters.append(info.iters)
            worst = max(worst, max(info.relres[r] for r in range(nrhs)))
        torch.cuda.synchronize()
        tseq = time.perf_counter() - t0
        agree = float(torch.linalg.norm(xb - xseq) / torch.linalg.norm(xseq))
        del xseq
        log("sequential solves %.1fs (batch vs sequential %.2e)" % (tseq, agree))
        solve = {"batch_graph": {"seconds": tb, "seconds_incl_setup": tb + setup_s,
                                 "relres_max": max(max(i.relres[r] for r in range(nrhs))
                                                   for i in infos)},
                 "sequential": {"seconds": tseq, "seconds_incl_setup": tseq + setup_s,
                                "relres_max": worst, "iters_per_frequency": iters},
                 "speedup_batch_vs_sequential": tseq / tb,
                 "solution_rel_diff_batch_vs_sequential": agree}
    cpu = None
    if not a.no_cpu:
        tsum, cores = 0.0, 1
        for f, d, pl in zip(freqs, depth, plans):
            ts, cores, _ = oracle_matvec_seconds(V, Fc, 2 * np.pi * f / C_AIR, d, pl.orders())
            tsum += ts[0]
        cpu = {"value": len(plans) / tsum, "unit": "matvecs/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "sample": ("one complete FP64 FMM matvec of one right-hand side at each of the "
                          "%d frequencies (same meshes, trees and orders) by the C/OpenMP "
                          "oracle: %.1fs in total, not extrapolated" % (len(plans), tsum))}
    tms, (fb, lb, flb, shb) = best
    ach = flb / (tms * 1e-3) / 1e12
    return {
        "metric": "FMM matvecs/s + time-to-1e-6 GMRES solve, mixed vs FP64, 1 GPU",
        "value": mv32, "unit": "matvecs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * rates["mixed"] / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (mixed: FP32 translations/near field, FP64 recurrences and Arnoldi)",
        "data": "synthetic perturbed ellipsoid and plane waves generated in-process",
        "config": {"workload": a.workload, "triangles": int(n), "frequencies_hz": freqs,
                   "depths": depth, "nrhs_per_frequency": nrhs, "parallelism": "1 GPU",
                   "orders": [[int(q) for q in p.orders()[2:d + 1]] for p, d in zip(plans, depth)]},
        "fp64_matvecs_per_s": mv64, "speedup_mixed_vs_fp64_matvec": mv32 / mv64,
        "setup_s": setup_s,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": ach / PEAK["fp32"], "traffic": None,
                     "kernel": "M2L level %d at %.0f Hz, FP32 matrix-core kernel (eager pass)"
                               % (lb, fb), "launch_ms": tms, "tile_shape": list(shb)},
        "matvec_rcr_flops": flops_all,
        "matvec_achieved_tflops": {"fp32": flops_all / rates["mixed"] * a.steps / 1e12,
                                   "fp64": flops_all / rates["fp64"] * a.steps / 1e12},
        "solve": solve,
        "cpu_baseline": cpu,
    }


def run_table(a, torch, F, log):
    """GPU FP64 and mixed matvecs/s beside the CPU oracle (median of 5 after
    one warm-up) on the 1,280-, 20,480- and 81,920-triangle spheres."""
    rows = []
    for name, lev, r, f, depth, s in (("sphere-l3-1k (1,280 triangles)", 3, 0.1, 1000.0, 3, 5),
                                      ("sphere-l5-4k (20,480 triangles)", 5, 0.1, 4000.0, 0, 5),
                                      ("sphere-l6-5k (81,920 triangles)", 6, 0.3, 5000.0, 6, 5)):
        k = 2 * np.pi * f / C_AIR
        mesh = F.Mesh.icosphere(lev, r)
        V, Fc = mesh.arrays()
        tree = F.Tree(mesh, s=s, depth=depth)
        L = tree.info()[1]
        plan = F.Plan(tree, k, precision=F.FMB_MIXED, nrhs_max=1, **plan_kwargs(a))
        x = torch.randn(1, Fc.shape[0], dtype=torch.complex128, device="cuda")
        y = torch.empty_like(x)
        gpu = {}
        for pn, prec in (("fp64", F.FMB_FP64), ("mixed", F.FMB_MIXED)):
            dt, _, _ = timed_matvecs(torch, plan, x, y, prec, 20, 3, [])
            gpu[pn] = 20 / dt
        ts, cores, _ = oracle_matvec_seconds(V, Fc, k, L, plan.orders(), reps=5, warmup=1)
        rows.append({"workload": name, "depth": L,
                     "orders": [int(q) for q in plan.orders()[2:L + 1]],
                     "gpu_fp64_matvecs_per_s": gpu["fp64"], "gpu_mixed_matvecs_per_s": gpu["mixed"],
                     "cpu_oracle_matvecs_per_s": 1.0 / float(np.median(ts)),
                     "cpu_oracle_seconds": ts, "cpu_threads": cores})
        log(json.dumps(rows[-1]))
        del plan
    return {"table": "GPU vs CPU oracle per configuration (one right-hand side)",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sphere-l7-10k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--field-points", type=int, default=16384,
                    help="field points off the sphere after the solves (0: none)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fixed tile-shape rule instead of timing the shapes at plan creation")
    ap.add_argument("--split-bf16", action="store_true",
                    help="split-BF16 rotations in the FP32 pass's LDS-mode translations")
    ap.add_argument("--inner-eps", type=float, default=0.5,
                    help="truncation eps of the refined solve's inner plan (0: no refined solve)")
    ap.add_argument("--inner-tol", type=float, default=1e-3,
                    help="relative tolerance of each inner solve of the refined solve")
    ap.add_argument("--ldso", action="store_true",
                    help="fixed tile-shape rule with the operators in LDS where they fit")
    ap.add_argument("--m2l-tile", type=int, default=256,
                    help="M2L work in tile-major order over this many target boxes (0: "
                         "offset-major)")
    ap.add_argument("--serial-levels", action="store_true",
                    help="all translations on one stream (per-level times without overlap)")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file (default: a fresh file, so that setup_s times "
                         "the tuning and setup_warm_s the cached shapes)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the FMM-BEM path here is single-GPU "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    if a.tune_cache is None:
        import tempfile
        a.tune_cache = os.path.join(tempfile.mkdtemp(prefix="fmbem_bench_"), "rcr_tiles.txt")
    os.environ["FMBEM_TUNE_CACHE"] = a.tune_cache
    import torch
    import fmbem_gpu_c_amd as F

    w = WORKLOADS[a.workload]
    log = (lambda *s: print(*s, file=sys.stderr, flush=True)) if a.verbose else (lambda *s: None)
    if w["kind"] == "sphere":
        out = run_sphere(a, w, torch, F, log)
    elif w["kind"] == "batch":
        out = run_batch(a, w, torch, F, log)
    else:
        out = run_table(a, torch, F, log)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
