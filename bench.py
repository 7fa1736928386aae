"""Benchmark: angular unknowns swept per second on one MI355X.

Workload (one "step" = one full sweep psi = L^{-1} q through the C ABI, with q
and psi resident in HBM): unit cube, 64^3 hexes, DG p=1 Gauss-Lobatto (8 dofs),
level-symmetric S8 (80 directions), per-element lognormal sigma_t (E = 1,
log-sd 1, seed 0), c ~ U[0.5, 0.99], Q = 1 per steradian, vacuum boundaries,
FP64.  Also reported: the plain GPU source iteration to a 1e-8 FP64 scalar-flux
residual on the c = 0.9999, log-sd 1.5 medium, and the CPU baseline (the
oracle's level-synchronous OpenMP sweep on the host cores).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UNKNOWN_3D_P1 = 28.0   # 8 (q) + 8 (psi) + 8*d*nf/n (upwind traces) at the Mult boundary


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, seed, reps):
    """The oracle's FP64 sweep of the same workload on the host cores."""
    import oracle
    from allium_amd import problems as P
    import allium_amd as A

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cfg = P.problem_3d_hetero(n=n, seed=seed)
    om, w = A.level_symmetric(8, 3)
    t0 = time.perf_counter()
    pr = oracle.Problem(3, cfg["nel"], cfg["vertices"], 1, om, w, cfg["sigma_t"])
    q = pr.form_source(Q=1.0)
    setup = time.perf_counter() - t0
    psi = np.empty(pr.size)
    pr.sweep(q, psi, threads=threads)          # warm-up
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        pr.sweep(q, psi, threads=threads)
        times.append(time.perf_counter() - t)
    tm = float(np.median(times))
    host = host_cpu()
    return {"value": pr.size / tm, "unit": "angular unknowns/s", "cores": threads, "kind": "port",
            "threads_used": threads, "nproc": host["nproc"], "physical_cores": host["physical_cores"],
            "sample": f"oracle FP64 level-synchronous OpenMP sweep of the same {n}^3 S8 p=1 workload on {threads} "
                      f"OpenMP threads, median of {reps} sweeps after 1 warm-up (setup {setup:.1f}s excluded); "
                      f"host {host['model']}, nproc {host['nproc']}, {host['physical_cores']} physical cores "
                      f"({host['sockets']} sockets x {host['cores_per_socket']} cores, "
                      f"{host['threads_per_core']} threads per core)",
            "seconds_per_sweep": tm}


def host_cpu():
    """lscpu's model and core counts, and nproc (the CPUs this process may use)."""
    info = {"model": "unknown", "sockets": None, "cores_per_socket": None, "threads_per_core": None,
            "physical_cores": None, "nproc": None}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        kv = {l.split(":", 1)[0].strip(): l.split(":", 1)[1].strip() for l in txt.splitlines() if ":" in l}
        info["model"] = kv.get("Model name", "unknown")
        for key, name in (("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
                          ("Thread(s) per core", "threads_per_core")):
            if kv.get(key, "").isdigit():
                info[name] = int(kv[key])
        if info["sockets"] and info["cores_per_socket"]:
            info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    except Exception:
        pass
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except Exception:
        info["nproc"] = len(os.sched_getaffinity(0))
    return info


def measure_traffic(args, tuning):
    """HBM bytes per sweep from two rocprofv3 --pmc child runs (FETCH_SIZE,
    then WRITE_SIZE), each scaled by what the counter reads for a device copy
    of known size recorded in the same run.  Returns (bytes, detail) or
    (None, reason)."""
    import csv
    import shutil
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    copy_mib = 256
    tun = f"{tuning['mode']},{tuning['block']},{tuning['lanes']}"
    box = []
    if "box" in tuning:
        b = tuning["box"]
        box = ["--box", f"{b['bx']},{b['by']},{b['bz']},{b['dirs_per_group']},{b['threads']},{b['prefetch']}"]
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="allium_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run",
               "--", sys.executable, os.path.abspath(__file__), "--n", str(args.n), "--seed", str(args.seed),
               "--steps", "2", "--warmup", "0", "--no-cpu", "--no-si", "--no-trt", "--no-config3", "--no-config1", "--no-traffic",
               "--no-fp32-check", "--tuning", tun,
               "--copy-mib", str(copy_mib), "--precision", args.precision, *box]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=os.environ.get("TMPDIR", "/tmp"))
        path = os.path.join(d, "run_counter_collection.csv")
        if r.returncode != 0 or not os.path.exists(path):
            return None, f"rocprofv3 {counter} run failed ({r.returncode})"
        sweep_kib, copy_kib = 0.0, 0.0
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if row["Counter_Name"] != counter:
                    continue
                if "sweep" in row["Kernel_Name"]:
                    sweep_kib += float(row["Counter_Value"])
                elif "copyBuffer" in row["Kernel_Name"]:
                    copy_kib = max(copy_kib, float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if copy_kib <= 0:
            return None, "no calibration copy in the counter run"
        factor = copy_mib * 1024.0 / copy_kib
        out[counter] = (sweep_kib * 1024.0 * factor / 2.0, factor)   # 2 timed sweeps
    total = out["FETCH_SIZE"][0] + out["WRITE_SIZE"][0]
    detail = {"fetch_bytes": out["FETCH_SIZE"][0], "write_bytes": out["WRITE_SIZE"][0],
              "fetch_scale": out["FETCH_SIZE"][1], "write_scale": out["WRITE_SIZE"][1]}
    return total, detail


def si_run(n, seed, maxit, learned=True, tuning=None, dsa_legs=True):
    """Wall time to a 1e-8 FP64 scalar-flux residual on the 64^3 p=1 S8 c=0.9999
    medium (lognormal sigma_t, log-sd 1.5, Q=1): plain GPU source iteration
    against source iteration with MIP DSA, GMRES(50) and FGMRES(50) on the same
    operator, the learned correction in source iteration (FP64 and FP32-refined
    sweeps, and falling back to DSA) and as FGMRES's preconditioner.  The
    source is HBM-resident before each timed solve.  The network is trained
    first on the GPU from seeded 32^3 media (not part of the timed solves; its
    time is reported), as is the DSA setup.  The network is trained on 200
    seeded 32^3 media: plain-iteration pairs, then two rounds with pairs from
    iterations that use it."""
    import allium_amd as A
    from allium_amd import cnn
    from allium_amd import problems as P

    cfg = P.problem_3d_hetero(n=n, seed=seed, log_sd=1.5, c_const=0.9999)
    m = A.Mesh(cfg["nel"])
    Linv = A.InverseAdvectionOperator(m, 1, 8, cfg["sigma_t"])
    if tuning is None:
        Linv.set_launch(-1)
    else:
        if "box" in tuning:
            b = tuning["box"]
            Linv.set_box_tiling(b["bx"], b["by"], b["bz"], dirs_per_group=b["dirs_per_group"], threads=b["threads"],
                                prefetch=b["prefetch"])
        Linv.set_tuning(tuning["mode"], tuning["block"], tuning["lanes"])
    q = A.DeviceVector.from_host(Linv.form_source(Q=1.0))

    def solve(krylov=None, **kw):
        A.synchronize()
        t = time.perf_counter()
        if krylov:
            phi, _, it, norms = A.krylov_solve(Linv, cfg["sigma_s"], q, solver=krylov, kdim=50, abstol=1e-8,
                                               maxit=300, want_psi=False, **kw)
            wall = time.perf_counter() - t
            res = {"solver": krylov, "kdim": 50, "iterations": it, "final_residual": float(norms[-1]),
                   "wall_s": wall, "converged": bool(norms[-1] <= 1e-8)}
            if "accel" in kw:
                res["learned_steps"] = A.krylov_solve.last_accel_steps
            return phi, res
        raw = np.zeros(maxit)
        phi, _, it, norms = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=maxit, want_psi=False,
                                               raw_norms=raw, **kw)
        wall = time.perf_counter() - t
        return phi, {"iterations": it, "final_residual": float(norms[-1]), "wall_s": wall,
                     "converged": bool(norms[-1] < 1e-8),
                     "final_unpreconditioned_residual": float(raw[it - 1])}

    phi0, plain = solve()
    out = {"plain": plain}

    def compare(name, **kw):
        phi1, res = solve(**kw)
        res["rel_diff_vs_plain"] = float(np.linalg.norm(phi1 - phi0) / np.linalg.norm(phi0))
        res["speedup_vs_plain"] = plain["wall_s"] / res["wall_s"]
        out[name] = res
        return res

    compare("gmres_fp64", krylov="gmres")
    compare("fgmres_fp64", krylov="fgmres")
    dsa = None
    if dsa_legs:
        # MIP diffusion synthetic acceleration (setup on the host, not timed in the solve)
        t = time.perf_counter()
        dsa = A.DSA(Linv, cfg["sigma_s"])
        dsa_setup = time.perf_counter() - t
        res = compare("dsa_fp64", accel=dsa)
        info = dsa.info()
        res.update(setup_s=dsa_setup, levels=info["level_size"],
                   cg_iterations_per_application=info["cg_iterations"] / max(1, info["applications"]),
                   dsa_gpu_s=info["seconds"])
    if not learned:
        return out
    compare("fp32_refined", precision=A.FP32_REFINED)
    net, info = cnn.fit(n=32, media=200, rounds=2, steps=1500, seed0=1000, sn=8, log=log)
    out["training"] = info
    for name, prec in (("learned_fp64", A.FP64), ("learned_fp32_refined", A.FP32_REFINED)):
        compare(name, accel=net, precision=prec)
    if dsa is not None:
        compare("learned_fp32_refined_dsa_fallback", accel=net, precision=A.FP32_REFINED, fallback=dsa)
    compare("fgmres_learned_fp64", krylov="fgmres", accel=net)
    return out


def config1_run(threads):
    """chive-style 2D steady one-group SI on the unit square, 32x32 quads, p=1,
    S8 (40 dirs), sigma_t=1, c=0.9, Q=1, vacuum, to a 1e-8 residual: the GPU
    (plain and with MIP DSA) and the CPU oracle (OpenMP) on the same input."""
    import oracle
    import allium_amd as A
    from allium_amd import problems as P

    cfg = P.problem_2d_cfg1()
    om, w = A.level_symmetric(8, 2)
    Linv = A.InverseAdvectionOperator(A.Mesh(cfg["nel"], vertices=cfg["vertices"]), 1, (om, w), cfg["sigma_t"])
    Linv.set_launch(-1)
    q = Linv.form_source(Q=cfg["Q"])
    out = {}
    for name, kw in (("gpu_plain", {}), ("gpu_dsa", {"accel": A.DSA(Linv, cfg["sigma_s"])})):
        A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=1000, want_psi=False, **kw)   # warm-up
        A.synchronize()
        t = time.perf_counter()
        phi, _, it, norms = A.source_iteration(Linv, cfg["sigma_s"], q, abstol=1e-8, maxit=1000, want_psi=False, **kw)
        out[name] = {"iterations": it, "wall_s": time.perf_counter() - t, "final_residual": float(norms[-1])}
        if name == "gpu_plain":
            phi_gpu = phi
    pr = oracle.Problem(2, cfg["nel"], cfg["vertices"], 1, om, w, cfg["sigma_t"])
    qo = pr.form_source(Q=cfg["Q"])
    t = time.perf_counter()
    phi_o, _, it_o, norms_o = pr.source_iteration(cfg["sigma_s"], qo, abstol=1e-8, maxit=1000, threads=threads)
    out["cpu_oracle"] = {"iterations": it_o, "wall_s": time.perf_counter() - t, "threads": threads,
                         "final_residual": float(norms_o[-1])}
    out["rel_diff_gpu_vs_cpu"] = float(np.linalg.norm(phi_gpu - phi_o) / np.linalg.norm(phi_o))
    out["unknowns"] = int(Linv.size)
    return out


def trt_run(steps=20):
    """Multigroup linearized backward-Euler TRT, dimensionless: unit cube 32^3
    hex, p=1, S6 (48 dirs), 16 log-spaced groups on [1e-2, 1e1],
    sigma_g = 10/(e_g^3 sqrt T), c_v = 1, T0 = 0.01, Planckian T = 1 inflow on
    x = 0, dt = 1e-3, 20 steps."""
    import allium_amd as A
    from allium_amd import trt

    g = trt.TRT((32, 32, 32), sn=6, groups=16, emin=1e-2, emax=1e1, opacity_coef=10.0, cv=1.0, dt=1e-3, T0=0.01,
                T_inflow=1.0, inflow_sides=1, si_reltol=1e-8)
    g.set_launch(-1)
    A.synchronize()
    t = time.perf_counter()
    its = []
    for _ in range(steps):
        info = g.step()
        its.append(info["transport_iterations"])
    wall = time.perf_counter() - t
    unknowns = 16 * 48 * 32 ** 3 * 8
    return {"steps": steps, "wall_s": wall, "ms_per_step": wall / steps * 1e3, "transport_iterations": its,
            "sweeps": info["sweeps"], "sweep_seconds": info["sweep_seconds"],
            "unknowns_per_sweep": unknowns, "sweep_unknowns_per_s": unknowns * info["sweeps"] / info["sweep_seconds"],
            "T_min": info["T_min"], "T_max": info["T_max"], "radiation_energy": info["radiation_energy"]}


def config3_run(seed=0, reps=3, instances=12):
    """Uneven batch in one launch: 12 independent 3D problems, hex meshes n^3
    with n drawn from {16..48} by seed, interior nodes jittered by +-0.05h
    (with S12, 0.2h and some 0.1h meshes give cyclic sweep graphs, where the
    reference aborts), DG p=2
    (27 dofs), S12 (168 dirs), lognormal sigma_t (log-sd 1), FP64 and FP32 LU.
    Throughput of batched sweeps with an isotropic unit source."""
    import allium_amd as A
    from allium_amd import problems as P

    rng = np.random.default_rng(seed)
    ns = [int(v) for v in rng.integers(16, 49, instances)]
    t0 = time.perf_counter()
    ops, isos, psis = [], [], []
    for i, n in enumerate(ns):
        nel = (n, n, n)
        r = np.random.default_rng(seed * 1000 + i)
        verts = P.jittered_vertices(nel, 0.05, r)
        sig = P.lognormal_sigma(n ** 3, 1.0, r).reshape(1, -1)
        op = A.InverseAdvectionOperator(A.Mesh(nel, vertices=verts), 2, 12, sig)
        ops.append(op)
        isos.append(A.DeviceVector.from_host(np.ones(op.phi_size)))
        psis.append(A.DeviceVector(op.size))
    setup = time.perf_counter() - t0
    batch = A.SweepBatch(ops)
    unknowns = sum(op.size for op in ops)
    out = {"meshes": ns, "unknowns_per_batch_sweep": unknowns, "setup_s": setup,
           "levels_max": max(op.nlevels for op in ops), "levels_sum": sum(op.nlevels for op in ops)}
    psis32 = [A.DeviceVector(op.size) for op in ops]
    for name, prec, dst in (("fp64", A.FP64, psis), ("fp32_lu", A.FP32_REFINED, psis32)):
        for op in ops:
            op.set_precision(prec)
        batch.Mult(None, dst, isos)
        t_before, n_before = batch.timing()
        for _ in range(reps):
            batch.Mult(None, dst, isos)
        t_after, n_after = batch.timing()
        dt = (t_after - t_before) / (n_after - n_before)
        out[name] = {"ms_per_batch_sweep": dt * 1e3, "unknowns_per_s": unknowns / dt}
    for op in ops:
        op.set_precision(A.FP64)
    # FP32 LU against the FP64 sweep, every instance
    worst = 0.0
    for d64, d32 in zip(psis, psis32):
        a, b = d64.to_host(), d32.to_host()
        if not (np.isfinite(a).all() and np.isfinite(b).all()):
            out["error"] = "non-finite angular flux"
        worst = max(worst, float(np.abs(b - a).max() / np.abs(a).max()))
        del a, b
    out["fp32_lu"]["rel_err_vs_fp64"] = worst
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--launch", type=int, default=-1,
                    help="-1 autotune, 0 levels, 1 graph, 3 direction sweeps, 4 box sweeps, 5 box sweeps in a graph")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-config1", action="store_true", help="skip the 2D chive-style SI (GPU and CPU oracle)")
    ap.add_argument("--no-si", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="no network training and no learned legs")
    ap.add_argument("--no-dsa", action="store_true", help="skip the DSA legs of the 64^3 solves")
    ap.add_argument("--no-trt", action="store_true", help="skip the 20-step multigroup TRT run")
    ap.add_argument("--no-config3", action="store_true", help="skip the batched p=2 jittered instances")
    ap.add_argument("--no-fp32-check", action="store_true", help="skip the FP32 element-solve sweep of the workload")
    ap.add_argument("--si-maxit", type=int, default=3000)
    ap.add_argument("--tuning", default=None, help="mode,block,lanes (skips the autotuner)")
    ap.add_argument("--box", default=None, help="bx,by,bz,dirs_per_group,threads,prefetch of box sweeps (modes 4/5)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic child runs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="device-to-device copy of this many MiB before timing (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding (angles/groups over "
                                   "RCCL) is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import allium_amd as A
    from allium_amd import problems as P

    log("device:", A.device_name())
    cfg = P.problem_3d_hetero(n=args.n, seed=args.seed)
    t0 = time.perf_counter()
    m = A.Mesh(cfg["nel"])
    Linv = A.InverseAdvectionOperator(m, 1, 8, cfg["sigma_t"],
                                      precision=A.FP32_REFINED if args.precision == "fp32" else A.FP64)
    log(f"setup {time.perf_counter() - t0:.1f}s: {Linv.size} unknowns, {Linv.nlevels} levels")
    if args.tuning:
        md, bl, ln = (int(v) for v in args.tuning.split(","))
        if args.box:
            bx, by, bz, nd, th, pf = (int(v) for v in args.box.split(","))
            Linv.set_box_tiling(bx, by, bz, dirs_per_group=nd, threads=th, prefetch=pf)
        Linv.set_tuning(md, bl, ln)
    else:
        Linv.set_launch(args.launch)
    tuning = Linv.tuning()
    if tuning["mode"] >= 4:
        tuning["box"] = Linv.box_tiling()
    log(f"tuning {tuning}")
    traffic, traffic_detail = (None, "skipped")
    if not args.no_traffic:
        traffic, traffic_detail = measure_traffic(args, tuning)
        log(f"traffic per sweep: {traffic} ({traffic_detail})")
    q = Linv.form_source(Q=1.0)
    dq = A.DeviceVector.from_host(q)
    dpsi = A.DeviceVector(q.size)
    del q
    if args.copy_mib:
        nd = args.copy_mib * (1 << 20) // 8
        ca, cb = A.DeviceVector(nd), A.DeviceVector(nd)
        cb.copy_from(ca)
        A.synchronize()
        del ca, cb
    for _ in range(args.warmup):
        Linv.Mult(dq, dpsi)
    A.synchronize()
    lib = A._lib.load()
    import ctypes
    ms = ctypes.c_double()
    A.check(lib.allium_event_timer(0, None))
    for _ in range(args.steps):
        Linv.Mult(dq, dpsi)
    A.check(lib.allium_event_timer(1, ctypes.byref(ms)))
    A.synchronize()
    t_sweep = ms.value / 1e3 / args.steps
    unknowns = Linv.size
    value = unknowns / t_sweep
    achieved = BYTES_PER_UNKNOWN_3D_P1 * unknowns / t_sweep / 1e9
    psi_h = dpsi.to_host()
    if not np.isfinite(psi_h).all():
        raise SystemExit("non-finite angular flux")
    log(f"sweep {t_sweep * 1e3:.3f} ms -> {value:.3e} unknowns/s, {achieved:.0f} GB/s algorithmic")
    fp32 = None
    if args.precision == "fp64" and not args.no_fp32_check:
        # the same sweep with FP32 element solves: its error against the FP64 sweep, and its time
        Linv.set_precision(A.FP32_REFINED)
        Linv.Mult(dq, dpsi)
        A.synchronize()
        A.check(lib.allium_event_timer(0, None))
        for _ in range(3):
            Linv.Mult(dq, dpsi)
        A.check(lib.allium_event_timer(1, ctypes.byref(ms)))
        psi32 = dpsi.to_host()
        fp32 = {"rel_err_vs_fp64": float(np.abs(psi32 - psi_h).max() / np.abs(psi_h).max()),
                "ms_per_sweep": ms.value / 3}
        del psi32
        Linv.set_precision(A.FP64)
        log(f"FP32 element solves: {fp32}")

    out = {
        "metric": "angular unknowns swept/s",
        "value": value,
        "unit": "angular unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_sweep * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-lu+f64-refine",
        "data": "synthetic (seeded lognormal sigma_t, uniform c)",
        "config": {"workload": f"3D heterogeneous one-group sweep: unit cube {args.n}^3 hex, DG p=1, S8 "
                               f"(80 dirs), lognormal sigma_t (E=1, log-sd 1), c~U[0.5,0.99], Q=1, vacuum",
                   "mesh": [args.n] * 3, "fe_order": 1, "sn": 8, "groups": 1, "seed": args.seed,
                   "tuning": tuning, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_unit": "bytes per sweep (FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "traffic_detail": traffic_detail,
                     "algorithmic_bytes_per_sweep": BYTES_PER_UNKNOWN_3D_P1 * Linv.size,
                     "bytes_per_unknown": BYTES_PER_UNKNOWN_3D_P1},
    }
    if fp32 is not None:
        out["fp32_element_solves"] = fp32
    if not args.no_si:
        try:
            out["source_iteration"] = si_run(args.n, args.seed, args.si_maxit, learned=not args.no_learned,
                                             tuning=tuning if args.tuning else None, dsa_legs=not args.no_dsa)
            log("SI:", out["source_iteration"])
        except Exception as exc:  # reported, not hidden
            out["source_iteration"] = {"error": str(exc)}
    if not args.no_config3:
        try:
            out["batch_p2_config3"] = config3_run(args.seed)
            log("config 3 batch:", out["batch_p2_config3"])
        except Exception as exc:  # reported, not hidden
            out["batch_p2_config3"] = {"error": str(exc)}
    if not args.no_trt:
        try:
            out["trt_config4"] = trt_run()
            log("TRT:", {k: v for k, v in out["trt_config4"].items() if k != "transport_iterations"})
        except Exception as exc:  # reported, not hidden
            out["trt_config4"] = {"error": str(exc)}
    if not args.no_config1 and not args.no_cpu:
        try:
            out["chive_config1"] = config1_run(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
            log("config 1:", out["chive_config1"])
        except Exception as exc:  # reported, not hidden
            out["chive_config1"] = {"error": str(exc)}
    if not args.no_cpu:
        cb = cpu_baseline(args.n, args.seed, args.cpu_reps)
        log(f"cpu baseline {cb['value']:.3e} unknowns/s on {cb['cores']} threads")
        # cores: the threads used; the host's physical cores and nproc each under their own name
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "threads_used",
                                                  "nproc", "physical_cores")}
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
