"""Benchmark of the Heat3ds linear-solve path on one MI355X.

A step is one pass of the hot path: calRHS! + PBiCGSTAB! with the 5-sweep
red-black Gauss-Seidel preconditioner to ε = 1e-4, FP64, on the modelA
stack at 240 x 240 x 30 interior cells (the reference's README/benchmark
case: 208 iterations with the reference's sequential sums, 223 with the
exactly rounded sums of this implementation), with all fields resident in
HBM.  θ is reset to the initial state before every step (untimed).

Printed JSON line (one, on stdout):
  value      algorithmic HBM GB/s of the whole solve:
             560 B/cell/iteration (FP64, the reference's op sequence: ten
             sweeps, two SpMV, five dots, BiCG1!, two Triad!, BICG2! and
             two fills) x cells x iterations / step time; value_fused is
             the same on the 480 B basis of the fused minimum (no fills)
  parity     the same case with the reference's sequential grid sums
             (H3D_SUM_SEQUENTIAL): its iteration count, time, the README
             residual lines and whether θ's SHA-256 equals the committed
             fixture of the oracle's sequential run; max_abs_err_K is the
             largest |θ| difference (K, whole field) of the default solve
             (exactly rounded sums) from it
  roofline   dominant kernel: on grids the resident preconditioner serves
             (240 x 240 x 30) the one-launch PBiCGSTAB! loop,
             k_rb_bicg_solve<double> (every iteration of the solve in one
             launch; 544 B per interior cell and iteration in the
             reference's per-operation units), timed as one launch of 200
             iterations (HIP events on the context stream, h3d_time_kernel
             id 11), per iteration; else one red-black GS colour
             (k_rb_sweep, 16 B per cell, HIP events around 200 launches);
             traffic = HBM bytes per iteration (per launch) from the
             committed rocprofv3 FETCH_SIZE and WRITE_SIZE figures
             (profiles/traffic.json)
  cpu_baseline  the oracle (OpenMP C restatement of the reference
             algorithm) on a bounded sample of the same workload
  fp32_refined  the same case solved to a true relative residual of 1e-8
             with FP64 and with FP32-inner/FP64-refinement, and the speed-up;
             the same with CG! (FP32 CG inner solves)
  c3         480 x 480 x 60 with the seeded power map: the same solves
             (PBiCGSTAB! and CG!, FP64 and FP32-refined) to 1e-8 and their
             maximum temperature difference (K) from an FP64 solve to 1e-12
  cg         CG! + GS(5) on the same 240 x 240 x 30 case to eps = 1e-6
  c5_step    one implicit-Euler step (dt = 1 ms) of the 960 x 960 x 120
             transient case from 300 K, FP64 PBiCGSTAB! to 1e-8: iterations,
             time and the reference-equivalent HBM rate; the same step with
             FP32-refined CG!

Profiling workloads (no JSON line; run under rocprofv3):
  --profile sweep [--nx 240|480]   the modelA case, a 3-iteration solve,
             three plain copies of one H-cell FP64 field (the counters'
             calibration), then 20 full red-black sweeps (40 launches of
             k_rb_sweep<double>)
  --profile solve [--nx 240|480]   the headline solve, once, for kernel
             traces
  --profile resident   the modelA 240 x 240 x 30 case, a 3-iteration solve,
             the copy calibration, 20 launches of PBiCGSTAB!'s first
             resident launch (k_rb_resident<double, 2>, three-launch
             iteration), then the one-launch loop (k_rb_bicg_solve<double>)
             for 2 and for 200 iterations
  --profile kernels [--nx 240|480]  the modelA case, a 3-iteration solve,
             then 20 launches each of one full FP64 sweep, FP64 CalcAX!,
             FP64 CalcAX! with its dot, the FP32 sweep and FP32 CalcAX!
             (the h3d_time_kernel loops), for per-kernel counters
  --profile transient [--solver cg --precision fp32]   the 960 x 960 x 120
             power-map case, one implicit-Euler step's first 40 iterations,
             the copy calibration, then 10 full sweeps
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_ITER = {"fp64": 560.0, "fp32": 280.0}
FUSED_BYTES_PER_CELL_ITER = 480.0  # the same op sequence without the two fills
SWEEP_BYTES_PER_CELL = 32.0  # x, λ, rhs read + x written, FP64, one full red-black sweep
# the first resident launch of a PBiCGSTAB! iteration in the reference's
# per-operation units (FP64): BiCG1! 32 + five full sweeps 160 + CalcAX! 24
# + the (q, r0) dot 16 bytes per interior cell
RESIDENT_K1_BYTES_PER_CELL = 32.0 + 5 * SWEEP_BYTES_PER_CELL + 24.0 + 16.0
# the rest of the iteration: Triad! 24 + five sweeps 160 + CalcAX! 24 + the
# (t, t) and (t, s) dots 8 + 16 + BICG2! 32 + Triad! 24 + the (r, r) and
# (r, r0) dots 8 + 16; one whole iteration = 544 bytes per interior cell
RESIDENT_K2_BYTES_PER_CELL = 24.0 + 5 * SWEEP_BYTES_PER_CELL + 24.0 + 8.0 + 16.0 + 32.0 + 24.0 + 8.0 + 16.0
RESIDENT_ITER_BYTES_PER_CELL = RESIDENT_K1_BYTES_PER_CELL + RESIDENT_K2_BYTES_PER_CELL


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=240)
    ap.add_argument("--nz", type=int, default=30)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-refined", action="store_true", help="skip the FP32-refinement comparison")
    ap.add_argument("--no-c3", action="store_true", help="skip the 480x480x60 precision comparison")
    ap.add_argument("--no-c5", action="store_true", help="skip the 960x960x120 transient step")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG! solve")
    ap.add_argument("--no-parity", action="store_true", help="skip the reference-order parity solve")
    ap.add_argument("--cpu-iters", type=int, default=200)
    ap.add_argument("--profile", choices=("sweep", "solve", "transient", "kernels", "resident"), default=None,
                    help="run a profiling workload instead of the benchmark")
    ap.add_argument("--solver", choices=("pbicgstab", "cg"), default="pbicgstab")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """(model name, physical cores of the host, CPUs this process may run on)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    if phys is not None:
        cores.add((phys, core))
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, len(cores) or (os.cpu_count() or 1), avail


def cpu_baseline(nx, nz, iters):
    """Oracle PBiCGSTAB+GS on the same case for a bounded number of
    iterations, on the host cores this process may use, with threaded grid
    sums (the reference's par = "thread" path); returns GB/s in the same
    algorithmic unit."""
    from oracle import oracle as O
    O.build()
    exe = O.EXE
    model, physical, avail = host_cpu()
    # one thread per physical core this process may use; OMP_NUM_THREADS,
    # where the environment sets it, is the share of the host given to it
    cores = min(physical, avail)
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    cmd = [exe, str(nx), str(nx), str(nz), "pbicgstab", "gs", "1e-30", "1", str(cores)]
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores",
               ORC_SUM_MODE="2")
    # a bounded sample: the first iters iterations of the solve
    env["ORC_MAXITR"] = str(iters)
    t0 = time.perf_counter()
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, check=True).stdout
    wall = time.perf_counter() - t0
    solve_s = None
    itr = None
    for line in out.splitlines():
        if line.startswith("geometry_s="):
            solve_s = float(line.split()[1].split("=")[1])
        if line.startswith("converged="):
            itr = int(line.split()[1].split("=")[1])
    cells = nx * nx * nz
    gbs = BYTES_PER_CELL_ITER["fp64"] * cells * itr / solve_s / 1e9
    return {"value": round(gbs, 4), "unit": "GB/s", "cores": cores, "kind": "port",
            "cpu_model": model, "host_physical_cores": physical, "cpus_available": avail,
            "sample": "oracle (OpenMP C restatement of the reference algorithm, threaded sweeps and grid sums) "
                      "PBiCGSTAB+GS FP64 at %dx%dx%d, first %d iterations, %.1f s solve (%.1f s wall incl. "
                      "geometry), %d threads pinned one per core" % (nx, nx, nz, itr, solve_s, wall, cores),
            "s_per_iteration": round(solve_s / itr, 4)}


# kernels timed on the large grids: h3d_time_kernel id, name, algorithmic
# bytes per interior cell (steady, transient) of one launch set (SURVEY's
# unit: x, λ, rhs read and x written per full red-black sweep; x, λ read and
# ax written for CalcAX!, plus one word per fused dot; a transient operator
# adds ρ·cp), and the precision's word size
KERNELS = ((0, "sweep_fp64 (both colours)", 32.0, 40.0, 8), (1, "calc_ax_fp64", 24.0, 32.0, 8),
           (2, "calc_ax_dot_fp64 (with its finalize)", 32.0, 40.0, 8), (3, "sweep_fp32 (both colours)", 32.0, 40.0, 4),
           (4, "calc_ax_fp32", 24.0, 32.0, 4), (5, "sweep2_fp64 (two-colour launch)", 32.0, 40.0, 8),
           (6, "sweep2_fp32 (two-colour launch)", 32.0, 40.0, 4))


def kernel_fracs(wk, cells, steady, dt, reps, two_colour):
    """Mean launch time (HIP events on the context stream, best of three
    sets) and fraction of the 8 TB/s HBM peak in the algorithmic unit."""
    out = {}
    for which, name, b_st, b_tr, word in KERNELS:
        if which >= 5 and not two_colour:
            continue
        # the two-colour launches (the default for FP32 on deep grids; FP64
        # keeps per-colour launches) with 60-plane chunks
        wk.set_sweep_chunk(60 if which >= 5 else -1)
        ms = min(wk.time_kernel(which, reps, dt, steady) for _ in range(3))
        bpc = (b_st if steady else b_tr) * word / 8.0
        achieved = bpc * cells / (ms * 1e-3) / 1e9
        out[name] = {"us": round(ms * 1e3, 2), "alg_bytes_per_cell": bpc, "achieved_GBs": round(achieved, 1),
                     "frac": round(achieved / HBM_PEAK_GBS, 4)}
    wk.set_sweep_chunk(-1)
    return out


def precision_c3(np, h2_amd, L, WorkBuffers, load_model):
    """480x480x60, seeded power map: FP64 and FP32-refined solves to 1e-8
    against an FP64 solve to 1e-12."""
    log("building model 480x480x60 (power map)")
    m = h2_amd.build_model(480, 480, 60, power="map")
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    th0 = wk.download("theta")

    def run(prec, tol, solver=L.SOLVER_PBICGSTAB):
        wk.upload("theta", th0)
        wk.fill("b", 0.0)
        t0 = time.perf_counter()
        wk.calc_rhs(10000.0, True)
        conv, itr, _ = wk.solve(solver, 10000.0, tol, 20000, L.SMOOTHER_GS, True, precision=prec)
        return time.perf_counter() - t0, conv, itr, wk.download("theta")

    _, _, itr_ref, ref = run(L.FP64, 1e-12)
    out = {"tol": 1e-8, "reference": "FP64 PBiCGSTAB to 1e-12 (%d iterations)" % itr_ref}
    for name, prec, solver in (("fp64", L.FP64, L.SOLVER_PBICGSTAB),
                               ("fp32_refined", L.FP32_REFINED, L.SOLVER_PBICGSTAB),
                               ("cg_fp64", L.FP64, L.SOLVER_CG),
                               ("cg_fp32_refined", L.FP32_REFINED, L.SOLVER_CG)):
        t, conv, itr, th = run(prec, 1e-8, solver)
        out[name] = {"ms": round(t * 1e3, 2), "itr": itr, "converged": conv,
                     "max_abs_err_K": float("%.3e" % np.max(np.abs(th - ref)))}
        log("c3 %s: %.1f ms, itr %d, converged %s" % (name, t * 1e3, itr, conv))
    # FP32 storage of a vector like the correction: ||A(fl32(d) - d)|| / ||A d||
    # for d = θ* - θ0, the relative error a matvec of a rounded preconditioned
    # direction carries before any arithmetic
    d = ref - th0
    wk.upload("pcg_p", d)
    wk.calc_ax("pcg_q", "pcg_p")
    wk.upload("pcg_p_", d.astype(np.float32).astype(np.float64))
    wk.calc_ax("pcg_s", "pcg_p_")
    wk.triad("pcg_s_", "pcg_q", "pcg_s", -1.0)
    noise = float(np.sqrt(wk.fdot1("pcg_s_") / wk.fdot1("pcg_q")))
    f32 = out["fp32_refined"]
    out["configured_form"] = {
        "form": "FP32 PBiCGSTAB+GS(5) inner solves, inner tol 1e-4, FP64 refinement to 1e-8",
        "meets_config": bool(f32["converged"] and f32["max_abs_err_K"] <= 1e-5),
        "fp32_matvec_noise": float("%.4e" % noise),
        "substitute": "cg_fp32_refined"}
    out["kernels"] = kernel_fracs(wk, 480 * 480 * 60, True, 10000.0, 100, False)
    out["speedup_fp32_refined"] = round(out["fp64"]["ms"] / out["fp32_refined"]["ms"], 3)
    out["speedup_cg_fp32_refined_vs_cg_fp64"] = round(out["cg_fp64"]["ms"] / out["cg_fp32_refined"]["ms"], 3)
    out["speedup_cg_fp32_refined_vs_fp64"] = round(out["fp64"]["ms"] / out["cg_fp32_refined"]["ms"], 3)
    wk.close()
    return out


def transient_step_c5(h2_amd, L, WorkBuffers, load_model):
    """One implicit-Euler step of the 960x960x120 power-map case, FP64."""
    log("building model 960x960x120 (power map)")
    m = h2_amd.build_model(960, 960, 120, power="map")
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    dt = 1e-3
    th0 = wk.download("theta")

    def one(solver, prec):
        wk.upload("theta", th0)
        wk.fill("b", 0.0)
        t0 = time.perf_counter()
        wk.calc_rhs(dt, False)
        conv, itr, _ = wk.solve(solver, dt, 1e-8, 20000, L.SMOOTHER_GS, False, precision=prec)
        return time.perf_counter() - t0, conv, itr

    t, conv, itr = one(L.SOLVER_PBICGSTAB, L.FP64)
    cells = 960 * 960 * 120
    # a transient iteration reads ap0's class table, no extra array: 560 B/cell
    out = {"dt": dt, "tol": 1e-8, "itr": itr, "converged": conv, "s": round(t, 3),
           "ms_per_iteration": round(t / max(itr, 1) * 1e3, 3),
           "gbs": round(BYTES_PER_CELL_ITER["fp64"] * cells * itr / t / 1e9, 1)}
    log("c5 step: %.2f s, itr %d" % (t, itr))
    t32, conv32, itr32 = one(L.SOLVER_CG, L.FP32_REFINED)
    out["cg_fp32_refined"] = {"s": round(t32, 3), "inner_itr": itr32, "converged": conv32,
                              "speedup_vs_fp64_pbicgstab": round(t / t32, 3)}
    log("c5 step CG FP32-refined: %.2f s, itr %d" % (t32, itr32))
    out["kernels"] = kernel_fracs(wk, cells, False, dt, 20, True)
    wk.close()
    return out


def profile_workload(args):
    """The workloads the committed rocprofv3 summaries were taken on."""
    import h2_amd
    from h2_amd import _lib as L
    from h2_amd.heat3ds import WorkBuffers, load_model
    if args.profile in ("sweep", "solve", "kernels", "resident"):
        nx = args.nx
        nz = 30 * max(1, nx // 240)
        m = h2_amd.build_model(nx, nx, nz)
        wk = WorkBuffers(*m.shape)
        load_model(wk, m)
        if args.profile == "kernels":
            wk.calc_rhs(10000.0, True)
            wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-30, 3, L.SMOOTHER_GS, True)
            for which in (0, 1, 2, 3, 4):
                print("kernel %d ms %.4f" % (which, wk.time_kernel(which, 20, 10000.0, True)), flush=True)
        elif args.profile == "resident":
            wk.calc_rhs(10000.0, True)
            wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-30, 3, L.SMOOTHER_GS, True)
            for _ in range(3):
                wk.copy("pcg_p", "theta")  # calibration: a plain copy of H*8 bytes
            print("resident first launch ms", wk.time_kernel(9, 20, 10000.0, True), flush=True)
            print("one-launch loop ms per iteration", wk.time_kernel(11, 200, 10000.0, True), flush=True)
        elif args.profile == "sweep":
            wk.calc_rhs(10000.0, True)
            wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-30, 3, L.SMOOTHER_GS, True)
            for _ in range(3):
                wk.copy("pcg_p", "theta")  # calibration: a plain copy of H*8 bytes
            print("sweep ms", wk.time_kernel(0, 20, 10000.0, True), flush=True)
        else:
            wk.fill("b", 0.0)
            wk.calc_rhs(10000.0, True)
            conv, itr, _ = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, args.eps, 20000, L.SMOOTHER_GS, True)
            print("solve: itr %d, device ms %.3f" % (itr, wk.last_solve_ms()), flush=True)
        wk.close()
        return
    solver = L.SOLVER_CG if args.solver == "cg" else L.SOLVER_PBICGSTAB
    precision = L.FP32_REFINED if args.precision == "fp32" else L.FP64
    m = h2_amd.build_model(960, 960, 120, power="map")
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    wk.fill("b", 0.0)
    wk.calc_rhs(1e-3, False)
    conv, itr, res0 = wk.solve(solver, 1e-3, 1e-8, 40, L.SMOOTHER_GS, False, precision=precision)
    print("solver", args.solver, "precision", args.precision, "iterations", itr, "device ms", wk.last_solve_ms(),
          flush=True)
    for _ in range(3):
        wk.copy("pcg_p", "theta")  # calibration: a plain copy of H*8 bytes
    print("sweep ms", wk.time_kernel(0, 10, 1e-3, False), flush=True)
    wk.close()


def parity_block(np, L, wk, step, nx, nz, eps, th_exact):
    """The headline case with the reference's sequential grid sums, and the
    default solve's temperature difference from it."""
    import hashlib
    wk.set_sum_order(L.SUM_SEQUENTIAL)
    try:
        t, dev_ms, conv, itr, res0 = step()
        th_ref = wk.download("theta")
        hist = wk.history()
    finally:
        wk.set_sum_order(L.SUM_EXACT)
    out = {"sum_order": "sequential (k, j, i) running sums, as the reference's SequentialEx",
           "itr": itr, "converged": bool(conv), "ms": round(t * 1e3, 3), "device_ms": round(dev_ms, 3),
           "res0": repr(res0),
           "max_abs_err_K": float("%.4e" % np.max(np.abs(th_exact - th_ref)))}
    gold = os.path.join(ROOT, "tests", "golden", "c2_pbicgstab_seq.npz")
    pins = os.path.join(ROOT, "tests", "golden", "readme_pins.json")
    if (nx, nz, eps) == (240, 30, 1e-4) and os.path.exists(gold) and os.path.exists(pins):
        g = np.load(gold)
        sha = hashlib.sha256(np.asfortranarray(th_ref).tobytes(order="F")).hexdigest()
        out["theta_sha256_equals_oracle_sequential_run"] = sha == str(g["theta_sha256"])
        with open(pins) as f:
            p = json.load(f)
        out["readme_lines"] = {k: "%.14E" % hist[int(k) - 1] for k in p["history"] if int(k) <= len(hist)}
        out["readme_lines_match"] = all(out["readme_lines"].get(k) == v for k, v in p["history"].items())
        out["readme_res0_match"] = repr(res0) == p["res0"]
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X, and the "
              "multi-GPU (Y-slab) decomposition is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.profile:
        profile_workload(args)
        return
    import numpy as np
    import h2_amd
    from h2_amd import _lib as L
    from h2_amd.heat3ds import WorkBuffers, load_model

    nx, nz = args.nx, args.nz
    log("building model %dx%dx%d" % (nx, nx, nz))
    m = h2_amd.build_model(nx, nx, nz)
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    th0 = wk.download("theta")
    dt = 10000.0
    cells = nx * nx * nz

    def step(precision=L.FP64, tol=args.eps, solver=L.SOLVER_PBICGSTAB):
        wk.upload("theta", th0)
        wk.fill("b", 0.0)
        t0 = time.perf_counter()
        wk.calc_rhs(dt, True)
        conv, itr, res0 = wk.solve(solver, dt, tol, 20000, L.SMOOTHER_GS, True, precision)
        t1 = time.perf_counter()
        return t1 - t0, wk.last_solve_ms(), conv, itr, res0

    for w in range(args.warmup):
        r = step()
        log("warmup %d: %.3f ms, itr %d" % (w, r[0] * 1e3, r[3]))
    times, dev_ms, itrs = [], [], []
    for s in range(args.steps):
        r = step()
        times.append(r[0])
        dev_ms.append(r[1])
        itrs.append(r[3])
        log("step %d: %.3f ms wall, %.3f ms device, itr %d, converged %s" % (s, r[0] * 1e3, r[1], r[3], r[2]))
    total = sum(times)
    ms_per_step = total / args.steps * 1e3
    itr = itrs[-1]
    value = BYTES_PER_CELL_ITER["fp64"] * cells * itr / (ms_per_step * 1e-3) / 1e9
    value_fused = FUSED_BYTES_PER_CELL_ITER * cells * itr / (ms_per_step * 1e-3) / 1e9
    th_exact = wk.download("theta")
    parity = None
    if not args.no_parity:
        parity = parity_block(np, L, wk, step, nx, nz, args.eps, th_exact)
        log("parity: sequential sums %d iterations, %.1f ms, max |dθ| %.3e K"
            % (parity["itr"], parity["ms"], parity["max_abs_err_K"]))

    # dominant kernel, HIP events on the context stream
    resident = wk.resident_workgroups()
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    tfig = {}
    if os.path.exists(tpath):
        with open(tpath) as f:
            tfig = json.load(f)
    extra = {}
    if resident > 0:
        # ms per iteration of one launch running 200 iterations
        launch_ms = min(wk.time_kernel(11, 200, dt, True) for _ in range(3))
        per_launch_bytes = RESIDENT_ITER_BYTES_PER_CELL * cells
        name = ("k_rb_bicg_solve<double> (PBiCGSTAB!'s whole loop in one launch: per iteration BiCG1!, GS(5) "
                "with the iterate in LDS, CalcAX!, Triad!, GS(5), CalcAX!, BICG2!, Triad!, five dots through "
                "in-launch grid barriers; %d co-resident workgroups; per iteration)" % resident)
        traffic = tfig.get("k_rb_bicg_solve_%dx%dx%d" % (nx, nx, nz), {}).get("bytes_per_iteration")
        sweep_ms = wk.time_kernel(0, 200, dt, True)  # the per-colour launches, for reference
        extra["resident_k1_launch_ms"] = round(min(wk.time_kernel(9, 200, dt, True) for _ in range(3)), 5)
    else:
        sweep_ms = wk.time_kernel(0, 200, dt, True)  # both colours
        launch_ms = sweep_ms / 2.0
        per_launch_bytes = SWEEP_BYTES_PER_CELL / 2.0 * cells
        name = "k_rb_sweep<double> (one red-black colour, colour-split layout)"
        traffic = tfig.get("k_rb_sweep_%dx%dx%d" % (nx, nx, nz), {}).get("bytes_per_launch")
    achieved = per_launch_bytes / (launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "launch_ms": round(launch_ms, 5),
                "algorithmic_bytes_per_launch": per_launch_bytes,
                "per_colour_sweep_launch_ms": round(sweep_ms / 2.0, 5)}
    roofline.update(extra)

    refined = None
    if not args.no_refined:
        r64 = step(L.FP64, 1e-8)
        refined = {"tol": 1e-8, "fp64_ms": round(r64[0] * 1e3, 3), "fp64_itr": r64[3]}
        try:
            r32 = step(L.FP32_REFINED, 1e-8)
            refined.update({"fp32_refined_ms": round(r32[0] * 1e3, 3), "fp32_refined_inner_itr": r32[3],
                            "speedup": round(r64[0] / r32[0], 3), "fp32_refined_converged": r32[2],
                            "outer_residuals": [float("%.3e" % v) for v in wk.history()]})
        except L.H3DError as e:
            refined["fp32_refined_error"] = str(e)
        # the same pair with CG! as the solver (FP32 CG inner solves)
        c64 = step(L.FP64, 1e-8, L.SOLVER_CG)
        c32 = step(L.FP32_REFINED, 1e-8, L.SOLVER_CG)
        refined.update({"cg_fp64_ms": round(c64[0] * 1e3, 3), "cg_fp64_itr": c64[3],
                        "cg_fp32_refined_ms": round(c32[0] * 1e3, 3), "cg_fp32_refined_inner_itr": c32[3],
                        "cg_fp32_refined_converged": c32[2], "speedup_cg": round(c64[0] / c32[0], 3),
                        "speedup_cg_refined_vs_fp64_pbicgstab": round(r64[0] / c32[0], 3)})
    # CG! + GS(5) to 1e-6 on the same case
    cg = None
    if not args.no_cg:
        wk.upload("theta", th0)
        wk.fill("b", 0.0)
        t0 = time.perf_counter()
        wk.calc_rhs(dt, True)
        conv, itr_cg, _ = wk.solve(L.SOLVER_CG, dt, 1e-6, 20000, L.SMOOTHER_GS, True)
        cg = {"eps": 1e-6, "ms": round((time.perf_counter() - t0) * 1e3, 3), "itr": itr_cg, "converged": conv}
    wk.close()

    c3 = None
    if not args.no_c3:
        c3 = precision_c3(np, h2_amd, L, WorkBuffers, load_model)
    c5 = None
    if not args.no_c5:
        c5 = transient_step_c5(h2_amd, L, WorkBuffers, load_model)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(nx, nz, args.cpu_iters)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}

    line = {"metric": "solve wall-time & HBM GB/s, 240²×30; FP32+refine vs FP64 at 1e-8 rel resid",
            "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (modelA geometry built on the host)",
            "config": {"workload": "modelA %dx%dx%d steady PBiCGSTAB+GS(5) eps=%g, calRHS!+solve"
                                   % (nx, nx, nz, args.eps), "cells": cells, "iterations": itr,
                       "device_solve_ms": round(sum(dev_ms) / len(dev_ms), 3)},
            "value_fused": round(value_fused, 2), "parity": parity,
            "roofline": roofline, "cpu_baseline": cpu, "fp32_refined": refined, "cg": cg, "c3": c3,
            "c5_step": c5}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
