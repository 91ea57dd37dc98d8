"""Headline benchmark: MG-PCG solve to 1e-8 relative residual on the 384x192x192 trilinear-hex
cantilever (43,022,595 DoF), FP32-preconditioned (mixed) vs all-FP64, plus fine-level K.u GDoF/s.

One step = one full MG-PCG solve from a zero initial guess (Galerkin hierarchy setup, coarse FP64
factorisation and all PCG iterations included, as in every iteration of a design loop) with the load
and densities already resident in HBM; setup and PCG-only times are reported beside it.
Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", default="384,192,192")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--ku-reps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lbl", action="store_true", help="skip the layer-by-layer batched-vs-sequential timing")
    ap.add_argument("--cpu-configs", action="store_true",
                    help="also time the CPU restatement on the workloads C0-C3 (about a minute of host time)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0))))
    return ap.parse_args()


def _host_cpu():
    """(model name, logical CPUs of the host, CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def _oracle_for_this_host():
    """The oracle is timed on the host it runs on: build a copy tuned for this CPU (-march=native) into a
    temporary directory; fall back to the in-tree, portable build if that fails."""
    d = tempfile.mkdtemp(prefix="vf_oracle_")
    out = os.path.join(d, "liboracle.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"OUT={out}", "ARCHFLAGS=-march=native"],
                       check=True, capture_output=True, timeout=300)
        return out, "built on the timing host with -O3 -march=native"
    except Exception as e:  # noqa: BLE001
        return None, f"in-tree portable build (-march=x86-64-v3; host build failed: {e})"


def cpu_baseline(ne, rho_value, threads, gpu_solutions):
    """The oracle (CPU restatement of the reference algorithm, FP64, OpenMP) on the same workload: one full
    MG-PCG solve from zero (Galerkin setup, coarse Cholesky and every PCG iteration) timed end to end, the
    solution compared with the GPU solutions, and the oracle's K.u throughput."""
    import numpy as np

    from oracle import oracle as orc
    from voxelfem_amd import workloads as wl
    from voxelfem_amd.mg_levels import optMGLevel

    path, how = _oracle_for_this_host()
    if path:
        orc.use_library(path)
    orc.set_threads(threads)
    s = orc.Sim(ne)
    s.set_etensor(1.0, 0.3)
    s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
    s.set_densities(np.full(int(np.prod(ne)), rho_value))
    s.set_dirichlet(wl.cantilever_mask(ne))
    f = wl.cantilever_load(ne)
    s.set_forces(f)
    mg = orc.MG(s, optMGLevel(ne, 3))
    t0 = time.perf_counter()
    u, its = mg.pcg(np.zeros_like(f), f, 1000, 1e-8)
    t_solve = time.perf_counter() - t0
    c_ref = 0.5 * float(np.sum(f * u))
    umax = float(np.abs(u).max())
    errs = {}
    for name, ug in gpu_solutions.items():
        cg = 0.5 * float(np.sum(f * ug))
        errs[name] = {"u_rel_err": float(np.abs(ug - u).max()) / umax, "compliance_rel_err": abs(cg - c_ref) / abs(c_ref)}
    # K.u throughput of the CPU restatement
    x = np.random.default_rng(0).normal(size=f.shape)
    s.apply_K(x)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        s.apply_K(x)
    t_ku = (time.perf_counter() - t0) / reps
    nn = int(np.prod([n + 1 for n in ne]))
    model, ncpu, avail = _host_cpu()
    return {
        "value": t_solve,
        "unit": "s",
        "cores": orc.get_threads(),
        "kind": "port",
        "sample": (f"oracle FP64 MG-PCG, one full solve from zero (setup + {its} iterations) of the same "
                   f"{ne[0]}x{ne[1]}x{ne[2]} cantilever to 1e-8, {orc.get_threads()} OpenMP threads on '{model}' "
                   f"({avail} of {ncpu} logical CPUs available to the process); oracle {how}"),
        "iterations": its,
        "compliance": c_ref,
        "gpu_vs_oracle": errs,
        "ku_gdofs": 3 * nn / t_ku / 1e9,
    }


def cpu_config_baselines(threads):
    """The oracle on the workloads C0-C3 as their GPU reports (voxelfem_amd.configs) run them: C0 one FP64 solve to
    1e-8 plus the sensitivity; C1 the FP64 PCG to 1e-8; C2 the filtered MBB solve to 1e-12 plus the gradient
    through the filter chain; C3 the 40 zero-guess layer solves to 1e-8 with their gradients. Seconds, PCG
    iterations and final relative residuals."""
    import numpy as np

    from oracle import oracle as orc
    from voxelfem_amd import workloads as wl
    from voxelfem_amd.mg_levels import optMGLevel

    orc.set_threads(threads)

    def sim(ne, rho, mask, f):
        s = orc.Sim(ne)
        s.set_etensor(1.0, 0.3)
        s.set_law(0, 1.0, 1e-9, 3.0, 3.0)
        s.set_densities(rho)
        s.set_dirichlet(mask)
        if f is not None:
            s.set_forces(f)
        return s

    def solve(s, ne, f, tol):
        mg = orc.MG(s, optMGLevel(ne, 3))
        u, its, hist = mg.pcg(np.zeros_like(f), f, 3000, tol, 1, 1, False, hist_len=3000)
        return u, its, float(hist[-1] / np.linalg.norm(f)) if its else 0.0

    out = {}
    ne = wl.GRIDS["cantilever_64x32x32"]
    f = wl.cantilever_load(ne)
    s = sim(ne, wl.noise_density(ne, 0), wl.cantilever_mask(ne), f)
    t0 = time.perf_counter()
    u, its, rr = solve(s, ne, f, 1e-8)
    s.compliance_gradient(u)
    out["c0"] = {"s": time.perf_counter() - t0, "pcg_iterations": its, "final_rel_residual": rr}
    print(f"cpu baseline c0: {out['c0']}", file=sys.stderr, flush=True)
    ne = wl.GRIDS["cantilever_256x128x128"]
    f = wl.cantilever_load(ne)
    s = sim(ne, wl.smoothed_density(ne, 1), wl.cantilever_mask(ne), f)
    t0 = time.perf_counter()
    u, its, rr = solve(s, ne, f, 1e-8)
    out["c1"] = {"s": time.perf_counter() - t0, "pcg_iterations": its, "final_rel_residual": rr}
    print(f"cpu baseline c1: {out['c1']}", file=sys.stderr, flush=True)
    ne = wl.GRIDS["mbb_128x64x64"]
    n = int(np.prod(ne))
    x0 = np.random.default_rng(2).uniform(0.2, 0.4, n)
    f = wl.mbb_load(ne)
    chain = orc.Chain(ne, 2.5, True, 8.0)
    t0 = time.perf_counter()
    _, p = chain.forward(x0)
    s = sim(ne, p, wl.mbb_mask(ne), f)
    u, its, rr = solve(s, ne, f, 1e-12)
    chain.backprop(s.compliance_gradient(u), x0)
    out["c2"] = {"s": time.perf_counter() - t0, "pcg_iterations": its, "final_rel_residual": rr}
    print(f"cpu baseline c2: {out['c2']}", file=sys.stderr, flush=True)
    ne = wl.GRIDS["lbl_160x80x80"]
    nny, nnz = ne[1] + 1, ne[2] + 1
    iy = (np.arange(int(np.prod([k + 1 for k in ne]))) // nnz) % nny
    s = sim(ne, wl.smoothed_density(ne, 3), np.where(iy == 0, 7, 0).astype(np.uint8), None)
    s.set_gravity([0.0, -1.0, 0.0])
    mg = orc.MG(s, optMGLevel(ne, 3))
    t0 = time.perf_counter()
    _, _, _, its = mg.lbl_run_zero(2, 1000, 1e-8)
    out["c3"] = {"s": time.perf_counter() - t0, "pcg_iterations_total": int(np.sum(its)), "layer_solves": len(its)}
    print(f"cpu baseline c3: {out['c3']}", file=sys.stderr, flush=True)
    model, ncpu, avail = _host_cpu()
    out["host"] = f"{orc.get_threads()} OpenMP threads on '{model}' ({avail} of {ncpu} logical CPUs available)"
    return out


def lbl_batched_vs_sequential(pv, precision):
    """Layer-by-layer workload 160x80x80 (40 layers of 2 along the build axis, clamped base, self-weight,
    smoothed density of seed 3): the 40 zero-guess layer solves and their gradients, run batched and one
    after the other; seconds of the second of two runs each."""
    import numpy as np

    from voxelfem_amd import workloads as wl
    from voxelfem_amd.mg_levels import optMGLevel

    ne = wl.GRIDS["lbl_160x80x80"]
    full = wl.configure(pv, ne, None, wl.smoothed_density(ne, 3))
    sim = full.getIntermediateFabricationShape(1.0, False, pv.InterpolationLaw.SIMP)
    mg = sim.multigridSolver(optMGLevel(ne, 3))
    ev = pv.LayerByLayerEvaluator(sim)
    ev.selectInitMethod("zero")
    ev.precision = precision
    out = {}
    for name, batched in (("sequential", False), ("batched", True)):
        ev.batched = batched
        ev.run(mg, True, 2, 1000, 1e-8)
        ev.run(mg, True, 2, 1000, 1e-8)
        out[name] = {"s": ev.lastRunSeconds, "objective": ev.objective(),
                     "pcg_iterations": int(np.sum(ev.perLayerIterations()))}
    out["speedup"] = out["sequential"]["s"] / out["batched"]["s"]
    out["objective_rel_diff"] = abs(out["batched"]["objective"] - out["sequential"]["objective"]) / abs(
        out["sequential"]["objective"])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(see the sharding paragraph of the design document)", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import voxelfem_amd
    from voxelfem_amd import capi
    from voxelfem_amd import workloads as wl

    pv = voxelfem_amd.pyVoxelFEM
    ne = tuple(int(x) for x in args.grid.split(","))
    nn = int(np.prod([n + 1 for n in ne]))
    nel = int(np.prod(ne))
    levels = voxelfem_amd.optMGLevel(ne, 3)
    rho_value = 0.3
    tps = wl.configure(pv, ne, wl.CANTILEVER_BC, np.full(nel, rho_value))
    mg = tps.multigridSolver(levels)
    mg.useGraphs = True  # steady-state PCG iterations replayed from a captured HIP graph
    sim_h, mg_h = tps._handle(), mg._handle()
    dev = torch.device("cuda:0")
    f = torch.zeros(3 * nn, dtype=torch.float64, device=dev)
    capi.check(capi.lib().vf_sim_build_load(sim_h, f.data_ptr(), capi.VF_DEVICE))
    x = torch.zeros_like(f)
    torch.cuda.synchronize()

    L = capi.lib()

    def setup(finish=False):
        # Galerkin hierarchy (level-1 operator, assembled block rows, coarsest band Cholesky and inverse). The
        # factorization is issued from a worker thread onto a second stream and overlaps the first cycle's
        # descent inside a step; `finish` completes it (for timing the setup alone).
        capi.check(L.vf_mg_update_stiffness(mg_h))
        if finish:
            capi.check(L.vf_mg_finish_update(mg_h))

    def solve(prec):
        x.zero_()
        return capi.pcg(mg_h, x.data_ptr(), f.data_ptr(), 1000, args.tol, 1, 1, False, prec)

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    results = {}
    solutions = {}
    for name, prec in (("fp64", capi.VF_PREC_FP64), ("mixed", capi.VF_PREC_FP32)):
        for _ in range(args.warmup):
            setup()
            solve(prec)
        # one step = hierarchy setup + PCG solve, as in every design iteration of an optimisation; the fine
        # K d launches of the timed steps are timed where they run (HIP events on the solver's stream)
        capi.set_kernel_timing(mg_h, True)
        dt, stats = timed(lambda: (setup(), solve(prec))[1], args.steps)
        kd_n, kd_ms = capi.kernel_timing(mg_h)
        capi.set_kernel_timing(mg_h, False)
        comp = 0.5 * float(torch.dot(f, x).item())
        solutions[name] = x.view(3, nn).t().contiguous().cpu().numpy()
        t_setup, _ = timed(lambda: setup(True), 1)
        t_solve, _ = timed(lambda: solve(prec), args.steps)
        results[name] = {"s": dt, "setup_s": t_setup, "solve_s": t_solve, "iterations": int(stats.iterations),
                         "rel_residual": stats.residual_norm / stats.rhs_norm, "compliance": comp,
                         "kd_launches": kd_n, "kd_avg_ms": kd_ms / max(kd_n, 1)}

    # fine-level K.u kernel: the FP64 K d of the timed mixed steps' PCG iterations (the headline region),
    # and for comparison the same kernel back to back on resident data
    tuned = {"fp64": capi.tuning(mg_h, capi.VF_PREC_FP64), "fp32": capi.tuning(mg_h, capi.VF_PREC_FP32)}
    tuned1 = {"fp64": capi.tuning_level1(mg_h, capi.VF_PREC_FP64),
              "fp32": capi.tuning_level1(mg_h, capi.VF_PREC_FP32)}
    ku_ms64 = capi.bench_mg_kernel(mg_h, 0, 0, args.ku_reps, capi.VF_PREC_FP64)
    ku_ms32 = capi.bench_mg_kernel(mg_h, 0, 0, args.ku_reps, capi.VF_PREC_FP32)
    bytes64 = 48.0 * nn + 8.0 * nel
    flops = 1152.0 * nel
    t64 = results["mixed"]["kd_avg_ms"] * 1e-3
    t_hbm, t_flop = bytes64 / (PEAK_HBM_GBS * 1e9), flops / (PEAK_FP64_TFLOPS * 1e12)
    if t_flop >= t_hbm:
        roof = {"bound": "mfma", "achieved": flops / t64 / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": bytes64 / t64 / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated), if recorded for this
    # kernel and grid
    roof["traffic"] = None
    tfile = os.path.join(ROOT, "profiles", "ku_walsh_384x192x192_traffic.json")
    if tuned["fp64"][0] == "walsh" and os.path.exists(tfile):
        rec = json.load(open(tfile))
        if tuple(rec.get("grid", ())) == ne:
            roof["traffic"] = rec["traffic_bytes"]
            roof["traffic_source"] = ("not measured in this run: rocprofv3 FETCH_SIZE + WRITE_SIZE of the same kernel "
                                      "on this grid, calibrated against a copy kernel, recorded in "
                                      "profiles/ku_walsh_384x192x192_traffic.json")
    # the compute side of the roofline of the reference's dense element formulation (F = 1152 flop per
    # element in FP64; MI355X's FP64 vector and matrix peaks are the same 78.6 TFLOP/s)
    if roof["bound"] == "mfma":
        roof["bound_note"] = ("FP64 compute bound of the reference formulation (1152 flop/element); the kernel runs "
                              "on the vector units, where the FP64 peak equals the matrix-core peak")
    # the same launch against the HBM roof alone (the Walsh-basis kernel does ~5x fewer flops than F)
    roof["hbm_frac"] = bytes64 / t64 / 1e9 / PEAK_HBM_GBS
    roof["kernel"] = "fine K.u, FP64, " + {"tiled": "k_ku_tiled<double,0> (LDS-tiled node rows)",
                                           "walsh": "k_ku_walsh<double,0> (element-centric, Walsh basis)",
                                           "gather": "k_fine_apply<double,0> (register-gather)"}[tuned["fp64"][0]]
    roof["avg_us"] = t64 * 1e6
    roof["launches"] = results["mixed"]["kd_launches"]
    roof["where"] = "K d of every PCG iteration of the timed mixed steps (events on the solver's stream)"
    roof["isolated_avg_us"] = ku_ms64 * 1e3  # back-to-back launches on resident data (caches warm)
    roof["isolated_frac"] = flops / (ku_ms64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS if roof["bound"] == "mfma" \
        else bytes64 / (ku_ms64 * 1e-3) / 1e9 / PEAK_HBM_GBS
    # fine smoothing sweep (8 colours): 2 x (576 + 72) flops per node update in the dense count
    sweep = {}
    for pname, prec, peak in (("fp64", capi.VF_PREC_FP64, PEAK_FP64_TFLOPS), ("fp32", capi.VF_PREC_FP32,
                                                                               PEAK_FP32_TFLOPS)):
        ms = capi.bench_mg_kernel(mg_h, 2, 0, args.ku_reps, prec)
        tf = 1296.0 * nn / (ms * 1e-3) / 1e12
        sweep[pname] = {"avg_us": ms * 1e3, "tflops": tf, "frac_of_peak": tf / peak}

    line = {
        "metric": "MG-PCG solve s to 1e-8 rel. residual @43M DoF, mixed vs FP64; K·u GDoF/s",
        "value": results["mixed"]["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results["mixed"]["s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (PCG) / f32 (V-cycle)",
        "data": "synthetic",
        "config": {"workload": f"cantilever {ne[0]}x{ne[1]}x{ne[2]} Q1 hex, uniform density 0.3, E0=1, "
                               f"Emin=1e-9, nu=0.3, SIMP p=3", "dof": 3 * nn, "mg_levels": levels,
                   "smoother": "8-colour block Gauss-Seidel, 1 pre + 1 post sweep", "tol": args.tol},
        "fp64_step_s": results["fp64"]["s"],
        "mixed_step_s": results["mixed"]["s"],
        "speedup_mixed_vs_fp64": results["fp64"]["s"] / results["mixed"]["s"],
        "setup_s": {"fp64": results["fp64"]["setup_s"], "mixed": results["mixed"]["setup_s"]},
        "pcg_only_s": {"fp64": results["fp64"]["solve_s"], "mixed": results["mixed"]["solve_s"]},
        "speedup_pcg_only": results["fp64"]["solve_s"] / results["mixed"]["solve_s"],
        "fine_sweep": sweep,
        "fine_kernels": {p: {"K.u": t[0], "sweep": t[1], "tuner_ms": [round(v, 4) for v in t[2]]}
                         for p, t in tuned.items()},
        "level1_operator": {p: {"residual": t[0], "sweep": t[1], "tuner_ms": [round(v, 4) for v in t[2]]}
                            for p, t in tuned1.items()},
        "pcg_iterations": {"fp64": results["fp64"]["iterations"], "mixed": results["mixed"]["iterations"]},
        "final_rel_residual": {"fp64": results["fp64"]["rel_residual"], "mixed": results["mixed"]["rel_residual"]},
        "compliance": {"fp64": results["fp64"]["compliance"], "mixed": results["mixed"]["compliance"],
                       "rel_diff": abs(results["mixed"]["compliance"] - results["fp64"]["compliance"])
                       / abs(results["fp64"]["compliance"])},
        "ku_gdofs": {"fp64": 3 * nn / t64 / 1e9, "fp32": 3 * nn / (ku_ms32 * 1e-3) / 1e9},
        "roofline": roof,
    }
    if not args.no_lbl:
        line["lbl_160x80x80"] = {"fp64": lbl_batched_vs_sequential(pv, "fp64"),
                                 "mixed": lbl_batched_vs_sequential(pv, "fp32")}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(ne, rho_value, args.cpu_threads, solutions)
    if args.cpu_configs:
        if args.no_cpu_baseline:
            from oracle import oracle as orc

            path, _ = _oracle_for_this_host()
            if path:
                orc.use_library(path)
        line["cpu_baseline_configs"] = cpu_config_baselines(args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
