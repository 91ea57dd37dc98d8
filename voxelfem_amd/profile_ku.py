"""Driver for counter-based profiling of the fine-level K u (run it under rocprofv3 --pmc):

    python -m voxelfem_amd.profile_ku [--grid 384,192,192] [--variant 3] [--reps 3] [--which 0] [--prec fp64]
                                      [--level 0]

It builds the cantilever workload's hierarchy, launches the chosen fine-level kernel family `reps` times
through vf_bench_mg_kernel (variant: 1 register-gather, 2 LDS-tiled, 3 Walsh-basis K u with the tiled
smoother, 4 Walsh-basis K u with the structured gather smoother; which: 0 K u, 2 one smoothing sweep; level:
the hierarchy level, whose kernel family the autotuner picks on levels >= 1), then
runs a 2 GiB device fill and a 2 GiB device copy whose byte counts are known, to calibrate FETCH_SIZE /
WRITE_SIZE.
"""
import argparse

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", default="384,192,192")
    ap.add_argument("--variant", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--which", type=int, default=0, choices=[0, 2])
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--level", type=int, default=0)
    args = ap.parse_args(argv)
    import torch

    import voxelfem_amd
    from voxelfem_amd import capi
    from voxelfem_amd import workloads as wl

    ne = tuple(int(x) for x in args.grid.split(","))
    tps = wl.configure(voxelfem_amd.pyVoxelFEM, ne, wl.CANTILEVER_BC, np.full(int(np.prod(ne)), 0.3))
    mg = tps.multigridSolver(voxelfem_amd.optMGLevel(ne, 3))
    h = mg._handle()
    capi.check(capi.lib().vf_mg_update_stiffness(h))
    capi.set_kernel_variant(args.variant)
    prec = capi.VF_PREC_FP32 if args.prec == "fp32" else capi.VF_PREC_FP64
    ms = capi.bench_mg_kernel(h, args.which, args.level, args.reps, prec)
    capi.set_kernel_variant(0)
    what = "K u" if args.which == 0 else "sweep"
    print(f"{what} {args.prec} level {args.level} variant {args.variant}: {ms * 1e3:.1f} us per call", flush=True)
    a = torch.ones(2**28, dtype=torch.float64, device="cuda:0")
    b = torch.zeros_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"calibration: fill and copy of {a.numel() * 8} bytes", flush=True)


if __name__ == "__main__":
    main()
