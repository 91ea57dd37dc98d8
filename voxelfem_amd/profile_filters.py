"""Timing of the grid-changing and build-direction density filters on device-resident data:

    python -m voxelfem_amd.profile_filters [--cells 384,192,192] [--planar 384,192] [--reps 20]

UpsampleFilter(2) maps the vertex grid of a half-resolution design ((nx/2+1) x (ny/2+1) x (nz/2+1)) onto the
(nx+1) x (ny+1) x (nz+1) vertices, VertexToCellFilter those vertices onto the nx x ny x nz cells (the chain
UpsampleFilter -> VertexToCellFilter that gives a design grid coarser than the elements), each forward and
adjoint; LangelaarFilter runs on the planar grid. Every entry is timed with HIP events on the stream the kernels
are launched on (the default stream: the C entries are called with stream = null), after one warm-up call, and
reported with its algorithmic bytes (each value read once, each value written once) and the fraction of the
8 TB/s HBM peak that those bytes give.
"""
import argparse
import ctypes
import json

HBM_PEAK_GBS = 8000.0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", default="384,192,192")
    ap.add_argument("--planar", default="384,192")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    import torch

    import voxelfem_amd  # noqa: F401  (loads the library next to torch's HIP runtime)
    from voxelfem_amd import capi

    L = capi.lib()
    dev = torch.device("cuda:0")
    cells = [int(v) for v in args.cells.split(",")]
    verts = [c + 1 for c in cells]
    coarse = [c // 2 + 1 for c in cells]
    i64 = ctypes.c_int64

    def buf(shape, fill=None):
        n = 1
        for s in shape:
            n *= s
        t = torch.empty(n, dtype=torch.float64, device=dev)
        if fill is None:
            t.uniform_(0.05, 0.95)
        else:
            t.fill_(fill)
        return t

    def timed(fn):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(args.reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.reps * 1e3  # microseconds

    rows = []

    def report(name, us, nbytes):
        gbs = nbytes / (us * 1e-6) / 1e9
        rows.append({"kernel": name, "us": round(us, 2), "alg_MB": round(nbytes / 1e6, 2),
                     "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)})
        print(f"{name:34s} {us:9.1f} us  {nbytes / 1e6:8.1f} MB  {gbs:7.1f} GB/s  {gbs / HBM_PEAK_GBS:.3f} of HBM",
              flush=True)

    cd = (i64 * 3)(*coarse)
    vd = (i64 * 3)(*verts)
    x_c, x_v, x_e = buf(coarse), buf(verts), buf(cells)
    g_v, g_c = buf(verts), buf(coarse)
    nc, nv, ne = x_c.numel(), x_v.numel(), x_e.numel()
    up = lambda: capi.check(L.vf_upsample_filter(3, cd, 2, x_c.data_ptr(), x_v.data_ptr(), capi.VF_DEVICE, None))
    upb = lambda: capi.check(L.vf_upsample_filter_backprop(3, cd, 2, g_v.data_ptr(), g_c.data_ptr(), capi.VF_DEVICE,
                                                           None))
    v2c = lambda: capi.check(L.vf_vertex_to_cell(3, vd, x_v.data_ptr(), x_e.data_ptr(), capi.VF_DEVICE, None))
    v2cb = lambda: capi.check(L.vf_vertex_to_cell_backprop(3, vd, x_e.data_ptr(), g_v.data_ptr(), capi.VF_DEVICE,
                                                           None))
    report(f"upsample x2 {'x'.join(map(str, coarse))}", timed(up), 8 * (nc + nv))
    report("upsample x2 backprop", timed(upb), 8 * (nv + nc))
    report(f"vertex-to-cell {'x'.join(map(str, verts))}", timed(v2c), 8 * (nv + ne))
    report("vertex-to-cell backprop", timed(v2cb), 8 * (ne + nv))

    pd = [int(v) for v in args.planar.split(",")]
    pdims = (i64 * 2)(*pd)
    p_in, p_out, p_smax = buf(pd), buf(pd), buf(pd)
    p_g, p_scr, p_res = buf(pd), buf(pd), buf(pd)
    np_ = p_in.numel()
    lg = lambda: capi.check(L.vf_langelaar(pdims, p_in.data_ptr(), p_out.data_ptr(), p_smax.data_ptr(), None))
    lgb = lambda: capi.check(L.vf_langelaar_backprop(pdims, p_g.data_ptr(), p_in.data_ptr(), p_out.data_ptr(),
                                                     p_smax.data_ptr(), p_scr.data_ptr(), p_res.data_ptr(), None))
    report(f"langelaar {pd[0]}x{pd[1]} ({pd[1]} layers)", timed(lg), 8 * 3 * np_)
    report("langelaar backprop", timed(lgb), 8 * 6 * np_)
    print(json.dumps({"filters": rows}))


if __name__ == "__main__":
    main()
