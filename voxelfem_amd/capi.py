"""ctypes view of the C ABI (include/voxelfem.h) for callers that keep data resident on the device
(torch tensors, bench.py). The library is the one pyVoxelFEM already loaded."""
import ctypes

from . import LIBRARY_PATH

VF_HOST, VF_DEVICE = 0, 1
VF_PREC_FP64, VF_PREC_FP32 = 0, 1

_lib = None


class PcgStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("residual_norm", ctypes.c_double),
                ("rhs_norm", ctypes.c_double), ("seconds", ctypes.c_double)]


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIBRARY_PATH)
        L.vf_last_error.restype = ctypes.c_char_p
        L.vf_version.restype = ctypes.c_char_p
        vp, i64, dbl, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        L.vf_mg_pcg.argtypes = [vp, vp, vp, i64, dbl, i64, i64, i32, i32, i32, vp, vp, i32, ctypes.POINTER(PcgStats)]
        L.vf_bench_apply_K.argtypes = [vp, i64, i32, ctypes.POINTER(dbl)]
        L.vf_bench_mg_kernel.argtypes = [vp, i32, i64, i64, i32, ctypes.POINTER(dbl)]
        L.vf_set_kernel_variant.argtypes = [i32]
        L.vf_mg_tuning.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(dbl)]
        L.vf_mg_tuning_level1.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(dbl)]
        L.vf_mg_set_kernel_timing.argtypes = [vp, i32]
        L.vf_mg_kernel_timing.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(dbl)]
        L.vf_sim_set_densities.argtypes = [vp, vp, i32]
        L.vf_sim_build_load.argtypes = [vp, vp, i32]
        L.vf_sim_compliance_gradient.argtypes = [vp, vp, vp, i32]
        L.vf_sim_dot.argtypes = [vp, vp, vp, i32, i32, ctypes.POINTER(dbl)]
        L.vf_sim_apply_K.argtypes = [vp, vp, vp, i32]
        L.vf_smoothing_filter.argtypes = [ctypes.POINTER(i64), dbl, i32, vp, vp, i32, vp]
        L.vf_projection.argtypes = [dbl, i64, vp, vp, i32, vp]
        L.vf_upsample_filter.argtypes = [i64, ctypes.POINTER(i64), i64, vp, vp, i32, vp]
        L.vf_upsample_filter_backprop.argtypes = [i64, ctypes.POINTER(i64), i64, vp, vp, i32, vp]
        L.vf_vertex_to_cell.argtypes = [i64, ctypes.POINTER(i64), vp, vp, i32, vp]
        L.vf_vertex_to_cell_backprop.argtypes = [i64, ctypes.POINTER(i64), vp, vp, i32, vp]
        L.vf_langelaar.argtypes = [ctypes.POINTER(i64), vp, vp, vp, vp]
        L.vf_langelaar_backprop.argtypes = [ctypes.POINTER(i64), vp, vp, vp, vp, vp, vp, vp]
        L.vf_projection_backprop.argtypes = [dbl, i64, vp, vp, vp, i32, vp]
        L.vf_mg_update_stiffness.argtypes = [vp]
        L.vf_mg_finish_update.argtypes = [vp]
        L.vf_mean.argtypes = [i64, vp, i32, vp, ctypes.POINTER(dbl)]
        L.vf_mean_noalloc.argtypes = [i64, vp, vp, vp, i32]
        L.vf_mg_set_overlap_setup.argtypes = [vp, i32]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().vf_last_error().decode())


def pcg(mg_handle, x_ptr, b_ptr, max_iter, tol, mg_iterations=1, smoothing=1, fmg=False, precision=VF_PREC_FP64,
        dirichlet_satisfied=False):
    st = PcgStats()
    check(lib().vf_mg_pcg(ctypes.c_void_p(mg_handle), ctypes.c_void_p(x_ptr), ctypes.c_void_p(b_ptr),
                          int(max_iter), float(tol), int(mg_iterations), int(smoothing), int(bool(fmg)),
                          int(precision), int(bool(dirichlet_satisfied)), None, None, VF_DEVICE,
                          ctypes.byref(st)))
    return st


def bench_mg_kernel(mg_handle, which, level, reps, precision=VF_PREC_FP64):
    """which: 0 K u, 1 residual, 2 one forward smoothing sweep; returns the average ms per call."""
    ms = ctypes.c_double()
    check(lib().vf_bench_mg_kernel(ctypes.c_void_p(mg_handle), int(which), int(level), int(reps), int(precision),
                                   ctypes.byref(ms)))
    return ms.value


def set_kernel_variant(variant):
    """0: autotuned per solver and precision (default); 1: register-gather; 2: LDS-tiled fine kernels;
    3: Walsh-basis K u with the LDS-tiled smoother; 4: Walsh-basis K u with the structured gather smoother."""
    check(lib().vf_set_kernel_variant(int(variant)))


def tuning(mg_handle, precision=VF_PREC_FP64):
    """(K u family, smoother family, timings ms [gather, tiled, Walsh K u; gather, tiled, structured-gather
    sweep])."""
    ch = (ctypes.c_int * 2)()
    ms = (ctypes.c_double * 6)()
    check(lib().vf_mg_tuning(ctypes.c_void_p(mg_handle), int(precision), ch, ms))
    return ("gather", "tiled", "walsh")[ch[0]], ("gather", "tiled", "gather-structured")[ch[1]], list(ms)


def tuning_level1(mg_handle, precision=VF_PREC_FP64):
    """(residual form, smoother form, timings ms [residual: blocks, vector, matrix; sweep: blocks, vector,
    matrix]) of level 1; a form is "blocks", "on-the-fly" (vector units) or "on-the-fly-mfma"."""
    ch = (ctypes.c_int * 2)()
    ms = (ctypes.c_double * 6)()
    check(lib().vf_mg_tuning_level1(ctypes.c_void_p(mg_handle), int(precision), ch, ms))
    forms = ("blocks", "on-the-fly", "on-the-fly-mfma")
    return forms[ch[0]], forms[ch[1]], list(ms)


def set_kernel_timing(mg_handle, enable):
    """Start (True, totals reset) or stop timing the fine K d launches of PCG where they run."""
    check(lib().vf_mg_set_kernel_timing(ctypes.c_void_p(mg_handle), 1 if enable else 0))


def kernel_timing(mg_handle):
    """(number of timed K d launches, their summed milliseconds)."""
    n = ctypes.c_int64()
    ms = ctypes.c_double()
    check(lib().vf_mg_kernel_timing(ctypes.c_void_p(mg_handle), ctypes.byref(n), ctypes.byref(ms)))
    return n.value, ms.value


def bench_apply_K(sim_handle, reps, precision=VF_PREC_FP64):
    ms = ctypes.c_double()
    check(lib().vf_bench_apply_K(ctypes.c_void_p(sim_handle), int(reps), int(precision), ctypes.byref(ms)))
    return ms.value
