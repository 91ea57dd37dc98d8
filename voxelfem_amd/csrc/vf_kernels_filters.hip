// vf_kernels_filters.hip -- the grid-changing and build-direction density filters of
// TopologyOptimizationFilter.hh: UpsampleFilter (419-525), VertexToCellFilter (527-597) and, on planar grids,
// LangelaarFilter (600-712). One work-item per output value (or per input value for the adjoints, gathered,
// so every sum has one writer and a fixed order); grids are C order with the last index fastest.
#include <cmath>

#include "vf_internal.hpp"

namespace vf {
namespace {

constexpr int TPB = 256;
inline unsigned nblk(size_t n) { return (unsigned)std::max<size_t>(1, (n + TPB - 1) / TPB); }

struct Dims {
    int nd;
    int n[3];
};
__device__ __forceinline__ void unflatten(const Dims &d, long i, int idx[3]) {
    for (int k = d.nd - 1; k >= 0; --k) {
        idx[k] = (int)(i % d.n[k]);
        i /= d.n[k];
    }
}
__device__ __forceinline__ long flatten(const Dims &d, const int idx[3]) {
    long f = 0;
    for (int k = 0; k < d.nd; ++k) f = f * d.n[k] + idx[k];
    return f;
}

// UpsampleFilter::apply, degree 1: fine node j lies in the coarse cell c = min(j / f, in - 2) per axis, at the
// local coordinate t = (j - c f) / f; its value is the multilinear interpolant of the cell's 2^N corner values
__global__ void __launch_bounds__(TPB) k_upsample(Dims din, Dims dout, int f, const double *__restrict__ in,
                                                  double *__restrict__ out) {
    const long j = (long)blockIdx.x * TPB + threadIdx.x;
    long nout = 1;
    for (int k = 0; k < dout.nd; ++k) nout *= dout.n[k];
    if (j >= nout) return;
    int jf[3], c[3];
    double t[3];
    unflatten(dout, j, jf);
    for (int k = 0; k < din.nd; ++k) {
        c[k] = min(jf[k] / f, din.n[k] - 2);
        t[k] = (double)(jf[k] - c[k] * f) / (double)f;
    }
    double v = 0;
    for (int corner = 0; corner < (1 << din.nd); ++corner) {
        int ci[3];
        double w = 1;
        for (int k = 0; k < din.nd; ++k) {
            const int bit = (corner >> (din.nd - 1 - k)) & 1;
            ci[k] = c[k] + bit;
            w *= bit ? t[k] : 1.0 - t[k];
        }
        v += w * in[flatten(din, ci)];
    }
    out[j] = v;
}

// UpsampleFilter::backprop: coarse node n_c gathers the fine nodes of its hat's support (n_f = f n_c, +-(f - 1)
// inside the grid) weighted by prod_k (1 - |n_v - n_f|_k / f)
__global__ void __launch_bounds__(TPB) k_upsample_bp(Dims din, Dims dout, int f, const double *__restrict__ gout,
                                                     double *__restrict__ gin) {
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    long nin = 1;
    for (int k = 0; k < din.nd; ++k) nin *= din.n[k];
    if (i >= nin) return;
    int nc[3], lo[3], hi[3];
    unflatten(din, i, nc);
    long count = 1;
    for (int k = 0; k < din.nd; ++k) {
        const int nf = nc[k] * f;
        lo[k] = nc[k] > 0 ? nf - (f - 1) : nf;
        hi[k] = nc[k] < din.n[k] - 1 ? nf + (f - 1) : nf;
        count *= hi[k] - lo[k] + 1;
    }
    double s = 0;
    for (long q = 0; q < count; ++q) {
        int v[3];
        long r = q;
        double w = 1;
        for (int k = din.nd - 1; k >= 0; --k) {
            const int len = hi[k] - lo[k] + 1;
            v[k] = lo[k] + (int)(r % len);
            r /= len;
        }
        for (int k = 0; k < din.nd; ++k) w *= 1.0 - (double)abs(v[k] - nc[k] * f) / (double)f;
        s += w * gout[flatten(dout, v)];
    }
    gin[i] = s;
}

// VertexToCellFilter: cell value = 2^-N sum of its corner vertices; adjoint: vertex = 2^-N sum of its cells
__global__ void __launch_bounds__(TPB) k_v2c(Dims din, Dims dout, const double *__restrict__ in,
                                             double *__restrict__ out) {
    const long e = (long)blockIdx.x * TPB + threadIdx.x;
    long nout = 1;
    for (int k = 0; k < dout.nd; ++k) nout *= dout.n[k];
    if (e >= nout) return;
    int c[3];
    unflatten(dout, e, c);
    double s = 0;
    for (int corner = 0; corner < (1 << din.nd); ++corner) {
        int v[3];
        for (int k = 0; k < din.nd; ++k) v[k] = c[k] + ((corner >> (din.nd - 1 - k)) & 1);
        s += in[flatten(din, v)];
    }
    out[e] = s * (din.nd == 2 ? 0.25 : 0.125);
}
__global__ void __launch_bounds__(TPB) k_v2c_bp(Dims din, Dims dout, const double *__restrict__ gout,
                                                double *__restrict__ gin) {
    const long vi = (long)blockIdx.x * TPB + threadIdx.x;
    long nin = 1;
    for (int k = 0; k < din.nd; ++k) nin *= din.n[k];
    if (vi >= nin) return;
    int v[3];
    unflatten(din, vi, v);
    double s = 0;
    for (int corner = 0; corner < (1 << din.nd); ++corner) {
        int e[3];
        bool ok = true;
        for (int k = 0; k < din.nd; ++k) {
            e[k] = v[k] - 1 + ((corner >> (din.nd - 1 - k)) & 1);
            ok = ok && e[k] >= 0 && e[k] < dout.n[k];
        }
        if (ok) s += gout[flatten(dout, e)];
    }
    gin[vi] = s * (din.nd == 2 ? 0.25 : 0.125);
}

// LangelaarFilter on a planar grid (nx x ny, y the build direction): layer y of the output is
// smin(in, smax(out of the supporting voxels (x - 1 .. x + 1, y - 1))), smax = (sum out^P)^(1/Q),
// smin(a, b) = (a + b - sqrt((a - b)^2 + eps) + sqrt(eps)) / 2; layer 0 is copied
constexpr double LG_EPS = 1e-4, LG_P = 40.0, LG_Q = 40.0 - 1.58;
__device__ __forceinline__ double lg_smin(double a, double b) {
    return 0.5 * (a + b - pow((a - b) * (a - b) + LG_EPS, 0.5) + pow(LG_EPS, 0.5));
}
__device__ __forceinline__ double lg_dsmin_dx1(double a, double b) {
    return 0.5 * (1 - (a - b) * pow((a - b) * (a - b) + LG_EPS, -0.5));
}
__device__ __forceinline__ double lg_dsmin_dx2(double a, double b) {
    return 0.5 * (1 + (a - b) * pow((a - b) * (a - b) + LG_EPS, -0.5));
}
__device__ __forceinline__ double lg_sum(int nx, int ny, const double *v, int x, int y) {
    // supporting region of (x, y): (x, y - 1), then (x - 1, y - 1), (x + 1, y - 1) inside the grid
    double s = pow(v[(long)x * ny + y - 1], LG_P);
    if (x > 0) s += pow(v[(long)(x - 1) * ny + y - 1], LG_P);
    if (x + 1 < nx) s += pow(v[(long)(x + 1) * ny + y - 1], LG_P);
    return s;
}
__global__ void __launch_bounds__(TPB) k_langelaar_layer(int nx, int ny, int y, const double *__restrict__ in,
                                                         double *__restrict__ out, double *__restrict__ smax) {
    const int x = blockIdx.x * TPB + threadIdx.x;
    if (x >= nx) return;
    const long i = (long)x * ny + y;
    if (y == 0) {
        out[i] = in[i];
        return;
    }
    const double m = pow(lg_sum(nx, ny, out, x, y), 1.0 / LG_Q);
    smax[i] = m;
    out[i] = lg_smin(in[i], m);
}
// multipliers of layer y (top layer first): lambda_k = g_k + sum over the voxels i of layer y + 1 supported by
// k of lambda_i dsmin/dx2(vars_i, smax_i) dsmax_i/dfiltered_k
__global__ void __launch_bounds__(TPB) k_langelaar_lambda(int nx, int ny, int y, const double *__restrict__ g,
                                                          const double *__restrict__ vars,
                                                          const double *__restrict__ filtered,
                                                          const double *__restrict__ smax, double *__restrict__ lam) {
    const int x = blockIdx.x * TPB + threadIdx.x;
    if (x >= nx) return;
    const long k = (long)x * ny + y;
    double l = g[k];
    if (y + 1 < ny) {
        // the voxels above whose support holds (x, y), in the reference's order of accumulation (x ascending)
        const int xs[3] = {x - 1, x, x + 1};
        for (int q = 0; q < 3; ++q) {
            const int xi = xs[q];
            if (xi < 0 || xi >= nx) continue;
            const long i = (long)xi * ny + y + 1;
            const double sum = lg_sum(nx, ny, filtered, xi, y + 1);
            const double dsmax = LG_P * pow(filtered[k], LG_P - 1) / LG_Q * pow(sum, 1 / LG_Q - 1);
            l += lam[i] * (lg_dsmin_dx2(vars[i], smax[i]) * dsmax);
        }
    }
    lam[k] = l;
}
__global__ void __launch_bounds__(TPB) k_langelaar_out(int nx, int ny, const double *__restrict__ lam,
                                                       const double *__restrict__ vars,
                                                       const double *__restrict__ smax, double *__restrict__ out) {
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= (long)nx * ny) return;
    out[i] = i % ny == 0 ? lam[i] : lam[i] * lg_dsmin_dx1(vars[i], smax[i]);
}

Dims make_dims(int nd, const int64_t *n) {
    Dims d;
    d.nd = nd;
    for (int k = 0; k < 3; ++k) d.n[k] = k < nd ? (int)n[k] : 1;
    return d;
}
size_t count(const Dims &d) {
    size_t c = 1;
    for (int k = 0; k < d.nd; ++k) c *= (size_t)d.n[k];
    return c;
}

}  // namespace

void upsample_dims(int nd, const int64_t *in, int factor, int64_t *out) {
    for (int k = 0; k < nd; ++k) out[k] = (in[k] - 1) * factor + 1;
}

void launch_upsample(int nd, const int64_t *inDims, int factor, const double *in, double *out, bool adjoint,
                     hipStream_t s) {
    int64_t od[3];
    upsample_dims(nd, inDims, factor, od);
    const Dims din = make_dims(nd, inDims), dout = make_dims(nd, od);
    if (!adjoint) hipLaunchKernelGGL(k_upsample, dim3(nblk(count(dout))), dim3(TPB), 0, s, din, dout, factor, in, out);
    else hipLaunchKernelGGL(k_upsample_bp, dim3(nblk(count(din))), dim3(TPB), 0, s, din, dout, factor, in, out);
    VF_HIP(hipGetLastError());
}

void launch_vertex_to_cell(int nd, const int64_t *inDims, const double *in, double *out, bool adjoint, hipStream_t s) {
    int64_t od[3];
    for (int k = 0; k < nd; ++k) od[k] = inDims[k] - 1;
    const Dims din = make_dims(nd, inDims), dout = make_dims(nd, od);
    if (!adjoint) hipLaunchKernelGGL(k_v2c, dim3(nblk(count(dout))), dim3(TPB), 0, s, din, dout, in, out);
    else hipLaunchKernelGGL(k_v2c_bp, dim3(nblk(count(din))), dim3(TPB), 0, s, din, dout, in, out);
    VF_HIP(hipGetLastError());
}

void launch_langelaar(int nx, int ny, const double *in, double *out, double *smax, hipStream_t s) {
    for (int y = 0; y < ny; ++y)
        hipLaunchKernelGGL(k_langelaar_layer, dim3(nblk(nx)), dim3(TPB), 0, s, nx, ny, y, in, out, smax);
    VF_HIP(hipGetLastError());
}

void launch_langelaar_backprop(int nx, int ny, const double *g, const double *vars, const double *filtered,
                               const double *smax, double *lam, double *out, hipStream_t s) {
    for (int y = ny - 1; y >= 0; --y)
        hipLaunchKernelGGL(k_langelaar_lambda, dim3(nblk(nx)), dim3(TPB), 0, s, nx, ny, y, g, vars, filtered, smax,
                           lam);
    hipLaunchKernelGGL(k_langelaar_out, dim3(nblk((size_t)nx * ny)), dim3(TPB), 0, s, nx, ny, lam, vars, smax, out);
    VF_HIP(hipGetLastError());
}

}  // namespace vf
