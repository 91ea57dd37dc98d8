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

// Grids of the upsample and vertex-to-cell kernels as three extents n0 x n1 x n2 (C order, n2 fastest); a
// planar grid is 1 x nx x ny. A workgroup covers BZ values along n2 and BY along n1; blockIdx.x enumerates
// (i0, n1 block, n2 block) with n2 blocks fastest, so a wavefront reads and writes contiguous runs.
constexpr int BZ = 64, BY = 4;
struct Shape {
    int n[3];
    int nb1, nb2;
};
Shape make_shape(int nd, const int64_t *d) {
    Shape s;
    s.n[0] = nd == 3 ? (int)d[0] : 1;
    s.n[1] = (int)d[nd - 2];
    s.n[2] = (int)d[nd - 1];
    s.nb1 = (s.n[1] + BY - 1) / BY;
    s.nb2 = (s.n[2] + BZ - 1) / BZ;
    return s;
}
unsigned shape_blocks(const Shape &s) { return (unsigned)((size_t)s.n[0] * s.nb1 * s.nb2); }
__device__ __forceinline__ bool decode(const Shape &s, int i[3]) {
    const int b = blockIdx.x;
    const int b2 = b % s.nb2, r = b / s.nb2;
    const int b1 = r % s.nb1;
    i[0] = r / s.nb1;
    i[1] = b1 * BY + (int)threadIdx.y;
    i[2] = b2 * BZ + (int)threadIdx.x;
    return i[1] < s.n[1] && i[2] < s.n[2];
}
__device__ __forceinline__ int flat(const Shape &s, const int i[3]) { return (i[0] * s.n[1] + i[1]) * s.n[2] + i[2]; }

// UpsampleFilter::apply, degree 1 (ND = 2 or 3 axes, the last ND of the three): fine node j lies in the coarse
// cell c = min(j / f, in - 2) per axis, at the local coordinate t = (j - c f) / f; its value is the multilinear
// interpolant of the cell's 2^ND corner values, summed corner by corner (first axis most significant)
// (j / f from (j + 1/2) / f in FP64, exact for any grid that fits an int; t = r * (1/f), which equals the
// reference's r / f for power-of-two factors and is within an ulp of it otherwise)
template <int ND>
__global__ void __launch_bounds__(BZ *BY) k_upsample(Shape in, Shape out, int f, double invf,
                                                     const double *__restrict__ a, double *__restrict__ b) {
    constexpr int A0 = 3 - ND;
    int j[3];
    if (!decode(out, j)) return;
    int c[3] = {0, 0, 0};
    double t[3] = {0, 0, 0};
#pragma unroll
    for (int k = A0; k < 3; ++k) {
        c[k] = min((int)(((double)j[k] + 0.5) * invf), in.n[k] - 2);
        t[k] = (double)(j[k] - c[k] * f) * invf;
    }
    double v = 0;
#pragma unroll
    for (int corner = 0; corner < (1 << ND); ++corner) {
        int ci[3] = {0, 0, 0};
        double w = 1;
#pragma unroll
        for (int k = A0; k < 3; ++k) {
            const int bit = (corner >> (2 - k)) & 1;
            ci[k] = c[k] + bit;
            w *= bit ? t[k] : 1.0 - t[k];
        }
        v += w * a[flat(in, ci)];
    }
    b[flat(out, j)] = v;
}

// UpsampleFilter::backprop: coarse node n_c gathers the fine nodes of its hat's support (n_f = f n_c, +-(f - 1)
// inside the grid, C order) weighted by prod_k (1 - |n_v - n_f|_k / f). For F = 2..4 the (2F - 1)^ND offsets are
// unrolled (factor 2 in 3D, 2..4 in 2D, so that the values fit in registers): every value is loaded (a clamped
// index outside the grid) before the in-order sum, which skips the offsets outside the grid.
template <int ND, int F>
__global__ void __launch_bounds__(BZ *BY) k_upsample_bp_f(Shape in, Shape out, const double *__restrict__ gout,
                                                          double *__restrict__ gin) {
    constexpr int A0 = 3 - ND, W = 2 * F - 1, NV = ND == 3 ? W * W * W : W * W;
    int nc[3];
    if (!decode(in, nc)) return;
    double wt[3][W];
    bool ok[3][W];
    int idx[3][W];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int o = 0; o < W; ++o) {
            const int v = nc[k] * F + o - (F - 1);
            ok[k][o] = k < A0 ? o == F - 1 : v >= 0 && v < out.n[k];
            idx[k][o] = ok[k][o] ? v : 0;
            wt[k][o] = 1.0 - (double)abs(o - (F - 1)) / (double)F;
        }
    double val[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int o2 = q % W, o1 = (q / W) % W, o0 = ND == 3 ? q / (W * W) : F - 1;
        const int v[3] = {idx[0][o0], idx[1][o1], idx[2][o2]};
        val[q] = gout[flat(out, v)];
    }
    double s = 0;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int o2 = q % W, o1 = (q / W) % W, o0 = ND == 3 ? q / (W * W) : F - 1;
        double w = 1;
        if (ND == 3) w *= wt[0][o0];
        w *= wt[1][o1];
        w *= wt[2][o2];
        const double t = s + w * val[q];
        s = ok[0][o0] && ok[1][o1] && ok[2][o2] ? t : s;
    }
    gin[flat(in, nc)] = s;
}
template <int ND>
__global__ void __launch_bounds__(BZ *BY) k_upsample_bp(Shape in, Shape out, int f, const double *__restrict__ gout,
                                                        double *__restrict__ gin) {
    constexpr int A0 = 3 - ND;
    int nc[3];
    if (!decode(in, nc)) return;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int k = A0; k < 3; ++k) {
        const int nf = nc[k] * f;
        lo[k] = nc[k] > 0 ? nf - (f - 1) : nf;
        hi[k] = nc[k] < in.n[k] - 1 ? nf + (f - 1) : nf;
    }
    double s = 0;
    int v[3];
    for (v[0] = lo[0]; v[0] <= hi[0]; ++v[0])
        for (v[1] = lo[1]; v[1] <= hi[1]; ++v[1])
            for (v[2] = lo[2]; v[2] <= hi[2]; ++v[2]) {
                double w = 1;
#pragma unroll
                for (int k = A0; k < 3; ++k) w *= 1.0 - (double)abs(v[k] - nc[k] * f) / (double)f;
                s += w * gout[flat(out, v)];
            }
    gin[flat(in, nc)] = s;
}

// VertexToCellFilter: cell value = 2^-ND sum of its corner vertices; adjoint: vertex = 2^-ND sum of its cells
template <int ND>
__global__ void __launch_bounds__(BZ *BY) k_v2c(Shape in, Shape out, const double *__restrict__ a,
                                                double *__restrict__ b) {
    constexpr int A0 = 3 - ND;
    int c[3];
    if (!decode(out, c)) return;
    double s = 0;
#pragma unroll
    for (int corner = 0; corner < (1 << ND); ++corner) {
        int v[3] = {c[0], c[1], c[2]};
#pragma unroll
        for (int k = A0; k < 3; ++k) v[k] += (corner >> (2 - k)) & 1;
        s += a[flat(in, v)];
    }
    b[flat(out, c)] = s * (ND == 2 ? 0.25 : 0.125);
}
template <int ND>
__global__ void __launch_bounds__(BZ *BY) k_v2c_bp(Shape in, Shape out, const double *__restrict__ gout,
                                                   double *__restrict__ gin) {
    constexpr int A0 = 3 - ND;
    int vi[3];
    if (!decode(in, vi)) return;
    // every cell value is loaded first (a clamped index for the cells outside the grid), then summed in order
    double val[1 << ND];
    bool okc[1 << ND];
#pragma unroll
    for (int corner = 0; corner < (1 << ND); ++corner) {
        int e[3] = {vi[0], vi[1], vi[2]};
        bool ok = true;
#pragma unroll
        for (int k = A0; k < 3; ++k) {
            e[k] += ((corner >> (2 - k)) & 1) - 1;
            ok = ok && e[k] >= 0 && e[k] < out.n[k];
        }
        okc[corner] = ok;
        const int fi = flat(out, e);
        val[corner] = gout[ok ? fi : 0];
    }
    double s = 0;
#pragma unroll
    for (int corner = 0; corner < (1 << ND); ++corner) s = okc[corner] ? s + val[corner] : s;
    gin[flat(in, vi)] = s * (ND == 2 ? 0.25 : 0.125);
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
// The recurrence runs in one workgroup when a layer fits in LDS (nx <= LG_MAXX): layer y - 1 (forward) or y + 1
// (backprop) is kept in LDS and every layer costs two barriers instead of a launch; the next layer's inputs are
// loaded while the current one is formed. v^40 and v^39 are formed by repeated squaring.
constexpr int LG_TPB = 1024, LG_MAXX = 2048, LG_PER = LG_MAXX / LG_TPB;
__device__ __forceinline__ double lg_pow40(double v) {
    const double v2 = v * v, v4 = v2 * v2, v5 = v4 * v, v10 = v5 * v5, v20 = v10 * v10;
    return v20 * v20;
}
__device__ __forceinline__ double lg_pow39(double v) {
    const double v2 = v * v, v4 = v2 * v2, v8 = v4 * v4, v16 = v8 * v8, v32 = v16 * v16;
    return v32 * v4 * v2 * v;
}
__global__ void __launch_bounds__(LG_TPB) k_langelaar_fused(int nx, int ny, const double *__restrict__ in,
                                                            double *__restrict__ out, double *__restrict__ smax) {
    __shared__ double prev[LG_MAXX];
    const int t = threadIdx.x;
    double nxt[LG_PER];
#pragma unroll
    for (int q = 0; q < LG_PER; ++q) {
        const int x = t + q * LG_TPB;
        if (x < nx) {
            const double v = in[(long)x * ny];
            out[(long)x * ny] = v;
            prev[x] = v;
            if (ny > 1) nxt[q] = in[(long)x * ny + 1];
        }
    }
    __syncthreads();
    for (int y = 1; y < ny; ++y) {
        double r[LG_PER];
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            const int x = t + q * LG_TPB;
            if (x < nx) {
                const double cur = nxt[q];
                if (y + 1 < ny) nxt[q] = in[(long)x * ny + y + 1];
                // supporting region (x, y - 1), (x - 1, y - 1), (x + 1, y - 1), summed in that order
                double sum = lg_pow40(prev[x]);
                if (x > 0) sum += lg_pow40(prev[x - 1]);
                if (x + 1 < nx) sum += lg_pow40(prev[x + 1]);
                const double m = pow(sum, 1.0 / LG_Q);
                const long i = (long)x * ny + y;
                smax[i] = m;
                r[q] = lg_smin(cur, m);
                out[i] = r[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            const int x = t + q * LG_TPB;
            if (x < nx) prev[x] = r[q];
        }
        __syncthreads();
    }
}
// The backprop's per-voxel factors, in parallel over the grid (they depend on the apply's cached values only):
// h_i = dsmin/dx2(vars_i, smax_i) pow(sum over the supporting region of filtered, 1/Q - 1) and
// o_i = dsmin/dx1(vars_i, smax_i) for the voxels of layers >= 1 (h = 0, o = 1 on layer 0)
__global__ void __launch_bounds__(TPB) k_langelaar_factors(int nx, int ny, const double *__restrict__ vars,
                                                           const double *__restrict__ filtered,
                                                           const double *__restrict__ smax, double *__restrict__ h,
                                                           double *__restrict__ o) {
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= (long)nx * ny) return;
    const int x = (int)(i / ny), y = (int)(i - (long)x * ny);
    if (y == 0) {
        h[i] = 0;
        o[i] = 1;
        return;
    }
    double sum = lg_pow40(filtered[i - 1]);
    if (x > 0) sum += lg_pow40(filtered[i - ny - 1]);
    if (x + 1 < nx) sum += lg_pow40(filtered[i + ny - 1]);
    h[i] = lg_dsmin_dx2(vars[i], smax[i]) * pow(sum, 1 / LG_Q - 1);
    o[i] = lg_dsmin_dx1(vars[i], smax[i]);
}
// multipliers from the top layer down, lambda_k = g_k + sum over the voxels i above that k supports (x - 1, x,
// x + 1) of lambda_i (h_i (P filtered_k^39 / Q)); out = lambda o (o read from out, where the factors kernel put it)
__global__ void __launch_bounds__(LG_TPB) k_langelaar_bp_fused(int nx, int ny, const double *__restrict__ g,
                                                               const double *__restrict__ filtered,
                                                               const double *__restrict__ h, double *out) {
    __shared__ double lamA[LG_MAXX], hA[LG_MAXX];
    const int t = threadIdx.x;
    double ng[LG_PER], nf[LG_PER], nh[LG_PER], no[LG_PER];
    auto load = [&](int y) {
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            const int x = t + q * LG_TPB;
            if (x < nx) {
                const long k = (long)x * ny + y;
                ng[q] = g[k];
                nf[q] = filtered[k];
                nh[q] = h[k];
                no[q] = out[k];
            }
        }
    };
    load(ny - 1);
    for (int y = ny - 1; y >= 0; --y) {
        double gk[LG_PER], fk[LG_PER], hk[LG_PER], ok[LG_PER];
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            gk[q] = ng[q];
            fk[q] = nf[q];
            hk[q] = nh[q];
            ok[q] = no[q];
        }
        if (y > 0) load(y - 1);
        double lam[LG_PER];
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            const int x = t + q * LG_TPB;
            if (x < nx) {
                double l = gk[q];
                if (y + 1 < ny) {
                    const double F = LG_P * lg_pow39(fk[q]) / LG_Q;
                    if (x > 0) l += lamA[x - 1] * (hA[x - 1] * F);
                    l += lamA[x] * (hA[x] * F);
                    if (x + 1 < nx) l += lamA[x + 1] * (hA[x + 1] * F);
                }
                lam[q] = l;
                out[(long)x * ny + y] = y == 0 ? l : l * ok[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LG_PER; ++q) {
            const int x = t + q * LG_TPB;
            if (x < nx) {
                lamA[x] = lam[q];
                hA[x] = hk[q];
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(TPB) k_langelaar_out(int nx, int ny, const double *__restrict__ lam,
                                                       const double *__restrict__ vars,
                                                       const double *__restrict__ smax, double *__restrict__ out) {
    const long i = (long)blockIdx.x * TPB + threadIdx.x;
    if (i >= (long)nx * ny) return;
    out[i] = i % ny == 0 ? lam[i] : lam[i] * lg_dsmin_dx1(vars[i], smax[i]);
}

}  // namespace

void upsample_dims(int nd, const int64_t *in, int factor, int64_t *out) {
    for (int k = 0; k < nd; ++k) out[k] = (in[k] - 1) * factor + 1;
}

void launch_upsample(int nd, const int64_t *inDims, int factor, const double *in, double *out, bool adjoint,
                     hipStream_t s) {
    int64_t od[3];
    upsample_dims(nd, inDims, factor, od);
    const Shape si = make_shape(nd, inDims), so = make_shape(nd, od);
    const dim3 blk(BZ, BY);
    const double invf = 1.0 / (double)factor;
    if (!adjoint) {
        if (nd == 3) hipLaunchKernelGGL(k_upsample<3>, dim3(shape_blocks(so)), blk, 0, s, si, so, factor, invf, in, out);
        else hipLaunchKernelGGL(k_upsample<2>, dim3(shape_blocks(so)), blk, 0, s, si, so, factor, invf, in, out);
    } else {
        const dim3 g(shape_blocks(si));
        if (nd == 3 && factor == 2) hipLaunchKernelGGL((k_upsample_bp_f<3, 2>), g, blk, 0, s, si, so, in, out);
        else if (nd == 2 && factor == 2) hipLaunchKernelGGL((k_upsample_bp_f<2, 2>), g, blk, 0, s, si, so, in, out);
        else if (nd == 2 && factor == 3) hipLaunchKernelGGL((k_upsample_bp_f<2, 3>), g, blk, 0, s, si, so, in, out);
        else if (nd == 2 && factor == 4) hipLaunchKernelGGL((k_upsample_bp_f<2, 4>), g, blk, 0, s, si, so, in, out);
        else if (nd == 3) hipLaunchKernelGGL(k_upsample_bp<3>, g, blk, 0, s, si, so, factor, in, out);
        else hipLaunchKernelGGL(k_upsample_bp<2>, g, blk, 0, s, si, so, factor, in, out);
    }
    VF_HIP(hipGetLastError());
}

void launch_vertex_to_cell(int nd, const int64_t *inDims, const double *in, double *out, bool adjoint, hipStream_t s) {
    int64_t od[3];
    for (int k = 0; k < nd; ++k) od[k] = inDims[k] - 1;
    const Shape si = make_shape(nd, inDims), so = make_shape(nd, od);
    const dim3 blk(BZ, BY);
    if (nd == 3) {
        if (!adjoint) hipLaunchKernelGGL(k_v2c<3>, dim3(shape_blocks(so)), blk, 0, s, si, so, in, out);
        else hipLaunchKernelGGL(k_v2c_bp<3>, dim3(shape_blocks(si)), blk, 0, s, si, so, in, out);
    } else {
        if (!adjoint) hipLaunchKernelGGL(k_v2c<2>, dim3(shape_blocks(so)), blk, 0, s, si, so, in, out);
        else hipLaunchKernelGGL(k_v2c_bp<2>, dim3(shape_blocks(si)), blk, 0, s, si, so, in, out);
    }
    VF_HIP(hipGetLastError());
}

void launch_langelaar(int nx, int ny, const double *in, double *out, double *smax, hipStream_t s) {
    if (nx <= LG_MAXX) {
        hipLaunchKernelGGL(k_langelaar_fused, dim3(1), dim3(LG_TPB), 0, s, nx, ny, in, out, smax);
        VF_HIP(hipGetLastError());
        return;
    }
    for (int y = 0; y < ny; ++y)
        hipLaunchKernelGGL(k_langelaar_layer, dim3(nblk(nx)), dim3(TPB), 0, s, nx, ny, y, in, out, smax);
    VF_HIP(hipGetLastError());
}

void launch_langelaar_backprop(int nx, int ny, const double *g, const double *vars, const double *filtered,
                               const double *smax, double *lam, double *out, hipStream_t s) {
    if (nx <= LG_MAXX) {
        // lam holds the factors h, out the factors o until the recurrence replaces them; the multipliers stay in LDS
        hipLaunchKernelGGL(k_langelaar_factors, dim3(nblk((size_t)nx * ny)), dim3(TPB), 0, s, nx, ny, vars, filtered,
                           smax, lam, out);
        hipLaunchKernelGGL(k_langelaar_bp_fused, dim3(1), dim3(LG_TPB), 0, s, nx, ny, g, filtered, lam, out);
        VF_HIP(hipGetLastError());
        return;
    }
    for (int y = ny - 1; y >= 0; --y)
        hipLaunchKernelGGL(k_langelaar_lambda, dim3(nblk(nx)), dim3(TPB), 0, s, nx, ny, y, g, vars, filtered, smax,
                           lam);
    hipLaunchKernelGGL(k_langelaar_out, dim3(nblk((size_t)nx * ny)), dim3(TPB), 0, s, nx, ny, lam, vars, smax, out);
    VF_HIP(hipGetLastError());
}

}  // namespace vf
