// vf_kernels_vec.hip -- PCG vector kernels with deterministic reductions, Dirichlet handling,
// Young's modulus interpolation, compliance sensitivity, self-weight loads, density filters and
// the optimality-criterion update, for gfx950.
//
// Reductions are two-pass and order-fixed (per-workgroup partials in a fixed tree, then one
// workgroup sums the partials in a fixed order), so every solve is bit-reproducible.
#include "vf_internal.hpp"

namespace vf {
namespace {

constexpr int TPB = 256;
constexpr int ITEMS = 8;  // elements per work-item in the active-region kernels
inline unsigned nblocks(long count) { return (unsigned)((count + TPB - 1) / TPB); }

// The active region of a level: for component c and x-slab ix, the contiguous node range
// [ix*sx, ix*sx + (ydet + margin)*nz). Launch: grid.y = 3 * nn0, grid.x covers one slab.
struct Active {
    int rowLen;  // (ydet + margin) * nz
    int sx;      // nn1 * nz
    int nn0;
    int N;
};
inline Active active_of(const Grid &g, int margin) {
    Active a;
    a.rowLen = min(g.ydet + margin, g.nn[1]) * g.nn[2];
    a.sx = g.nn[1] * g.nn[2];
    a.nn0 = g.nn[0];
    a.N = g.nnodes;
    return a;
}
inline dim3 active_grid(const Active &a) {
    return dim3((unsigned)((a.rowLen + TPB * ITEMS - 1) / (TPB * ITEMS)), (unsigned)(3 * a.nn0));
}
__device__ __forceinline__ size_t active_base(const Active &a) {
    const int c = blockIdx.y / a.nn0, ix = blockIdx.y - c * a.nn0;
    return (size_t)c * a.N + (size_t)ix * a.sx;
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < TPB / 64; ++i) r += sh[i];
    return r;
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fill_active(Active a, T *x, T v) {
    const size_t base = active_base(a);
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) x[base + i] = v;
    }
}

template <typename Td, typename Ts>
__global__ void __launch_bounds__(TPB) k_copy_active(Active a, const Ts *__restrict__ s, Td *__restrict__ d) {
    const size_t base = active_base(a);
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) d[base + i] = (Td)s[base + i];
    }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fill(size_t n, T *x, T v) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) x[i] = v;
}

template <typename Td, typename Ts>
__global__ void __launch_bounds__(TPB) k_convert(size_t n, const Ts *__restrict__ s, Td *__restrict__ d) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) d[i] = (Td)s[i];
}

__global__ void __launch_bounds__(TPB) k_dot_active(Active a, const double *__restrict__ x,
                                                    const double *__restrict__ y, double *__restrict__ partials) {
    __shared__ double sh[TPB / 64];
    const size_t base = active_base(a);
    double s = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) s += x[base + i] * y[base + i];
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(TPB) k_dot2_active(Active a, const double *__restrict__ x,
                                                     const double *__restrict__ y, const double *__restrict__ z,
                                                     double *__restrict__ partials, int stride) {
    __shared__ double sh[TPB / 64];
    const size_t base = active_base(a);
    double s = 0, t = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double yy = y[base + i];
            s += x[base + i] * yy;
            t += z[base + i] * yy;
        }
    }
    s = block_sum(s, sh);
    __syncthreads();
    t = block_sum(t, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
        partials[stride + blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(TPB) k_dot_full(size_t n, const double *__restrict__ x,
                                                  const double *__restrict__ y, double *__restrict__ partials) {
    __shared__ double sh[TPB / 64];
    double s = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const size_t i = ((size_t)blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < n) s += x[i] * (y ? y[i] : 1.0);
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Fixed-order sum of `count` partials (per set; set k starts at partials + k*stride): 1024 work-items
// each add a fixed strided subset, then a fixed tree. Writes out[set].
constexpr int FTPB = 1024;
__global__ void __launch_bounds__(FTPB) k_final_sum(const double *__restrict__ partials, int count, int stride,
                                                    int nsets, double *__restrict__ out) {
    __shared__ double sh[FTPB / 64];
    for (int set = 0; set < nsets; ++set) {
        const double *p = partials + (size_t)set * stride;
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        int i = threadIdx.x;
        for (; i + 3 * FTPB < count; i += 4 * FTPB) {
            s0 += p[i];
            s1 += p[i + FTPB];
            s2 += p[i + 2 * FTPB];
            s3 += p[i + 3 * FTPB];
        }
        for (; i < count; i += FTPB) s0 += p[i];
        double s = (s0 + s1) + (s2 + s3);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0) sh[w] = s;
        __syncthreads();
        if (threadIdx.x < 64) {
            double t = threadIdx.x < FTPB / 64 ? sh[threadIdx.x] : 0.0;
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            if (threadIdx.x == 0) out[set] = t;
        }
        __syncthreads();
    }
}

// ---- PCG (MultigridSolver.hh:1119-1143). Device scalars: sc[0]=<r,s>  sc[1]=<Ad_prev,s>
// sc[2]=previous <r,s>  sc[3]=<d,Ad>  sc[4]=<r,r>  sc[5]=alpha of the last update
__global__ void __launch_bounds__(TPB) k_pcg_direction(Active a, const double *__restrict__ s,
                                                       double *__restrict__ d, const double *__restrict__ sc,
                                                       int mode) {
    // mode 0: Fletcher-Reeves beta = rz / rz_old ; mode 1: Polak-Ribiere (flexible) beta =
    // <r_k - r_{k-1}, s_k>/rz_old = -alpha_prev <Ad_prev, s_k> / rz_old
    const double beta = mode == 0 ? sc[0] / sc[2] : -sc[5] * sc[1] / sc[2];
    const size_t base = active_base(a);
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) d[base + i] = s[base + i] + beta * d[base + i];
    }
}

__global__ void __launch_bounds__(TPB) k_pcg_update(Active a, double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ d, const double *__restrict__ Ad,
                                                    const double *__restrict__ sc, double *__restrict__ partials) {
    __shared__ double sh[TPB / 64];
    const double alpha = sc[0] / sc[3];
    const size_t base = active_base(a);
    double s = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            x[base + i] += alpha * d[base + i];
            const double rn = r[base + i] - alpha * Ad[base + i];
            r[base + i] = rn;
            s += rn * rn;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// Fused forms used when the iterate is not needed between iterations (no host callback): the
// preconditioned vector s is read in the precision the cycle produced it (TS), x += alpha d is deferred
// into the next direction update, and the residual update also writes the single-precision right-hand
// side of the next cycle. The arithmetic of each entry is the same as in the unfused kernels.
template <typename TS>
__global__ void __launch_bounds__(TPB) k_pcg_dots_s(Active a, const double *__restrict__ r, const TS *__restrict__ s,
                                                    const double *__restrict__ adp, double *__restrict__ partials,
                                                    int stride) {
    __shared__ double sh[TPB / 64];
    const size_t base = active_base(a);
    double rs = 0, as = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double sv = (double)s[base + i];
            rs += r[base + i] * sv;
            if (adp) as += adp[base + i] * sv;
        }
    }
    rs = block_sum(rs, sh);
    if (adp) {
        __syncthreads();
        as = block_sum(as, sh);
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.y * gridDim.x + blockIdx.x] = rs;
        if (adp) partials[stride + blockIdx.y * gridDim.x + blockIdx.x] = as;
    }
}

// first: d = s. Otherwise x += alpha_prev d, then d = s + beta d (beta as in k_pcg_direction).
template <typename TS>
__global__ void __launch_bounds__(TPB) k_pcg_direction_x(Active a, const TS *__restrict__ s, double *__restrict__ d,
                                                         double *__restrict__ x, const double *__restrict__ sc,
                                                         int mode, int first) {
    const size_t base = active_base(a);
    if (first) {
        for (int k = 0; k < ITEMS; ++k) {
            const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
            if (i < a.rowLen) d[base + i] = (double)s[base + i];
        }
        return;
    }
    const double alpha = sc[5];
    const double beta = mode == 0 ? sc[0] / sc[2] : -sc[5] * sc[1] / sc[2];
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double dd = d[base + i];
            x[base + i] += alpha * dd;
            d[base + i] = (double)s[base + i] + beta * dd;
        }
    }
}

// r -= alpha Ad with <r,r> partials; r32 (optional) receives float(r)
__global__ void __launch_bounds__(TPB) k_pcg_residual(Active a, double *__restrict__ r, const double *__restrict__ Ad,
                                                      const double *__restrict__ sc, double *__restrict__ partials,
                                                      float *__restrict__ r32) {
    __shared__ double sh[TPB / 64];
    const double alpha = sc[0] / sc[3];
    const size_t base = active_base(a);
    double s = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double rn = r[base + i] - alpha * Ad[base + i];
            r[base + i] = rn;
            if (r32) r32[base + i] = (float)rn;
            s += rn * rn;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// x += alpha d with alpha = sc[5] (the deferred update of the last iteration)
__global__ void __launch_bounds__(TPB) k_pcg_finish_x(Active a, double *__restrict__ x, const double *__restrict__ d,
                                                      const double *__restrict__ sc) {
    const double alpha = sc[5];
    const size_t base = active_base(a);
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) x[base + i] += alpha * d[base + i];
    }
}

__global__ void k_pcg_scalars(double *sc, int op) {
    if (threadIdx.x != 0) return;
    if (op == 0) sc[2] = sc[0];                 // previous <r,s> <- <r,s>
    else if (op == 1) sc[5] = sc[0] / sc[3];    // alpha of this iteration
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_zero_dirichlet(int count, const int *__restrict__ nodes,
                                                        const uint8_t *__restrict__ masks, int N, T *__restrict__ u) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= count) return;
    const int n = nodes[i];
    const uint8_t m = masks[i];
    if (m & 1) u[n] = T(0);
    if (m & 2) u[N + n] = T(0);
    if (m & 4) u[2 * N + n] = T(0);
}

__global__ void __launch_bounds__(TPB) k_enforce_dirichlet(int count, const int *__restrict__ nodes,
                                                           const uint8_t *__restrict__ masks, int N,
                                                           const double *__restrict__ vals, double *__restrict__ u) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= count) return;
    const int n = nodes[i];
    const uint8_t m = masks[i];
    for (int c = 0; c < 3; ++c)
        if (m & (1 << c)) u[c * N + n] = vals[3 * i + c];
}

// ---- physics
// E = Emin + rho^gamma (E0 - Emin) (SIMP) or Emin + rho (E0 - Emin)/(1 + q (1 - rho)) (RAMP);
// 0 on masked layers (TensorProductSimulator.hh:2055-2060, 2088-2102)
__global__ void __launch_bounds__(TPB) k_young(Grid g, const double *__restrict__ rho, double *__restrict__ E,
                                               float *__restrict__ E32, int law, double E0, double Emin,
                                               double gamma, double q, int maskActive) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ey = (e / g.ne[2]) % g.ne[1];
    const double r = rho[e];
    double v;
    if (maskActive && ey >= g.ymask) v = 0.0;
    else if (law == 0) v = Emin + pow(r, gamma) * (E0 - Emin);
    else v = Emin + r * (E0 - Emin) / (1 + q * (1 - r));
    E[e] = v;
    E32[e] = (float)v;
}

__global__ void __launch_bounds__(TPB) k_zero_young_layers(Grid g, int y0, int y1, double *E, float *E32) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ey = (e / g.ne[2]) % g.ne[1];
    if (ey >= y0 && ey < y1) {
        E[e] = 0.0;
        E32[e] = 0.0f;
    }
}

__device__ __forceinline__ void gather_elem(const Grid &g, const double *__restrict__ u, int ex, int ey, int ez,
                                            double ue[24]) {
    const int N = g.nnodes;
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const int n = ((ex + ((l >> 2) & 1)) * g.nn[1] + ey + ((l >> 1) & 1)) * g.nn[2] + ez + (l & 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * l + c] = u[c * N + n];
    }
}

__device__ __forceinline__ double quad_form(const double *__restrict__ K0, const double ue[24]) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < 24; ++a) {
        double r = 0;
#pragma unroll
        for (int b = 0; b < 24; ++b) r += K0[a * 24 + b] * ue[b];
        s += ue[a] * r;
    }
    return s;
}

// g_e = -1/2 dE/drho u_e^T K0 u_e (+ self-weight term) (TensorProductSimulator.hh:972-1040)
__global__ void __launch_bounds__(TPB) k_compliance_gradient(Grid g, const double *__restrict__ K0,
                                                             const double *__restrict__ rho,
                                                             const double *__restrict__ u, double *__restrict__ out,
                                                             int law, double E0, double Emin, double gamma, double q,
                                                             double g0, double g1, double g2, double vol,
                                                             int accumulate) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ez = e % g.ne[2], ey = (e / g.ne[2]) % g.ne[1], ex = e / (g.ne[1] * g.ne[2]);
    if (ey >= g.ymask) {
        if (!accumulate) out[e] = 0.0;
        return;
    }
    double ue[24];
    gather_elem(g, u, ex, ey, ez, ue);
    const double r = rho[e];
    const double dE = law == 0 ? gamma * pow(r, gamma - 1.0) * (E0 - Emin)
                               : (1 + q) * (E0 - Emin) / ((1 + q * (1 - r)) * (1 + q * (1 - r)));
    double val = -0.5 * dE * quad_form(K0, ue);
    if (g0 != 0 || g1 != 0 || g2 != 0) {
        for (int l = 0; l < 8; ++l) val += 0.125 * (g0 * ue[3 * l] + g1 * ue[3 * l + 1] + g2 * ue[3 * l + 2]) * vol;
    }
    if (accumulate) out[e] += val;
    else out[e] = val;
}

__global__ void __launch_bounds__(TPB) k_energy_density(Grid g, const double *__restrict__ K0,
                                                        const double *__restrict__ E, const double *__restrict__ u,
                                                        double *__restrict__ out) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ez = e % g.ne[2], ey = (e / g.ne[2]) % g.ne[1], ex = e / (g.ne[1] * g.ne[2]);
    double ue[24];
    gather_elem(g, u, ex, ey, ez, ue);
    out[e] = 0.5 * E[e] * quad_form(K0, ue);
}

// f_n += g * sum over incident unmasked elements of rho_e V / 8 (TensorProductSimulator.hh:1277-1286)
__global__ void __launch_bounds__(TPB) k_self_weight(Grid g, const double *__restrict__ rho, double g0, double g1,
                                                     double g2, double vol, double *__restrict__ f) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int nz = g.nn[2], ix = n / (g.nn[1] * nz), iy = (n / nz) % g.nn[1], iz = n % nz;
    double w = 0;
    for (int e = 0; e < 8; ++e) {
        const int ex = ix - ((e >> 2) & 1), ey = iy - ((e >> 1) & 1), ez = iz - (e & 1);
        if ((unsigned)ex >= (unsigned)g.ne[0] || (unsigned)ey >= (unsigned)g.ne[1] || (unsigned)ez >= (unsigned)g.ne[2])
            continue;
        if (ey >= g.ymask) continue;
        w += 0.125 * rho[(ex * g.ne[1] + ey) * g.ne[2] + ez] * vol;
    }
    f[n] += g0 * w;
    f[g.nnodes + n] += g1 * w;
    f[2 * g.nnodes + n] += g2 * w;
}

// ---- filters (TopologyOptimizationFilter.hh:189-400)
__device__ __forceinline__ int reflect_index(int i, int s) {
    while (i < 0 || i >= s) {
        if (i >= s) i = 2 * s - i - 1;
        if (i < 0) i = -i - 1;
    }
    return i;
}

__global__ void __launch_bounds__(TPB) k_smoothing(int d0, int d1, int d2, int R, double rmin, int type, double W,
                                                   const double *__restrict__ in, double *__restrict__ out) {
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= d0 * d1 * d2) return;
    const int k = t % d2, j = (t / d2) % d1, i = t / (d1 * d2);
    double s = 0;
    for (int a = -R; a <= R; ++a) {
        const int ii = reflect_index(i + a, d0);
        for (int b = -R; b <= R; ++b) {
            const int jj = reflect_index(j + b, d1);
            const size_t rowb = ((size_t)ii * d1 + jj) * d2;
            for (int c = -R; c <= R; ++c) {
                const double w = type == 1 ? rmin - sqrt((double)(a * a + b * b + c * c)) : 1.0;
                if (w <= 0) continue;
                s += w * in[rowb + reflect_index(k + c, d2)];
            }
        }
    }
    out[t] = s / W;
}

__global__ void __launch_bounds__(TPB) k_projection(double beta, double th, size_t n, const double *__restrict__ in,
                                                    double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = (th + tanh(beta * (in[i] - 0.5))) / (2 * th);
}

__global__ void __launch_bounds__(TPB) k_projection_bp(double beta, double scale, size_t n,
                                                       const double *__restrict__ g, const double *__restrict__ v,
                                                       double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double t = tanh(beta * (v[i] - 0.5));
    out[i] = g[i] * (1.0 - t * t) * scale;
}

// OCOptimizer ceval update (OptimalityCriterion.hh:66-72)
__global__ void __launch_bounds__(TPB) k_oc_update(size_t n, const double *__restrict__ x0,
                                                   const double *__restrict__ dJ, const double *__restrict__ dc,
                                                   double lambda, double m, double p, double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double x = x0[i];
    double r = x * pow(dJ[i] / (dc[i] * lambda), p);
    // The reference's clamps propagate NaN (std::max/std::min with the ratio as first operand, Eigen's
    // cwiseMax/cwiseMin likewise) and a NaN ratio then keeps x0; fmax/fmin would drop it, so test it first.
    // An infinite ratio is clamped to x0 +- m, as in the reference.
    if (isnan(r)) r = x;
    else r = fmin(fmax(fmin(fmax(r, x - m), x + m), 0.0), 1.0);
    out[i] = r;
}


// ---- layer-by-layer batches as one stacked problem. The partial structures of a batch lie side by side
// along x, segment k over the node planes [k P, k P + W] of the stacked grid (P = W + 2^L element columns,
// W the design's width); between segments, 2^L - 1 planes of void elements with constrained nodes. Segment
// states: 0 iterating, 1 converged at the last check (the deferred x update is still due), 2 done.
constexpr int SEG_ACTIVE = 0, SEG_FINISHING = 1;
__device__ __forceinline__ int seg_of(int ix, int P, int nseg) { return min(ix / P, nseg - 1); }

// Fixed-order sums of partials per segment: grid.x = segment. Rows layout (Active kernels): partial
// (c nn0 + ix) gx + bx for component c, node plane ix, block bx; a segment sums its planes' rows.
// Chunk layout (Walsh K u): partial chunk * tiles + tile, chunk c over planes [c xc, c xc + xc), xc | P.
// nsets sets, set j at partials + j setStride, written to sc[16 segment + off + j].
__global__ void __launch_bounds__(TPB) k_seg_sum(const double *__restrict__ partials, int layout, int a0, int a1,
                                                 int a2, int P, int nseg, int nsets, size_t setStride,
                                                 double *__restrict__ sc, int off) {
    __shared__ double sh[TPB / 64];
    const int k = blockIdx.x;
    for (int set = 0; set < nsets; ++set) {
        const double *p = partials + set * setStride;
        double v = 0;
        if (layout == 0) {  // rows: a0 = gx, a1 = nn0
            const int gx = a0, nn0 = a1;
            const int ix0 = k * P, ix1 = k == nseg - 1 ? nn0 : min((k + 1) * P, nn0);
            const int per = (ix1 - ix0) * gx, count = 3 * per;
            for (int i = threadIdx.x; i < count; i += TPB) {
                const int c = i / per, rem = i - c * per;
                v += p[(size_t)(c * nn0 + ix0) * gx + rem];
            }
        } else {  // chunks: a0 = tiles, a1 = xc, a2 = number of chunks
            const int tiles = a0, xc = a1, nch = a2;
            const int c0 = k * P / xc, c1 = k == nseg - 1 ? nch : min((k + 1) * P / xc, nch);
            const int count = (c1 - c0) * tiles;
            for (int i = threadIdx.x; i < count; i += TPB) v += p[(size_t)c0 * tiles + i];
        }
        v = block_sum(v, sh);
        if (threadIdx.x == 0) sc[16 * k + off + set] = v;
        __syncthreads();
    }
}

// op 0: previous <r,s> <- <r,s>; op 1: alpha = <r,s> / <d,Ad> (segments still iterating only)
__global__ void k_seg_scalars(double *sc, const int *__restrict__ state, int nseg, int op) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg || state[k] != SEG_ACTIVE) return;
    double *c = sc + 16 * k;
    if (op == 0) c[2] = c[0];
    else c[5] = c[0] / c[3];
}

// k_pcg_direction_x per segment; a segment that converged at the last check only takes its deferred
// x += alpha d
template <typename TS>
__global__ void __launch_bounds__(TPB) k_seg_direction_x(Active a, int P, int nseg, const int *__restrict__ state,
                                                         const TS *__restrict__ s, double *__restrict__ d,
                                                         double *__restrict__ x, const double *__restrict__ scAll,
                                                         int mode, int first) {
    const int ix = blockIdx.y % a.nn0, seg = seg_of(ix, P, nseg), st = state[seg];
    if (st > SEG_FINISHING) return;
    const double *sc = scAll + 16 * seg;
    const size_t base = active_base(a);
    if (st == SEG_FINISHING) {
        const double alpha = sc[5];
        for (int k = 0; k < ITEMS; ++k) {
            const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
            if (i < a.rowLen) x[base + i] += alpha * d[base + i];
        }
        return;
    }
    if (first) {
        for (int k = 0; k < ITEMS; ++k) {
            const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
            if (i < a.rowLen) d[base + i] = (double)s[base + i];
        }
        return;
    }
    const double alpha = sc[5];
    const double beta = mode == 0 ? sc[0] / sc[2] : -sc[5] * sc[1] / sc[2];
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double dd = d[base + i];
            x[base + i] += alpha * dd;
            d[base + i] = (double)s[base + i] + beta * dd;
        }
    }
}

// k_pcg_residual per segment (iterating segments only; the others leave r and write zero partials)
__global__ void __launch_bounds__(TPB) k_seg_residual(Active a, int P, int nseg, const int *__restrict__ state,
                                                      double *__restrict__ r, const double *__restrict__ Ad,
                                                      const double *__restrict__ scAll, double *__restrict__ partials,
                                                      float *__restrict__ r32) {
    __shared__ double sh[TPB / 64];
    const int ix = blockIdx.y % a.nn0, seg = seg_of(ix, P, nseg);
    if (state[seg] != SEG_ACTIVE) {  // uniform over the workgroup (one plane)
        if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = 0.0;
        return;
    }
    const double *sc = scAll + 16 * seg;
    const double alpha = sc[0] / sc[3];
    const size_t base = active_base(a);
    double s = 0;
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) {
            const double rn = r[base + i] - alpha * Ad[base + i];
            r[base + i] = rn;
            if (r32) r32[base + i] = (float)rn;
            s += rn * rn;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// the deferred x += alpha d of the segments that converged at the last check
__global__ void __launch_bounds__(TPB) k_seg_finish_x(Active a, int P, int nseg, const int *__restrict__ state,
                                                      double *__restrict__ x, const double *__restrict__ d,
                                                      const double *__restrict__ scAll) {
    const int ix = blockIdx.y % a.nn0, seg = seg_of(ix, P, nseg);
    if (state[seg] != SEG_FINISHING) return;
    const double alpha = scAll[16 * seg + 5];
    const size_t base = active_base(a);
    for (int k = 0; k < ITEMS; ++k) {
        const int i = (blockIdx.x * ITEMS + k) * TPB + threadIdx.x;
        if (i < a.rowLen) x[base + i] += alpha * d[base + i];
    }
}

// stacked element fields: segment k holds the design's elements (x - k P, y, z) below its first masked row
// top[k]; void (zero) elsewhere
template <typename T>
__global__ void __launch_bounds__(TPB) k_stack_elems(Grid gs, Grid go, int P, int nseg, const int *__restrict__ top,
                                                     const T *__restrict__ src, T *__restrict__ dst) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= gs.nelems) return;
    const int ez = e % gs.ne[2], ey = (e / gs.ne[2]) % gs.ne[1], ex = e / (gs.ne[1] * gs.ne[2]);
    const int k = min(ex / P, nseg - 1), lx = ex - k * P;
    T v = T(0);
    if (lx < go.ne[0] && ey < top[k]) v = src[((size_t)lx * go.ne[1] + ey) * go.ne[2] + ez];
    dst[e] = v;
}
// stacked node field (3 components): segment k holds the design's nodes (x - k P, y, z) up to its first
// detached row top[k] + 1; zero in the gaps and on the detached rows
__global__ void __launch_bounds__(TPB) k_stack_nodes(Grid gs, Grid go, int P, int nseg, const int *__restrict__ top,
                                                     const double *__restrict__ src, double *__restrict__ dst) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= gs.nnodes) return;
    const int nz = gs.nn[2], ix = n / (gs.nn[1] * nz), iy = (n / nz) % gs.nn[1], iz = n % nz;
    const int k = min(ix / P, nseg - 1), lx = ix - k * P;
    const bool in = lx < go.nn[0] && iy <= top[k];
    const size_t m = ((size_t)lx * go.nn[1] + iy) * go.nn[2] + iz;
    for (int c = 0; c < 3; ++c) dst[(size_t)c * gs.nnodes + n] = in ? src[(size_t)c * go.nnodes + m] : 0.0;
}
// per-node masks of one level of the stacked grid: every component of a node in the void columns (local x
// beyond the segment width Wl) or on a row at or above its segment's first detached row ydet[k] is held
// (mask 7); the design's own constraints are already in the mask
__global__ void __launch_bounds__(TPB) k_stack_mask(Grid g, int Pl, int Wl, int nseg, const int *__restrict__ ydet,
                                                    uint8_t *__restrict__ mask) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int nz = g.nn[2], ix = n / (g.nn[1] * nz), iy = (n / nz) % g.nn[1];
    const int k = min(ix / Pl, nseg - 1), lx = ix - k * Pl;
    if (lx > Wl || iy >= ydet[k]) mask[n] = 7;
}
// v = 0 on every node whose mask holds all three components
template <typename T>
__global__ void __launch_bounds__(TPB) k_zero_held(int N, const uint8_t *__restrict__ mask, T *__restrict__ v) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= N || mask[n] != 7) return;
    v[n] = T(0);
    v[N + n] = T(0);
    v[2 * (size_t)N + n] = T(0);
}
// g[e] += the stacked gradient of each segment in turn (segment order = layer order), for the element
// rows below the segment's first masked row (masked elements contribute nothing)
__global__ void __launch_bounds__(TPB) k_fold_gradient(Grid go, Grid gs, int P, int nseg, const int *__restrict__ top,
                                                       const double *__restrict__ gst, double *__restrict__ g) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= go.nelems) return;
    const int ez = e % go.ne[2], ey = (e / go.ne[2]) % go.ne[1], ex = e / (go.ne[1] * go.ne[2]);
    double acc = g[e];
    for (int k = 0; k < nseg; ++k)
        if (ey < top[k]) acc += gst[((size_t)(k * P + ex) * gs.ne[1] + ey) * gs.ne[2] + ez];
    g[e] = acc;
}

}  // namespace

// ------------------------------------------------------------------ launchers
template <typename T>
void launch_fill_active(const Grid &g, int margin, T *x, T value, hipStream_t s) {
    const Active a = active_of(g, margin);
    if (a.rowLen <= 0) return;
    hipLaunchKernelGGL(k_fill_active<T>, active_grid(a), dim3(TPB), 0, s, a, x, value);
    VF_HIP(hipGetLastError());
}
template <typename T>
void launch_fill(size_t n, T *x, T value, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks(n)), dim3(TPB), 0, s, n, x, value);
    VF_HIP(hipGetLastError());
}
template <typename Td, typename Ts>
void launch_convert(size_t n, const Ts *src, Td *dst, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL((k_convert<Td, Ts>), dim3(nblocks(n)), dim3(TPB), 0, s, n, src, dst);
    VF_HIP(hipGetLastError());
}
template <typename Td, typename Ts>
void launch_copy_active(const Grid &g, int margin, const Ts *src, Td *dst, hipStream_t s) {
    const Active a = active_of(g, margin);
    if (a.rowLen <= 0) return;
    hipLaunchKernelGGL((k_copy_active<Td, Ts>), active_grid(a), dim3(TPB), 0, s, a, src, dst);
    VF_HIP(hipGetLastError());
}

int reduction_partials(const Grid &g) {
    const Active a = active_of(g, 0);
    const dim3 gr = active_grid(a);
    return (int)(gr.x * gr.y);
}
int reduction_partials_full(size_t n) { return (int)((n + TPB * ITEMS - 1) / (TPB * ITEMS)); }

void launch_dot_active(const Grid &g, const double *a, const double *b, double *partials, double *out, hipStream_t s) {
    const Active ac = active_of(g, 0);
    if (ac.rowLen <= 0) { launch_fill<double>(1, out, 0.0, s); return; }
    const dim3 gr = active_grid(ac);
    hipLaunchKernelGGL(k_dot_active, gr, dim3(TPB), 0, s, ac, a, b, partials);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, s, partials, (int)(gr.x * gr.y), 0, 1, out);
    VF_HIP(hipGetLastError());
}

void launch_dot2_active(const Grid &g, const double *a, const double *b, const double *c, double *partials,
                        double *out2, hipStream_t s) {
    const Active ac = active_of(g, 0);
    if (ac.rowLen <= 0) { launch_fill<double>(2, out2, 0.0, s); return; }
    const dim3 gr = active_grid(ac);
    const int cnt = (int)(gr.x * gr.y);
    hipLaunchKernelGGL(k_dot2_active, gr, dim3(TPB), 0, s, ac, a, b, c, partials, cnt);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, s, partials, cnt, cnt, 2, out2);
    VF_HIP(hipGetLastError());
}

void launch_final_sum(const double *partials, int count, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, s, partials, count, 0, 1, out);
    VF_HIP(hipGetLastError());
}

// out[0] = in[0] / d (IEEE division on the device, the same rounding as the host's).
__global__ void k_div_scalar(const double *__restrict__ in, double d, double *__restrict__ out) {
    if (threadIdx.x == 0) out[0] = in[0] / d;
}
void launch_div_scalar(const double *in, double d, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_div_scalar, dim3(1), dim3(64), 0, s, in, d, out);
    VF_HIP(hipGetLastError());
}

void launch_dot_full(size_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s) {
    if (!n) { launch_fill<double>(1, out, 0.0, s); return; }
    const int cnt = reduction_partials_full(n);
    hipLaunchKernelGGL(k_dot_full, dim3(cnt), dim3(TPB), 0, s, n, a, b, partials);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, s, partials, cnt, 0, 1, out);
    VF_HIP(hipGetLastError());
}

void launch_pcg_direction(const Grid &g, const double *s, double *d, const double *sc, int mode, hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) return;
    hipLaunchKernelGGL(k_pcg_direction, active_grid(a), dim3(TPB), 0, st, a, s, d, sc, mode);
    VF_HIP(hipGetLastError());
}

template <typename TS>
void launch_pcg_dots_s(const Grid &g, const double *r, const TS *s, const double *adp, double *partials,
                       double *out, hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) { launch_fill<double>(adp ? 2 : 1, out, 0.0, st); return; }
    const dim3 gr = active_grid(a);
    const int cnt = (int)(gr.x * gr.y);
    hipLaunchKernelGGL(k_pcg_dots_s<TS>, gr, dim3(TPB), 0, st, a, r, s, adp, partials, cnt);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, st, partials, cnt, cnt, adp ? 2 : 1, out);
    VF_HIP(hipGetLastError());
}
template <typename TS>
void launch_pcg_direction_x(const Grid &g, const TS *s, double *d, double *x, const double *sc, int mode, int first,
                            hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) return;
    hipLaunchKernelGGL(k_pcg_direction_x<TS>, active_grid(a), dim3(TPB), 0, st, a, s, d, x, sc, mode, first);
    VF_HIP(hipGetLastError());
}
void launch_pcg_residual(const Grid &g, double *r, const double *Ad, const double *sc, double *partials,
                         double *rr_out, float *r32, hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) { launch_fill<double>(1, rr_out, 0.0, st); return; }
    const dim3 gr = active_grid(a);
    hipLaunchKernelGGL(k_pcg_residual, gr, dim3(TPB), 0, st, a, r, Ad, sc, partials, r32);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, st, partials, (int)(gr.x * gr.y), 0, 1, rr_out);
    VF_HIP(hipGetLastError());
}
void launch_pcg_finish_x(const Grid &g, double *x, const double *d, const double *sc, hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) return;
    hipLaunchKernelGGL(k_pcg_finish_x, active_grid(a), dim3(TPB), 0, st, a, x, d, sc);
    VF_HIP(hipGetLastError());
}
template void launch_pcg_dots_s<double>(const Grid &, const double *, const double *, const double *, double *,
                                        double *, hipStream_t);
template void launch_pcg_dots_s<float>(const Grid &, const double *, const float *, const double *, double *,
                                       double *, hipStream_t);
template void launch_pcg_direction_x<double>(const Grid &, const double *, double *, double *, const double *, int,
                                             int, hipStream_t);
template void launch_pcg_direction_x<float>(const Grid &, const float *, double *, double *, const double *, int, int,
                                            hipStream_t);

void launch_pcg_update(const Grid &g, double *x, double *r, const double *d, const double *Ad, const double *sc,
                       double *partials, double *rr_out, hipStream_t st) {
    const Active a = active_of(g, 0);
    if (a.rowLen <= 0) { launch_fill<double>(1, rr_out, 0.0, st); return; }
    const dim3 gr = active_grid(a);
    hipLaunchKernelGGL(k_pcg_update, gr, dim3(TPB), 0, st, a, x, r, d, Ad, sc, partials);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(FTPB), 0, st, partials, (int)(gr.x * gr.y), 0, 1, rr_out);
    VF_HIP(hipGetLastError());
}

void launch_pcg_scalars(double *sc, int op, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(64), 0, st, sc, op);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_zero_dirichlet(int count, const int *nodes, const uint8_t *masks, int nnodes, T *u, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_zero_dirichlet<T>, dim3(nblocks(count)), dim3(TPB), 0, s, count, nodes, masks, nnodes, u);
    VF_HIP(hipGetLastError());
}

void launch_enforce_dirichlet(int count, const int *nodes, const uint8_t *masks, int nnodes, const double *vals,
                              double *u, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_enforce_dirichlet, dim3(nblocks(count)), dim3(TPB), 0, s, count, nodes, masks, nnodes, vals, u);
    VF_HIP(hipGetLastError());
}

void launch_young(const Grid &g, const double *rho, double *E, float *E32, int law, double E0, double Emin,
                  double gamma, double q, int maskActive, hipStream_t s) {
    hipLaunchKernelGGL(k_young, dim3(nblocks(g.nelems)), dim3(TPB), 0, s, g, rho, E, E32, law, E0, Emin, gamma, q,
                       maskActive);
    VF_HIP(hipGetLastError());
}

void launch_zero_young_layers(const Grid &g, int y0, int y1, double *E, float *E32, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_young_layers, dim3(nblocks(g.nelems)), dim3(TPB), 0, s, g, y0, y1, E, E32);
    VF_HIP(hipGetLastError());
}

void launch_compliance_gradient(const Grid &g, const double *K0, const double *rho, const double *u, double *out,
                                int law, double E0, double Emin, double gamma, double q, const double *grav, double vol,
                                int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_compliance_gradient, dim3(nblocks(g.nelems)), dim3(TPB), 0, s, g, K0, rho, u, out, law, E0,
                       Emin, gamma, q, grav[0], grav[1], grav[2], vol, accumulate);
    VF_HIP(hipGetLastError());
}

void launch_energy_density(const Grid &g, const double *K0, const double *E, const double *u, double *out,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_energy_density, dim3(nblocks(g.nelems)), dim3(TPB), 0, s, g, K0, E, u, out);
    VF_HIP(hipGetLastError());
}

void launch_self_weight(const Grid &g, const double *rho, const double *grav, double vol, double *f, hipStream_t s) {
    hipLaunchKernelGGL(k_self_weight, dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, rho, grav[0], grav[1], grav[2], vol, f);
    VF_HIP(hipGetLastError());
}

namespace {
// ---- constantStrainLoad (TensorProductSimulator.hh:1119-1145 with the element load of 84-113): element e
// contributes rho_e * V * int_ref eps(phi_l e_c) : sigma = rho_e * sum_d sigma[c][d] s_d(l) V / (4 h_d) to
// component c of its local node l (s_d(l) = +-1 the side of l along axis d; the derivative of the Q1 shape
// function integrates to s_d / (4 h_d) times the cell volume). Nodes sum their elements in increasing
// element index order (local node 7 of the lowest element first), as the reference's serial accumulation.
__global__ void __launch_bounds__(TPB) k_constant_strain_load(Grid g, const double *__restrict__ rho, StressCoef sc,
                                                              double *__restrict__ f) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int nz = g.nn[2], ix = n / (g.nn[1] * nz), iy = (n / nz) % g.nn[1], iz = n % nz;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int l = 7; l >= 0; --l) {
        const int i = (l >> 2) & 1, j = (l >> 1) & 1, k = l & 1;
        const int ex = ix - i, ey = iy - j, ez = iz - k;
        if ((unsigned)ex >= (unsigned)g.ne[0] || (unsigned)ey >= (unsigned)g.ne[1] || (unsigned)ez >= (unsigned)g.ne[2])
            continue;
        const double r = rho[(ex * g.ne[1] + ey) * g.ne[2] + ez];
        const double sx = i ? 1.0 : -1.0, sy = j ? 1.0 : -1.0, sz = k ? 1.0 : -1.0;
        a0 += r * (sc.w[0][0] * sx + sc.w[0][1] * sy + sc.w[0][2] * sz);
        a1 += r * (sc.w[1][0] * sx + sc.w[1][1] * sy + sc.w[1][2] * sz);
        a2 += r * (sc.w[2][0] * sx + sc.w[2][1] * sy + sc.w[2][2] * sz);
    }
    f[n] = a0;
    f[g.nnodes + n] = a1;
    f[2 * g.nnodes + n] = a2;
}

// ---- sampleNodalField (TensorProductSimulator.hh:1161-1179): the query point is clamped to the domain
// and located in its element (getElementNDIndex, 1731-1752: a coordinate within 1e-10 of the far face
// maps to the last element); the element's trilinear interpolant is then evaluated, as in the reference,
// at the clamped point's coordinates q themselves (Element::interpolate receives q, not the reference
// coordinates of q in the element): value = sum_l u[node(e, l)] * prod_d (bit_d(l) ? q_d : 1 - q_d).
__global__ void __launch_bounds__(TPB) k_sample_nodal(Grid g, double lo0, double lo1, double lo2, double hi0,
                                                      double hi1, double hi2, double h0, double h1, double h2,
                                                      const double *__restrict__ u, int ncomp,
                                                      const double *__restrict__ p, int np, double *__restrict__ out,
                                                      int *__restrict__ bad) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= np) return;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2}, h[3] = {h0, h1, h2};
    double q[3];
    int e[3];
    for (int d = 0; d < 3; ++d) {
        q[d] = fmin(fmax(p[3 * i + d], lo[d]), hi[d]);
        const double fi = (q[d] - lo[d]) / h[d];
        if (fabs(fi - g.ne[d]) < 1e-10) e[d] = g.ne[d] - 1;
        else e[d] = (int)fi;
        if (e[d] < 0 || e[d] >= g.ne[d]) { atomicOr(bad, 1); e[d] = 0; }
    }
    double w[8];
    for (int l = 0; l < 8; ++l) {
        const int a = (l >> 2) & 1, b = (l >> 1) & 1, c = l & 1;
        w[l] = ((c ? q[2] : 1.0 - q[2]) * (b ? q[1] : 1.0 - q[1])) * (a ? q[0] : 1.0 - q[0]);
    }
    for (int k = 0; k < ncomp; ++k) {
        double v = 0;
        for (int l = 0; l < 8; ++l) {
            const int nx = e[0] + ((l >> 2) & 1), ny = e[1] + ((l >> 1) & 1), nzz = e[2] + (l & 1);
            const double un = u[(size_t)k * g.nnodes + (nx * g.nn[1] + ny) * g.nn[2] + nzz];
            v = l == 0 ? un * w[0] : v + un * w[l];
        }
        out[(size_t)k * np + i] = v;
    }
}
}  // namespace

void launch_constant_strain_load(const Grid &g, const double *rho, const StressCoef &sc, double *f, hipStream_t s) {
    hipLaunchKernelGGL(k_constant_strain_load, dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, rho, sc, f);
    VF_HIP(hipGetLastError());
}

void launch_sample_nodal(const Grid &g, const double lo[3], const double hi[3], const double h[3], const double *u,
                         int ncomp, const double *p, int np, double *out, int *bad, hipStream_t s) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_sample_nodal, dim3(nblocks(np)), dim3(TPB), 0, s, g, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2],
                       h[0], h[1], h[2], u, ncomp, p, np, out, bad);
    VF_HIP(hipGetLastError());
}

namespace {
__global__ void __launch_bounds__(TPB) k_axpy_plain(size_t n, double a, const double *__restrict__ x,
                                                    double *__restrict__ y) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
}  // namespace

void launch_axpy_plain(size_t n, double a, const double *x, double *y, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_axpy_plain, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, n, a, x, y);
    VF_HIP(hipGetLastError());
}

void launch_smoothing_filter(const int dims[3], double rmin, int type, const double *in, double *out, hipStream_t s) {
    const int R = (int)ceil(rmin) - 1;
    double W = 0;
    for (int a = -R; a <= R; ++a)
        for (int b = -R; b <= R; ++b)
            for (int c = -R; c <= R; ++c) {
                const double w = type == 1 ? rmin - sqrt((double)(a * a + b * b + c * c)) : 1.0;
                if (w > 0) W += w;
            }
    const long n = (long)dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(k_smoothing, dim3(nblocks(n)), dim3(TPB), 0, s, dims[0], dims[1], dims[2], R, rmin, type, W, in,
                       out);
    VF_HIP(hipGetLastError());
}

void launch_projection(double beta, size_t n, const double *in, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_projection, dim3(nblocks(n)), dim3(TPB), 0, s, beta, tanh(0.5 * beta), n, in, out);
    VF_HIP(hipGetLastError());
}

void launch_projection_backprop(double beta, size_t n, const double *g, const double *vars, double *out,
                                hipStream_t s) {
    const double scale = 1.0 / (2 * tanh(0.5 * beta) / beta);
    hipLaunchKernelGGL(k_projection_bp, dim3(nblocks(n)), dim3(TPB), 0, s, beta, scale, n, g, vars, out);
    VF_HIP(hipGetLastError());
}

void launch_oc_update(size_t n, const double *x0, const double *dJ, const double *dc, double lambda, double move,
                      double p, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_oc_update, dim3(nblocks(n)), dim3(TPB), 0, s, n, x0, dJ, dc, lambda, move, p, out);
    VF_HIP(hipGetLastError());
}

template void launch_fill_active<double>(const Grid &, int, double *, double, hipStream_t);
template void launch_fill_active<float>(const Grid &, int, float *, float, hipStream_t);
template void launch_fill<double>(size_t, double *, double, hipStream_t);
template void launch_fill<float>(size_t, float *, float, hipStream_t);
template void launch_fill<int>(size_t, int *, int, hipStream_t);
template void launch_convert<float, double>(size_t, const double *, float *, hipStream_t);
template void launch_convert<double, float>(size_t, const float *, double *, hipStream_t);
template void launch_copy_active<double, double>(const Grid &, int, const double *, double *, hipStream_t);
template void launch_copy_active<float, double>(const Grid &, int, const double *, float *, hipStream_t);
template void launch_copy_active<double, float>(const Grid &, int, const float *, double *, hipStream_t);
template void launch_zero_dirichlet<double>(int, const int *, const uint8_t *, int, double *, hipStream_t);
template void launch_zero_dirichlet<float>(int, const int *, const uint8_t *, int, float *, hipStream_t);

// ---- stacked layer-by-layer batches
void launch_seg_sum_rows(const Grid &g, const double *partials, int nsets, size_t setStride, int P, int nseg,
                         double *sc, int off, hipStream_t s) {
    const Active a = active_of(g, 0);
    const dim3 gr = active_grid(a);
    hipLaunchKernelGGL(k_seg_sum, dim3(nseg), dim3(TPB), 0, s, partials, 0, (int)gr.x, a.nn0, 0, P, nseg, nsets,
                       setStride, sc, off);
    VF_HIP(hipGetLastError());
}
void launch_seg_sum_chunks(const double *partials, int tiles, int xchunk, int nchunks, int P, int nseg, double *sc,
                           int off, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_sum, dim3(nseg), dim3(TPB), 0, s, partials, 1, tiles, xchunk, nchunks, P, nseg, 1,
                       (size_t)0, sc, off);
    VF_HIP(hipGetLastError());
}
int active_partials(const Grid &g) {
    const Active a = active_of(g, 0);
    const dim3 gr = active_grid(a);
    return (int)(gr.x * gr.y);
}
void launch_seg_dot(const Grid &g, const double *a, const double *b, double *partials, int P, int nseg, double *sc,
                    int off, hipStream_t s) {
    const Active ac = active_of(g, 0);
    const dim3 gr = active_grid(ac);
    hipLaunchKernelGGL(k_dot_active, gr, dim3(TPB), 0, s, ac, a, b, partials);
    launch_seg_sum_rows(g, partials, 1, 0, P, nseg, sc, off, s);
}
template <typename TS>
void launch_seg_dots_s(const Grid &g, const double *r, const TS *sv, const double *adp, double *partials, int P,
                       int nseg, double *sc, hipStream_t s) {
    const Active a = active_of(g, 0);
    const dim3 gr = active_grid(a);
    const int cnt = (int)(gr.x * gr.y);
    hipLaunchKernelGGL(k_pcg_dots_s<TS>, gr, dim3(TPB), 0, s, a, r, sv, adp, partials, cnt);
    launch_seg_sum_rows(g, partials, adp ? 2 : 1, (size_t)cnt, P, nseg, sc, 0, s);
}
void launch_seg_scalars(double *sc, const int *state, int nseg, int op, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_scalars, dim3((nseg + 63) / 64), dim3(64), 0, s, sc, state, nseg, op);
    VF_HIP(hipGetLastError());
}
template <typename TS>
void launch_seg_direction_x(const Grid &g, const TS *sv, double *d, double *x, const double *sc, const int *state,
                            int P, int nseg, int mode, int first, hipStream_t s) {
    const Active a = active_of(g, 0);
    hipLaunchKernelGGL(k_seg_direction_x<TS>, active_grid(a), dim3(TPB), 0, s, a, P, nseg, state, sv, d, x, sc, mode,
                       first);
    VF_HIP(hipGetLastError());
}
void launch_seg_residual(const Grid &g, double *r, const double *Ad, double *sc, const int *state, int P, int nseg,
                         double *partials, float *r32, hipStream_t s) {
    const Active a = active_of(g, 0);
    hipLaunchKernelGGL(k_seg_residual, active_grid(a), dim3(TPB), 0, s, a, P, nseg, state, r, Ad, sc, partials, r32);
    launch_seg_sum_rows(g, partials, 1, 0, P, nseg, sc, 4, s);
}
void launch_seg_finish_x(const Grid &g, double *x, const double *d, const double *sc, const int *state, int P,
                         int nseg, hipStream_t s) {
    const Active a = active_of(g, 0);
    hipLaunchKernelGGL(k_seg_finish_x, active_grid(a), dim3(TPB), 0, s, a, P, nseg, state, x, d, sc);
    VF_HIP(hipGetLastError());
}
template <typename T>
void launch_stack_elems(const Grid &gs, const Grid &go, int P, int nseg, const int *top, const T *src, T *dst,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_stack_elems<T>, dim3(nblocks(gs.nelems)), dim3(TPB), 0, s, gs, go, P, nseg, top, src, dst);
    VF_HIP(hipGetLastError());
}
void launch_stack_nodes(const Grid &gs, const Grid &go, int P, int nseg, const int *top, const double *src,
                        double *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_stack_nodes, dim3(nblocks(gs.nnodes)), dim3(TPB), 0, s, gs, go, P, nseg, top, src, dst);
    VF_HIP(hipGetLastError());
}
void launch_fold_gradient(const Grid &go, const Grid &gs, int P, int nseg, const int *top, const double *gst,
                          double *g, hipStream_t s) {
    hipLaunchKernelGGL(k_fold_gradient, dim3(nblocks(go.nelems)), dim3(TPB), 0, s, go, gs, P, nseg, top, gst, g);
    VF_HIP(hipGetLastError());
}
void launch_stack_mask(const Grid &g, int Pl, int Wl, int nseg, const int *ydet, uint8_t *mask, hipStream_t s) {
    hipLaunchKernelGGL(k_stack_mask, dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, Pl, Wl, nseg, ydet, mask);
    VF_HIP(hipGetLastError());
}
template <typename T>
void launch_zero_held(const Grid &g, const uint8_t *mask, T *v, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_held<T>, dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g.nnodes, mask, v);
    VF_HIP(hipGetLastError());
}
template void launch_zero_held<double>(const Grid &, const uint8_t *, double *, hipStream_t);
template void launch_zero_held<float>(const Grid &, const uint8_t *, float *, hipStream_t);
template void launch_seg_dots_s<double>(const Grid &, const double *, const double *, const double *, double *, int,
                                        int, double *, hipStream_t);
template void launch_seg_dots_s<float>(const Grid &, const double *, const float *, const double *, double *, int,
                                       int, double *, hipStream_t);
template void launch_seg_direction_x<double>(const Grid &, const double *, double *, double *, const double *,
                                             const int *, int, int, int, int, hipStream_t);
template void launch_seg_direction_x<float>(const Grid &, const float *, double *, double *, const double *,
                                            const int *, int, int, int, int, hipStream_t);
template void launch_stack_elems<double>(const Grid &, const Grid &, int, int, const int *, const double *, double *,
                                         hipStream_t);
template void launch_stack_elems<float>(const Grid &, const Grid &, int, int, const int *, const float *, float *,
                                        hipStream_t);

}  // namespace vf
