// vf_planar.hip -- planar (2D, degrees [1,1]) simulator and multigrid solver on the device (FP64).
//
// The reference instantiates its whole stack for N = 2 (python_bindings/VoxelFEM.cc:301-308): bilinear
// quads, 2^2 = 4 colours, 3^2 = 9 neighbour blocks per node row. Planar grids are small (the reference's 2D
// demos use 384 x 192 and 256 x 128 elements, 0.15 M DoF), so the kernels here are one work-item per node
// or per element entry, FP64 throughout: a planar V-cycle is bound by launch latency, not by HBM or the
// vector units, and the full fine operator (E_e K0 with 64 doubles of K0) stays in L2.
//
// Levels >= 1 hold element matrices (the Galerkin products of the level below, MultigridSolver.hh:704-819),
// which the operator, residual and smoother gather per node: the same operators as the reference's
// on-the-fly level 1 and assembled block rows of levels >= 2, up to the order of the sums. The coarsest
// level is factored by the band Cholesky shared with the 3D path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "vf_planar.hpp"

namespace vf {
namespace planar {
namespace {

constexpr int TPB = 256;
constexpr int64_t LAYER_NONE = INT64_MAX;
constexpr int NPART = 512;  // partial sums of a dot product

inline unsigned nblk(size_t n) { return (unsigned)std::max<size_t>(1, (n + TPB - 1) / TPB); }
void sync(hipStream_t s) { VF_HIP(hipStreamSynchronize(s)); }
template <typename T>
void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
    if (!n) return;
    VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    VF_HIP(hipStreamSynchronize(s));
}
template <typename T>
void d2h(T *dst, const T *src, size_t n, hipStream_t s) {
    if (!n) return;
    VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
}
template <typename T>
void d2d(T *dst, const T *src, size_t n, hipStream_t s) {
    if (!n || dst == src) return;
    VF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

// ---- element stiffness (TensorProductSimulator.hh:67-80): 2-point tensor Gauss rule (exact for Q1),
// isotropic tensor lambda = E nu / ((1 + nu)(1 - nu)), mu = E / (2 (1 + nu)) (the (N - 1) nu form of the
// isotropic tensor at N = 2: plane stress)
void element_K0(double E, double nu, const double h[2], double K0[64]) {
    const double lam = E * nu / ((1 + nu) * (1 - nu)), mu = E / (2 * (1 + nu));
    const double gp[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
    const double w = 0.25 * h[0] * h[1];
    std::fill(K0, K0 + 64, 0.0);
    for (int qx = 0; qx < 2; ++qx)
        for (int qy = 0; qy < 2; ++qy) {
            double g[4][2];
            for (int l = 0; l < 4; ++l) {
                const int i = (l >> 1) & 1, j = l & 1;
                const double vx = i ? gp[qx] : 1 - gp[qx], vy = j ? gp[qy] : 1 - gp[qy];
                const double dx = (i ? 1.0 : -1.0) / h[0], dy = (j ? 1.0 : -1.0) / h[1];
                g[l][0] = dx * vy;
                g[l][1] = vx * dy;
            }
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 2; ++c)
                    for (int b = 0; b < 4; ++b)
                        for (int d = 0; d < 2; ++d) {
                            double v = lam * g[a][c] * g[b][d] + mu * g[a][d] * g[b][c];
                            if (c == d) v += mu * (g[a][0] * g[b][0] + g[a][1] * g[b][1]);
                            K0[(2 * a + c) * 8 + 2 * b + d] += w * v;
                        }
        }
}

// ---- kernels -------------------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_young2(PGrid g, const double *__restrict__ rho, double *__restrict__ E, int law,
                                                double E0, double Emin, double gamma, double q, int maskActive) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const double r = rho[e];
    double v;
    if (maskActive && e % g.ne[1] >= g.ymask) v = 0.0;
    else if (law == 0) v = Emin + pow(r, gamma) * (E0 - Emin);
    else v = Emin + r * (E0 - Emin) / (1 + q * (1 - r));
    E[e] = v;
}

__global__ void __launch_bounds__(TPB) k_zero_young_rows2(PGrid g, int y0, int y1, double *__restrict__ E) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ey = e % g.ne[1];
    if (ey >= y0 && ey < y1) E[e] = 0.0;
}

// row 2 ln + c of the element matrix of element e times the element's displacements, for the node (ix, iy)
// at local corner ln of e
template <bool FINE>
__device__ __forceinline__ void elem_rows(const PGrid &g, const double *__restrict__ K0, const double *__restrict__ E,
                                          const double *__restrict__ Ke, const double *__restrict__ u, int ex, int ey,
                                          int ln, double out[2], double M[4]) {
    const int e = ex * g.ne[1] + ey;
    const double *K = FINE ? K0 : Ke + (size_t)e * 64;
    const double s = FINE ? E[e] : 1.0;
    double ue[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int nm = (ex + ((m >> 1) & 1)) * g.nn[1] + ey + (m & 1);
        ue[2 * m] = u[nm];
        ue[2 * m + 1] = u[g.nnodes + nm];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double t = 0;
#pragma unroll
        // K is symmetric: the fine level reads column 2 ln + c of K0 (cache-resident, the oracle's order of the
        // products); coarse levels read row 2 ln + c, whose 8 doubles are contiguous (rows 2 ln and 2 ln + 1 fill
        // two 64-byte lines), instead of one line per entry of a column
        for (int a = 0; a < 8; ++a) t += (FINE ? K[a * 8 + 2 * ln + c] : K[(2 * ln + c) * 8 + a]) * ue[a];
        out[c] += s * t;
        if (M) {
            M[2 * c] += s * K[(2 * ln + c) * 8 + 2 * ln];
            M[2 * c + 1] += s * K[(2 * ln + c) * 8 + 2 * ln + 1];
        }
    }
}

template <bool FINE>
__device__ __forceinline__ void node_Ku(const PGrid &g, const double *K0, const double *E, const double *Ke,
                                        const double *u, int ix, int iy, double f[2], double *M) {
#pragma unroll
    for (int i = 1; i >= 0; --i)
#pragma unroll
        for (int j = 1; j >= 0; --j) {
            const int ex = ix - i, ey = iy - j;
            if (ex < 0 || ey < 0 || ex >= g.ne[0] || ey >= g.ne[1]) continue;
            elem_rows<FINE>(g, K0, E, Ke, u, ex, ey, 2 * i + j, f, M);
        }
}

// y (=, +=, -=) K u over the rows of non-detached nodes; detached rows are zeroed when zeroInit
template <bool FINE>
__global__ void __launch_bounds__(TPB) k_apply2(PGrid g, const double *__restrict__ K0, const double *__restrict__ E,
                                                const double *__restrict__ Ke, const double *__restrict__ u,
                                                double *__restrict__ y, int zeroInit, int negate) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    if (iy >= g.ydet) {
        if (zeroInit) y[n] = y[g.nnodes + n] = 0.0;
        return;
    }
    double f[2] = {0, 0};
    node_Ku<FINE>(g, K0, E, Ke, u, ix, iy, f, nullptr);
    const double b0 = zeroInit ? 0.0 : y[n], b1 = zeroInit ? 0.0 : y[g.nnodes + n];
    y[n] = negate ? b0 - f[0] : b0 + f[0];
    y[g.nnodes + n] = negate ? b1 - f[1] : b1 + f[1];
}

// computeResidual (MultigridSolver.hh:527-541): r = b - K u on non-detached rows, r = b on the next four
// rows, Dirichlet components zeroed on every node
template <bool FINE>
__global__ void __launch_bounds__(TPB) k_residual2(PGrid g, const double *__restrict__ K0, const double *__restrict__ E,
                                                   const double *__restrict__ Ke, const uint8_t *__restrict__ dmask,
                                                   const double *__restrict__ u, const double *__restrict__ b,
                                                   double *__restrict__ r) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    const uint8_t dm = dmask[n];
    if (iy < g.ydet) {
        double f[2] = {0, 0};
        node_Ku<FINE>(g, K0, E, Ke, u, ix, iy, f, nullptr);
        r[n] = (dm & 1) ? 0.0 : b[n] - f[0];
        r[g.nnodes + n] = (dm & 2) ? 0.0 : b[g.nnodes + n] - f[1];
    } else if (iy < g.ydet + 4) {
        r[n] = (dm & 1) ? 0.0 : b[n];
        r[g.nnodes + n] = (dm & 2) ? 0.0 : b[g.nnodes + n];
    } else {
        if (dm & 1) r[n] = 0.0;
        if (dm & 2) r[g.nnodes + n] = 0.0;
    }
}

// computeResidual on a coarse level with four work-items per node (one per incident element, summed in
// node_Ku's order by the first of the four, as in k_smooth2_x4): the same values as k_residual2<false>
__global__ void __launch_bounds__(TPB) k_residual2_x4(PGrid g, const double *__restrict__ Ke,
                                                     const uint8_t *__restrict__ dmask, const double *__restrict__ u,
                                                     const double *__restrict__ b, double *__restrict__ r) {
    const int tt = blockIdx.x * TPB + threadIdx.x, n = tt >> 2, p = tt & 3;
    if (n >= g.nnodes) return;  // all four work-items of a node
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    double v[2] = {0, 0};
    if (iy < g.ydet) {
        const int i = 1 - (p >> 1), j = 1 - (p & 1);
        const int ex = ix - i, ey = iy - j;
        if (ex >= 0 && ey >= 0 && ex < g.ne[0] && ey < g.ne[1])
            elem_rows<false>(g, nullptr, nullptr, Ke, u, ex, ey, 2 * i + j, v, nullptr);
    }
    const int q0 = (int)(threadIdx.x & 63) & ~3;
    double f[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double a = __shfl(v[k], q0, 64);
        a += __shfl(v[k], q0 + 1, 64);
        a += __shfl(v[k], q0 + 2, 64);
        a += __shfl(v[k], q0 + 3, 64);
        f[k] = a;
    }
    if (p) return;
    const uint8_t dm = dmask[n];
    if (iy < g.ydet) {
        r[n] = (dm & 1) ? 0.0 : b[n] - f[0];
        r[g.nnodes + n] = (dm & 2) ? 0.0 : b[g.nnodes + n] - f[1];
    } else if (iy < g.ydet + 4) {
        r[n] = (dm & 1) ? 0.0 : b[n];
        r[g.nnodes + n] = (dm & 2) ? 0.0 : b[g.nnodes + n];
    } else {
        if (dm & 1) r[n] = 0.0;
        if (dm & 2) r[g.nnodes + n] = 0.0;
    }
}

// one colour of the multicolour block Gauss-Seidel sweep (MultigridSolver.hh:347-379, 408-458): nodes
// (ox + 2 i, oy + 2 j) of the non-detached rows; the 2 x 2 block solve for free nodes, the point sweep in
// component order (reversed backwards) for partially constrained ones, fully constrained nodes kept
__device__ __forceinline__ void gs_update2(const PGrid &g, uint8_t dm, const double f[2], const double M[4],
                                           double *__restrict__ u, const double *__restrict__ b, int n, int forward) {
    const double r0 = b[n] - f[0], r1 = b[g.nnodes + n] - f[1];
    double d0, d1;
    if (dm) {
        const double fr0 = (dm & 1) ? 0.0 : 1.0, fr1 = (dm & 2) ? 0.0 : 1.0;
        if (forward) {
            d0 = r0 * (fr0 / M[0]);
            d1 = (r1 - M[2] * d0) * (fr1 / M[3]);
        } else {
            d1 = r1 * (fr1 / M[3]);
            d0 = (r0 - M[1] * d1) * (fr0 / M[0]);
        }
    } else {
        const double id = 1.0 / (M[0] * M[3] - M[1] * M[2]);
        d0 = (M[3] * r0 - M[1] * r1) * id;
        d1 = (M[0] * r1 - M[2] * r0) * id;
    }
    u[n] += d0;
    u[g.nnodes + n] += d1;
}
template <bool FINE>
__global__ void __launch_bounds__(TPB) k_smooth2(PGrid g, const double *__restrict__ K0, const double *__restrict__ E,
                                                 const double *__restrict__ Ke, const uint8_t *__restrict__ dmask,
                                                 double *__restrict__ u, const double *__restrict__ b, int ox, int oy,
                                                 int forward) {
    const int mx = (g.nn[0] - ox + 1) / 2, my = (g.ydet - oy + 1) / 2;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (my <= 0 || t >= mx * my) return;
    const int ix = ox + 2 * (t / my), iy = oy + 2 * (t % my);
    const int n = ix * g.nn[1] + iy;
    const uint8_t dm = dmask[n];
    if (dm == 3) return;
    double f[2] = {0, 0}, M[4] = {0, 0, 0, 0};
    node_Ku<FINE>(g, K0, E, Ke, u, ix, iy, f, M);
    gs_update2(g, dm, f, M, u, b, n, forward);
}
// The same colour step on a coarse level with four work-items per node, one per incident element in node_Ku's
// order ((1,1), (1,0), (0,1), (0,0)): four times the work-items in flight and a quarter of the dependent gathers
// each. The first of the four sums the contributions in that order (the coarse products carry no modulus, so
// the sums equal node_Ku's bit for bit) and updates the node.
__global__ void __launch_bounds__(TPB) k_smooth2_x4(PGrid g, const double *__restrict__ Ke,
                                                    const uint8_t *__restrict__ dmask, double *__restrict__ u,
                                                    const double *__restrict__ b, int ox, int oy, int forward) {
    const int mx = (g.nn[0] - ox + 1) / 2, my = (g.ydet - oy + 1) / 2;
    const int tt = blockIdx.x * TPB + threadIdx.x, t = tt >> 2, p = tt & 3;
    if (my <= 0 || t >= mx * my) return;  // all four work-items of a node
    const int ix = ox + 2 * (t / my), iy = oy + 2 * (t % my);
    const int n = ix * g.nn[1] + iy;
    const uint8_t dm = dmask[n];
    if (dm == 3) return;  // all four work-items of a node
    const int i = 1 - (p >> 1), j = 1 - (p & 1);
    const int ex = ix - i, ey = iy - j;
    double v[6] = {0, 0, 0, 0, 0, 0};  // f[2], M[4] of this element
    if (ex >= 0 && ey >= 0 && ex < g.ne[0] && ey < g.ne[1])
        elem_rows<false>(g, nullptr, nullptr, Ke, u, ex, ey, 2 * i + j, v, v + 2);
    const int q0 = (int)(threadIdx.x & 63) & ~3;
    double acc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double a = __shfl(v[k], q0, 64);
        a += __shfl(v[k], q0 + 1, 64);
        a += __shfl(v[k], q0 + 2, 64);
        a += __shfl(v[k], q0 + 3, 64);
        acc[k] = a;
    }
    if (p) return;
    gs_update2(g, dm, acc, acc + 2, u, b, n, forward);
}

// restriction (MultigridSolver.hh:216-275): the transpose of the bilinear interpolation over the fine
// non-detached rows; detached coarse rows are zero
__global__ void __launch_bounds__(TPB) k_restrict2(PGrid f, PGrid c, const double *__restrict__ rf,
                                                   double *__restrict__ rc) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= c.nnodes) return;
    const int ix = n / c.nn[1], iy = n % c.nn[1];
    double v0 = 0, v1 = 0;
    if (iy < c.ydet) {
        for (int ox = -1; ox <= 1; ++ox) {
            const int fx = 2 * ix + ox;
            if (fx < 0 || fx >= f.nn[0]) continue;
            for (int oy = -1; oy <= 1; ++oy) {
                const int fy = 2 * iy + oy;
                if (fy < 0 || fy >= f.ydet) continue;
                const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0);
                const int m = fx * f.nn[1] + fy;
                v0 += w * rf[m];
                v1 += w * rf[f.nnodes + m];
            }
        }
    }
    rc[n] = v0;
    rc[c.nnodes + n] = v1;
}

// interpolation (MultigridSolver.hh:178-214): fine (+)= P coarse, over the non-detached fine rows when
// accumulating, over every row otherwise
__global__ void __launch_bounds__(TPB) k_interp2(PGrid f, PGrid c, const double *__restrict__ xc,
                                                 double *__restrict__ xf, int accumulate) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= f.nnodes) return;
    const int ix = n / f.nn[1], iy = n % f.nn[1];
    if (iy >= (accumulate ? f.ydet : f.nn[1])) return;
    const int id[2] = {ix, iy};
    int e[2];
    double w[2][2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        e[d] = min(id[d] / 2, c.ne[d]);
        const int lf = id[d] - 2 * e[d];
        w[d][0] = lf ? 0.5 : 1.0;
        w[d][1] = lf ? 0.5 : 0.0;
    }
    double v0 = 0, v1 = 0;
#pragma unroll
    for (int lc = 0; lc < 4; ++lc) {
        const int a = (lc >> 1) & 1, bb = lc & 1;
        const double coeff = w[0][a] * w[1][bb];
        if (coeff == 0) continue;
        const int m = (e[0] + a) * c.nn[1] + e[1] + bb;
        v0 += coeff * xc[m];
        v1 += coeff * xc[c.nnodes + m];
    }
    if (accumulate) {
        xf[n] += v0;
        xf[f.nnodes + n] += v1;
    } else {
        xf[n] = v0;
        xf[f.nnodes + n] = v1;
    }
}

__global__ void __launch_bounds__(TPB) k_zero_dir2(int nn, const uint8_t *__restrict__ dmask, double *__restrict__ u) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= nn) return;
    const uint8_t dm = dmask[n];
    if (dm & 1) u[n] = 0.0;
    if (dm & 2) u[nn + n] = 0.0;
}
__global__ void __launch_bounds__(TPB) k_enforce_dir2(int nn, const uint8_t *__restrict__ dmask,
                                                      const double *__restrict__ vals, double *__restrict__ u) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= nn) return;
    const uint8_t dm = dmask[n];
    if (dm & 1) u[n] = vals[n];
    if (dm & 2) u[nn + n] = vals[nn + n];
}
// x = 0 on the rows iy < rows (both components)
__global__ void __launch_bounds__(TPB) k_zero_rows2(PGrid g, int rows, double *__restrict__ x) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes || n % g.nn[1] >= rows) return;
    x[n] = x[g.nnodes + n] = 0.0;
}
// x = 0 on the rows iy >= g.ydet
__global__ void __launch_bounds__(TPB) k_zero_detached2(PGrid g, double *__restrict__ x) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes || n % g.nn[1] < g.ydet) return;
    x[n] = x[g.nnodes + n] = 0.0;
}
// PCG vector updates over the non-detached rows: mode 0: d = s + beta d; mode 1: x += alpha d, r -= alpha Ad
__global__ void __launch_bounds__(TPB) k_pcg_update2(PGrid g, int mode, double coef, const double *__restrict__ a,
                                                     double *__restrict__ d, double *__restrict__ x,
                                                     double *__restrict__ r, const double *__restrict__ Ad) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes || n % g.nn[1] >= g.ydet) return;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int i = c * g.nnodes + n;
        if (mode == 0) {
            d[i] = a[i] + coef * d[i];
        } else {
            x[i] += coef * d[i];
            r[i] -= coef * Ad[i];
        }
    }
}

// dot products: partial sums per workgroup (grid-stride over nodes), then one workgroup adds the partials
__global__ void __launch_bounds__(TPB) k_dot_part2(PGrid g, int masked, const double *__restrict__ a,
                                                   const double *__restrict__ b, double *__restrict__ part) {
    __shared__ double sh[TPB];
    double s = 0;
    for (int n = blockIdx.x * TPB + threadIdx.x; n < g.nnodes; n += gridDim.x * TPB) {
        if (masked && n % g.nn[1] >= g.ydet) continue;
        s += a[n] * b[n] + a[g.nnodes + n] * b[g.nnodes + n];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(TPB) k_dot_sum2(int np, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double sh[TPB];
    double s = 0;
    for (int i = threadIdx.x; i < np; i += TPB) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// level-1 element matrices: sum over the four fine children of E_f Chat_fi (MultigridSolver.hh:724-742);
// masked coarse elements are zero. One work-item per (element, entry).
__global__ void __launch_bounds__(TPB) k_coarsen_l1(PGrid c, PGrid f, const double *__restrict__ Ef,
                                                    const double *__restrict__ Chat, double *__restrict__ Ke) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)c.nelems * 64) return;
    const int e = (int)(t / 64), k = (int)(t % 64);
    const int ex = e / c.ne[1], ey = e % c.ne[1];
    double v = 0;
    if (ey < c.ymask) {
#pragma unroll
        for (int fi = 0; fi < 4; ++fi) {
            const int fe = (2 * ex + ((fi >> 1) & 1)) * f.ne[1] + 2 * ey + (fi & 1);
            v += Ef[fe] * Chat[fi * 64 + k];
        }
    }
    Ke[t] = v;
}
// level l >= 2: Ke_c = sum_fi I_fi^T Ke_f(child fi) I_fi with I[2i+c, 2j+d] = phi_fi(i, j) delta_cd
// (MultigridSolver.hh:704-722, 744-819); phi[fi * 16 + i * 4 + j]
__global__ void __launch_bounds__(TPB) k_coarsen2(PGrid c, PGrid f, const double *__restrict__ Kf,
                                                  const double *__restrict__ phi, double *__restrict__ Kc) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)c.nelems * 64) return;
    const int e = (int)(t / 64), k = (int)(t % 64);
    const int ex = e / c.ne[1], ey = e % c.ne[1];
    const int a = k / 8, bcol = k % 8, i2 = a / 2, c1 = a % 2, j2 = bcol / 2, d1 = bcol % 2;
    double v = 0;
    if (ey < c.ymask) {
        for (int fi = 0; fi < 4; ++fi) {
            const double *K = Kf + (size_t)((2 * ex + ((fi >> 1) & 1)) * f.ne[1] + 2 * ey + (fi & 1)) * 64;
            const double *p = phi + fi * 16;
            for (int i = 0; i < 4; ++i) {
                const double pi = p[i * 4 + i2];
                if (pi == 0) continue;
                double s = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) s += K[(2 * i + c1) * 8 + 2 * j + d1] * p[j * 4 + j2];
                v += pi * s;
            }
        }
    }
    Kc[t] = v;
}

// lower band of the reduced coarsest operator (row DoF, neighbour slot): red[3 n + c] < 0 for fixed or
// detached DoFs; block (I, I - q) of the band at B + (I (kb + 1) + q) 4096, row-major
__global__ void __launch_bounds__(TPB) k_band_assemble2(PGrid g, int fine, const double *__restrict__ K0,
                                                        const double *__restrict__ E, const double *__restrict__ Ke,
                                                        const int *__restrict__ red, int kb, double *__restrict__ B) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)g.nnodes * 18) return;
    const int row = (int)(t % (2L * g.nnodes)), slot = (int)(t / (2L * g.nnodes));
    const int n = row / 2, a = row % 2;
    const int rr = red[3 * n + a];
    if (rr < 0) return;
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    const int jx = ix + slot / 3 - 1, jy = iy + slot % 3 - 1;
    if (jx < 0 || jy < 0 || jx >= g.nn[0] || jy >= g.nn[1]) return;
    const int m = jx * g.nn[1] + jy;
    double Bv[2] = {0, 0};
    for (int ex = max(max(ix, jx) - 1, 0); ex <= min(min(ix, jx), g.ne[0] - 1); ++ex)
        for (int ey = max(max(iy, jy) - 1, 0); ey <= min(min(iy, jy), g.ne[1] - 1); ++ey) {
            const int e = ex * g.ne[1] + ey;
            const int ln = 2 * (ix - ex) + (iy - ey), lm = 2 * (jx - ex) + (jy - ey);
            const double *K = fine ? K0 : Ke + (size_t)e * 64;
            const double s = fine ? E[e] : 1.0;
            for (int c = 0; c < 2; ++c) Bv[c] += s * K[(2 * ln + a) * 8 + 2 * lm + c];
        }
    for (int c = 0; c < 2; ++c) {
        const int cc = red[3 * m + c];
        if (cc < 0 || cc > rr) continue;
        const int I = rr / 64, J = cc / 64;
        B[((size_t)I * (kb + 1) + (I - J)) * 4096 + (rr % 64) * 64 + cc % 64] = Bv[c];
    }
}

// element kernels (TensorProductSimulator.hh:972-1073, 1269-1288); integrated Q1 shape function = 1/4
__device__ __forceinline__ void gather_elem2(const PGrid &g, const double *u, int ex, int ey, double ue[8]) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int nd = (ex + ((l >> 1) & 1)) * g.nn[1] + ey + (l & 1);
        ue[2 * l] = u[nd];
        ue[2 * l + 1] = u[g.nnodes + nd];
    }
}
__device__ __forceinline__ double quad_form2(const double *K0, const double ue[8]) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double r = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) r += K0[a * 8 + b] * ue[b];
        s += ue[a] * r;
    }
    return s;
}
__global__ void __launch_bounds__(TPB) k_grad2(PGrid g, const double *__restrict__ K0, const double *__restrict__ rho,
                                               const double *__restrict__ u, double *__restrict__ out, int law,
                                               double E0, double Emin, double gamma, double q, double g0, double g1,
                                               double area, int accumulate) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    const int ex = e / g.ne[1], ey = e % g.ne[1];
    if (ey >= g.ymask) {
        if (!accumulate) out[e] = 0.0;
        return;
    }
    double ue[8];
    gather_elem2(g, u, ex, ey, ue);
    const double r = rho[e];
    const double dE = law == 0 ? gamma * pow(r, gamma - 1.0) * (E0 - Emin)
                               : (1 + q) * (E0 - Emin) / ((1 + q * (1 - r)) * (1 + q * (1 - r)));
    double val = -0.5 * dE * quad_form2(K0, ue);
    if (g0 != 0 || g1 != 0)
        for (int l = 0; l < 4; ++l) val += 0.25 * (g0 * ue[2 * l] + g1 * ue[2 * l + 1]) * area;
    if (accumulate) out[e] += val;
    else out[e] = val;
}
__global__ void __launch_bounds__(TPB) k_energy2(PGrid g, const double *__restrict__ K0, const double *__restrict__ E,
                                                 const double *__restrict__ u, double *__restrict__ out) {
    const int e = blockIdx.x * TPB + threadIdx.x;
    if (e >= g.nelems) return;
    double ue[8];
    gather_elem2(g, u, e / g.ne[1], e % g.ne[1], ue);
    out[e] = 0.5 * E[e] * quad_form2(K0, ue);
}
__global__ void __launch_bounds__(TPB) k_self_weight2(PGrid g, const double *__restrict__ rho, double g0, double g1,
                                                      double area, double *__restrict__ f) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    double w = 0;
    for (int e = 0; e < 4; ++e) {
        const int ex = ix - ((e >> 1) & 1), ey = iy - (e & 1);
        if (ex < 0 || ey < 0 || ex >= g.ne[0] || ey >= g.ne[1] || ey >= g.ymask) continue;
        w += 0.25 * rho[ex * g.ne[1] + ey] * area;
    }
    f[n] += g0 * w;
    f[g.nnodes + n] += g1 * w;
}

// constantStrainLoad: f = sum_e rho_e int_e B^T C eps, the corner weights w[c][d] = sigma_cd A / (2 h_d)
// carrying the sign of the corner along axis d
__global__ void __launch_bounds__(TPB) k_const_strain2(PGrid g, const double *__restrict__ rho, double w00, double w01,
                                                       double w10, double w11, double *__restrict__ f) {
    const int n = blockIdx.x * TPB + threadIdx.x;
    if (n >= g.nnodes) return;
    const int ix = n / g.nn[1], iy = n % g.nn[1];
    double a0 = 0, a1 = 0;
    for (int l = 3; l >= 0; --l) {
        const int i = (l >> 1) & 1, j = l & 1, ex = ix - i, ey = iy - j;
        if (ex < 0 || ey < 0 || ex >= g.ne[0] || ey >= g.ne[1]) continue;
        const double r = rho[ex * g.ne[1] + ey], sx = i ? 1.0 : -1.0, sy = j ? 1.0 : -1.0;
        a0 += r * (w00 * sx + w01 * sy);
        a1 += r * (w10 * sx + w11 * sy);
    }
    f[n] = a0;
    f[g.nnodes + n] = a1;
}
// sampleNodalField (TensorProductSimulator.hh:1161-1179) at N = 2: the point is clamped to the domain and
// located in its element (a coordinate within 1e-10 of the far face maps to the last element); the element's
// bilinear interpolant is evaluated at the clamped coordinates q themselves, as the reference does
__global__ void __launch_bounds__(TPB) k_sample2(PGrid g, double lo0, double lo1, double hi0, double hi1, double h0,
                                                 double h1, const double *__restrict__ u, int ncomp,
                                                 const double *__restrict__ p, int np, double *__restrict__ out,
                                                 int *__restrict__ bad) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= np) return;
    const double lo[2] = {lo0, lo1}, hi[2] = {hi0, hi1}, h[2] = {h0, h1};
    double q[2];
    int e[2];
    for (int d = 0; d < 2; ++d) {
        q[d] = fmin(fmax(p[2 * i + d], lo[d]), hi[d]);
        const double fi = (q[d] - lo[d]) / h[d];
        if (fabs(fi - g.ne[d]) < 1e-10) e[d] = g.ne[d] - 1;
        else e[d] = (int)fi;
        if (e[d] < 0 || e[d] >= g.ne[d]) {
            atomicOr(bad, 1);
            e[d] = 0;
        }
    }
    double w[4];
    for (int l = 0; l < 4; ++l) w[l] = (l & 1 ? q[1] : 1.0 - q[1]) * ((l >> 1) & 1 ? q[0] : 1.0 - q[0]);
    for (int k = 0; k < ncomp; ++k) {
        double v = 0;
        for (int l = 0; l < 4; ++l) {
            const double un = u[(size_t)k * g.nnodes + (e[0] + ((l >> 1) & 1)) * g.nn[1] + e[1] + (l & 1)];
            v = l == 0 ? un * w[0] : v + un * w[l];
        }
        out[(size_t)k * np + i] = v;
    }
}

// SmoothingFilter at N = 2 (TopologyOptimizationFilter.hh:297-400): normalised reflect-BC convolution over
// the offsets |a|, |b| <= R (Linear weight rmin - |(a, b)|, kept where > 0; rmin = radius + 1)
__device__ __forceinline__ int reflect2(int i, int s) {
    while (i < 0 || i >= s) {
        if (i >= s) i = 2 * s - i - 1;
        if (i < 0) i = -i - 1;
    }
    return i;
}
__global__ void __launch_bounds__(TPB) k_smoothing2(int d0, int d1, int R, double rmin, int type, double W,
                                                    const double *__restrict__ in, double *__restrict__ out) {
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= d0 * d1) return;
    const int i = t / d1, j = t % d1;
    double s = 0;
    for (int a = -R; a <= R; ++a) {
        const int ii = reflect2(i + a, d0);
        for (int b = -R; b <= R; ++b) {
            const double w = type == 1 ? rmin - sqrt((double)(a * a + b * b)) : 1.0;
            if (w <= 0) continue;
            s += w * in[(size_t)ii * d1 + reflect2(j + b, d1)];
        }
    }
    out[t] = s / W;
}

void check_launch() { VF_HIP(hipGetLastError()); }

PGrid make_pgrid(const int64_t ne[2]) {
    PGrid g;
    for (int d = 0; d < 2; ++d) {
        g.ne[d] = (int)ne[d];
        g.nn[d] = (int)ne[d] + 1;
    }
    g.nnodes = g.nn[0] * g.nn[1];
    g.nelems = g.ne[0] * g.ne[1];
    g.ydet = g.nn[1];
    g.ymask = g.ne[1];
    return g;
}

// first masked element row and first detached node row of a level with element height hy at mask height H
void level_masks(double H, double hy, bool active, int64_t &firstMasked, int64_t &firstDetached) {
    if (!active) {
        firstMasked = firstDetached = LAYER_NONE;
        return;
    }
    firstMasked = (int64_t)std::ceil(H / hy - 1e-10);
    firstDetached = firstMasked + 1;
}

}  // namespace

// ================================================================================================ PSim
PSim::PSim(const int64_t ne_[2], const double lo_[2], const double hi_[2], hipStream_t s) : stream(s) {
    for (int d = 0; d < 2; ++d) {
        if (ne_[d] < 1) throw std::runtime_error("grid must have at least one element per dimension");
        ne[d] = ne_[d];
        lo[d] = lo_[d];
        hi[d] = hi_[d];
        h[d] = (hi[d] - lo[d]) / (double)ne[d];
        if (!(h[d] > 0)) throw std::runtime_error("domain must have positive extent");
    }
    if ((ne[0] + 1) * (ne[1] + 1) * 3 >= INT32_MAX) throw std::runtime_error("grid too large for 32-bit node indexing");
    g = make_pgrid(ne);
    firstMasked = firstDetached = LAYER_NONE;
    maskHeight = std::numeric_limits<double>::infinity();
    K0.resize(64);
    rho.resize(nel());
    E.resize(nel());
    launch_fill<double>(nel(), rho.data(), 0.0, stream);
    hmask.assign(nn(), 0);
    hvals.assign(ndof(), 0.0);
    dmask.resize(nn());
    dvals.resize(ndof());
    h2d(dmask.data(), hmask.data(), nn(), stream);
    h2d(dvals.data(), hvals.data(), ndof(), stream);
    hforce.assign(ndof(), 0.0);
    force.resize(ndof());
    launch_fill<double>(ndof(), force.data(), 0.0, stream);
    partials.resize(NPART);
    scal.resize(16);
    update_K0();
    update_E();
    sync(stream);
}

void PSim::refresh_grid() {
    g.ydet = (int)std::min<int64_t>(firstDetached, g.nn[1]);
    g.ymask = (int)std::min<int64_t>(firstMasked, g.ne[1]);
}
void PSim::update_K0() {
    element_K0(young, poisson, h, K0h);
    h2d(K0.data(), K0h, 64, stream);
    ++version;
}
void PSim::update_E() {
    refresh_grid();
    hipLaunchKernelGGL(k_young2, dim3(nblk(nel())), dim3(TPB), 0, stream, g, rho.data(), E.data(), law, E0, Emin,
                       gamma, q, maskHeight < hi[1] ? 1 : 0);
    check_launch();
    ++version;
}
void PSim::set_material(double Ey, double nu) {
    young = Ey;
    poisson = nu;
    update_K0();
}
void PSim::set_stretchings(const double hh[2]) {
    for (int d = 0; d < 2; ++d)
        if (!(hh[d] > 0)) throw std::runtime_error("element stretchings must be positive");
    h[0] = hh[0];
    h[1] = hh[1];
    update_K0();
}
void PSim::set_law(int law_, double E0_, double Emin_, double gamma_, double q_) {
    law = law_;
    E0 = E0_;
    Emin = Emin_;
    gamma = gamma_;
    q = q_;
    update_E();
}
void PSim::set_densities(const double *r, bool onDevice) {
    if (onDevice) d2d(rho.data(), r, nel(), stream);
    else h2d(rho.data(), r, nel(), stream);
    update_E();
}
void PSim::set_dirichlet(const uint8_t *mask, const double *values) {
    for (size_t n = 0; n < nn(); ++n) hmask[n] = mask[n] & 3;
    for (size_t i = 0; i < ndof(); ++i) hvals[i] = values ? values[i] : 0.0;
    h2d(dmask.data(), hmask.data(), nn(), stream);
    h2d(dvals.data(), hvals.data(), ndof(), stream);
    ++bcVersion;
}
void PSim::set_forces(const double *f) {
    hasForce = false;
    for (size_t i = 0; i < ndof(); ++i) {
        hforce[i] = f[i];
        hasForce = hasForce || f[i] != 0;
    }
    h2d(force.data(), hforce.data(), ndof(), stream);
}
// setFabricationMaskHeight (TensorProductSimulator.hh:286-310)
void PSim::set_mask_height(double H, bool updateYoung) {
    if (H < 0 || H > hi[1])
        throw std::runtime_error("Fabrication height (" + std::to_string(H) + ") has to be in between 0 and " +
                                 std::to_string(hi[1]));
    maskHeight = H;
    firstMasked = (int64_t)std::ceil(maskHeight / h[1] - 1e-10);
    firstDetached = firstMasked + 1;
    refresh_grid();
    if (updateYoung) update_E();
}
// decrementFabricationMaskHeightByLayer (TensorProductSimulator.hh:311-324)
void PSim::remove_top_layers(int64_t count) {
    maskHeight -= (double)count * h[1];
    firstMasked -= count;
    firstDetached = firstMasked + 1;
    refresh_grid();
    hipLaunchKernelGGL(k_zero_young_rows2, dim3(nblk(nel())), dim3(TPB), 0, stream, g, (int)firstMasked,
                       (int)(firstMasked + count), E.data());
    check_launch();
    ++version;
}
void PSim::apply_K(const double *u, double *y, bool zeroInit, bool negate) {
    refresh_grid();
    hipLaunchKernelGGL(k_apply2<true>, dim3(nblk(nn())), dim3(TPB), 0, stream, g, K0.data(), E.data(), nullptr, u, y,
                       zeroInit ? 1 : 0, negate ? 1 : 0);
    check_launch();
}
void PSim::build_load(double *f) {
    refresh_grid();
    d2d(f, force.data(), ndof(), stream);
    if (grav[0] != 0 || grav[1] != 0) {
        hipLaunchKernelGGL(k_self_weight2, dim3(nblk(nn())), dim3(TPB), 0, stream, g, rho.data(), grav[0], grav[1],
                           h[0] * h[1], f);
        check_launch();
    }
}
void PSim::compliance_gradient(const double *u, double *gr, bool accumulate) {
    refresh_grid();
    hipLaunchKernelGGL(k_grad2, dim3(nblk(nel())), dim3(TPB), 0, stream, g, K0.data(), rho.data(), u, gr, law, E0, Emin,
                       gamma, q, grav[0], grav[1], h[0] * h[1], accumulate ? 1 : 0);
    check_launch();
}
void PSim::energy_density(const double *u, double *out) {
    hipLaunchKernelGGL(k_energy2, dim3(nblk(nel())), dim3(TPB), 0, stream, g, K0.data(), E.data(), u, out);
    check_launch();
}
double PSim::dot(const double *a, const double *b, bool masked) {
    refresh_grid();
    const int np = (int)std::min<size_t>(NPART, nblk(nn()));
    hipLaunchKernelGGL(k_dot_part2, dim3(np), dim3(TPB), 0, stream, g, masked ? 1 : 0, a, b, partials.data());
    hipLaunchKernelGGL(k_dot_sum2, dim3(1), dim3(TPB), 0, stream, np, partials.data(), scal.data());
    check_launch();
    double v;
    d2h(&v, scal.data(), 1, stream);
    return v;
}
void PSim::zero_dirichlet(double *u) {
    hipLaunchKernelGGL(k_zero_dir2, dim3(nblk(nn())), dim3(TPB), 0, stream, g.nnodes, dmask.data(), u);
    check_launch();
}
void PSim::enforce_dirichlet(double *u) {
    hipLaunchKernelGGL(k_enforce_dir2, dim3(nblk(nn())), dim3(TPB), 0, stream, g.nnodes, dmask.data(), dvals.data(), u);
    check_launch();
}
void PSim::zero_detached(double *u) {
    refresh_grid();
    hipLaunchKernelGGL(k_zero_detached2, dim3(nblk(nn())), dim3(TPB), 0, stream, g, u);
    check_launch();
}
void PSim::constant_strain_load(const double eps[4], double *F) {
    // sigma = lambda* tr(eps) I + 2 mu eps with the plane-stress lambda* = E nu / (1 - nu^2)
    const double lam = young * poisson / ((1 + poisson) * (1 - poisson)), mu = young / (2 * (1 + poisson));
    const double tr = eps[0] + eps[3], A = h[0] * h[1];
    double w[2][2];
    for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 2; ++d) w[c][d] = ((c == d ? lam * tr : 0.0) + 2 * mu * eps[2 * c + d]) * A / (2.0 * h[d]);
    hipLaunchKernelGGL(k_const_strain2, dim3(nblk(nn())), dim3(TPB), 0, stream, g, rho.data(), w[0][0], w[0][1], w[1][0],
                       w[1][1], F);
    check_launch();
}
void PSim::sample(const double *u, int ncomp, const double *p, int np, double *out, int *bad) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_sample2, dim3(nblk(np)), dim3(TPB), 0, stream, g, lo[0], lo[1], hi[0], hi[1], h[0], h[1], u,
                       ncomp, p, np, out, bad);
    check_launch();
}
double *PSim::stage(int which, size_t n) {
    DBuf<double> &b = which == 0 ? stA : (which == 1 ? stB : stC);
    if (b.size() < n) b.resize(n);
    return b.data();
}

// ================================================================================================= PMG
PMG::PMG(PSim *s, int levels) : sim(s), L(levels) {
    if (levels < 0) throw std::runtime_error("numCoarseningLevels must be >= 0");
    auto l0 = std::make_unique<PLevel>();
    l0->g = sim->g;
    l0->h[0] = sim->h[0];
    l0->h[1] = sim->h[1];
    l0->hmask = sim->hmask;
    lv.push_back(std::move(l0));
    int64_t ne[2] = {sim->ne[0], sim->ne[1]};
    for (int l = 1; l <= L; ++l) {
        for (int d = 0; d < 2; ++d) {
            if (ne[d] % 2)
                throw std::runtime_error(
                    "Grid size currently must be divisible by 2^numCoarseningLevels (nonuniform coarsening not yet "
                    "implemented)");
            ne[d] /= 2;
        }
        auto c = std::make_unique<PLevel>();
        c->g = make_pgrid(ne);
        for (int d = 0; d < 2; ++d) c->h[d] = (sim->hi[d] - sim->lo[d]) / (double)ne[d];
        // Dirichlet coarsening (MultigridSolver.hh:58-103): even fine index i -> coarse i / 2; odd -> both
        // (i - 1) / 2 and (i + 1) / 2. Computed once, from the constraints present now, as the reference does.
        const PLevel &f = *lv.back();
        c->hmask.assign(c->g.nnodes, 0);
        for (int n = 0; n < f.g.nnodes; ++n) {
            const uint8_t m = f.hmask[n];
            if (!m) continue;
            const int id[2] = {n / f.g.nn[1], n % f.g.nn[1]};
            int a0[2], a1[2];
            for (int d = 0; d < 2; ++d) {
                if (id[d] % 2 == 0) a0[d] = a1[d] = id[d] / 2;
                else { a0[d] = (id[d] - 1) / 2; a1[d] = (id[d] + 1) / 2; }
            }
            for (int a = a0[0]; a <= a1[0]; ++a)
                for (int b = a0[1]; b <= a1[1]; ++b) c->hmask[a * c->g.nn[1] + b] |= m;
        }
        c->dmask.resize(c->g.nnodes);
        h2d(c->dmask.data(), c->hmask.data(), c->hmask.size(), st());
        c->Ke.resize((size_t)c->g.nelems * 64);
        lv.push_back(std::move(c));
    }
    for (auto &p : lv) {
        const size_t n = 2 * (size_t)p->g.nnodes;
        p->x.resize(n);
        p->b.resize(n);
        p->r.resize(n);
        launch_fill<double>(n, p->x.data(), 0.0, st());
        launch_fill<double>(n, p->b.data(), 0.0, st());
        launch_fill<double>(n, p->r.data(), 0.0, st());
    }
    // coarse shape functions at the nodes of child fi (MultigridSolver.hh:664-687), phi[fi][fine l][coarse l]
    for (int fi = 0; fi < 4; ++fi)
        for (int lf = 0; lf < 4; ++lf)
            for (int lc = 0; lc < 4; ++lc) {
                double v = 1;
                for (int d = 0; d < 2; ++d) {
                    const int sh = 1 - d;
                    const double x = 0.5 * (((fi >> sh) & 1) + ((lf >> sh) & 1));
                    v *= ((lc >> sh) & 1) ? x : 1 - x;
                }
                phi[fi * 16 + lf * 4 + lc] = v;
            }
    dChat.resize(256);
    dphi.resize(64);
    h2d(dphi.data(), phi, 64, st());
    update_chat();
    flag.resize(1);
    sync(st());
}

// coarsenedFineK0s (MultigridSolver.hh:116-120): Chat_fi = I_fi^T K0 I_fi
void PMG::update_chat() {
    for (int fi = 0; fi < 4; ++fi)
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                double v = 0;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j)
                        v += phi[fi * 16 + i * 4 + a / 2] * sim->K0h[(2 * i + a % 2) * 8 + 2 * j + b % 2] *
                             phi[fi * 16 + j * 4 + b / 2];
                Chat[fi * 64 + a * 8 + b] = v;
            }
    h2d(dChat.data(), Chat, 256, st());
}

void PMG::sync_masks() {
    sim->refresh_grid();
    lv[0]->g = sim->g;
    const bool active = sim->maskHeight < std::numeric_limits<double>::infinity();
    for (int l = 1; l <= L; ++l) {
        int64_t fm, fd;
        level_masks(sim->maskHeight, lv[l]->h[1], active, fm, fd);
        lv[l]->g.ydet = (int)std::min<int64_t>(fd, lv[l]->g.nn[1]);
        lv[l]->g.ymask = (int)std::min<int64_t>(fm, lv[l]->g.ne[1]);
    }
}

void PMG::update_stiffness() {
    sync_masks();
    update_chat();  // the material or the stretchings may have changed since construction
    for (int l = 1; l <= L; ++l) {
        const PLevel &c = *lv[l], &f = *lv[l - 1];
        const size_t work = (size_t)c.g.nelems * 64;
        if (l == 1)
            hipLaunchKernelGGL(k_coarsen_l1, dim3(nblk(work)), dim3(TPB), 0, st(), c.g, f.g, sim->E.data(), dChat.data(),
                               lv[1]->Ke.data());
        else
            hipLaunchKernelGGL(k_coarsen2, dim3(nblk(work)), dim3(TPB), 0, st(), c.g, f.g, f.Ke.data(), dphi.data(),
                               lv[l]->Ke.data());
        check_launch();
    }
    factor_coarsest();
    stiffReady = true;
    stiffVersion = sim->version;
    stiffBc = sim->bcVersion;
    maskedRowsAtBuild = sim->firstMasked;
}

void PMG::ensure_stiffness() {
    sync_masks();
    if (!stiffReady || stiffVersion != sim->version || stiffBc != sim->bcVersion ||
        maskedRowsAtBuild != sim->firstMasked)
        update_stiffness();
}

// coarsest direct solve (TensorProductSimulator.hh:1181-1230): the reduced operator without fixed and
// detached DoFs, factored by the band Cholesky (its explicit inverse applied in the solves)
void PMG::factor_coarsest() {
    PLevel &c = *lv[L];
    const PGrid &g = c.g;
    const std::vector<uint8_t> &hm = L == 0 ? sim->hmask : c.hmask;
    if (L == 0)
        for (size_t n = 0; n < sim->nn(); ++n)
            for (int k = 0; k < 2; ++k)
                if ((hm[n] >> k) & 1 && sim->hvals[k * sim->nn() + n] != 0)
                    throw std::runtime_error("Nonzero Dirichlet constraints currently unsupported");
    red.assign(3 * (size_t)g.nnodes, -1);
    redToFull.clear();
    for (int n = 0; n < g.nnodes; ++n) {
        const int iy = n % g.nn[1];
        for (int k = 0; k < 2; ++k) {
            if (((hm[n] >> k) & 1) || iy >= g.ydet) continue;
            red[3 * (size_t)n + k] = (int)redToFull.size();
            redToFull.push_back(3 * n + k);
        }
    }
    nred = (int)redToFull.size();
    factored = false;
    if (nred == 0) {
        factored = true;
        return;
    }
    const int ld = (nred + 63) / 64 * 64;
    if (ld > 24576)
        throw std::runtime_error("direct solve of " + std::to_string(nred) +
                                 " unknowns exceeds the coarse-solver limit (24576); use more coarsening levels");
    int bw = 0;
    for (int n = 0; n < g.nnodes; ++n) {
        const int ix = n / g.nn[1], iy = n % g.nn[1];
        for (int k = 0; k < 2; ++k) {
            const int rr = red[3 * (size_t)n + k];
            if (rr < 0) continue;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    const int jx = ix + dx, jy = iy + dy;
                    if (jx < 0 || jy < 0 || jx >= g.nn[0] || jy >= g.nn[1]) continue;
                    for (int k2 = 0; k2 < 2; ++k2) {
                        const int cc = red[3 * (size_t)(jx * g.nn[1] + jy) + k2];
                        if (cc >= 0 && cc < rr) bw = std::max(bw, rr - cc);
                    }
                }
        }
    }
    kb = std::max(1, (bw + 63) / 64);
    dRed.resize(red.size());
    dRedToFull.resize(redToFull.size());
    h2d(dRed.data(), red.data(), red.size(), st());
    h2d(dRedToFull.data(), redToFull.data(), redToFull.size(), st());
    const size_t bsz = band_storage(nred, kb);
    B.resize(bsz);
    Linv.resize((size_t)(ld / 64) * 4096);
    X.resize((size_t)ld * ld);
    XT.resize((size_t)ld * ld);
    tin.resize(ld);
    tout.resize(ld);
    launch_fill<double>(bsz, B.data(), 0.0, st());
    hipLaunchKernelGGL(k_band_assemble2, dim3(nblk((size_t)g.nnodes * 18)), dim3(TPB), 0, st(), g, L == 0 ? 1 : 0,
                       sim->K0.data(), sim->E.data(), L == 0 ? nullptr : c.Ke.data(), dRed.data(), kb, B.data());
    check_launch();
    band_factor_inverse(nred, kb, B.data(), Linv.data(), X.data(), XT.data(), flag.data(), st());
    int f = 0;
    d2h(&f, flag.data(), 1, st());
    if (f) throw std::runtime_error("coarsest-level operator is not positive definite");
    factored = true;
}

void PMG::apply_K(int l, const double *u, double *y, bool zeroInit, bool negate) {
    const PGrid &g = lv[l]->g;
    if (l == 0)
        hipLaunchKernelGGL(k_apply2<true>, dim3(nblk(g.nnodes)), dim3(TPB), 0, st(), g, sim->K0.data(), sim->E.data(),
                           nullptr, u, y, zeroInit ? 1 : 0, negate ? 1 : 0);
    else
        hipLaunchKernelGGL(k_apply2<false>, dim3(nblk(g.nnodes)), dim3(TPB), 0, st(), g, nullptr, nullptr,
                           lv[l]->Ke.data(), u, y, zeroInit ? 1 : 0, negate ? 1 : 0);
    check_launch();
}
void PMG::residual(int l, const double *u, const double *b, double *r) {
    const PGrid &g = lv[l]->g;
    const uint8_t *dm = l == 0 ? sim->dmask.data() : lv[l]->dmask.data();
    if (l == 0)
        hipLaunchKernelGGL(k_residual2<true>, dim3(nblk(g.nnodes)), dim3(TPB), 0, st(), g, sim->K0.data(),
                           sim->E.data(), nullptr, dm, u, b, r);
    else
        hipLaunchKernelGGL(k_residual2_x4, dim3(nblk(4 * (size_t)g.nnodes)), dim3(TPB), 0, st(), g,
                           lv[l]->Ke.data(), dm, u, b, r);
    check_launch();
}
void PMG::smooth(int l, double *u, const double *b, bool forward) {
    const PGrid &g = lv[l]->g;
    const uint8_t *dm = l == 0 ? sim->dmask.data() : lv[l]->dmask.data();
    for (int t = 0; t < 4; ++t) {
        const int color = forward ? t : 3 - t;
        const int ox = (color >> 1) & 1, oy = color & 1;
        const int cnt = ((g.nn[0] - ox + 1) / 2) * std::max(0, (g.ydet - oy + 1) / 2);
        if (cnt <= 0) continue;
        if (l == 0)
            hipLaunchKernelGGL(k_smooth2<true>, dim3(nblk(cnt)), dim3(TPB), 0, st(), g, sim->K0.data(), sim->E.data(),
                               nullptr, dm, u, b, ox, oy, forward ? 1 : 0);
        else
            hipLaunchKernelGGL(k_smooth2_x4, dim3(nblk(4 * (size_t)cnt)), dim3(TPB), 0, st(), g, lv[l]->Ke.data(), dm,
                               u, b, ox, oy, forward ? 1 : 0);
        check_launch();
    }
}
void PMG::restrict_to(int l, const double *fine, double *coarse) {
    hipLaunchKernelGGL(k_restrict2, dim3(nblk(lv[l + 1]->g.nnodes)), dim3(TPB), 0, st(), lv[l]->g, lv[l + 1]->g, fine,
                       coarse);
    check_launch();
}
void PMG::interpolate(int l, const double *coarse, double *fine, bool accumulate) {
    hipLaunchKernelGGL(k_interp2, dim3(nblk(lv[l]->g.nnodes)), dim3(TPB), 0, st(), lv[l]->g, lv[l + 1]->g, coarse, fine,
                       accumulate ? 1 : 0);
    check_launch();
}
void PMG::coarse_solve(const double *b, double *x) {
    if (!factored) factor_coarsest();
    launch_coarse_chol_apply<double>(nred, X.data(), XT.data(), dRedToFull.data(), lv[L]->g.nnodes, b, x, tin.data(),
                                     tout.data(), st(), 2);
}
void PMG::zero_dirichlet(int l, double *u) {
    const uint8_t *dm = l == 0 ? sim->dmask.data() : lv[l]->dmask.data();
    hipLaunchKernelGGL(k_zero_dir2, dim3(nblk(lv[l]->g.nnodes)), dim3(TPB), 0, st(), lv[l]->g.nnodes, dm, u);
    check_launch();
}

// vcycle / fullMultigrid (MultigridSolver.hh:587-658)
void PMG::vcycle(int l, int nu, bool residualSystem) {
    PLevel &p = *lv[l];
    if (l == L) {
        coarse_solve(p.b.data(), p.x.data());
        return;
    }
    if (residualSystem || l > 0) zero_dirichlet(l, p.x.data());
    else sim->enforce_dirichlet(p.x.data());
    for (int i = 0; i < nu; ++i) smooth(l, p.x.data(), p.b.data(), true);
    residual(l, p.x.data(), p.b.data(), p.r.data());
    PLevel &c = *lv[l + 1];
    restrict_to(l, p.r.data(), c.b.data());
    hipLaunchKernelGGL(k_zero_rows2, dim3(nblk(c.g.nnodes)), dim3(TPB), 0, st(), c.g,
                       std::min(c.g.ydet + 4, c.g.nn[1]), c.x.data());
    check_launch();
    vcycle(l + 1, nu, true);
    interpolate(l, c.x.data(), p.x.data(), true);
    for (int i = 0; i < nu; ++i) smooth(l, p.x.data(), p.b.data(), !symmetricGS);
}
void PMG::fmg(int l, int nu, bool residualSystem) {
    PLevel &p = *lv[l];
    if (l == L) {
        coarse_solve(p.b.data(), p.x.data());
        return;
    }
    restrict_to(l, p.b.data(), lv[l + 1]->b.data());
    fmg(l + 1, nu, residualSystem);
    interpolate(l, lv[l + 1]->x.data(), p.x.data(), false);
    vcycle(l, nu, residualSystem);
}

// applyPreconditionerInv (MultigridSolver.hh:577-580)
void PMG::precondition(const double *r, double *s, int mgIter, int nu, bool fullMG) {
    sync_masks();
    const size_t n = level_n(0);
    if (nu == 0) {
        d2d(s, r, n, st());
        return;
    }
    ensure_stiffness();
    PLevel &p = *lv[0];
    d2d(p.b.data(), r, n, st());
    launch_fill<double>(n, p.x.data(), 0.0, st());
    for (int i = 0; i < mgIter; ++i) {
        if (fullMG && i == 0) fmg(0, nu, true);
        else vcycle(0, nu, true);
    }
    d2d(s, p.x.data(), n, st());
}

// MultigridSolver::solve (MultigridSolver.hh:546-575)
void PMG::solve(double *x, const double *f, int numSteps, int nu, bool stiffnessUpdated, bool zeroDirichlet,
                bool fullMG, int (*cb)(void *, int64_t, const double *), void *user) {
    sync_masks();
    if (!stiffnessUpdated) update_stiffness();
    else ensure_stiffness();
    if (numSteps == 0) return;
    const size_t n = level_n(0);
    std::vector<double> hx;
    PLevel &p = *lv[0];
    if (L == 0) {
        coarse_solve(f, x);
        if (cb) {
            hx.resize(n);
            d2h(hx.data(), x, n, st());
            if (cb(user, 0, hx.data())) throw std::runtime_error("solve aborted by callback");
        }
        return;
    }
    d2d(p.x.data(), x, n, st());
    d2d(p.b.data(), f, n, st());
    for (int i = 0; i < numSteps; ++i) {
        if (fullMG && i == 0) fmg(0, nu, zeroDirichlet);
        else vcycle(0, nu, zeroDirichlet);
        if (cb) {
            hx.resize(n);
            d2h(hx.data(), p.x.data(), n, st());
            if (cb(user, i, hx.data())) throw std::runtime_error("solve aborted by callback");
        }
    }
    d2d(x, p.x.data(), n, st());
    sync(st());
}

// preconditionedConjugateGradient (MultigridSolver.hh:1046-1152), FP64
int64_t PMG::pcg(double *x, const double *b, int64_t maxIter, double tol, int mgIter, int nu, bool fullMG,
                 bool dirichletSatisfied, int (*cb)(void *, int64_t, const double *, const double *), void *user,
                 double *residualNorm, double *rhsNorm) {
    sync_masks();
    const size_t n = level_n(0);
    const PGrid &g = lv[0]->g;
    pr.resize(n);
    ps.resize(n);
    pd.resize(n);
    pAd.resize(n);
    double *r = pr.data(), *s = ps.data(), *d = pd.data(), *Ad = pAd.data();
    if (L == 0) {
        residual(0, x, b, r);
        const double rr = sim->dot(r, r, false), bn = sim->dot(b, b, false);
        if (residualNorm) *residualNorm = std::sqrt(rr);
        if (rhsNorm) *rhsNorm = std::sqrt(bn);
        if (rr < tol * tol * bn) return 0;
        ensure_stiffness();
        coarse_solve(b, x);
        sync(st());
        return 1;
    }
    if (!dirichletSatisfied) sim->enforce_dirichlet(x);
    const double bnorm = sim->dot(b, b, true);
    residual(0, x, b, r);
    double rr = sim->dot(r, r, true);
    if (std::isnan(rr)) throw NanError("NaN encountered");
    double rz = 0;
    bool haveD = false, updated = false;
    int64_t it = 0;
    std::vector<double> hx, hr;
    while (it < maxIter && rr > tol * tol * bnorm) {
        ++it;
        if (mgIter > 0) {
            if (!updated) {
                update_stiffness();
                updated = true;
            }
            precondition(r, s, mgIter, nu, fullMG);
        } else {
            d2d(s, r, n, st());
        }
        sim->zero_dirichlet(s);
        const double rzOld = rz;
        rz = sim->dot(r, s, true);
        if (haveD) {
            hipLaunchKernelGGL(k_pcg_update2, dim3(nblk(g.nnodes)), dim3(TPB), 0, st(), g, 0, rz / rzOld, s, d, nullptr,
                               nullptr, nullptr);
            check_launch();
        } else {
            d2d(d, s, n, st());
            haveD = true;
        }
        apply_K(0, d, Ad, true, false);
        sim->zero_dirichlet(Ad);
        const double alpha = rz / sim->dot(d, Ad, true);
        hipLaunchKernelGGL(k_pcg_update2, dim3(nblk(g.nnodes)), dim3(TPB), 0, st(), g, 1, alpha, nullptr, d, x, r, Ad);
        check_launch();
        rr = sim->dot(r, r, true);
        if (std::isnan(rr)) throw NanError("NaN encountered at iteration" + std::to_string(it));
        if (cb) {
            hx.resize(n);
            hr.resize(n);
            d2h(hx.data(), x, n, st());
            d2h(hr.data(), r, n, st());
            if (cb(user, it, hx.data(), hr.data())) throw std::runtime_error("PCG aborted by callback");
        }
    }
    if (residualNorm) *residualNorm = std::sqrt(rr);
    if (rhsNorm) *rhsNorm = std::sqrt(bnorm);
    sync(st());
    return it;
}

void launch_smoothing_filter_2d(const int dims[2], double rmin, int type, const double *in, double *out,
                                hipStream_t s) {
    const int R = (int)std::ceil(rmin) - 1;
    double W = 0;
    for (int a = -R; a <= R; ++a)
        for (int b = -R; b <= R; ++b) {
            const double w = type == 1 ? rmin - std::sqrt((double)(a * a + b * b)) : 1.0;
            if (w > 0) W += w;
        }
    const size_t n = (size_t)dims[0] * dims[1];
    hipLaunchKernelGGL(k_smoothing2, dim3(nblk(n)), dim3(TPB), 0, s, dims[0], dims[1], R, rmin, type, W, in, out);
    check_launch();
}

}  // namespace planar
}  // namespace vf
