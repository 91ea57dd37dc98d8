// vf_kernels_ops.hip -- level operators of the multigrid hierarchy on gfx950:
// matrix-free fine K u, level-1 on-the-fly Galerkin operator, assembled 27-point block rows,
// multicolour block Gauss-Seidel and the grid transfers.
//
// Thread mapping: one work-item per output node, z (the contiguous index) fastest, so a
// 64-wide wavefront reads 64 consecutive doubles of each component plane. Element matrices
// (K0, the eight coarsened K0s) are indexed with compile-time offsets from a uniform pointer,
// which the compiler turns into scalar loads shared by the whole wavefront.
#include "vf_internal.hpp"

#include <climits>

namespace vf {
namespace {

constexpr int TPB = 256;

__device__ __forceinline__ int eidx(const Grid &g, int ex, int ey, int ez) { return (ex * g.ne[1] + ey) * g.ne[2] + ez; }

template <typename T>
__device__ __forceinline__ void store_out(T *y, const T *b, int N, int n, const T f[3], int mode, uint8_t dm) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T v;
        if (mode == AM_SET) v = f[c];
        else if (mode == AM_ADD) v = y[c * N + n] + f[c];
        else if (mode == AM_SUB) v = y[c * N + n] - f[c];
        else v = b[c * N + n] - f[c];
        if (dm & (1 << c)) v = T(0);
        y[c * N + n] = v;
    }
}

// Gather the 3x3x3 neighbourhood of node (ix,iy,iz); out-of-grid values read as zero.
template <typename T>
__device__ __forceinline__ void gather27(const Grid &g, const T *__restrict__ u, int ix, int iy, int iz,
                                         T un[27][3]) {
    const int N = g.nnodes, sx = g.nn[1] * g.nn[2], sy = g.nn[2];
    const int n = (ix * g.nn[1] + iy) * g.nn[2] + iz;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) {
                const int s = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
                const bool ok = (unsigned)(ix + dx) < (unsigned)g.nn[0] && (unsigned)(iy + dy) < (unsigned)g.nn[1] &&
                                (unsigned)(iz + dz) < (unsigned)g.nn[2];
                const int m = n + dx * sx + dy * sy + dz;
#pragma unroll
                for (int c = 0; c < 3; ++c) un[s][c] = ok ? u[c * N + m] : T(0);
            }
}

// neighbour slot of local node m of the element whose local index of the centre node is (i,j,k)
__device__ __forceinline__ constexpr int slot_of(int m, int i, int j, int k) {
    return (((m >> 2) & 1) - i + 1) * 9 + (((m >> 1) & 1) - j + 1) * 3 + ((m & 1) - k + 1);
}

template <typename T>
__device__ __forceinline__ void inv3_apply(const T M[9], const T r[3], T out[3]) {
    const T c00 = M[4] * M[8] - M[5] * M[7];
    const T c01 = M[5] * M[6] - M[3] * M[8];
    const T c02 = M[3] * M[7] - M[4] * M[6];
    const T det = M[0] * c00 + M[1] * c01 + M[2] * c02;
    const T id = T(1) / det;
    const T i00 = c00 * id, i01 = (M[2] * M[7] - M[1] * M[8]) * id, i02 = (M[1] * M[5] - M[2] * M[4]) * id;
    const T i10 = c01 * id, i11 = (M[0] * M[8] - M[2] * M[6]) * id, i12 = (M[2] * M[3] - M[0] * M[5]) * id;
    const T i20 = c02 * id, i21 = (M[1] * M[6] - M[0] * M[7]) * id, i22 = (M[0] * M[4] - M[1] * M[3]) * id;
    out[0] = i00 * r[0] + i01 * r[1] + i02 * r[2];
    out[1] = i10 * r[0] + i11 * r[1] + i12 * r[2];
    out[2] = i20 * r[0] + i21 * r[1] + i22 * r[2];
}

// u_n += update from (b - S') and M (MultigridSolver.hh:347-378): block solve for free nodes,
// sequential masked point updates for partially constrained nodes.
template <typename T>
__device__ __forceinline__ void gs_update(T *u, int N, int n, const T bms[3], const T M[9], uint8_t dm, int forward) {
    T ud[3];
    if (dm) {
        ud[0] = ud[1] = ud[2] = T(0);
        for (int t = 0; t < 3; ++t) {
            const int i = forward ? t : 2 - t;
            const T fr = (dm & (1 << i)) ? T(0) : T(1);
            ud[i] = (bms[i] - (M[3 * i] * ud[0] + M[3 * i + 1] * ud[1] + M[3 * i + 2] * ud[2])) * (fr / M[4 * i]);
        }
    } else {
        inv3_apply(M, bms, ud);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c * N + n] += ud[c];
}

// ------------------------------------------------------------------ fine level (K0, E)
// Sum over the 8 incident elements of E_e * K0[:, 3 ln + c]^T u_e (K0 symmetric, column access
// as in TPSStencils.hh:453-471); optionally also the diagonal block M = sum E_e K0[3ln.., 3ln..].
template <typename T, bool WANT_M>
__device__ __forceinline__ void fine_row(const Grid &g, const T *__restrict__ K0, const T *__restrict__ E, int ix,
                                         int iy, int iz, const T un[27][3], T f[3], T M[9]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
        const int ex = ix - i, ey = iy - j, ez = iz - k;
        const bool ok = (unsigned)ex < (unsigned)g.ne[0] && (unsigned)ey < (unsigned)g.ne[1] &&
                        (unsigned)ez < (unsigned)g.ne[2];
        if (!ok) continue;
        const T Ee = E[eidx(g, ex, ey, ez)];
        const int ln = 4 * i + 2 * j + k;
        T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int s = slot_of(m, i, j, k);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T um = un[s][d];
                const int row = (3 * m + d) * 24 + 3 * ln;
                a0 += K0[row + 0] * um;
                a1 += K0[row + 1] * um;
                a2 += K0[row + 2] * um;
            }
        }
        f[0] += Ee * a0;
        f[1] += Ee * a1;
        f[2] += Ee * a2;
        if (WANT_M) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int c2 = 0; c2 < 3; ++c2) M[3 * c + c2] += Ee * K0[(3 * ln + c) * 24 + 3 * ln + c2];
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_fine_apply(Grid g, const T *__restrict__ K0, const T *__restrict__ E,
                                                    const T *__restrict__ u, T *__restrict__ y,
                                                    const T *__restrict__ b, const uint8_t *__restrict__ dmask) {
    const int nz = g.nn[2], slab = g.ydet * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = r / nz, iz = r - iy * nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    T un[27][3];
    gather27(g, u, ix, iy, iz, un);
    T f[3] = {0, 0, 0};
    T M[9];
    fine_row<T, false>(g, K0, E, ix, iy, iz, un, f, M);
    store_out(y, b, g.nnodes, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

// Rows of detached nodes of an AM_SET product are zero (TPSStencils.hh:717-727).
template <typename T>
__global__ void __launch_bounds__(TPB) k_zero_detached(Grid g, T *__restrict__ y) {
    const int nz = g.nn[2], rows = g.nn[1] - g.ydet, slab = rows * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = g.ydet + r / nz, iz = r % nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    y[n] = T(0);
    y[g.nnodes + n] = T(0);
    y[2 * g.nnodes + n] = T(0);
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fine_gs(Grid g, const T *__restrict__ K0, const T *__restrict__ E,
                                                 T *__restrict__ u, const T *__restrict__ b,
                                                 const uint8_t *__restrict__ dmask, int cx, int cy, int cz,
                                                 int forward) {
    const int mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int jx = t / (my * mz), r = t - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = g.nnodes, n = (ix * g.nn[1] + iy) * g.nn[2] + iz;
    const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
    if (dm == 7) return;
    T un[27][3];
    gather27(g, u, ix, iy, iz, un);
    T f[3] = {0, 0, 0};
    T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    fine_row<T, true>(g, K0, E, ix, iy, iz, un, f, M);
    const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
    gs_update(u, N, n, bms, M, dm, forward);
}

// ------------------------------------------------------------------ level 1 (fine E, Chat)
// Coarse element e with centre-node local index ln: Ke = sum_fi E_f Chat_fi (MultigridSolver.hh:483-491).
template <typename T, bool WANT_M, bool COLS>
__device__ __forceinline__ void l1_row(const Grid &gc, const Grid &gf, const T *__restrict__ Chat,
                                       const T *__restrict__ Ef, int ix, int iy, int iz, const T un[27][3], T f[3],
                                       T M[9]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
        const int ex = ix - i, ey = iy - j, ez = iz - k;
        const bool ok = (unsigned)ex < (unsigned)gc.ne[0] && (unsigned)ey < (unsigned)gc.ne[1] &&
                        (unsigned)ez < (unsigned)gc.ne[2];
        if (!ok) continue;
        const int ln = 4 * i + 2 * j + k;
        T ue[24];
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) ue[3 * m + d] = un[slot_of(m, i, j, k)][d];
        const int fbase = eidx(gf, 2 * ex, 2 * ey, 2 * ez);
#pragma unroll
        for (int fi = 0; fi < 8; ++fi) {
            const T Efi = Ef[fbase + ((fi >> 2) & 1) * gf.ne[1] * gf.ne[2] + ((fi >> 1) & 1) * gf.ne[2] + (fi & 1)];
            const T *C = Chat + fi * 576;
            T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
            for (int bb = 0; bb < 24; ++bb) {
                if (COLS) {  // Ke[:, 3ln+c]^T u (smoother form, MultigridSolver.hh:310-314)
                    a0 += C[bb * 24 + 3 * ln + 0] * ue[bb];
                    a1 += C[bb * 24 + 3 * ln + 1] * ue[bb];
                    a2 += C[bb * 24 + 3 * ln + 2] * ue[bb];
                } else {     // Ke[3ln+c, :] u (operator form, MultigridSolver.hh:488-491)
                    a0 += C[(3 * ln + 0) * 24 + bb] * ue[bb];
                    a1 += C[(3 * ln + 1) * 24 + bb] * ue[bb];
                    a2 += C[(3 * ln + 2) * 24 + bb] * ue[bb];
                }
            }
            f[0] += Efi * a0;
            f[1] += Efi * a1;
            f[2] += Efi * a2;
            if (WANT_M) {
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int c2 = 0; c2 < 3; ++c2) M[3 * c + c2] += Efi * C[(3 * ln + c) * 24 + 3 * ln + c2];
            }
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_l1_apply(Grid gc, Grid gf, const T *__restrict__ Chat,
                                                  const T *__restrict__ Ef, const T *__restrict__ u,
                                                  T *__restrict__ y, const T *__restrict__ b,
                                                  const uint8_t *__restrict__ dmask) {
    const int nz = gc.nn[2], slab = gc.ydet * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= gc.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = r / nz, iz = r - iy * nz;
    const int n = (ix * gc.nn[1] + iy) * nz + iz;
    T un[27][3];
    gather27(gc, u, ix, iy, iz, un);
    T f[3] = {0, 0, 0};
    T M[9];
    l1_row<T, false, false>(gc, gf, Chat, Ef, ix, iy, iz, un, f, M);
    store_out(y, b, gc.nnodes, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_l1_gs(Grid gc, Grid gf, const T *__restrict__ Chat,
                                               const T *__restrict__ Ef, T *__restrict__ u, const T *__restrict__ b,
                                               const uint8_t *__restrict__ dmask, int cx, int cy, int cz,
                                               int forward) {
    const int mx = (gc.nn[0] - cx + 1) / 2, my = (gc.ydet - cy + 1) / 2, mz = (gc.nn[2] - cz + 1) / 2;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int jx = t / (my * mz), r = t - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = gc.nnodes, n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
    const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
    if (dm == 7) return;
    T un[27][3];
    gather27(gc, u, ix, iy, iz, un);
    T f[3] = {0, 0, 0};
    T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    l1_row<T, true, true>(gc, gf, Chat, Ef, ix, iy, iz, un, f, M);
    const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
    gs_update(u, N, n, bms, M, dm, forward);
}

// ---- level-1 K u element by element in the Walsh basis (reflection-symmetric element matrix).
// With the Walsh basis W[l, k] = prod_{axes a in k} s_a(l) (s_a(l) = +-1 the side of local node l along a,
// per displacement component), the child matrices Chat_fi = R_fi Chat_0 R_fi (R_fi: reflection across the
// axes of fi, which permutes the nodes and flips the reflected displacement components) become sign flips
// of Chat_0 in that basis, so the coarse element matrix sum_fi E_fi Chat_fi reads
//   (W^T K_E W)[(k, a), (k', c)] = Cw[(k, a), (k', c)] Ehat[k ^ k' ^ e_a ^ e_c],
//   Cw = W^T Chat_0 W,  Ehat[d] = sum_fi (-1)^popcount(d & fi) E_fi,  e_a = the bit of axis a,
// and K_E u_E = W (Cw o Ehat) W^T u_E / 64 costs about 800 operations for all 24 outputs instead of
// 8 x 576 for the child products (W W^T = 8 I; the 1/64 is folded into the table Cw).
// A workgroup owns a tile of WX x WY x WZ nodes, evaluates every element touching it into LDS and sums
// each owned node's eight element contributions in a fixed order (no atomics: deterministic).
constexpr int WX = 4, WY = 8, WZ = 8, WNT = WX * WY * WZ;
constexpr int WEX = WX + 1, WEY = WY + 1, WEZ = WZ + 1, WNE = WEX * WEY * WEZ;
static_assert(WNT == TPB, "one owned node per work-item");

// v: the element's nodal values [l][a] on entry, K_E v on exit
template <typename T>
__device__ __forceinline__ void l1_walsh_element(const T *__restrict__ Cw, const T Ef8[8], T v[8][3]) {
    T Eh[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) Eh[i] = Ef8[i];
    // Ehat = H E (H[d][fi] = (-1)^popcount(d & fi)), and w = W^T v, one axis at a time
#pragma unroll
    for (int bit = 4; bit >= 1; bit >>= 1)
#pragma unroll
        for (int lo = 0; lo < 8; ++lo) {
            if (lo & bit) continue;
            const int hi = lo | bit;
            const T e0 = Eh[lo], e1 = Eh[hi];
            Eh[lo] = e0 + e1;
            Eh[hi] = e0 - e1;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T x0 = v[lo][a], x1 = v[hi][a];
                v[lo][a] = x0 + x1;
                v[hi][a] = x1 - x0;
            }
        }
    // out[k][a] = sum_d Ehat[d] sum_c Cw[(k, a), (k', c)] w[k'][c] with k' = k ^ d ^ e_a ^ e_c; rows and
    // columns of the constant mode k = 0 vanish (rigid translations)
    T out[8][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) out[0][a] = T(0);
#pragma unroll
    for (int k = 1; k < 8; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T acc = T(0);
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                T inner = T(0);
                bool any = false;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int kp = k ^ d ^ (4 >> a) ^ (4 >> c);
                    if (kp == 0) continue;
                    inner += Cw[(3 * k + a) * 24 + 3 * kp + c] * v[kp][c];
                    any = true;
                }
                if (any) acc += Eh[d] * inner;
            }
            out[k][a] = acc;
        }
    // back to the nodes: v = W out
#pragma unroll
    for (int bit = 4; bit >= 1; bit >>= 1)
#pragma unroll
        for (int lo = 0; lo < 8; ++lo) {
            if (lo & bit) continue;
            const int hi = lo | bit;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T x0 = out[lo][a], x1 = out[hi][a];
                out[lo][a] = x0 - x1;
                out[hi][a] = x0 + x1;
            }
        }
#pragma unroll
    for (int l = 0; l < 8; ++l)
#pragma unroll
        for (int a = 0; a < 3; ++a) v[l][a] = out[l][a];
}

// y (mode) K u over the nodes of rows below ydet (detached rows are left to the launcher); grid: one
// workgroup per node tile
template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_l1_apply_w(Grid gc, Grid gf, const T *__restrict__ Cw,
                                                    const T *__restrict__ Ef, const T *__restrict__ u,
                                                    T *__restrict__ y, const T *__restrict__ b,
                                                    const uint8_t *__restrict__ dmask, int ntY, int ntZ) {
    __shared__ T sY[WNE * 24];
    const int N = gc.nnodes, sx = gc.nn[1] * gc.nn[2], sy = gc.nn[2];
    const int fz = gf.ne[2], fyz = gf.ne[1] * gf.ne[2];
    const int tz = blockIdx.x % ntZ, r0 = blockIdx.x / ntZ, ty = r0 % ntY, tx = r0 / ntY;
    const int x0 = tx * WX, y0 = ty * WY, z0 = tz * WZ;
    const int t = threadIdx.x;
    for (int el = t; el < WNE; el += TPB) {
        const int lx = el / (WEY * WEZ), rr = el - lx * WEY * WEZ, ly = rr / WEZ, lz = rr - ly * WEZ;
        const int ex = x0 - 1 + lx, ey = y0 - 1 + ly, ez = z0 - 1 + lz;
        T v[8][3];
        if ((unsigned)ex < (unsigned)gc.ne[0] && (unsigned)ey < (unsigned)gc.ne[1] && (unsigned)ez < (unsigned)gc.ne[2]) {
            const int n0 = (ex * gc.nn[1] + ey) * gc.nn[2] + ez;
            T Ef8[8];
            const int fb = eidx(gf, 2 * ex, 2 * ey, 2 * ez);
#pragma unroll
            for (int l = 0; l < 8; ++l) {
                const int nm = n0 + ((l >> 2) & 1) * sx + ((l >> 1) & 1) * sy + (l & 1);
#pragma unroll
                for (int a = 0; a < 3; ++a) v[l][a] = u[(size_t)a * N + nm];
                Ef8[l] = Ef[fb + ((l >> 2) & 1) * fyz + ((l >> 1) & 1) * fz + (l & 1)];
            }
            l1_walsh_element(Cw, Ef8, v);
        } else {
#pragma unroll
            for (int l = 0; l < 8; ++l)
#pragma unroll
                for (int a = 0; a < 3; ++a) v[l][a] = T(0);
        }
#pragma unroll
        for (int l = 0; l < 8; ++l)
#pragma unroll
            for (int a = 0; a < 3; ++a) sY[el * 24 + 3 * l + a] = v[l][a];
    }
    __syncthreads();
    const int ox = t / (WY * WZ), orr = t - ox * WY * WZ, oy = orr / WZ, oz = orr - oy * WZ;
    const int ix = x0 + ox, iy = y0 + oy, iz = z0 + oz;
    if (ix >= gc.nn[0] || iy >= gc.ydet || iz >= gc.nn[2]) return;
    T f[3] = {0, 0, 0};
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        // the element with this node at local corner l: local element coordinates (o + 1 - bit)
        const int el = ((ox + 1 - ((l >> 2) & 1)) * WEY + (oy + 1 - ((l >> 1) & 1))) * WEZ + (oz + 1 - (l & 1));
#pragma unroll
        for (int a = 0; a < 3; ++a) f[a] += sY[el * 24 + 3 * l + a];
    }
    const int n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
    store_out(y, b, N, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

// Level-1 node row from the fine moduli when the element matrix is reflection-symmetric (box element,
// orthotropic tensor): Chat_fi[(l, a), (m, c)] = S_fi(a) S_fi(c) Chat_0[(l ^ fi, a), (m ^ fi, c)] with
// S_fi(a) = -1 iff fi has the bit of axis a set. With g = ln ^ fi and m' = m ^ fi the row of the centre
// node (local index ln = e in element e) reads
//   f[a] = sum_g sum_e E_{e, e^g} S(a) sum_{m', c} Chat_0[(g, a), (m', c)] S(c) u_{e, m' ^ e ^ g, c},
// so each Chat_0 entry (a uniform, scalar operand) serves the eight elements in turn, the scalar working set
// is one child's 24 x 24 matrix, and the signs are free source negations. Absent elements read zero moduli.
// The row is split over two work-items: work-item h of a pair takes the four elements with x bit i = h.
// Its element loop stays unrolled over (j, k); the runtime x bit is folded away by loading the element's
// node values in x-mirrored order when h = 1 (node m' holds local node m' ^ 4h, x component negated), so
// Chat_0 rows, register indices and signs stay compile-time, and the remaining runtime sign (-1)^h is
// applied once to the x-row of f and the x-off-diagonal entries of M. The pair sums its halves with one
// lane exchange: twice the work-items per colour and half the work each, so a colour's launch has twice as
// many waves in flight (the smoother's launches cover one colour, an eighth of the nodes, each).
template <typename T, bool WANT_M>
__device__ __forceinline__ void l1_row_refl_half(const Grid &gc, const Grid &gf, const T *__restrict__ C0,
                                                 const T *__restrict__ Ef, const T *__restrict__ u, int ix, int iy,
                                                 int iz, int h, T f[3], T M[9]) {
    const int N = gc.nnodes, sx = gc.nn[1] * gc.nn[2], sy = gc.nn[2];
    const int fz = gf.ne[2], fyz = gf.ne[1] * gf.ne[2];
    const int n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
    const int sxh = h ? -sx : sx;
    const int fyzA = h ? fyz : 0, fyzB = h ? 0 : fyz;  // x offset of a child whose reflected x bit is 0 / 1
    const T hs = h ? T(-1) : T(1);
    const int ex = ix - h;
    const bool xin = (unsigned)ex < (unsigned)gc.ne[0];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int j = (e >> 1) & 1, k = e & 1;
        const int ey = iy - j, ez = iz - k;
        const bool ok = xin && (unsigned)ey < (unsigned)gc.ne[1] && (unsigned)ez < (unsigned)gc.ne[2];
        if (!ok) continue;
        T ue[8][3];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            // local node m ^ 4h of the element at x offset (m_x - h) = (m >> 2 & 1) * (h ? -1 : 1)
            const int nm = n + ((m >> 2) & 1) * sxh + (((m >> 1) & 1) - j) * sy + ((m & 1) - k);
            ue[m][0] = hs * u[nm];
            ue[m][1] = u[(size_t)N + nm];
            ue[m][2] = u[(size_t)2 * N + nm];
        }
        const int fb = eidx(gf, 2 * ex, 2 * ey, 2 * ez);
#pragma unroll
        for (int g = 0; g < 8; ++g) {
            const int fis = (g & 4) | ((e ^ g) & 3);  // the child index with its x bit taken from g
            const T Efi = Ef[fb + ((g >> 2) & 1 ? fyzB : fyzA) + ((fis >> 1) & 1) * fz + (fis & 1)];
            T t[3] = {0, 0, 0};
#pragma unroll
            for (int mp = 0; mp < 8; ++mp)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T uv = ((fis >> (2 - c)) & 1) ? -ue[mp ^ fis][c] : ue[mp ^ fis][c];
#pragma unroll
                    for (int a = 0; a < 3; ++a) t[a] += C0[(3 * g + a) * 24 + 3 * mp + c] * uv;
                }
#pragma unroll
            for (int a = 0; a < 3; ++a) f[a] += (((fis >> (2 - a)) & 1) ? -Efi : Efi) * t[a];
            if (WANT_M) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const bool neg = (bool)((fis >> (2 - a)) & 1) != (bool)((fis >> (2 - c)) & 1);
                        M[3 * a + c] += (neg ? -Efi : Efi) * C0[(3 * g + a) * 24 + 3 * g + c];
                    }
            }
        }
    }
    f[0] *= hs;
    if (WANT_M) {
        M[1] *= hs;
        M[2] *= hs;
        M[3] *= hs;
        M[6] *= hs;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] += __shfl_xor(f[a], 1, 64);
    if (WANT_M)
#pragma unroll
        for (int q = 0; q < 9; ++q) M[q] += __shfl_xor(M[q], 1, 64);
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_l1_gs_r2(Grid gc, Grid gf, const T *__restrict__ C0,
                                                  const T *__restrict__ Ef, T *__restrict__ u, const T *__restrict__ b,
                                                  const uint8_t *__restrict__ dmask, int cx, int cy, int cz,
                                                  int forward) {
    const int mx = (gc.nn[0] - cx + 1) / 2, my = (gc.ydet - cy + 1) / 2, mz = (gc.nn[2] - cz + 1) / 2;
    const int tt = blockIdx.x * TPB + threadIdx.x, t = tt >> 1, h = tt & 1;
    if (t >= mx * my * mz) return;  // both work-items of a pair
    const int jx = t / (my * mz), r = t - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = gc.nnodes, n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
    const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
    if (dm == 7) return;  // both work-items of a pair
    T f[3] = {0, 0, 0};
    T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    l1_row_refl_half<T, true>(gc, gf, C0, Ef, u, ix, iy, iz, h, f, M);
    if (h) return;
    const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
    gs_update(u, N, n, bms, M, dm, forward);
}

// ------------------------------------------------------------------ assembled block rows
// Colour-major block storage (see BlockSlot). The neighbours are visited GROUP at a time with the
// accumulators pinned after each group, which bounds how many of the 243 block entries and 81 u values
// the compiler keeps in flight (and so the register count) while still overlapping memory latency.
// Large levels use groups of 3 (occupancy); small levels, whose launches are latency-bound, groups of 9
// (three dependent memory rounds per row instead of nine).
template <typename T, bool WANT_M, int GROUP = 3>
__device__ __forceinline__ void block_row(const Grid &g, const T *__restrict__ blocks, const BlockSlot &bs, int ix,
                                          int iy, int iz, const T *__restrict__ u, T f[3], T M[9]) {
    const int N = g.nnodes, sx = g.nn[1] * g.nn[2], sy = g.nn[2];
    const int n = (ix * g.nn[1] + iy) * g.nn[2] + iz;
    const T *__restrict__ bp = blocks + bs.base + bs.j;
    const size_t st = (size_t)bs.cnt;
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        const int dx = s / 9 - 1, dy = (s / 3) % 3 - 1, dz = s % 3 - 1;
        const bool ok = (unsigned)(ix + dx) < (unsigned)g.nn[0] && (unsigned)(iy + dy) < (unsigned)g.nn[1] &&
                        (unsigned)(iz + dz) < (unsigned)g.nn[2];
        const int m = n + dx * sx + dy * sy + dz;
        T un[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) un[c] = ok ? u[(size_t)c * N + m] : T(0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T acc = 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T B = bp[(size_t)(s * 9 + 3 * a + c) * st];
                acc += B * un[c];
                if (WANT_M && s == 13) M[3 * a + c] = B;
            }
            f[a] += acc;
        }
        if (s % GROUP == GROUP - 1) {
            asm volatile("" : "+v"(f[0]), "+v"(f[1]), "+v"(f[2]));
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <typename T, int MODE, int GROUP>
__device__ __forceinline__ void block_apply_row(const Grid &g, const T *__restrict__ blocks, const T *__restrict__ u,
                                                T *__restrict__ y, const T *__restrict__ b,
                                                const uint8_t *__restrict__ dmask) {
    const int nz = g.nn[2], slab = g.ydet * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = r / nz, iz = r - iy * nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    const BlockSlot bs = block_slot(g, ix, iy, iz);
    T f[3] = {0, 0, 0};
    T M[9];
    block_row<T, false, GROUP>(g, blocks, bs, ix, iy, iz, u, f, M);
    store_out(y, b, g.nnodes, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

// Block-row product on a small level with four work-items per node (see k_block_gs_x4); work-item 0
// writes the row.
template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_block_apply_x4(Grid g, const T *__restrict__ blocks, const T *__restrict__ u,
                                                        T *__restrict__ y, const T *__restrict__ b,
                                                        const uint8_t *__restrict__ dmask) {
    const int nz = g.nn[2], slab = g.ydet * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    const int node = t >> 2, part = t & 3;
    const bool valid = node < g.nn[0] * slab;
    const int ix = valid ? node / slab : 0, r = valid ? node - ix * slab : 0, iy = r / nz, iz = r - iy * nz;
    const int N = g.nnodes, n = (ix * g.nn[1] + iy) * nz + iz;
    T f[3] = {0, 0, 0};
    if (valid && part < 3) {
        const BlockSlot bs = block_slot(g, ix, iy, iz);
        const T *__restrict__ bp = blocks + bs.base + bs.j;
        const size_t st = (size_t)bs.cnt;
        const int sx = g.nn[1] * nz;
        const int dx = part - 1;
        const bool okx = (unsigned)(ix + dx) < (unsigned)g.nn[0];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int s = 9 * part + q, dy = q / 3 - 1, dz = q % 3 - 1;
            const bool ok = okx && (unsigned)(iy + dy) < (unsigned)g.nn[1] && (unsigned)(iz + dz) < (unsigned)nz;
            const int m = n + dx * sx + dy * nz + dz;
            T un[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) un[c] = ok ? u[(size_t)c * N + m] : T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                T acc = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) acc += bp[(size_t)(s * 9 + 3 * a + c) * st] * un[c];
                f[a] += acc;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        f[c] += __shfl_xor(f[c], 1, 64);
        f[c] += __shfl_xor(f[c], 2, 64);
    }
    if (valid && part == 0) store_out(y, b, N, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_block_apply_g9(Grid g, const T *__restrict__ blocks, const T *__restrict__ u,
                                                        T *__restrict__ y, const T *__restrict__ b,
                                                        const uint8_t *__restrict__ dmask) {
    block_apply_row<T, MODE, 9>(g, blocks, u, y, b, dmask);
}

template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_block_apply(Grid g, const T *__restrict__ blocks, const T *__restrict__ u,
                                                     T *__restrict__ y, const T *__restrict__ b,
                                                     const uint8_t *__restrict__ dmask) {
    const int nz = g.nn[2], slab = g.ydet * nz;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = r / nz, iz = r - iy * nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    const BlockSlot bs = block_slot(g, ix, iy, iz);
    T f[3] = {0, 0, 0};
    T M[9];
    block_row<T, false>(g, blocks, bs, ix, iy, iz, u, f, M);
    store_out(y, b, g.nnodes, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
}

// One colour of the block Gauss-Seidel sweep on a small level, four work-items per node: work-item
// g < 3 of a node gathers the nine block rows of neighbour plane x + g - 1, the partial rows are summed
// across the four with two lane exchanges, and work-item 1 (which holds the diagonal block) updates the
// node. The per-node dependency chain is a third of the one-work-item form, which bounds these
// latency-limited launches.
template <typename T>
__global__ void __launch_bounds__(TPB) k_block_gs_x4(Grid g, const T *__restrict__ blocks, T *__restrict__ u,
                                                     const T *__restrict__ b, const uint8_t *__restrict__ dmask,
                                                     int cx, int cy, int cz, int forward) {
    const int mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    const int t = blockIdx.x * TPB + threadIdx.x;
    const int node = t >> 2, part = t & 3;
    const bool valid = node < mx * my * mz;
    const int jx = valid ? node / (my * mz) : 0, r = valid ? node - jx * my * mz : 0, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = g.nnodes, n = (ix * g.nn[1] + iy) * g.nn[2] + iz;
    const uint8_t dm = valid ? (dmask ? dmask[n] : uint8_t(0)) : uint8_t(7);
    T f[3] = {0, 0, 0};
    T M[9];
    if (dm != 7 && part < 3) {
        const BlockSlot bs = block_slot(g, ix, iy, iz);
        const T *__restrict__ bp = blocks + bs.base + bs.j;
        const size_t st = (size_t)bs.cnt;
        const int sx = g.nn[1] * g.nn[2], sy = g.nn[2];
        const int dx = part - 1;
        const bool okx = (unsigned)(ix + dx) < (unsigned)g.nn[0];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int s = 9 * part + q, dy = q / 3 - 1, dz = q % 3 - 1;
            const bool ok = okx && (unsigned)(iy + dy) < (unsigned)g.nn[1] && (unsigned)(iz + dz) < (unsigned)g.nn[2];
            const int m = n + dx * sx + dy * sy + dz;
            T un[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) un[c] = ok ? u[(size_t)c * N + m] : T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                T acc = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T B = bp[(size_t)(s * 9 + 3 * a + c) * st];
                    acc += B * un[c];
                    if (q == 4) M[3 * a + c] = B;  // slot 13 (the node itself) when part == 1
                }
                f[a] += acc;
            }
        }
    }
    // every lane takes part in the exchanges (idle ones contribute zeros)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        f[c] += __shfl_xor(f[c], 1, 64);
        f[c] += __shfl_xor(f[c], 2, 64);
    }
    if (dm != 7 && part == 1) {
        const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
        gs_update(u, N, n, bms, M, dm, forward);
    }
}

template <typename T, int GROUP = 3>
__global__ void __launch_bounds__(TPB) k_block_gs(Grid g, const T *__restrict__ blocks, T *__restrict__ u,
                                                  const T *__restrict__ b, const uint8_t *__restrict__ dmask, int cx,
                                                  int cy, int cz, int forward) {
    const int mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int jx = t / (my * mz), r = t - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = g.nnodes, n = (ix * g.nn[1] + iy) * g.nn[2] + iz;
    const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
    if (dm == 7) return;
    const BlockSlot bs = block_slot(g, ix, iy, iz);
    T f[3] = {0, 0, 0};
    T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    block_row<T, true, GROUP>(g, blocks, bs, ix, iy, iz, u, f, M);
    const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
    gs_update(u, N, n, bms, M, dm, forward);
}

// ------------------------------------------------------------------ cached element matrices
template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_elem_apply(Grid g, const T *__restrict__ Ke, const T *__restrict__ u,
                                                    T *__restrict__ y) {
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= g.nnodes) return;
    const int nz = g.nn[2], ix = t / (g.nn[1] * nz), iy = (t / nz) % g.nn[1], iz = t % nz;
    T un[27][3];
    gather27(g, u, ix, iy, iz, un);
    T f[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
        const int ex = ix - i, ey = iy - j, ez = iz - k;
        if (!((unsigned)ex < (unsigned)g.ne[0] && (unsigned)ey < (unsigned)g.ne[1] && (unsigned)ez < (unsigned)g.ne[2]))
            continue;
        const int ln = 4 * i + 2 * j + k;
        const T *K = Ke + (size_t)eidx(g, ex, ey, ez) * 576;
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T um = un[slot_of(m, i, j, k)][d];
#pragma unroll
                for (int c = 0; c < 3; ++c) f[c] += K[(3 * ln + c) * 24 + 3 * m + d] * um;
            }
    }
    store_out(y, (const T *)nullptr, g.nnodes, t, f, MODE, uint8_t(0));
}

// ------------------------------------------------------------------ transfers
// r_c(n) = sum_{o in {-1,0,1}^3} prod_d (1 - |o_d|/2) r_f(2n + o) over active fine nodes
// (MultigridSolver.hh:214-275); every detached coarse row is zeroed. The weights are separable, so a
// work-item owns one coarse (y, z) column and marches along x through its chunk of coarse planes: each
// fine plane's 3x3 in-plane sum is formed once and serves the coarse planes on both sides of it.
// xc (optional): the coarse iterate, zeroed in the same pass (the cycle starts the coarse level from zero)
template <typename T>
__global__ void __launch_bounds__(TPB) k_restrict(Grid gf, Grid gc, const T *__restrict__ rf, T *__restrict__ rc,
                                                  T *__restrict__ xc, int xchunk) {
    const int ncy = gc.nn[1], ncz = gc.nn[2];
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= ncy * ncz) return;
    const int cy = t / ncz, cz = t - cy * ncz;
    const int cx0 = blockIdx.y * xchunk, cx1 = min(cx0 + xchunk, gc.nn[0]);
    const bool act = cy < gc.ydet;
    const int Nf = gf.nnodes, Nc = gc.nnodes, nfy = gf.nn[1], nfz = gf.nn[2];
    auto plane = [&](int fx, T out[3]) {
        out[0] = out[1] = out[2] = T(0);
        if (!act || (unsigned)fx >= (unsigned)gf.nn[0]) return;
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy) {
            const int fy = 2 * cy + oy;
            if ((unsigned)fy >= (unsigned)gf.ydet) continue;
#pragma unroll
            for (int oz = -1; oz <= 1; ++oz) {
                const int fz = 2 * cz + oz;
                if ((unsigned)fz >= (unsigned)nfz) continue;
                const T w = T((oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0));
                const int m = (fx * nfy + fy) * nfz + fz;
                out[0] += w * rf[m];
                out[1] += w * rf[Nf + m];
                out[2] += w * rf[2 * Nf + m];
            }
        }
    };
    T prev[3];
    plane(2 * cx0 - 1, prev);
    for (int cx = cx0; cx < cx1; ++cx) {
        T a[3], b[3];
        plane(2 * cx, a);
        plane(2 * cx + 1, b);
        const int n = (cx * ncy + cy) * ncz + cz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rc[(size_t)c * Nc + n] = T(0.5) * prev[c] + a[c] + T(0.5) * b[c];
            if (xc) xc[(size_t)c * Nc + n] = T(0);
            prev[c] = b[c];
        }
    }
}

// x_f (+)= P x_c, trilinear (MultigridSolver.hh:162-212). accumulate: active fine rows only
// (accum_interpolation); otherwise every fine row is overwritten (interpolation). A work-item owns one
// fine (y, z) column and marches along x through its chunk of coarse elements: the bilinear (y, z)
// interpolant of each coarse plane is formed once and serves the fine planes on both sides of it.
template <typename T>
__global__ void __launch_bounds__(TPB) k_interpolate(Grid gf, Grid gc, const T *__restrict__ xc, T *__restrict__ xf,
                                                     int accumulate, int xchunk) {
    const int nfy = gf.nn[1], nfz = gf.nn[2];
    const int rows = accumulate ? gf.ydet : nfy;
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= rows * nfz) return;
    const int fy = t / nfz, fz = t - fy * nfz;
    const int ncy = gc.nn[1], ncz = gc.nn[2], Nc = gc.nnodes, Nf = gf.nnodes;
    // coarse rows / columns of the (y, z) interpolant and their weights (the second index is clamped
    // where its weight is zero)
    const int cy = min(fy >> 1, gc.ne[1]), cz = min(fz >> 1, gc.ne[2]);
    const bool oy = fy & 1, oz = fz & 1;
    const T wy1 = oy ? T(0.5) : T(0), wy0 = oy ? T(0.5) : T(1);
    const T wz1 = oz ? T(0.5) : T(0), wz0 = oz ? T(0.5) : T(1);
    const int cy1 = oy ? cy + 1 : cy, cz1 = oz ? cz + 1 : cz;
    const int o00 = cy * ncz + cz, o01 = cy * ncz + cz1, o10 = cy1 * ncz + cz, o11 = cy1 * ncz + cz1;
    auto inplane = [&](int cx, T out[3]) {
        const int pb = cx * ncy * ncz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T *p = xc + (size_t)c * Nc + pb;
            out[c] = wy0 * (wz0 * p[o00] + wz1 * p[o01]) + wy1 * (wz0 * p[o10] + wz1 * p[o11]);
        }
    };
    auto put = [&](int fx, const T v[3]) {
        const size_t n = ((size_t)fx * nfy + fy) * nfz + fz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (accumulate) xf[(size_t)c * Nf + n] += v[c];
            else xf[(size_t)c * Nf + n] = v[c];
        }
    };
    const int ne0 = gc.ne[0];
    const int e0 = blockIdx.y * xchunk, e1 = min(e0 + xchunk, ne0);
    T a[3];
    inplane(e0, a);
    for (int e = e0; e < e1; ++e) {
        T b[3];
        inplane(e + 1, b);
        put(2 * e, a);
        const T h[3] = {T(0.5) * (a[0] + b[0]), T(0.5) * (a[1] + b[1]), T(0.5) * (a[2] + b[2])};
        put(2 * e + 1, h);
        a[0] = b[0];
        a[1] = b[1];
        a[2] = b[2];
    }
    if (e1 == ne0) put(2 * ne0, a);  // the last fine plane, on the last coarse plane
}

inline unsigned nblocks(long count) { return (unsigned)((count + TPB - 1) / TPB); }


}  // namespace

// ------------------------------------------------------------------ launchers
#define VF_DISPATCH_MODE(KERNEL, T, mode, grid, ...)                                                        \
    switch (mode) {                                                                                         \
    case AM_SET: hipLaunchKernelGGL((KERNEL<T, AM_SET>), grid, dim3(TPB), 0, s, __VA_ARGS__); break;         \
    case AM_ADD: hipLaunchKernelGGL((KERNEL<T, AM_ADD>), grid, dim3(TPB), 0, s, __VA_ARGS__); break;         \
    case AM_SUB: hipLaunchKernelGGL((KERNEL<T, AM_SUB>), grid, dim3(TPB), 0, s, __VA_ARGS__); break;         \
    default: hipLaunchKernelGGL((KERNEL<T, AM_RESIDUAL>), grid, dim3(TPB), 0, s, __VA_ARGS__); break;        \
    }

template <typename T>
void launch_zero_detached(const Grid &g, T *y, hipStream_t s) {
    if (g.ydet >= g.nn[1]) return;
    const long cz = (long)g.nn[0] * (g.nn[1] - g.ydet) * g.nn[2];
    hipLaunchKernelGGL(k_zero_detached<T>, dim3(nblocks(cz)), dim3(TPB), 0, s, g, y);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_fine_apply(const Grid &g, const T *K0, const T *E, const T *u, T *y, const T *b, const uint8_t *dmask,
                       int mode, hipStream_t s) {
    const long count = (long)g.nn[0] * g.ydet * g.nn[2];
    if (count > 0) { VF_DISPATCH_MODE(k_fine_apply, T, mode, dim3(nblocks(count)), g, K0, E, u, y, b, dmask); }
    if (mode == AM_SET && g.ydet < g.nn[1]) {
        const long cz = (long)g.nn[0] * (g.nn[1] - g.ydet) * g.nn[2];
        hipLaunchKernelGGL(k_zero_detached<T>, dim3(nblocks(cz)), dim3(TPB), 0, s, g, y);
    }
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_l1_apply(const Grid &gc, const Grid &gf, const T *Chat, const T *Cw, const T *Ef, const T *u, T *y,
                     const T *b, const uint8_t *dmask, int mode, bool reflect, hipStream_t s) {
    const long count = (long)gc.nn[0] * gc.ydet * gc.nn[2];
    if (count > 0 && reflect) {
        if (!Cw) throw std::runtime_error("the Walsh-basis level-1 product needs its table");
        const int ntX = (gc.nn[0] + WX - 1) / WX, ntY = (gc.ydet + WY - 1) / WY, ntZ = (gc.nn[2] + WZ - 1) / WZ;
        const long tiles = (long)ntX * ntY * ntZ;
        if (tiles > INT_MAX) throw std::runtime_error("grid too large for the level-1 product");
        VF_DISPATCH_MODE(k_l1_apply_w, T, mode, dim3((unsigned)tiles), gc, gf, Cw, Ef, u, y, b, dmask, ntY, ntZ);
    } else if (count > 0) {
        VF_DISPATCH_MODE(k_l1_apply, T, mode, dim3(nblocks(count)), gc, gf, Chat, Ef, u, y, b, dmask);
    }
    if (mode == AM_SET && gc.ydet < gc.nn[1]) {
        const long cz = (long)gc.nn[0] * (gc.nn[1] - gc.ydet) * gc.nn[2];
        hipLaunchKernelGGL(k_zero_detached<T>, dim3(nblocks(cz)), dim3(TPB), 0, s, gc, y);
    }
    VF_HIP(hipGetLastError());
}

// levels below this many nodes use the latency-oriented grouping of block_row
constexpr int SMALL_LEVEL = 65536;
// levels below this many nodes smooth with four work-items per node (latency-bound launches); up to
// GS_G9_LEVEL nodes, one work-item per node with nine block rows requested at a time
constexpr int GS_X4_LEVEL = 1 << 19, GS_G9_LEVEL = 1 << 22;

template <typename T>
void launch_block_apply(const Grid &g, const T *blocks, const T *u, T *y, const T *b, const uint8_t *dmask, int mode,
                        hipStream_t s) {
    const long count = (long)g.nn[0] * g.ydet * g.nn[2];
    if (count > 0 && g.nnodes < GS_X4_LEVEL) {
        VF_DISPATCH_MODE(k_block_apply_x4, T, mode, dim3(nblocks(4 * count)), g, blocks, u, y, b, dmask);
    } else if (count > 0 && g.nnodes < SMALL_LEVEL) {
        VF_DISPATCH_MODE(k_block_apply_g9, T, mode, dim3(nblocks(count)), g, blocks, u, y, b, dmask);
    } else if (count > 0) {
        VF_DISPATCH_MODE(k_block_apply, T, mode, dim3(nblocks(count)), g, blocks, u, y, b, dmask);
    }
    if (mode == AM_SET && g.ydet < g.nn[1]) {
        const long cz = (long)g.nn[0] * (g.nn[1] - g.ydet) * g.nn[2];
        hipLaunchKernelGGL(k_zero_detached<T>, dim3(nblocks(cz)), dim3(TPB), 0, s, g, y);
    }
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_elem_apply(const Grid &g, const T *Ke, const T *u, T *y, int mode, hipStream_t s) {
    if (mode == AM_SET) hipLaunchKernelGGL((k_elem_apply<T, AM_SET>), dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, Ke, u, y);
    else if (mode == AM_ADD) hipLaunchKernelGGL((k_elem_apply<T, AM_ADD>), dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, Ke, u, y);
    else if (mode == AM_SUB) hipLaunchKernelGGL((k_elem_apply<T, AM_SUB>), dim3(nblocks(g.nnodes)), dim3(TPB), 0, s, g, Ke, u, y);
    else throw std::runtime_error("residual mode unsupported on cached element matrices");
    VF_HIP(hipGetLastError());
}

static inline long color_count(const Grid &g, int cx, int cy, int cz) {
    const long mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    return (mx > 0 && my > 0 && mz > 0) ? mx * my * mz : 0;
}

template <typename T>
void launch_fine_gs(const Grid &g, const T *K0, const T *E, T *u, const T *b, const uint8_t *dmask, int color,
                    int forward, hipStream_t s) {
    const int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    const long count = color_count(g, cx, cy, cz);
    if (!count) return;
    hipLaunchKernelGGL(k_fine_gs<T>, dim3(nblocks(count)), dim3(TPB), 0, s, g, K0, E, u, b, dmask, cx, cy, cz, forward);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_l1_gs(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, T *u, const T *b, const uint8_t *dmask,
                  int color, int forward, bool reflect, hipStream_t s) {
    const int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    const long count = color_count(gc, cx, cy, cz);
    if (!count) return;
    if (reflect)
        hipLaunchKernelGGL(k_l1_gs_r2<T>, dim3(nblocks(2 * count)), dim3(TPB), 0, s, gc, gf, Chat, Ef, u, b, dmask, cx,
                           cy, cz, forward);
    else
        hipLaunchKernelGGL(k_l1_gs<T>, dim3(nblocks(count)), dim3(TPB), 0, s, gc, gf, Chat, Ef, u, b, dmask, cx, cy,
                           cz, forward);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_block_gs(const Grid &g, const T *blocks, T *u, const T *b, const uint8_t *dmask, int color, int forward,
                     hipStream_t s) {
    const int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    const long count = color_count(g, cx, cy, cz);
    if (!count) return;
    if (g.nnodes < GS_X4_LEVEL)
        hipLaunchKernelGGL(k_block_gs_x4<T>, dim3(nblocks(4 * count)), dim3(TPB), 0, s, g, blocks, u, b, dmask, cx, cy,
                           cz, forward);
    else if (g.nnodes < GS_G9_LEVEL)
        hipLaunchKernelGGL((k_block_gs<T, 9>), dim3(nblocks(count)), dim3(TPB), 0, s, g, blocks, u, b, dmask, cx, cy, cz,
                           forward);
    else
        hipLaunchKernelGGL((k_block_gs<T, 3>), dim3(nblocks(count)), dim3(TPB), 0, s, g, blocks, u, b, dmask, cx, cy, cz,
                           forward);
    VF_HIP(hipGetLastError());
}

// x-chunks of a marching transfer: about TARGET workgroups in all, at least MINLEN planes per chunk; a
// small level (fewer than TARGET workgroups even at one plane each), whose launch is latency-bound, gets
// one plane per workgroup
static inline int transfer_chunks(unsigned planeBlocks, int len, int minLen) {
    constexpr unsigned TARGET = 2048;
    if ((long)planeBlocks * len <= (long)TARGET) return std::max(1, std::min(len, 65535));
    int c = (int)std::max(1u, TARGET / std::max(1u, planeBlocks));
    c = std::min(c, std::max(1, len / minLen));
    return std::max(1, std::min(c, 65535));
}
template <typename T>
void launch_restrict(const Grid &gf, const Grid &gc, const T *rf, T *rc, hipStream_t s, T *xzero) {
    if (gf.nn[0] != 2 * gc.nn[0] - 1 || gf.nn[1] != 2 * gc.nn[1] - 1 || gf.nn[2] != 2 * gc.nn[2] - 1)
        throw std::runtime_error("restriction: the fine grid is not the refinement of the coarse grid");
    const unsigned pb = nblocks((long)gc.nn[1] * gc.nn[2]);
    const int chunks = transfer_chunks(pb, gc.nn[0], 4);
    const int xchunk = (gc.nn[0] + chunks - 1) / chunks;
    const int nch = (gc.nn[0] + xchunk - 1) / xchunk;
    hipLaunchKernelGGL(k_restrict<T>, dim3(pb, nch), dim3(TPB), 0, s, gf, gc, rf, rc, xzero, xchunk);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_interpolate(const Grid &gf, const Grid &gc, const T *xc, T *xf, int accumulate, hipStream_t s) {
    if (gf.nn[0] != 2 * gc.nn[0] - 1 || gf.nn[1] != 2 * gc.nn[1] - 1 || gf.nn[2] != 2 * gc.nn[2] - 1)
        throw std::runtime_error("interpolation: the fine grid is not the refinement of the coarse grid");
    const long count = (long)(accumulate ? gf.ydet : gf.nn[1]) * gf.nn[2];
    if (!count || gc.ne[0] < 1) return;
    const unsigned pb = nblocks(count);
    const int chunks = transfer_chunks(pb, gc.ne[0], 2);
    const int xchunk = (gc.ne[0] + chunks - 1) / chunks;
    const int nch = (gc.ne[0] + xchunk - 1) / xchunk;
    hipLaunchKernelGGL(k_interpolate<T>, dim3(pb, nch), dim3(TPB), 0, s, gf, gc, xc, xf, accumulate, xchunk);
    VF_HIP(hipGetLastError());
}

#define VF_INST(T)                                                                                                     \
    template void launch_zero_detached<T>(const Grid &, T *, hipStream_t);                                             \
    template void launch_fine_apply<T>(const Grid &, const T *, const T *, const T *, T *, const T *, const uint8_t *,  \
                                       int, hipStream_t);                                                              \
    template void launch_l1_apply<T>(const Grid &, const Grid &, const T *, const T *, const T *, const T *, T *,      \
                                     const T *, const uint8_t *, int, bool, hipStream_t);                              \
    template void launch_block_apply<T>(const Grid &, const T *, const T *, T *, const T *, const uint8_t *, int,       \
                                        hipStream_t);                                                                  \
    template void launch_elem_apply<T>(const Grid &, const T *, const T *, T *, int, hipStream_t);                     \
    template void launch_fine_gs<T>(const Grid &, const T *, const T *, T *, const T *, const uint8_t *, int, int,      \
                                    hipStream_t);                                                                      \
    template void launch_l1_gs<T>(const Grid &, const Grid &, const T *, const T *, T *, const T *, const uint8_t *,    \
                                  int, int, bool, hipStream_t);                                                        \
    template void launch_block_gs<T>(const Grid &, const T *, T *, const T *, const uint8_t *, int, int, hipStream_t); \
    template void launch_restrict<T>(const Grid &, const Grid &, const T *, T *, hipStream_t, T *);                   \
    template void launch_interpolate<T>(const Grid &, const Grid &, const T *, T *, int, hipStream_t);

VF_INST(double)
VF_INST(float)

}  // namespace vf
