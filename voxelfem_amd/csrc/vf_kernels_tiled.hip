// vf_kernels_tiled.hip -- LDS-tiled fine-level operators for gfx950.
//
// * k_ku_tiled: matrix-free K u / residual (TPSStencils.hh:431-728). A workgroup owns a (y,z) tile
//   of the node plane and marches along x; every x-plane of u (3 component planes, one-node halo) is
//   copied once into LDS with coalesced loads and reused for the three output planes that touch it,
//   so HBM sees u about once per product.
//
// * k_gs_phase: the 8-colour block Gauss-Seidel sweep of MultigridSolver.hh:408-458 in two launches.
//   Colours are (x%2, y%2, z%2) applied in the order 0..7 (7..0 for the backward sweep). The four
//   colours of one x-parity only see the other parity's values, so one launch updates every plane of
//   one parity: the tile's in-plane values live in LDS and the four in-plane colours are applied in
//   order, each on the tile grown by the halo that the later colours read ((1,3), (1,2), (0,1), (0,0)
//   nodes in (y,z)). Halo nodes are recomputed from the same inputs by the neighbouring tiles, so the
//   result does not depend on the tiling. The launch reads its own planes from `uin` (state before the
//   sweep) and the neighbour planes from `unb`, and writes its planes to `uout` (`uout` != `uin`); the
//   first launch of a sweep uses unb = uin, the second unb = uout.
//
// Both kernels evaluate node rows from the structured element matrix described at lds_row; the host
// only selects them when the element matrix has that structure.
#include "vf_internal.hpp"

#include <algorithm>

namespace vf {
namespace {

__device__ __forceinline__ int split_lo(int k, int n, int parts) { return (int)(((long)k * n) / parts); }

// Element stiffness structure. For an axis-aligned box element and an orthotropic (in particular
// isotropic) elasticity tensor, K0[(l,c),(l^delta,d)] = sigma(l,c,d) * F[delta][c][d] with
// sigma = 1 for c == d and sigma = s_c(l) s_d(l) otherwise, s_a(l) = +-1 the side of local node l
// along axis a (F is row l = 0 of K0; the host checks the identity on the actual K0). A node row then
// needs only the nine F values of each neighbour offset delta = |o| (18 SGPRs) instead of 576 K0 entries.
// F is laid out as 8 groups of 16 (9 used) so each offset's group is one aligned scalar load.
// constant address space: uniform loads through such pointers become scalar loads
template <typename T>
using CT = __attribute__((address_space(4))) T;

// neighbour offsets s = 9 dx + 3 dy + dz (0..2 each, 1 = same coordinate) grouped by delta
__device__ constexpr int kNeighbourOrder[27] = {13, 12, 14, 10, 16, 9, 11, 15, 17, 4, 22, 3, 5, 21,
                                                 23, 1, 7, 19, 25, 0, 2, 6, 8, 18, 20, 24, 26};
__device__ __forceinline__ constexpr int delta_of(int s) {
    return ((s / 9 != 1) << 2) | (((s / 3) % 3 != 1) << 1) | (s % 3 != 1);
}

__device__ __forceinline__ constexpr int side(int l, int a) { return ((l >> (2 - a)) & 1) ? 1 : -1; }

// Row of the node at LDS neighbourhood pointers pm/p0/pp (planes x-1, x, x+1 at the node; component
// stride CS, row stride RZ) with the incident element moduli Ee[e], e = 4i+2j+k for the element at
// node - (i,j,k). f = (K u) at the node; with WANT_M also M = the node's 3x3 diagonal block.
// Per element e the products are accumulated over the element's nodes m in increasing order and the
// components d in order, then f = sum_e E_e a_e in e order.
template <typename T, bool WANT_M>
__device__ __forceinline__ void lds_row(const T *__restrict__ F, const T *pm, const T *p0, const T *pp, int CS,
                                        int RZ, const T Ee[8], T f[3], T M[9]) {
    T a[8][3];
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e][0] = a[e][1] = a[e][2] = T(0);
    // Software pipeline over the 27 neighbours, taken in groups of equal delta = |offset| so that each
    // group's nine F values are loaded once per row (8 scalar-load groups instead of 27): the three LDS
    // values of neighbour t+1 (and the F values of its group when it starts one) are requested while
    // neighbour t is being accumulated. LDS and scalar loads share one counter, so the wait for t is
    // issued explicitly before the loads of t+1 (otherwise the first use of t would also wait for t+1).
    // Scheduling barriers keep the three phases of a step in order.
    auto ldu = [&](int s, T v[3]) {
        const int dx = s / 9, dy = (s / 3) % 3, dz = s % 3;
        const T *p = dx == 0 ? pm : (dx == 1 ? p0 : pp);
        const int off = (dy - 1) * RZ + (dz - 1);
        v[0] = p[off];
        v[1] = p[CS + off];
        v[2] = p[2 * CS + off];
    };
    auto ldF = [&](int delta, T w[9]) {
        // the opaque copy of the table pointer stops loop-invariant code motion from hoisting all 72
        // F values out of the node loops (they would not fit in SGPRs and get spilled to VGPR lanes)
        const T *Fg = F + 16 * delta;
        asm volatile("" : "+s"(Fg));
        const CT<T> *Fp = (const CT<T> *)Fg;
#pragma unroll
        for (int q = 0; q < 9; ++q) w[q] = Fp[q];
    };
    T nu[3], nF[9], Fs[9];
    ldu(kNeighbourOrder[0], nu);
    ldF(delta_of(kNeighbourOrder[0]), nF);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int s = kNeighbourOrder[t];
        const int dx = s / 9, dy = (s / 3) % 3, dz = s % 3;
        const bool newGroup = t == 0 || delta_of(kNeighbourOrder[t - 1]) != delta_of(s);
        const bool nextGroup = t + 1 < 27 && delta_of(kNeighbourOrder[t + 1]) != delta_of(s);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt untouched
        __builtin_amdgcn_sched_barrier(0);
        T u[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) u[q] = nu[q];
        if (newGroup) {
#pragma unroll
            for (int q = 0; q < 9; ++q) Fs[q] = nF[q];
        }
        if (t + 1 < 27) ldu(kNeighbourOrder[t + 1], nu);
        if (nextGroup) ldF(delta_of(kNeighbourOrder[t + 1]), nF);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
            const int mx = dx - 1 + i, my = dy - 1 + j, mz = dz - 1 + k;
            if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int sg = c == d ? 1 : side(e, c) * side(e, d);
                    if (sg > 0) a[e][c] += Fs[3 * c + d] * u[d];
                    else a[e][c] -= Fs[3 * c + d] * u[d];
                }
        }
        // pin this neighbour's products here: without it the compiler sinks every accumulation chain to
        // the end of the row and keeps all 81 neighbour values live
#pragma unroll
        for (int e = 0; e < 8; ++e) asm volatile("" : "+v"(a[e][0]), "+v"(a[e][1]), "+v"(a[e][2]));
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        f[0] += Ee[e] * a[e][0];
        f[1] += Ee[e] * a[e][1];
        f[2] += Ee[e] * a[e][2];
    }
    if (WANT_M) {
        // M = sum_e E_e K0[(l,c),(l,d)] with l = e: F[0] with the sign pattern of each element
        T es = 0, exy = 0, exz = 0, eyz = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            es += Ee[e];
            exy += side(e, 0) * side(e, 1) > 0 ? Ee[e] : -Ee[e];
            exz += side(e, 0) * side(e, 2) > 0 ? Ee[e] : -Ee[e];
            eyz += side(e, 1) * side(e, 2) > 0 ? Ee[e] : -Ee[e];
        }
        const T *Fg = F;
        asm volatile("" : "+s"(Fg));
        const CT<T> *F0 = (const CT<T> *)Fg;
        M[0] = F0[0] * es;
        M[4] = F0[4] * es;
        M[8] = F0[8] * es;
        M[1] = F0[1] * exy;
        M[3] = F0[3] * exy;
        M[2] = F0[2] * exz;
        M[6] = F0[6] * exz;
        M[5] = F0[5] * eyz;
        M[7] = F0[7] * eyz;
    }
}

// ================================================================== K u
template <typename T, int MODE, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) k_ku_tiled(Grid g, const T *__restrict__ F, const T *__restrict__ E,
                                                 const T *__restrict__ u, T *__restrict__ y, const T *__restrict__ b,
                                                 const uint8_t *__restrict__ dmask, int ntY, int ntZ, int xchunk) {
    constexpr int RYM = TY + 2, RZM = TZ + 2, EYM = TY + 1, EZM = TZ + 1;
    constexpr int CS = RYM * RZM;
    static_assert(TY * TZ <= NT, "one output node per work-item");
    __shared__ T su[3][3 * CS];
    __shared__ T sE[2][EYM * EZM];
    const int N = g.nnodes, nx = g.nn[0], ny = g.nn[1], nz = g.nn[2];
    const int tz = blockIdx.x % ntZ, ty = blockIdx.x / ntZ;
    const int y0 = split_lo(ty, g.ydet, ntY), y1 = split_lo(ty + 1, g.ydet, ntY);
    const int z0 = split_lo(tz, nz, ntZ), z1 = split_lo(tz + 1, nz, ntZ);
    const int x0 = blockIdx.y * xchunk, x1 = min(x0 + xchunk, nx);
    const int OY = y1 - y0, OZ = z1 - z0;
    const int RY = OY + 2, RZ = OZ + 2, EY = OY + 1, EZ = OZ + 1;
    const int t = threadIdx.x;

    // loads iterate over the full-size tile (compile-time divisors) and skip what a partial tile lacks
    auto load_u = [&](int ix, int slot) {
        for (int w = t; w < 3 * CS; w += NT) {
            const int c = w / CS, r = w - c * CS, ly = r / RZM, lz = r - ly * RZM;
            if (ly >= RY || lz >= RZ) continue;
            const int gy = y0 - 1 + ly, gz = z0 - 1 + lz;
            T v = T(0);
            if ((unsigned)ix < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
                v = u[(size_t)c * N + ((size_t)ix * ny + gy) * nz + gz];
            su[slot][c * CS + ly * RZM + lz] = v;
        }
    };
    auto load_E = [&](int ex, int slot) {
        for (int w = t; w < EYM * EZM; w += NT) {
            const int ly = w / EZM, lz = w - ly * EZM;
            if (ly >= EY || lz >= EZ) continue;
            const int ey = y0 - 1 + ly, ez = z0 - 1 + lz;
            T v = T(0);
            if ((unsigned)ex < (unsigned)g.ne[0] && (unsigned)ey < (unsigned)g.ne[1] && (unsigned)ez < (unsigned)g.ne[2])
                v = E[((size_t)ex * g.ne[1] + ey) * g.ne[2] + ez];
            sE[slot][ly * EZM + lz] = v;
        }
    };

    // one output node per work-item (the tile never exceeds TY x TZ <= NT nodes)
    const bool act = t < OY * OZ;
    const int oy = act ? t / OZ : 0, oz = act ? t - oy * OZ : 0;
    const int li = (oy + 1) * RZM + oz + 1;
    load_u(x0 - 1, (x0 + 2) % 3);
    load_u(x0, x0 % 3);
    load_E(x0 - 1, (x0 + 1) & 1);
    for (int ix = x0; ix < x1; ++ix) {
        const int n = ((ix * ny) + y0 + oy) * nz + z0 + oz;
        // right-hand side and mask are requested before the tile loads so their latency overlaps
        uint8_t dm = 0;
        T bv[3] = {0, 0, 0};
        if (act) {
            if (dmask) dm = dmask[n];
            if (MODE == AM_RESIDUAL) {
                bv[0] = b[n];
                bv[1] = b[(size_t)N + n];
                bv[2] = b[(size_t)2 * N + n];
            }
        }
        load_u(ix + 1, (ix + 1) % 3);
        load_E(ix, ix & 1);
        __syncthreads();
        if (act) {
            const T *sm = su[(ix + 2) % 3], *s0 = su[ix % 3], *sp = su[(ix + 1) % 3];
            const T *em = sE[(ix + 1) & 1], *e0 = sE[ix & 1];
            T Ee[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
                Ee[e] = (i ? em : e0)[(oy + 1 - j) * EZM + (oz + 1 - k)];
            }
            T f[3] = {0, 0, 0};
            T M[9];
            lds_row<T, false>(F, sm + li, s0 + li, sp + li, CS, RZM, Ee, f, M);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T v = MODE == AM_RESIDUAL ? bv[c] - f[c] : f[c];
                if (dm & (1 << c)) v = T(0);
                y[(size_t)c * N + n] = v;
            }
        }
        __syncthreads();
    }
}

// Rows of detached nodes of a K u product are zero (TPSStencils.hh:717-727).
template <typename T>
__global__ void __launch_bounds__(256) k_zero_rows(Grid g, T *__restrict__ y) {
    const int nz = g.nn[2], rows = g.nn[1] - g.ydet, slab = rows * nz;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = g.ydet + r / nz, iz = r % nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    y[n] = T(0);
    y[g.nnodes + n] = T(0);
    y[(size_t)2 * g.nnodes + n] = T(0);
}

// ================================================================== GS, one x-parity per launch
template <typename T>
__device__ __forceinline__ void gs_point_update(T ud[3], const T bms[3], const T M[9], uint8_t dm, int forward) {
    if (dm) {
        // partially constrained node: sequential point updates of the free components
        ud[0] = ud[1] = ud[2] = T(0);
        for (int s = 0; s < 3; ++s) {
            const int i = forward ? s : 2 - s;
            const T fr = (dm & (1 << i)) ? T(0) : T(1);
            ud[i] = (bms[i] - (M[3 * i] * ud[0] + M[3 * i + 1] * ud[1] + M[3 * i + 2] * ud[2])) * (fr / M[4 * i]);
        }
    } else {
        const T c00 = M[4] * M[8] - M[5] * M[7];
        const T c01 = M[5] * M[6] - M[3] * M[8];
        const T c02 = M[3] * M[7] - M[4] * M[6];
        const T det = M[0] * c00 + M[1] * c01 + M[2] * c02;
        const T id = T(1) / det;
        const T i00 = c00 * id, i01 = (M[2] * M[7] - M[1] * M[8]) * id, i02 = (M[1] * M[5] - M[2] * M[4]) * id;
        const T i10 = c01 * id, i11 = (M[0] * M[8] - M[2] * M[6]) * id, i12 = (M[2] * M[3] - M[0] * M[5]) * id;
        const T i20 = c02 * id, i21 = (M[1] * M[6] - M[0] * M[7]) * id, i22 = (M[0] * M[4] - M[1] * M[3]) * id;
        ud[0] = i00 * bms[0] + i01 * bms[1] + i02 * bms[2];
        ud[1] = i10 * bms[0] + i11 * bms[1] + i12 * bms[2];
        ud[2] = i20 * bms[0] + i21 * bms[1] + i22 * bms[2];
    }
}

template <typename T, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT) k_gs_phase(Grid g, const T *__restrict__ F, const T *__restrict__ E,
                                                 const T *__restrict__ uin, const T *__restrict__ unb,
                                                 T *__restrict__ uout, const T *__restrict__ b,
                                                 const uint8_t *__restrict__ dmask, int parity, int forward,
                                                 int ntY, int ntZ, int pchunk) {
    constexpr int HY = 2, HZ = 4;                          // u halo: colour halo + one neighbour
    constexpr int RYM = TY + 2 * HY, RZM = TZ + 2 * HZ;
    constexpr int EYM = TY + 3, EZM = TZ + 7;              // element rows [y0-2, y1], cols [z0-4, z1+2]
    constexpr int CS = RYM * RZM;
    static_assert((TY / 2 + 1) * (TZ / 2 + 3) <= NT, "one node per work-item and colour position");
    __shared__ T su[3][3 * CS];  // rolling planes: ix-1, ix (updated in place), ix+1
    __shared__ T sE[2][EYM * EZM];
    const int N = g.nnodes, nx = g.nn[0], ny = g.nn[1], nz = g.nn[2];
    const int tz = blockIdx.x % ntZ, ty = blockIdx.x / ntZ;
    const int y0 = split_lo(ty, ny, ntY), y1 = split_lo(ty + 1, ny, ntY);
    const int z0 = split_lo(tz, nz, ntZ), z1 = split_lo(tz + 1, nz, ntZ);
    const int ry0 = y0 - HY, rz0 = z0 - HZ;
    const int RY = y1 - y0 + 2 * HY, RZ = z1 - z0 + 2 * HZ;
    const int ey0 = y0 - 2, ez0 = z0 - 4, EY = y1 - y0 + 3, EZ = z1 - z0 + 7;
    const int t = threadIdx.x;
    const int nplanes = (nx - parity + 1) / 2;
    const int k0 = blockIdx.y * pchunk, k1 = min(k0 + pchunk, nplanes);

    // loads iterate over the full-size tile (compile-time divisors) and skip what a partial tile lacks
    auto load_u = [&](const T *src, int ix, T *dst) {
        for (int w = t; w < 3 * CS; w += NT) {
            const int c = w / CS, r = w - c * CS, ly = r / RZM, lz = r - ly * RZM;
            if (ly >= RY || lz >= RZ) continue;
            const int gy = ry0 + ly, gz = rz0 + lz;
            T v = T(0);
            if ((unsigned)ix < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
                v = src[(size_t)c * N + ((size_t)ix * ny + gy) * nz + gz];
            dst[c * CS + ly * RZM + lz] = v;
        }
    };
    auto load_E = [&](int ex, T *dst) {
        for (int w = t; w < EYM * EZM; w += NT) {
            const int ly = w / EZM, lz = w - ly * EZM;
            if (ly >= EY || lz >= EZ) continue;
            const int ey = ey0 + ly, ez = ez0 + lz;
            T v = T(0);
            if ((unsigned)ex < (unsigned)g.ne[0] && (unsigned)ey < (unsigned)g.ne[1] && (unsigned)ez < (unsigned)g.ne[2])
                v = E[((size_t)ex * g.ne[1] + ey) * g.ne[2] + ez];
            dst[ly * EZM + lz] = v;
        }
    };

    // slots: neighbour planes ix-1 / ix+1 rotate through slots A/B (plane ix+1 of one step is plane
    // ix-1 of the next), slot C holds the plane being updated
    int sA = 0, sB = 1;
    const int sC = 2;
    if (k0 < k1) load_u(unb, parity + 2 * k0 - 1, su[sA]);
    for (int kp = k0; kp < k1; ++kp) {
        const int ix = parity + 2 * kp;
        // Each colour position updates at most one node per work-item (the host checks the node count
        // against NT). The right-hand side and Dirichlet mask of the four nodes are requested before the
        // tile loads so that their latency overlaps them.
        int pli[4], piy[4], piz[4];
        uint8_t pdm[4];
        T pb[4][3];
#pragma unroll
        for (int pos = 0; pos < 4; ++pos) {
            const int sub = forward ? pos : 3 - pos;
            const int cy = (sub >> 1) & 1, cz = sub & 1;
            const int hy = pos < 2 ? 1 : 0, hz = 3 - pos;
            // nodes of colour (cy, cz) in [y0-hy, y1+hy) x [z0-hz, z1+hz), clipped to the active rows
            int ya = max(y0 - hy, 0), za = max(z0 - hz, 0);
            const int yb = min(y1 + hy, g.ydet), zb = min(z1 + hz, nz);
            ya += (ya & 1) != cy;
            za += (za & 1) != cz;
            const int nr = yb > ya ? (yb - ya + 1) / 2 : 0, nc = zb > za ? (zb - za + 1) / 2 : 0;
            pdm[pos] = 7;
            pli[pos] = 0;
            piy[pos] = piz[pos] = 0;
            pb[pos][0] = pb[pos][1] = pb[pos][2] = T(0);
            if (t < nr * nc) {
                const int wr = t / nc;
                const int iy = ya + 2 * wr, iz = za + 2 * (t - wr * nc);
                const int n = ((ix * ny) + iy) * nz + iz;
                piy[pos] = iy;
                piz[pos] = iz;
                pli[pos] = (iy - ry0) * RZM + (iz - rz0);
                pdm[pos] = dmask ? dmask[n] : uint8_t(0);
                pb[pos][0] = b[n];
                pb[pos][1] = b[(size_t)N + n];
                pb[pos][2] = b[(size_t)2 * N + n];
            }
        }
        load_u(uin, ix, su[sC]);
        load_u(unb, ix + 1, su[sB]);
        load_E(ix - 1, sE[0]);
        load_E(ix, sE[1]);
        __syncthreads();
        const T *pm = su[sA], *pp = su[sB];
        T *pc = su[sC];
#pragma unroll
        for (int pos = 0; pos < 4; ++pos) {
            if (pdm[pos] != 7) {
                const int iy = piy[pos], iz = piz[pos], li = pli[pos];
                T Ee[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
                    Ee[e] = sE[1 - i][(iy - j - ey0) * EZM + (iz - k - ez0)];
                }
                T f[3] = {0, 0, 0};
                T M[9];
                lds_row<T, true>(F, pm + li, pc + li, pp + li, CS, RZM, Ee, f, M);
                const T bms[3] = {pb[pos][0] - f[0], pb[pos][1] - f[1], pb[pos][2] - f[2]};
                T ud[3];
                gs_point_update(ud, bms, M, pdm[pos], forward);
                pc[li] += ud[0];
                pc[CS + li] += ud[1];
                pc[2 * CS + li] += ud[2];
            }
            __syncthreads();
        }
        // the tile of plane ix, updated or not
        const int OY = y1 - y0, OZ = z1 - z0;
        for (int w = t; w < 3 * TY * TZ; w += NT) {
            const int c = w / (TY * TZ), r = w - c * TY * TZ, oy = r / TZ, oz = r - oy * TZ;
            if (oy >= OY || oz >= OZ) continue;
            uout[(size_t)c * N + ((size_t)ix * ny + y0 + oy) * nz + z0 + oz] = pc[c * CS + (oy + HY) * RZM + oz + HZ];
        }
        const int tmp = sA;  // plane ix+1 becomes the next step's plane ix'-1
        sA = sB;
        sB = tmp;
        __syncthreads();
    }
}

// ================================================================== GS, one colour per launch, register gather
// Row of a node whose 27 neighbour values sit in registers (un[s][c], s = 9 dx + 3 dy + dz, 0..2 each):
// the structured element matrix of lds_row, neighbours taken in groups of equal delta so each group's
// nine F values are loaded once.
template <typename T, bool WANT_M>
__device__ __forceinline__ void reg_row(const T *__restrict__ F, const T un[27][3], const T Ee[8], T f[3], T M[9]) {
    T a[8][3];
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e][0] = a[e][1] = a[e][2] = T(0);
    T Fs[9];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int s = kNeighbourOrder[t];
        const int dx = s / 9, dy = (s / 3) % 3, dz = s % 3;
        if (t == 0 || delta_of(kNeighbourOrder[t - 1]) != delta_of(s)) {
            const T *Fg = F + 16 * delta_of(s);
            asm volatile("" : "+s"(Fg));
            const CT<T> *Fp = (const CT<T> *)Fg;
#pragma unroll
            for (int q = 0; q < 9; ++q) Fs[q] = Fp[q];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
            const int mx = dx - 1 + i, my = dy - 1 + j, mz = dz - 1 + k;
            if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int sg = c == d ? 1 : side(e, c) * side(e, d);
                    if (sg > 0) a[e][c] += Fs[3 * c + d] * un[s][d];
                    else a[e][c] -= Fs[3 * c + d] * un[s][d];
                }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) asm volatile("" : "+v"(a[e][0]), "+v"(a[e][1]), "+v"(a[e][2]));
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        f[0] += Ee[e] * a[e][0];
        f[1] += Ee[e] * a[e][1];
        f[2] += Ee[e] * a[e][2];
    }
    if (WANT_M) {
        T es = 0, exy = 0, exz = 0, eyz = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            es += Ee[e];
            exy += side(e, 0) * side(e, 1) > 0 ? Ee[e] : -Ee[e];
            exz += side(e, 0) * side(e, 2) > 0 ? Ee[e] : -Ee[e];
            eyz += side(e, 1) * side(e, 2) > 0 ? Ee[e] : -Ee[e];
        }
        const T *Fg = F;
        asm volatile("" : "+s"(Fg));
        const CT<T> *F0 = (const CT<T> *)Fg;
        M[0] = F0[0] * es;
        M[4] = F0[4] * es;
        M[8] = F0[8] * es;
        M[1] = F0[1] * exy;
        M[3] = F0[3] * exy;
        M[2] = F0[2] * exz;
        M[6] = F0[6] * exz;
        M[5] = F0[5] * eyz;
        M[7] = F0[7] * eyz;
    }
}

template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    if constexpr (sizeof(T) == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return __builtin_bit_cast(T, v);
    } else {
        const int v = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
        return __builtin_bit_cast(T, v);
    }
}

// One colour of the 8-colour block Gauss-Seidel sweep (MultigridSolver.hh:408-458), one node per work-item.
// Interior nodes load their 81 neighbour values with buffer loads that share one per-lane offset (the
// neighbour and component offsets are scalar), so the gather costs no vector address arithmetic;
// nodes on the grid boundary take the bounds-checked path.
template <typename T>
__global__ void __launch_bounds__(256) k_gs_gather_f(Grid g, const T *__restrict__ F, const T *__restrict__ E,
                                                     T *__restrict__ u, const T *__restrict__ b,
                                                     const uint8_t *__restrict__ dmask, int cx, int cy, int cz,
                                                     int forward) {
    const int mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int jx = t / (my * mz), r = t - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int N = g.nnodes, nx = g.nn[0], ny = g.nn[1], nz = g.nn[2];
    const int sx = ny * nz, sy = nz;
    const int n = (ix * ny + iy) * nz + iz;
    const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
    if (dm == 7) return;
    T un[27][3];
    const bool interior = ix > 0 && ix < nx - 1 && iy > 0 && iy < ny - 1 && iz > 0 && iz < nz - 1;
    if (interior) {
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)u, 0, (int)((size_t)3 * N * sizeof(T)), 0x00020000);
        const int vo = (n - sx - sy - 1) * (int)sizeof(T);
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int o = (s / 9) * sx + ((s / 3) % 3) * sy + s % 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) un[s][c] = buffer_load<T>(rs, vo, (c * N + o) * (int)sizeof(T));
        }
    } else {
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int dx = s / 9 - 1, dy = (s / 3) % 3 - 1, dz = s % 3 - 1;
            const bool ok = (unsigned)(ix + dx) < (unsigned)nx && (unsigned)(iy + dy) < (unsigned)ny &&
                            (unsigned)(iz + dz) < (unsigned)nz;
            const int m = n + dx * sx + dy * sy + dz;
#pragma unroll
            for (int c = 0; c < 3; ++c) un[s][c] = ok ? u[(size_t)c * N + m] : T(0);
        }
    }
    T Ee[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int ex = ix - ((e >> 2) & 1), ey = iy - ((e >> 1) & 1), ez = iz - (e & 1);
        const bool ok = (unsigned)ex < (unsigned)g.ne[0] && (unsigned)ey < (unsigned)g.ne[1] &&
                        (unsigned)ez < (unsigned)g.ne[2];
        Ee[e] = ok ? E[((size_t)ex * g.ne[1] + ey) * g.ne[2] + ez] : T(0);
    }
    T f[3] = {0, 0, 0};
    T M[9];
    reg_row<T, true>(F, un, Ee, f, M);
    const T bms[3] = {b[n] - f[0], b[(size_t)N + n] - f[1], b[(size_t)2 * N + n] - f[2]};
    T ud[3];
    gs_point_update(ud, bms, M, dm, forward);
    u[n] += ud[0];
    u[(size_t)N + n] += ud[1];
    u[(size_t)2 * N + n] += ud[2];
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n = v;
        else
            n = 256;
    }
    return n;
}

// fine K u tile: 16 x 16 outputs, 256 work-items
constexpr int KU_TY = 16, KU_TZ = 16, KU_NT = 256;
// GS tile: 16 x 32 outputs, 192 work-items (the four colour positions cover 171/162/136/128 nodes)
constexpr int GS_TY = 16, GS_TZ = 32, GS_NT = 192;

}  // namespace

template <typename T>
void launch_fine_apply_tiled(const Grid &g, const T *F, const T *E, const T *u, T *y, const T *b,
                             const uint8_t *dmask, int mode, hipStream_t s) {
    if (mode != AM_SET && mode != AM_RESIDUAL) throw std::runtime_error("tiled K u supports the SET and RESIDUAL modes");
    if (g.ydet > 0) {
        const int ntY = ceil_div(g.ydet, KU_TY), ntZ = ceil_div(g.nn[2], KU_TZ);
        const int tiles = ntY * ntZ;
        // a few workgroups per CU; each x-chunk reloads one halo plane at each end
        const int want = 6 * cu_count();
        int chunks = std::max(1, std::min(g.nn[0], ceil_div(want, tiles)));
        const int xchunk = ceil_div(g.nn[0], chunks);
        chunks = ceil_div(g.nn[0], xchunk);
        if (chunks > 65535) throw std::runtime_error("grid too large for the tiled K u");
        const dim3 grid(tiles, chunks);
        if (mode == AM_RESIDUAL)
            hipLaunchKernelGGL((k_ku_tiled<T, AM_RESIDUAL, KU_TY, KU_TZ, KU_NT>), grid, dim3(KU_NT), 0, s, g, F, E,
                               u, y, b, dmask, ntY, ntZ, xchunk);
        else
            hipLaunchKernelGGL((k_ku_tiled<T, AM_SET, KU_TY, KU_TZ, KU_NT>), grid, dim3(KU_NT), 0, s, g, F, E, u,
                               y, b, dmask, ntY, ntZ, xchunk);
    }
    if (mode == AM_SET && g.ydet < g.nn[1]) {
        const long cz = (long)g.nn[0] * (g.nn[1] - g.ydet) * g.nn[2];
        hipLaunchKernelGGL(k_zero_rows<T>, dim3((unsigned)((cz + 255) / 256)), dim3(256), 0, s, g, y);
    }
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_fine_gs_phase(const Grid &g, const T *F, const T *E, const T *uin, const T *unb, T *uout, const T *b,
                          const uint8_t *dmask, int parity, int forward, hipStream_t s) {
    if (uin == uout) throw std::runtime_error("the two-phase smoother needs distinct input and output vectors");
    const int nplanes = (g.nn[0] - parity + 1) / 2;
    if (nplanes <= 0) return;
    const int ntY = ceil_div(g.nn[1], GS_TY), ntZ = ceil_div(g.nn[2], GS_TZ);
    const int tiles = ntY * ntZ;
    const int want = 6 * cu_count();
    int chunks = std::max(1, std::min(nplanes, ceil_div(want, tiles)));
    const int pchunk = ceil_div(nplanes, chunks);
    chunks = ceil_div(nplanes, pchunk);
    if (chunks > 65535) throw std::runtime_error("grid too large for the two-phase smoother");
    hipLaunchKernelGGL((k_gs_phase<T, GS_TY, GS_TZ, GS_NT>), dim3(tiles, chunks), dim3(GS_NT), 0, s, g, F, E, uin,
                       unb, uout, b, dmask, parity, forward, ntY, ntZ, pchunk);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_fine_gs_gather_f(const Grid &g, const T *F, const T *E, T *u, const T *b, const uint8_t *dmask, int color,
                             int forward, hipStream_t s) {
    const int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    const long mx = (g.nn[0] - cx + 1) / 2, my = (g.ydet - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    const long count = (mx > 0 && my > 0 && mz > 0) ? mx * my * mz : 0;
    if (!count) return;
    if ((size_t)3 * g.nnodes * sizeof(T) >= ((size_t)1 << 31)) throw std::runtime_error("field too large for 32-bit buffer offsets");
    hipLaunchKernelGGL(k_gs_gather_f<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, g, F, E, u, b, dmask,
                       cx, cy, cz, forward);
    VF_HIP(hipGetLastError());
}

template void launch_fine_gs_gather_f<double>(const Grid &, const double *, const double *, double *, const double *,
                                              const uint8_t *, int, int, hipStream_t);
template void launch_fine_gs_gather_f<float>(const Grid &, const float *, const float *, float *, const float *,
                                             const uint8_t *, int, int, hipStream_t);

template void launch_fine_apply_tiled<double>(const Grid &, const double *, const double *, const double *, double *,
                                              const double *, const uint8_t *, int, hipStream_t);
template void launch_fine_apply_tiled<float>(const Grid &, const float *, const float *, const float *, float *,
                                             const float *, const uint8_t *, int, hipStream_t);
template void launch_fine_gs_phase<double>(const Grid &, const double *, const double *, const double *,
                                           const double *, double *, const double *, const uint8_t *, int, int,
                                           hipStream_t);
template void launch_fine_gs_phase<float>(const Grid &, const float *, const float *, const float *, const float *,
                                          float *, const float *, const uint8_t *, int, int, hipStream_t);

}  // namespace vf
