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
#include <map>
#include <mutex>

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

// raw buffer load: a byte offset at or past the buffer's size reads as zero
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

// Colour position (0..3 inside an x-parity phase, in application order) of the in-plane neighbour
// (dy, dz) (0..2, 1 = same coordinate) of a node at position p: its y parity differs when dy != 1 and
// its z parity when dz != 1.
template <bool FWD>
__host__ __device__ constexpr int nbr_pos(int p, int dy, int dz) {
    const int sub = (FWD ? p : 3 - p) ^ ((dy != 1) << 1) ^ (dz != 1);
    return FWD ? sub : 3 - sub;
}
// in-plane neighbours (bit 3 dy + dz) whose colour is applied before position p (EARLIER) or not
template <bool FWD, bool EARLIER>
__host__ __device__ constexpr int nbr_mask(int p) {
    int m = 0;
    for (int dy = 0; dy < 3; ++dy)
        for (int dz = 0; dz < 3; ++dz)
            if ((nbr_pos<FWD>(p, dy, dz) < p) == EARLIER) m |= 1 << (3 * dy + dz);
    return m;
}

// Sum over the elements of one node plane that contain the in-plane neighbour offset (dy, dz) of
// X[j][k] (j, k = element offsets below the node), each term signed by the node's side along y (SY)
// and z (SZ): element j contains offset dy when dy == 1, or dy == 0 and j == 1, or dy == 2 and j == 0;
// the side is +1 for j == 1 and -1 for j == 0.
template <typename T, bool SY, bool SZ>
__device__ __forceinline__ T side_sum(const T X[2][2], int dy, int dz) {
    T r = T(0);
    bool first = true;
#pragma unroll
    for (int j = 1; j >= 0; --j) {
        if (!(dy == 1 || (dy == 0 && j == 1) || (dy == 2 && j == 0))) continue;
#pragma unroll
        for (int k = 1; k >= 0; --k) {
            if (!(dz == 1 || (dz == 0 && k == 1) || (dz == 2 && k == 0))) continue;
            const T v = ((SY && j == 0) != (SZ && k == 0)) ? -X[j][k] : X[j][k];
            r = first ? v : r + v;
            first = false;
        }
    }
    return r;
}

// Weighted neighbour sums of one x-plane of the separable-weight row (see w_row): for the in-plane
// offsets of group g (bit 1: dy != 1, bit 0: dz != 1) selected by MASK (bit 3 dy + dz) add
// W[c][d] u[d] to V[c][d]. Xp / Xs are the plane's element moduli combined along x without / with
// the node's x side (for the diagonal and yz weights / for the xy and xz weights). p points at the
// node in the plane (component stride CS, row stride RZ); zm / zp are the offsets of its z - 1 / z + 1
// neighbours (the plane rows are stored by z parity, so they depend on the node's parity).
template <typename T, int MASK>
__device__ __forceinline__ void plane_sums(const T *p, int CS, int RZ, int zm, int zp, const T Xp[2][2],
                                           const T Xs[2][2], int g, T V[9]) {
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) {
            if ((((dy != 1) << 1) | (dz != 1)) != g || !((MASK >> (3 * dy + dz)) & 1)) continue;
            const T w = side_sum<T, false, false>(Xp, dy, dz);
            const T wxy = side_sum<T, true, false>(Xs, dy, dz);
            const T wxz = side_sum<T, false, true>(Xs, dy, dz);
            const T wyz = side_sum<T, true, true>(Xp, dy, dz);
            const int off = (dy - 1) * RZ + (dz == 0 ? zm : (dz == 2 ? zp : 0));
            const T u0 = p[off], u1 = p[CS + off], u2 = p[2 * CS + off];
            V[0] += w * u0;
            V[4] += w * u1;
            V[8] += w * u2;
            V[1] += wxy * u1;
            V[3] += wxy * u0;
            V[2] += wxz * u2;
            V[6] += wxz * u0;
            V[5] += wyz * u2;
            V[7] += wyz * u1;
        }
}

template <typename T>
__device__ __forceinline__ void contract_group(const T *__restrict__ F, int delta, const T V[9], T f[3]) {
    const T *Fg = F + 16 * delta;
    asm volatile("" : "+s"(Fg));
    const CT<T> *Fp = (const CT<T> *)Fg;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] += Fp[3 * c] * V[3 * c] + Fp[3 * c + 1] * V[3 * c + 1] + Fp[3 * c + 2] * V[3 * c + 2];
}

// f += the in-plane part of a node row over the neighbours selected by MASK (X = E0 + E1, Xs = E1 - E0
// over the two element layers), one delta group at a time.
template <typename T, int MASK>
__device__ __forceinline__ void in_plane_part(const T *__restrict__ F, const T *p, int CS, int RZ, int zm, int zp,
                                              const T X[2][2], const T Xs[2][2], T f[3]) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        constexpr int kGroupMask[4] = {1 << 4, (1 << 3) | (1 << 5), (1 << 1) | (1 << 7), 1 | 4 | 64 | 256};
        if (!(MASK & kGroupMask[g])) continue;
        T V[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) V[q] = T(0);
        plane_sums<T, MASK>(p, CS, RZ, zm, zp, X, Xs, g, V);
        contract_group(F, g, V, f);
        asm volatile("" : "+v"(f[0]), "+v"(f[1]), "+v"(f[2]));
        __builtin_amdgcn_sched_barrier(0);
    }
}

// The value of v, hidden from the optimiser: arithmetic derived from it is recomputed where it is used
// instead of being hoisted out of the plane loop and kept live in registers.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// One x-parity of the 8-colour block Gauss-Seidel sweep (MultigridSolver.hh:408-458). A workgroup owns
// a (y, z) tile of every plane of that parity in its x-chunk and steps through them. Per plane:
//  1. the step's two new u planes and two element planes are loaded into LDS at the end of the previous
//     step (FP32: five workgroups share a CU and cover each other's loads) or requested into registers at
//     its start (FP64: two workgroups per CU);
//  2. for every node the four colours will update (up to three per work-item) the row part that the
//     colour steps do not change is evaluated in parallel and kept in registers as b - (K u)_fixed: the
//     two neighbour planes (not touched by this phase), the node itself, and the in-plane neighbours
//     whose colour comes later;
//  3. the four in-plane colours are applied in order; a colour adds only its in-plane neighbours of
//     earlier colours (0, 2, 6 and 8 of them), solves the node's 3x3 block and updates the LDS plane;
//  4. the tile's own nodes are written out and the prefetched registers are stored into LDS.
// Rows use the separable-weight form of w_row. Halo values outside the grid are zero; the tile loads
// go through buffer loads whose offsets are precomputed once per work-item (out-of-grid entries get an
// out-of-range offset, which the buffer unit returns as zero).
template <typename T, int TY, int TZ, int NT, bool FWD>
__global__ void __launch_bounds__(NT, sizeof(T) == 4 ? 5 : 1) k_gs_phase(Grid g, const T *__restrict__ F, const T *__restrict__ E,
                                                 const T *__restrict__ uin, const T *__restrict__ unb,
                                                 T *__restrict__ uout, const T *__restrict__ b,
                                                 const uint8_t *__restrict__ dmask, int parity, int ntY, int ntZ,
                                                 int pchunk, int zin) {
    constexpr int HY = 2, HZ = 4;                          // u halo: colour halo + one neighbour
    constexpr int RYM = TY + 2 * HY, RZM = TZ + 2 * HZ;
    constexpr int EYM = TY + 3, EZM = TZ + 8;              // element rows [y0-2, y1], cols [z0-4, z1+2] (+1)
    constexpr int CS = RYM * RZM, ES = EYM * EZM;
    // FP64: LDS rows hold the entries of even local z first, then the odd ones, so the nodes of one colour
    // (every other z) sit at consecutive addresses and 8-byte reads do not conflict in the LDS banks.
    // FP32 keeps natural rows (its reads conflict two-way at most; the split form costs it registers).
    constexpr bool SPLIT = sizeof(T) == 8;
    constexpr int HZ2 = RZM / 2, HEZ = EZM / 2;
    static_assert(RZM % 2 == 0 && EZM % 2 == 0, "parity-split rows");
    constexpr int NU = (3 * CS + NT - 1) / NT, NE = (ES + NT - 1) / NT;
    constexpr int MAXN = (TY / 2 + 1) * (TZ / 2 + 3) + (TY / 2 + 1) * (TZ / 2 + 2) + (TY / 2) * (TZ / 2 + 1) +
                         (TY / 2) * (TZ / 2);
    constexpr int OOB = 0x7fffffff;  // byte offset past every buffer (their sizes are below 2^31 - 2^30)
    static_assert((TY / 2 + 1) * (TZ / 2 + 3) <= NT, "at most one node per work-item and colour position");
    static_assert(MAXN <= 3 * NT, "at most three nodes per work-item and plane");
    static_assert(TY % 2 == 0 && TZ % 2 == 0, "even tile extents");
    static_assert(RYM < 256 && RZM < 256, "slot packing");
    __shared__ T su[3][3 * CS];  // planes ix-1, ix+1 (rotating) and ix (updated in place)
    __shared__ T sE[2][ES];      // element planes ix-1 (i = 1) and ix (i = 0)
    const int N = g.nnodes, nx = g.nn[0], ny = g.nn[1], nz = g.nn[2];
    const int tz = blockIdx.x % ntZ, ty = blockIdx.x / ntZ;
    const int y0 = split_lo(ty, ny, ntY), y1 = split_lo(ty + 1, ny, ntY);
    const int z0 = split_lo(tz, nz, ntZ), z1 = split_lo(tz + 1, nz, ntZ);
    const int ry0 = y0 - HY, rz0 = z0 - HZ;
    const int ey0 = y0 - 2, ez0 = z0 - 4;
    const int t = threadIdx.x;
    const int nplanes = (nx - parity + 1) / 2;
    const int k0 = blockIdx.y * pchunk, k1 = min(k0 + pchunk, nplanes);
    if (k0 >= k1) return;  // uniform over the workgroup

    // Node lists of the four colour positions (the tile grown by the halo the later colours read:
    // (1,3), (1,2), (0,1), (0,0) in (y,z)), clipped to the active rows; concatenated in position order.
    // Work-item t owns list entries t, t + NT, t + 2 NT (slots 0..2) for the whole launch.
    int base[5], ya[4], za[4], ncol[4];
    base[0] = 0;
#pragma unroll
    for (int pos = 0; pos < 4; ++pos) {
        const int sub = FWD ? pos : 3 - pos;
        const int cy = (sub >> 1) & 1, cz = sub & 1;
        const int hy = pos < 2 ? 1 : 0, hz = 3 - pos;
        int a = max(y0 - hy, 0), c = max(z0 - hz, 0);
        const int yb = min(y1 + hy, g.ydet), zb = min(z1 + hz, nz);
        a += (a & 1) != cy;
        c += (c & 1) != cz;
        const int nr = yb > a ? (yb - a + 1) / 2 : 0, nc = zb > c ? (zb - c + 1) / 2 : 0;
        ya[pos] = a;
        za[pos] = c;
        ncol[pos] = max(nc, 1);
        base[pos + 1] = base[pos] + nr * nc;
    }
    // slot r: (colour position + 1) << 16 | ly << 8 | lz in the LDS plane (0: no node); soff[r]: the
    // node's index inside its plane
    int slot[3], soff[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int q = r * NT + t;
        int pos = -1;
#pragma unroll
        for (int p = 0; p < 4; ++p)
            if (q >= base[p] && q < base[p + 1]) pos = p;
        slot[r] = 0;
        soff[r] = 0;
        if (pos >= 0) {
            int j = q, a = 0, c = 0, nc = 1;
#pragma unroll
            for (int p = 0; p < 4; ++p)
                if (pos == p) {
                    j = q - base[p];
                    a = ya[p];
                    c = za[p];
                    nc = ncol[p];
                }
            const int wr = j / nc;
            const int iy = a + 2 * wr, iz = c + 2 * (j - wr * nc);
            slot[r] = (pos + 1) << 16 | (iy - ry0) << 8 | (iz - rz0);
            soff[r] = iy * nz + iz;
        }
    }
    auto s_pos = [](int sl) { return (sl >> 16) - 1; };
    auto s_ly = [](int sl) { return (sl >> 8) & 255; };
    auto s_lz = [](int sl) { return sl & 255; };
    // split-row LDS indices of node (ly, lz) and element (ly, lz), and the z-neighbour offsets of a node
    auto uidx = [](int ly, int lz) { return ly * RZM + (SPLIT ? (lz & 1) * HZ2 + (lz >> 1) : lz); };
    auto eidx = [](int ly, int lz) { return ly * EZM + (SPLIT ? (lz & 1) * HEZ + (lz >> 1) : lz); };
    auto zoff_m = [](int lz) { return SPLIT ? ((lz & 1) ? -HZ2 : HZ2 - 1) : -1; };
    auto zoff_p = [](int lz) { return SPLIT ? ((lz & 1) ? 1 - HZ2 : HZ2) : 1; };

    // per-work-item byte offsets of its tile entries inside one plane (OOB outside the grid)
    int uo[NU], eo[NE];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int w = t + i * NT;
        const int c = w / CS, r = w - c * CS, ly = r / RZM, rr = r - ly * RZM, par = rr / HZ2;
        const int lz = SPLIT ? 2 * (rr - par * HZ2) + par : rr;
        const int gy = ry0 + ly, gz = rz0 + lz;
        const bool ok = w < 3 * CS && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz;
        uo[i] = ok ? (c * N + gy * nz + gz) * (int)sizeof(T) : OOB;
    }
#pragma unroll
    for (int i = 0; i < NE; ++i) {
        const int w = t + i * NT;
        const int ly = w / EZM, rr = w - ly * EZM, par = rr / HEZ;
        const int lz = SPLIT ? 2 * (rr - par * HEZ) + par : rr;
        const int gy = ey0 + ly, gz = ez0 + lz;
        const bool ok = w < ES && (unsigned)gy < (unsigned)g.ne[1] && (unsigned)gz < (unsigned)g.ne[2];
        eo[i] = ok ? (gy * g.ne[2] + gz) * (int)sizeof(T) : OOB;
    }
    // zin bit 0 / bit 1: uin / unb is known to be zero; an empty buffer range makes every load of it
    // return zero without touching memory
    const int ubytes = (int)((size_t)3 * N * sizeof(T));
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc((void *)uin, 0, (zin & 1) ? 0 : ubytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rnb =
        __builtin_amdgcn_make_buffer_rsrc((void *)unb, 0, (zin & 2) ? 0 : ubytes, 0x00020000);
    const int NEL = g.ne[0] * g.ne[1] * g.ne[2];
    const __amdgpu_buffer_rsrc_t rE =
        __builtin_amdgcn_make_buffer_rsrc((void *)E, 0, (int)((size_t)NEL * sizeof(T)), 0x00020000);
    const int planeB = ny * nz * (int)sizeof(T), eplaneB = g.ne[1] * g.ne[2] * (int)sizeof(T);

    auto fetch_u = [&](__amdgpu_buffer_rsrc_t rs, int ix, T (&v)[NU]) {
        if ((unsigned)ix < (unsigned)nx) {
            const int pb = ix * planeB;
#pragma unroll
            for (int i = 0; i < NU; ++i) v[i] = buffer_load<T>(rs, (int)((unsigned)uo[i] + (unsigned)pb), 0);
        } else {
#pragma unroll
            for (int i = 0; i < NU; ++i) v[i] = T(0);
        }
    };
    auto store_u = [&](const T (&v)[NU], T *dst) {
        const int tl = opaque(t);
#pragma unroll
        for (int i = 0; i < NU; ++i) {
            const int w = tl + i * NT;
            if ((i + 1) * NT <= 3 * CS || w < 3 * CS) dst[w] = v[i];
        }
    };
    auto fetch_E = [&](int ex, T (&v)[NE]) {
        if ((unsigned)ex < (unsigned)g.ne[0]) {
            const int pb = ex * eplaneB;
#pragma unroll
            for (int i = 0; i < NE; ++i) v[i] = buffer_load<T>(rE, (int)((unsigned)eo[i] + (unsigned)pb), 0);
        } else {
#pragma unroll
            for (int i = 0; i < NE; ++i) v[i] = T(0);
        }
    };
    auto store_E = [&](const T (&v)[NE], T *dst) {
        const int tl = opaque(t);
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int w = tl + i * NT;
            if ((i + 1) * NT <= ES || w < ES) dst[w] = v[i];
        }
    };

    int sA = 0, sB = 1;
    const int sC = 2;
    {
        const int ix = parity + 2 * k0;
        T v[NU];
        T ev[NE];
        fetch_u(rnb, ix - 1, v);
        store_u(v, su[sA]);
        fetch_u(rin, ix, v);
        store_u(v, su[sC]);
        fetch_u(rnb, ix + 1, v);
        store_u(v, su[sB]);
        fetch_E(ix - 1, ev);
        store_E(ev, sE[0]);
        fetch_E(ix, ev);
        store_E(ev, sE[1]);
    }
    __syncthreads();

    for (int kp = k0; kp < k1; ++kp) {
        const int ix = parity + 2 * kp;
        const bool more = kp + 1 < k1;
        // FP64: two workgroups per CU (LDS), too few to cover each other's loads, so the next step's planes
        // are requested here into registers and stored at the end of the step
        constexpr bool kPrefetch = sizeof(T) == 8;
        T pc[NU], pb[NU], pe0[NE], pe1[NE];
        if (kPrefetch && more) {
            fetch_u(rin, ix + 2, pc);
            fetch_u(rnb, ix + 3, pb);
            fetch_E(ix + 1, pe0);
            fetch_E(ix + 2, pe1);
        }
        // 2. b - (K u)_fixed for each owned node; its right-hand side and mask loads overlap the first
        // node's arithmetic
        T og[3][3];
        uint8_t dcur[3];
        const int nplane = ix * ny * nz;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            dcur[r] = 7;
            og[r][0] = og[r][1] = og[r][2] = T(0);
            if (slot[r]) {
                const int n = nplane + soff[r];
                dcur[r] = dmask ? dmask[n] : uint8_t(0);
                og[r][0] = b[n];
                og[r][1] = b[(size_t)N + n];
                og[r][2] = b[(size_t)2 * N + n];
            }
        }
        const T *pm = su[sA], *pp = su[sB];
        T *pcur = su[sC];
        auto fixed_part = [&](auto R) {
            constexpr int r = decltype(R)::value;
            if (slot[r] && dcur[r] != 7) {
                const int ly = s_ly(slot[r]), lz = s_lz(slot[r]);
                const int li = uidx(ly, lz), zm = zoff_m(lz), zp = zoff_p(lz);
                T e1[2][2], e0[2][2], e0n[2][2];
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        // element local (ly - j, lz - k): the y and z halos of the node and element tiles
                        // differ by exactly the element offset (HY = 2, HZ = 4)
                        const int ei = eidx(ly - j, lz - k);
                        e1[j][k] = sE[0][ei];  // x - 1 plane: elements i = 1, side +1
                        e0[j][k] = sE[1][ei];  // x + 1 plane: elements i = 0, side -1
                        e0n[j][k] = -e0[j][k];
                    }
                T f[3] = {0, 0, 0};
#pragma unroll
                for (int gg = 0; gg < 4; ++gg) {
                    T V[9];
#pragma unroll
                    for (int q = 0; q < 9; ++q) V[q] = T(0);
                    plane_sums<T, 511>(pm + li, CS, RZM, zm, zp, e1, e1, gg, V);
                    plane_sums<T, 511>(pp + li, CS, RZM, zm, zp, e0, e0n, gg, V);
                    contract_group(F, 4 | gg, V, f);
                    // one group at a time: keeps the neighbour values of later groups out of registers
                    asm volatile("" : "+v"(f[0]), "+v"(f[1]), "+v"(f[2]));
                    __builtin_amdgcn_sched_barrier(0);
                }
                T X[2][2], Xs[2][2];
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        X[j][k] = e0[j][k] + e1[j][k];
                        Xs[j][k] = e1[j][k] - e0[j][k];
                    }
                switch (s_pos(slot[r])) {
                    case 0: in_plane_part<T, nbr_mask<FWD, false>(0)>(F, pcur + li, CS, RZM, zm, zp, X, Xs, f); break;
                    case 1: in_plane_part<T, nbr_mask<FWD, false>(1)>(F, pcur + li, CS, RZM, zm, zp, X, Xs, f); break;
                    case 2: in_plane_part<T, nbr_mask<FWD, false>(2)>(F, pcur + li, CS, RZM, zm, zp, X, Xs, f); break;
                    default: in_plane_part<T, nbr_mask<FWD, false>(3)>(F, pcur + li, CS, RZM, zm, zp, X, Xs, f); break;
                }
                og[r][0] -= f[0];
                og[r][1] -= f[1];
                og[r][2] -= f[2];
            }
            asm volatile("" : "+v"(og[r][0]), "+v"(og[r][1]), "+v"(og[r][2]));
            __builtin_amdgcn_sched_barrier(0);
        };
        fixed_part(std::integral_constant<int, 0>());
        fixed_part(std::integral_constant<int, 1>());
        fixed_part(std::integral_constant<int, 2>());
        // 3. the four in-plane colours
        // each slot is tested separately (at most one holds a node of this colour) so that every
        // register array keeps constant indices
        auto colour_slot = [&](auto P, auto R) {
            constexpr int pos = decltype(P)::value, r = decltype(R)::value;
            if (s_pos(slot[r]) == pos && dcur[r] != 7) {
                const uint8_t dm = dcur[r];
                const int ly = s_ly(slot[r]), lz = s_lz(slot[r]);
                const T o[3] = {og[r][0], og[r][1], og[r][2]};
                const int li = uidx(ly, lz), zm = zoff_m(lz), zp = zoff_p(lz);
                T X[2][2], Xs[2][2];
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const int ei = eidx(ly - j, lz - k);
                        const T e1 = sE[0][ei], e0 = sE[1][ei];
                        X[j][k] = e0 + e1;
                        Xs[j][k] = e1 - e0;
                    }
                T f[3] = {0, 0, 0};
                constexpr int mask = nbr_mask<FWD, true>(pos);
                if constexpr (mask != 0) in_plane_part<T, mask>(F, pcur + li, CS, RZM, zm, zp, X, Xs, f);
                T M[9];
                {
                    const T w = side_sum<T, false, false>(X, 1, 1), wxy = side_sum<T, true, false>(Xs, 1, 1);
                    const T wxz = side_sum<T, false, true>(Xs, 1, 1), wyz = side_sum<T, true, true>(X, 1, 1);
                    const T *Fg = F;
                    asm volatile("" : "+s"(Fg));
                    const CT<T> *F0 = (const CT<T> *)Fg;
                    M[0] = F0[0] * w;
                    M[4] = F0[4] * w;
                    M[8] = F0[8] * w;
                    M[1] = F0[1] * wxy;
                    M[3] = F0[3] * wxy;
                    M[2] = F0[2] * wxz;
                    M[6] = F0[6] * wxz;
                    M[5] = F0[5] * wyz;
                    M[7] = F0[7] * wyz;
                }
                const T bms[3] = {o[0] - f[0], o[1] - f[1], o[2] - f[2]};
                T ud[3];
                gs_point_update(ud, bms, M, dm, FWD ? 1 : 0);
                pcur[li] += ud[0];
                pcur[CS + li] += ud[1];
                pcur[2 * CS + li] += ud[2];
            }
        };
        auto colour = [&](auto P) {
            colour_slot(P, std::integral_constant<int, 0>());
            colour_slot(P, std::integral_constant<int, 1>());
            colour_slot(P, std::integral_constant<int, 2>());
            __syncthreads();
        };
        colour(std::integral_constant<int, 0>());
        colour(std::integral_constant<int, 1>());
        colour(std::integral_constant<int, 2>());
        colour(std::integral_constant<int, 3>());
        // 4. the tile of plane ix, updated or not
        const int OY = y1 - y0, OZ = z1 - z0;
        for (int w = t; w < 3 * TY * TZ; w += NT) {
            const int c = w / (TY * TZ), r = w - c * TY * TZ, oy = r / TZ, oz = r - oy * TZ;
            if (oy >= OY || oz >= OZ) continue;
            uout[(size_t)c * N + ((size_t)ix * ny + y0 + oy) * nz + z0 + oz] = pcur[c * CS + uidx(oy + HY, oz + HZ)];
        }
        if (more) {
            // FP32: the next plane step's inputs are loaded now; the other workgroups on the CU keep its
            // SIMDs busy meanwhile
            if (!kPrefetch) {
                fetch_u(rin, ix + 2, pc);
                fetch_u(rnb, ix + 3, pb);
                fetch_E(ix + 1, pe0);
                fetch_E(ix + 2, pe1);
            }
            __syncthreads();
            // plane ix+1 becomes the next step's plane ix'-1; the freed slot takes plane ix'+1
            const int tmp = sA;
            sA = sB;
            sB = tmp;
            store_u(pc, su[sC]);
            store_u(pb, su[sB]);
            store_E(pe0, sE[0]);
            store_E(pe1, sE[1]);
            __syncthreads();
        }
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

// Row of a node in separable-weight form (same inputs and outputs as reg_row). With e = 4i+2j+k the
// element at node - (i,j,k) and s a neighbour at offset o, the structured element matrix gives
//   f[c] = sum_s sum_d F[delta(s)][c][d] W_s[c][d] u_s[d],
//   W_s[c][c] = sum_{e contains s} E_e,   W_s[c][d] = sum_{e contains s} E_e s_c(e) s_d(e)  (c != d).
// Element e contains s iff, per axis a, o_a = 0, or o_a = -1 and i_a = 1, or o_a = +1 and i_a = 0, so the
// four weight sets are separable 2 -> 3 transforms of the 2x2x2 moduli, per axis either "plain"
// (E1, E0 + E1, E0) or "signed" (E1, E1 - E0, -E0) for o_a = -1, 0, +1. Neighbours are taken plane by
// plane (x offset -1, +1, 0) and, inside a plane, in groups of equal delta; each group's weighted sums
// V[c][d] are contracted with its nine F values. About 470 operations per row instead of 8 x 72 products.
template <typename T, bool WANT_M>
__device__ __forceinline__ void w_row(const T *__restrict__ F, const T un[27][3], const T Ee[8], T f[3], T M[9]) {
    // z transforms per (i, j), index oz + 1
    T zp[2][2][3], zs[2][2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const T e0 = Ee[4 * i + 2 * j], e1 = Ee[4 * i + 2 * j + 1];
            zp[i][j][0] = e1;
            zp[i][j][1] = e0 + e1;
            zp[i][j][2] = e0;
            zs[i][j][0] = e1;
            zs[i][j][1] = e1 - e0;
            zs[i][j][2] = -e0;
        }
    // y transforms, index [i][oy + 1][oz + 1]: P plain(zp) for the diagonal, A signed(zp) for xy,
    // B plain(zs) for xz, C signed(zs) for yz
    T P[2][3][3], A[2][3][3], B[2][3][3], C[2][3][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int oz = 0; oz < 3; ++oz) {
            P[i][0][oz] = zp[i][1][oz];
            P[i][1][oz] = zp[i][0][oz] + zp[i][1][oz];
            P[i][2][oz] = zp[i][0][oz];
            A[i][0][oz] = zp[i][1][oz];
            A[i][1][oz] = zp[i][1][oz] - zp[i][0][oz];
            A[i][2][oz] = -zp[i][0][oz];
            B[i][0][oz] = zs[i][1][oz];
            B[i][1][oz] = zs[i][0][oz] + zs[i][1][oz];
            B[i][2][oz] = zs[i][0][oz];
            C[i][0][oz] = zs[i][1][oz];
            C[i][1][oz] = zs[i][1][oz] - zs[i][0][oz];
            C[i][2][oz] = -zs[i][0][oz];
        }
    constexpr int kPlane[3] = {0, 2, 1};
#pragma unroll
    for (int pi = 0; pi < 3; ++pi) {
        const int dx = kPlane[pi];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            T V[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) V[q] = T(0);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    if ((((dy != 1) << 1) | (dz != 1)) != g) continue;
                    T p, qxy, qxz, qyz;
                    if (dx == 0) {  // x offset -1: elements i = 1, s_x = +1
                        p = P[1][dy][dz];
                        qxy = A[1][dy][dz];
                        qxz = B[1][dy][dz];
                        qyz = C[1][dy][dz];
                    } else if (dx == 2) {  // x offset +1: elements i = 0, s_x = -1
                        p = P[0][dy][dz];
                        qxy = -A[0][dy][dz];
                        qxz = -B[0][dy][dz];
                        qyz = C[0][dy][dz];
                    } else {
                        p = P[0][dy][dz] + P[1][dy][dz];
                        qxy = A[1][dy][dz] - A[0][dy][dz];
                        qxz = B[1][dy][dz] - B[0][dy][dz];
                        qyz = C[0][dy][dz] + C[1][dy][dz];
                    }
                    const int s = 9 * dx + 3 * dy + dz;
                    const T u0 = un[s][0], u1 = un[s][1], u2 = un[s][2];
                    V[0] += p * u0;
                    V[4] += p * u1;
                    V[8] += p * u2;
                    V[1] += qxy * u1;
                    V[3] += qxy * u0;
                    V[2] += qxz * u2;
                    V[6] += qxz * u0;
                    V[5] += qyz * u2;
                    V[7] += qyz * u1;
                    if (WANT_M && s == 13) {
                        // the node's own 3x3 block: F[0] with the self weights
                        const T *Fg = F;
                        asm volatile("" : "+s"(Fg));
                        const CT<T> *F0 = (const CT<T> *)Fg;
                        M[0] = F0[0] * p;
                        M[4] = F0[4] * p;
                        M[8] = F0[8] * p;
                        M[1] = F0[1] * qxy;
                        M[3] = F0[3] * qxy;
                        M[2] = F0[2] * qxz;
                        M[6] = F0[6] * qxz;
                        M[5] = F0[5] * qyz;
                        M[7] = F0[7] * qyz;
                    }
                }
            const int delta = ((dx != 1) << 2) | g;
            const T *Fg = F + 16 * delta;
            asm volatile("" : "+s"(Fg));
            const CT<T> *Fp = (const CT<T> *)Fg;
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] += Fp[3 * c] * V[3 * c] + Fp[3 * c + 1] * V[3 * c + 1] + Fp[3 * c + 2] * V[3 * c + 2];
        }
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
    w_row<T, true>(F, un, Ee, f, M);
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
// GS tile: up to 16 x GS_TZ outputs, 256 work-items (the four colour positions cover at most 171/162/136/128
// nodes). FP32 uses 28 along z: a 193-node extent splits into seven tiles of at most 28 either way, and the
// smaller LDS arrays (31.4 KB) with at most 96 registers let five workgroups share a CU instead of four
// (sweep 487 -> 462 us at 384x192x192). FP64 stays at 32 (two workgroups per CU either way; 28 measured
// 2% slower there).
constexpr int GS_TY = 16, GS_NT = 256;
template <typename T>
constexpr int GS_TZ = sizeof(T) == 4 ? 28 : 32;

}  // namespace

int resident_workgroups(const void *kernel, int threads) {
    static std::mutex m;
    static std::map<std::pair<const void *, int>, int> cache;
    std::lock_guard<std::mutex> lock(m);
    const auto key = std::make_pair(kernel, threads);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) {
        (void)hipGetLastError();
        per = 1;
    }
    return cache[key] = per * cu_count();
}

int plane_chunks(int tiles, int n, int want, int resident) {
    int chunks = std::max(1, std::min(n, ceil_div(want, tiles)));
    if (tiles * chunks > resident) chunks = std::max(1, std::min(n, resident / tiles));
    // equal chunks; the count may drop when the rounded-up chunk length covers n in fewer pieces
    return ceil_div(n, ceil_div(n, chunks));
}

template <typename T>
void launch_fine_apply_tiled(const Grid &g, const T *F, const T *E, const T *u, T *y, const T *b,
                             const uint8_t *dmask, int mode, hipStream_t s) {
    if (mode != AM_SET && mode != AM_RESIDUAL) throw std::runtime_error("tiled K u supports the SET and RESIDUAL modes");
    if (g.ydet > 0) {
        const int ntY = ceil_div(g.ydet, KU_TY), ntZ = ceil_div(g.nn[2], KU_TZ);
        const int tiles = ntY * ntZ;
        // a few workgroups per CU; each x-chunk reloads one halo plane at each end
        const void *kern = mode == AM_RESIDUAL ? (const void *)k_ku_tiled<T, AM_RESIDUAL, KU_TY, KU_TZ, KU_NT>
                                               : (const void *)k_ku_tiled<T, AM_SET, KU_TY, KU_TZ, KU_NT>;
        const int chunks = plane_chunks(tiles, g.nn[0], 6 * cu_count(), resident_workgroups(kern, KU_NT));
        const int xchunk = ceil_div(g.nn[0], chunks);
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
                          const uint8_t *dmask, int parity, int forward, hipStream_t s, int zin) {
    if (uin == uout) throw std::runtime_error("the two-phase smoother needs distinct input and output vectors");
    const int nplanes = (g.nn[0] - parity + 1) / 2;
    if (nplanes <= 0) return;
    const int ntY = ceil_div(g.nn[1], GS_TY), ntZ = ceil_div(g.nn[2], GS_TZ<T>);
    const int tiles = ntY * ntZ;
    // as many workgroups as fit at once (two per CU in FP64, five in FP32) and no more; each chunk
    // of planes pays one unpipelined load of its first plane step
    const void *kern = forward ? (const void *)k_gs_phase<T, GS_TY, GS_TZ<T>, GS_NT, true>
                               : (const void *)k_gs_phase<T, GS_TY, GS_TZ<T>, GS_NT, false>;
    const int resident = resident_workgroups(kern, GS_NT);
    const int chunks = plane_chunks(tiles, nplanes, resident, resident);
    const int pchunk = ceil_div(nplanes, chunks);
    if (chunks > 65535) throw std::runtime_error("grid too large for the two-phase smoother");
    if ((size_t)3 * g.nnodes * sizeof(T) >= ((size_t)1 << 30) || (size_t)g.nelems * sizeof(T) >= ((size_t)1 << 30))
        throw std::runtime_error("field too large for the two-phase smoother's 32-bit buffer offsets");
    if (forward)
        hipLaunchKernelGGL((k_gs_phase<T, GS_TY, GS_TZ<T>, GS_NT, true>), dim3(tiles, chunks), dim3(GS_NT), 0, s, g, F, E,
                           uin, unb, uout, b, dmask, parity, ntY, ntZ, pchunk, zin);
    else
        hipLaunchKernelGGL((k_gs_phase<T, GS_TY, GS_TZ<T>, GS_NT, false>), dim3(tiles, chunks), dim3(GS_NT), 0, s, g, F,
                           E, uin, unb, uout, b, dmask, parity, ntY, ntZ, pchunk, zin);
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
                                           hipStream_t, int);
template void launch_fine_gs_phase<float>(const Grid &, const float *, const float *, const float *, const float *,
                                          float *, const float *, const uint8_t *, int, int, hipStream_t, int);

}  // namespace vf
