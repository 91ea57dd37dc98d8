// vf_kernels_setup.hip -- per-density-update setup of the multigrid hierarchy on gfx950:
// Galerkin coarsened element matrices (MultigridSolver.hh:704-819) and the assembly of the 27-point
// block rows of the intermediate levels (level 1 straight from the fine moduli). The coarsest level's
// factorisation lives in vf_kernels_chol.hip.
#include "vf_internal.hpp"

namespace vf {
namespace {

constexpr int TPB = 256;
inline unsigned nblocks(long count) { return (unsigned)((count + TPB - 1) / TPB); }

// ---- level-1 element matrices: Ke(e) = sum_fi E_f Chat_fi   (MultigridSolver.hh:737-743)
__global__ void __launch_bounds__(576) k_galerkin_l1(Grid gc, Grid gf, const double *__restrict__ Chat,
                                                     const double *__restrict__ Ef, double *__restrict__ Ke) {
    const int e = blockIdx.x;
    const int idx = threadIdx.x;
    const int ez = e % gc.ne[2], ey = (e / gc.ne[2]) % gc.ne[1], ex = e / (gc.ne[1] * gc.ne[2]);
    double acc = 0;
    if (ey < gc.ymask) {
#pragma unroll
        for (int fi = 0; fi < 8; ++fi) {
            const int fx = 2 * ex + ((fi >> 2) & 1), fy = 2 * ey + ((fi >> 1) & 1), fz = 2 * ez + (fi & 1);
            acc += Ef[(fx * gf.ne[1] + fy) * gf.ne[2] + fz] * Chat[fi * 576 + idx];
        }
    }
    Ke[(size_t)e * 576 + idx] = acc;
}

// ---- level-2 element matrices straight from the fine densities:
// Ke2(e) = sum_{k<64} E_{fine(e,k)} M2[k],  M2[8 fi2 + fi1] = P_fi2^T Chat_fi1 P_fi2.
// One workgroup of 576 work-items (one per matrix entry) handles EPB elements, so each M2 entry
// fetched from L2 is reused EPB times.
constexpr int EPB = 8;
__global__ void __launch_bounds__(576) k_galerkin_l2(Grid g2, Grid gf, const double *__restrict__ M2,
                                                     const double *__restrict__ Ef, double *__restrict__ Ke) {
    __shared__ double sE[EPB][64];
    const int e0 = blockIdx.x * EPB;
    const int t = threadIdx.x;
    for (int w = t; w < EPB * 64; w += 576) {
        const int j = w / 64, k = w % 64;
        const int e = e0 + j;
        double v = 0;
        if (e < g2.nelems) {
            const int ez = e % g2.ne[2], ey = (e / g2.ne[2]) % g2.ne[1], ex = e / (g2.ne[1] * g2.ne[2]);
            if (ey < g2.ymask) {
                const int f2 = k >> 3, f1 = k & 7;
                const int fx = 4 * ex + 2 * ((f2 >> 2) & 1) + ((f1 >> 2) & 1);
                const int fy = 4 * ey + 2 * ((f2 >> 1) & 1) + ((f1 >> 1) & 1);
                const int fz = 4 * ez + 2 * (f2 & 1) + (f1 & 1);
                v = Ef[(fx * gf.ne[1] + fy) * gf.ne[2] + fz];
            }
        }
        sE[j][k] = v;
    }
    __syncthreads();
    double acc[EPB];
#pragma unroll
    for (int j = 0; j < EPB; ++j) acc[j] = 0;
    for (int k = 0; k < 64; ++k) {
        const double m = M2[k * 576 + t];
#pragma unroll
        for (int j = 0; j < EPB; ++j) acc[j] += sE[j][k] * m;
    }
#pragma unroll
    for (int j = 0; j < EPB; ++j)
        if (e0 + j < g2.nelems) Ke[(size_t)(e0 + j) * 576 + t] = acc[j];
}

// ---- one more level of Galerkin coarsening: Kc(e) = sum_fi P_fi^T Kf(child fi) P_fi
// (accumulateCoarsenedStiffnessMatrix, MultigridSolver.hh:704-722). phi[fi][fine l][coarse l].
__global__ void __launch_bounds__(576) k_galerkin_recurse(Grid gc, Grid gf, const double *__restrict__ phi,
                                                          const double *__restrict__ Kf, double *__restrict__ Kc) {
    __shared__ double sK[576];
    __shared__ double sT[576];
    __shared__ double sphi[512];
    const int e = blockIdx.x;
    const int t = threadIdx.x;
    const int ez = e % gc.ne[2], ey = (e / gc.ne[2]) % gc.ne[1], ex = e / (gc.ne[1] * gc.ne[2]);
    if (t < 512) sphi[t] = phi[t];
    const int a = t / 24, bcol = t % 24;
    double y = 0;
    const bool masked = ey >= gc.ymask;
    for (int fi = 0; fi < 8 && !masked; ++fi) {
        const int fx = 2 * ex + ((fi >> 2) & 1), fy = 2 * ey + ((fi >> 1) & 1), fz = 2 * ez + (fi & 1);
        const size_t ce = (size_t)((fx * gf.ne[1] + fy) * gf.ne[2] + fz);
        __syncthreads();
        sK[t] = Kf[ce * 576 + t];
        __syncthreads();
        // T = Kf * I  : T[a][3j+d] = sum_i Kf[a][3i+d] phi[i][j]
        {
            const int j = bcol / 3, d = bcol % 3;
            double s = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) s += sK[a * 24 + 3 * i + d] * sphi[fi * 64 + i * 8 + j];
            sT[t] = s;
        }
        __syncthreads();
        // Kc += I^T T : Kc[3i2+c][b] += sum_i phi[i][i2] T[3i+c][b]
        {
            const int i2 = a / 3, c = a % 3;
            double s = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) s += sphi[fi * 64 + i * 8 + i2] * sT[(3 * i + c) * 24 + bcol];
            y += s;
        }
    }
    Kc[(size_t)e * 576 + t] = y;
}

// ---- block rows of an intermediate level from its element matrices.
// block(n, slot)[a][c] = sum over elements e containing n and m = n + off(slot) of Ke_e[3 l_n + a][3 l_m + c],
// accumulated in FP64 and stored in the precision of the cycle that reads it.
template <typename OutT>
__global__ void __launch_bounds__(TPB) k_assemble_blocks(Grid g, const double *__restrict__ Ke,
                                                         OutT *__restrict__ blocks) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)g.nnodes * 27) return;
    const int n = (int)(t % g.nnodes), slot = (int)(t / g.nnodes);
    const int nz = g.nn[2], ix = n / (g.nn[1] * nz), iy = (n / nz) % g.nn[1], iz = n % nz;
    const int off[3] = {slot / 9 - 1, (slot / 3) % 3 - 1, slot % 3 - 1};
    const int id[3] = {ix, iy, iz};
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool inside = true;
    for (int d = 0; d < 3; ++d) inside &= (unsigned)(id[d] + off[d]) < (unsigned)g.nn[d];
    if (inside) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = max(max(id[d] - 1, id[d] + off[d] - 1), 0);
            hi[d] = min(min(id[d], id[d] + off[d]), g.ne[d] - 1);
        }
        for (int ex = lo[0]; ex <= hi[0]; ++ex)
            for (int ey = lo[1]; ey <= hi[1]; ++ey)
                for (int ez = lo[2]; ez <= hi[2]; ++ez) {
                    const int ln = 4 * (ix - ex) + 2 * (iy - ey) + (iz - ez);
                    const int lm = 4 * (ix + off[0] - ex) + 2 * (iy + off[1] - ey) + (iz + off[2] - ez);
                    const double *K = Ke + (size_t)((ex * g.ne[1] + ey) * g.ne[2] + ez) * 576;
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) B[3 * a + c] += K[(3 * ln + a) * 24 + 3 * lm + c];
                }
    }
    const BlockSlot bs = block_slot(g, ix, iy, iz);
    for (int k = 0; k < 9; ++k) blocks[bs.base + (size_t)(slot * 9 + k) * bs.cnt + bs.j] = (OutT)B[k];
}

// ---- level-1 block rows straight from the fine moduli: the block of coarse node n and neighbour slot
// is sum over the coarse elements holding both nodes (unmasked: ey < ymask) of sum_fi E_child(fi) *
// Chat_fi[(3 ln + a), (3 lm + c)], i.e. the assembled form of the level-1 operator that
// MultigridSolver.hh:483-491 applies on the fly. One launch covers every (colour class, slot): a
// work-item owns one node of one class, so the colour-major output (BlockSlot) is written contiguously;
// the slot and the element offset fix ln and lm for the whole workgroup, so the Chat entries are uniform
// (scalar) loads and only the child moduli are per work-item. Sums in FP64, stored in OutT.
template <typename OutT>
__global__ void __launch_bounds__(TPB) k_assemble_l1_blocks(Grid gc, Grid gf, const double *__restrict__ Chat,
                                                            const double *__restrict__ Ef,
                                                            OutT *__restrict__ blocks) {
    const int slot = blockIdx.y, colour = blockIdx.z;
    const int cx = (colour >> 2) & 1, cy = (colour >> 1) & 1, cz = colour & 1;
    const int mx = (gc.nn[0] - cx + 1) / 2, my = (gc.nn[1] - cy + 1) / 2, mz = (gc.nn[2] - cz + 1) / 2;
    const int cnt = mx * my * mz;
    const int j = blockIdx.x * TPB + threadIdx.x;
    if (j >= cnt) return;
    const int jx = j / (my * mz), r = j - jx * my * mz, jy = r / mz, jz = r - jy * mz;
    const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
    const int ox = slot / 9 - 1, oy = (slot / 3) % 3 - 1, oz = slot % 3 - 1;
    const int eyEnd = min(gc.ne[1], gc.ymask);
    const int fyz = gf.ne[1] * gf.ne[2];
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = 0; e < 8; ++e) {
        // element at node - (i, jj, k); it holds m = n + o iff every local coordinate i + o stays in {0, 1}
        const int i = (e >> 2) & 1, jj = (e >> 1) & 1, k = e & 1;
        if (i + ox < 0 || i + ox > 1 || jj + oy < 0 || jj + oy > 1 || k + oz < 0 || k + oz > 1) continue;
        const int ex = ix - i, ey = iy - jj, ez = iz - k;
        if ((unsigned)ex >= (unsigned)gc.ne[0] || (unsigned)ey >= (unsigned)eyEnd || (unsigned)ez >= (unsigned)gc.ne[2])
            continue;
        const int ln = e, lm = 4 * (i + ox) + 2 * (jj + oy) + (k + oz);
        const size_t fb = ((size_t)(2 * ex) * gf.ne[1] + 2 * ey) * gf.ne[2] + 2 * ez;
        double Ec[8];
#pragma unroll
        for (int fi = 0; fi < 8; ++fi) Ec[fi] = Ef[fb + ((fi >> 2) & 1) * fyz + ((fi >> 1) & 1) * gf.ne[2] + (fi & 1)];
#pragma unroll
        for (int fi = 0; fi < 8; ++fi) {
            const double *C = Chat + fi * 576 + (3 * ln) * 24 + 3 * lm;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) B[3 * a + c] += Ec[fi] * C[a * 24 + c];
        }
    }
    size_t before = 0;
    for (int q = 0; q < colour; ++q) {
        const long qx = (gc.nn[0] - ((q >> 2) & 1) + 1) / 2, qy = (gc.nn[1] - ((q >> 1) & 1) + 1) / 2,
                   qz = (gc.nn[2] - (q & 1) + 1) / 2;
        before += (size_t)(qx * qy * qz);
    }
    OutT *out = blocks + before * 243 + (size_t)slot * 9 * cnt + j;
#pragma unroll
    for (int q = 0; q < 9; ++q) out[(size_t)q * cnt] = (OutT)B[q];
}

}  // namespace

void launch_galerkin_l1_elems(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, double *Ke,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_galerkin_l1, dim3(gc.nelems), dim3(576), 0, s, gc, gf, Chat, Ef, Ke);
    VF_HIP(hipGetLastError());
}

void launch_galerkin_l2_elems(const Grid &g2, const Grid &gf, const double *M2, const double *Ef, double *Ke,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_galerkin_l2, dim3((g2.nelems + EPB - 1) / EPB), dim3(576), 0, s, g2, gf, M2, Ef, Ke);
    VF_HIP(hipGetLastError());
}

void launch_galerkin_recurse(const Grid &gc, const Grid &gf, const double *phi, const double *Kf, double *Kc,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_galerkin_recurse, dim3(gc.nelems), dim3(576), 0, s, gc, gf, phi, Kf, Kc);
    VF_HIP(hipGetLastError());
}

template <typename OutT>
void launch_assemble_blocks(const Grid &g, const double *Ke, OutT *blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_assemble_blocks<OutT>, dim3(nblocks((long)g.nnodes * 27)), dim3(TPB), 0, s, g, Ke, blocks);
    VF_HIP(hipGetLastError());
}

template <typename OutT>
void launch_assemble_l1_blocks(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, OutT *blocks,
                               hipStream_t s) {
    const long maxClass = (long)((gc.nn[0] + 1) / 2) * ((gc.nn[1] + 1) / 2) * ((gc.nn[2] + 1) / 2);
    hipLaunchKernelGGL(k_assemble_l1_blocks<OutT>, dim3(nblocks(maxClass), 27, 8), dim3(TPB), 0, s, gc, gf, Chat, Ef,
                       blocks);
    VF_HIP(hipGetLastError());
}

template void launch_assemble_blocks<double>(const Grid &, const double *, double *, hipStream_t);
template void launch_assemble_blocks<float>(const Grid &, const double *, float *, hipStream_t);
template void launch_assemble_l1_blocks<double>(const Grid &, const Grid &, const double *, const double *, double *,
                                                hipStream_t);
template void launch_assemble_l1_blocks<float>(const Grid &, const Grid &, const double *, const double *, float *,
                                               hipStream_t);

}  // namespace vf
