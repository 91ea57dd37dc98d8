// vf_kernels_setup.hip -- per-density-update setup of the multigrid hierarchy on gfx950:
// Galerkin coarsened element matrices (MultigridSolver.hh:704-819), assembly of the 27-point
// block rows of the intermediate levels, assembly of the reduced coarsest operator and its
// FP64 inverse (replacing the CHOLMOD factorisation of TensorProductSimulator.hh:1198-1230),
// and the coarse solve as one dense FP64 matrix-vector product.
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
// block(n, slot)[a][c] = sum over elements e containing n and m = n + off(slot) of Ke_e[3 l_n + a][3 l_m + c].
__global__ void __launch_bounds__(TPB) k_assemble_blocks(Grid g, const double *__restrict__ Ke,
                                                         double *__restrict__ blocks) {
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
    for (int k = 0; k < 9; ++k) blocks[bs.base + (size_t)(slot * 9 + k) * bs.cnt + bs.j] = B[k];
}

// ---- level-1 block rows straight from the fine moduli: the block of coarse node n and neighbour slot
// is sum over the coarse elements holding both nodes (unmasked: ey < ymask) of sum_fi E_child(fi) *
// Chat_fi[(3 ln + a), (3 lm + c)], i.e. the assembled form of the level-1 operator that
// MultigridSolver.hh:483-491 applies on the fly. Colour-major output layout (BlockSlot).
__global__ void __launch_bounds__(TPB) k_assemble_l1_blocks(Grid gc, Grid gf, const double *__restrict__ Chat,
                                                            const double *__restrict__ Ef,
                                                            double *__restrict__ blocks) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)gc.nnodes * 27) return;
    const int n = (int)(t % gc.nnodes), slot = (int)(t / gc.nnodes);
    const int nz = gc.nn[2], ix = n / (gc.nn[1] * nz), iy = (n / nz) % gc.nn[1], iz = n % nz;
    const int off[3] = {slot / 9 - 1, (slot / 3) % 3 - 1, slot % 3 - 1};
    const int id[3] = {ix, iy, iz};
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool inside = true;
    for (int d = 0; d < 3; ++d) inside &= (unsigned)(id[d] + off[d]) < (unsigned)gc.nn[d];
    if (inside) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = max(max(id[d] - 1, id[d] + off[d] - 1), 0);
            hi[d] = min(min(id[d], id[d] + off[d]), gc.ne[d] - 1);
        }
        hi[1] = min(hi[1], gc.ymask - 1);
        for (int ex = lo[0]; ex <= hi[0]; ++ex)
            for (int ey = lo[1]; ey <= hi[1]; ++ey)
                for (int ez = lo[2]; ez <= hi[2]; ++ez) {
                    const int ln = 4 * (ix - ex) + 2 * (iy - ey) + (iz - ez);
                    const int lm = 4 * (ix + off[0] - ex) + 2 * (iy + off[1] - ey) + (iz + off[2] - ez);
                    for (int fi = 0; fi < 8; ++fi) {
                        const int fx = 2 * ex + ((fi >> 2) & 1), fy = 2 * ey + ((fi >> 1) & 1), fz = 2 * ez + (fi & 1);
                        const double Ee = Ef[(fx * gf.ne[1] + fy) * gf.ne[2] + fz];
                        const double *C = Chat + fi * 576 + (3 * ln) * 24 + 3 * lm;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) B[3 * a + c] += Ee * C[a * 24 + c];
                    }
                }
    }
    const BlockSlot bs = block_slot(gc, ix, iy, iz);
    for (int k = 0; k < 9; ++k) blocks[bs.base + (size_t)(slot * 9 + k) * bs.cnt + bs.j] = B[k];
}

// ---- reduced dense coarsest operator: one work-item per (row DoF, neighbour slot)
__global__ void __launch_bounds__(TPB) k_assemble_dense(Grid g, const double *__restrict__ Ke,
                                                        const int *__restrict__ red, int ld, double *__restrict__ A) {
    const long t = (long)blockIdx.x * TPB + threadIdx.x;
    if (t >= (long)g.nnodes * 27 * 3) return;
    const int row = (int)(t % (3L * g.nnodes)), slot = (int)(t / (3L * g.nnodes));
    const int rr = red[row];
    if (rr < 0) return;
    const int n = row / 3, a = row % 3;
    const int nz = g.nn[2], ix = n / (g.nn[1] * nz), iy = (n / nz) % g.nn[1], iz = n % nz;
    const int off[3] = {slot / 9 - 1, (slot / 3) % 3 - 1, slot % 3 - 1};
    const int id[3] = {ix, iy, iz};
    for (int d = 0; d < 3; ++d)
        if ((unsigned)(id[d] + off[d]) >= (unsigned)g.nn[d]) return;
    const int m = ((ix + off[0]) * g.nn[1] + iy + off[1]) * nz + iz + off[2];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = max(max(id[d] - 1, id[d] + off[d] - 1), 0);
        hi[d] = min(min(id[d], id[d] + off[d]), g.ne[d] - 1);
    }
    double B[3] = {0, 0, 0};
    for (int ex = lo[0]; ex <= hi[0]; ++ex)
        for (int ey = lo[1]; ey <= hi[1]; ++ey)
            for (int ez = lo[2]; ez <= hi[2]; ++ez) {
                const int ln = 4 * (ix - ex) + 2 * (iy - ey) + (iz - ez);
                const int lm = 4 * (ix + off[0] - ex) + 2 * (iy + off[1] - ey) + (iz + off[2] - ez);
                const double *K = Ke + (size_t)((ex * g.ne[1] + ey) * g.ne[2] + ez) * 576;
                for (int c = 0; c < 3; ++c) B[c] += K[(3 * ln + a) * 24 + 3 * lm + c];
            }
    for (int c = 0; c < 3; ++c) {
        const int cc = red[3 * m + c];
        if (cc >= 0) A[(size_t)rr * ld + cc] = B[c];
    }
}

__global__ void k_set_diag(int n0, int n1, int ld, double *A) {
    const int i = n0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1) A[(size_t)i * ld + i] = 1.0;
}

// ---- blocked Gauss-Jordan sweep (in place, returns -A^-1). BS x BS pivot blocks.
constexpr int BS = 64;

// Sweep the pivot block D = A[P,P] in LDS: D <- -D^-1 (scalar sweep operator), store +D^-1 in Dinv.
__global__ void __launch_bounds__(256) k_pivot_inverse(int p, int ld, const double *__restrict__ A,
                                                       double *__restrict__ Dinv) {
    __shared__ double D[BS][BS + 1];
    __shared__ double colk[BS], rowk[BS];
    const int t = threadIdx.x;
    const size_t base = (size_t)p * BS;
    for (int w = t; w < BS * BS; w += 256) D[w / BS][w % BS] = A[(base + w / BS) * ld + base + w % BS];
    __syncthreads();
    for (int k = 0; k < BS; ++k) {
        const double dinv = 1.0 / D[k][k];
        if (t < BS) {
            colk[t] = D[t][k];
            rowk[t] = D[k][t] * dinv;
        }
        __syncthreads();
        // each work-item owns a 4x4 sub-block: rows (t/16)*4.., columns (t%16)*4..
        const int i0 = (t >> 4) * 4, j0 = (t & 15) * 4;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int i = i0 + a;
            const double ci = colk[i];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = j0 + b;
                double v;
                if (i == k && j == k) v = -dinv;
                else if (i == k) v = rowk[j];
                else if (j == k) v = ci * dinv;
                else v = D[i][j] - ci * rowk[j];
                D[i][j] = v;
            }
        }
        __syncthreads();
    }
    for (int w = t; w < BS * BS; w += 256) Dinv[w] = -D[w / BS][w % BS];
}

// W[:, J] = Dinv * A[P, J] and CD[I, :] = A[I, P] * Dinv, Cold[I,:] = A[I,P] for every 64-tile J / I.
// blockIdx.y == 0: row panel (W); blockIdx.y == 1: column panel (Cold, CD).
__global__ void __launch_bounds__(256) k_panels(int p, int ld, const double *__restrict__ A,
                                                const double *__restrict__ Dinv, double *__restrict__ W,
                                                double *__restrict__ Cold, double *__restrict__ CD) {
    __shared__ double sD[BS][BS + 1];
    __shared__ double sX[BS][BS + 1];
    const int t = threadIdx.x, J = blockIdx.x;
    const size_t P0 = (size_t)p * BS, J0 = (size_t)J * BS;
    for (int w = t; w < BS * BS; w += 256) sD[w / BS][w % BS] = Dinv[w];
    if (blockIdx.y == 0) {
        for (int w = t; w < BS * BS; w += 256) sX[w / BS][w % BS] = A[(P0 + w / BS) * ld + J0 + w % BS];
        __syncthreads();
        // W[i][J0+j] = sum_k D[i][k] X[k][j]
        for (int w = t; w < BS * BS; w += 256) {
            const int i = w / BS, j = w % BS;
            double s = 0;
#pragma unroll 16
            for (int k = 0; k < BS; ++k) s += sD[i][k] * sX[k][j];
            W[(size_t)i * ld + J0 + j] = s;
        }
    } else {
        for (int w = t; w < BS * BS; w += 256) {
            const double v = A[(J0 + w / BS) * ld + P0 + w % BS];
            sX[w / BS][w % BS] = v;
            Cold[(J0 + w / BS) * BS + w % BS] = v;
        }
        __syncthreads();
        // CD[J0+i][j] = sum_k X[i][k] D[k][j]
        for (int w = t; w < BS * BS; w += 256) {
            const int i = w / BS, j = w % BS;
            double s = 0;
#pragma unroll 16
            for (int k = 0; k < BS; ++k) s += sX[i][k] * sD[k][j];
            CD[(J0 + i) * BS + j] = s;
        }
    }
}

// Tile (I,J) of the sweep update: A_IJ -= Cold_I W_J off the pivot row/column; the pivot row and
// column take W and CD, the pivot block takes -Dinv.
__global__ void __launch_bounds__(256) k_sweep_update(int p, int ld, double *__restrict__ A,
                                                      const double *__restrict__ Dinv, const double *__restrict__ W,
                                                      const double *__restrict__ Cold, const double *__restrict__ CD) {
    const int I = blockIdx.y, J = blockIdx.x, t = threadIdx.x;
    const size_t I0 = (size_t)I * BS, J0 = (size_t)J * BS;
    if (I == p && J == p) {
        for (int w = t; w < BS * BS; w += 256) A[(I0 + w / BS) * ld + J0 + w % BS] = -Dinv[w];
        return;
    }
    if (I == p) {
        for (int w = t; w < BS * BS; w += 256) A[(I0 + w / BS) * ld + J0 + w % BS] = W[(size_t)(w / BS) * ld + J0 + w % BS];
        return;
    }
    if (J == p) {
        for (int w = t; w < BS * BS; w += 256) A[(I0 + w / BS) * ld + J0 + w % BS] = CD[(I0 + w / BS) * BS + w % BS];
        return;
    }
    __shared__ double sC[BS][BS + 1];
    __shared__ double sW[BS][BS + 1];
    for (int w = t; w < BS * BS; w += 256) {
        sC[w / BS][w % BS] = Cold[(I0 + w / BS) * BS + w % BS];
        sW[w / BS][w % BS] = W[(size_t)(w / BS) * ld + J0 + w % BS];
    }
    __syncthreads();
    // each work-item owns a 4x4 sub-tile
    const int ti = (t / 16) * 4, tj = (t % 16) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0;
    for (int k = 0; k < BS; ++k) {
        double c[4], w[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a] = sC[ti + a][k];
#pragma unroll
        for (int b = 0; b < 4; ++b) w[b] = sW[k][tj + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += c[a] * w[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) A[(I0 + ti + a) * ld + J0 + tj + b] -= acc[a][b];
}

// ---- coarse solve: x_red = -(A_neg) b_red  with one wavefront per row
template <typename T>
__global__ void __launch_bounds__(TPB) k_coarse_gather(int nred, const int *__restrict__ redToFull, int nnodes,
                                                       const T *__restrict__ b, double *__restrict__ out) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= nred) return;
    const int f = redToFull[i];
    out[i] = (double)b[(f % 3) * nnodes + f / 3];
}

__global__ void __launch_bounds__(TPB) k_gemv_neg(int n, int ld, const double *__restrict__ A,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    const double *Ar = A + (size_t)row * ld;
    double s = 0;
    for (int j = lane; j < n; j += 64) s += Ar[j] * x[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[row] = -s;
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_coarse_scatter(int nred, const int *__restrict__ redToFull, int nnodes,
                                                        const double *__restrict__ in, T *__restrict__ x) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= nred) return;
    const int f = redToFull[i];
    x[(f % 3) * nnodes + f / 3] = (T)in[i];
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fill_plain(size_t n, T *x, T v) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) x[i] = v;
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

void launch_assemble_blocks(const Grid &g, const double *Ke, double *blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_assemble_blocks, dim3(nblocks((long)g.nnodes * 27)), dim3(TPB), 0, s, g, Ke, blocks);
    VF_HIP(hipGetLastError());
}

void launch_assemble_l1_blocks(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, double *blocks,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_assemble_l1_blocks, dim3(nblocks((long)gc.nnodes * 27)), dim3(TPB), 0, s, gc, gf, Chat, Ef,
                       blocks);
    VF_HIP(hipGetLastError());
}

void launch_assemble_dense(const Grid &g, const double *Ke, const int *red, int ld, double *A, hipStream_t s) {
    const size_t total = (size_t)ld * ld;
    hipLaunchKernelGGL(k_fill_plain<double>, dim3((unsigned)((total + TPB - 1) / TPB)), dim3(TPB), 0, s, total, A, 0.0);
    hipLaunchKernelGGL(k_assemble_dense, dim3(nblocks((long)g.nnodes * 81)), dim3(TPB), 0, s, g, Ke, red, ld, A);
    VF_HIP(hipGetLastError());
}

// A: ld x ld with ld a multiple of 64; rows/cols [n, ld) are padding (identity is written here).
void spd_inverse_inplace(int n, int ld, double *A, double *scratch, hipStream_t s) {
    if (ld % BS) throw std::runtime_error("spd_inverse_inplace: leading dimension must be a multiple of 64");
    if (ld > n) hipLaunchKernelGGL(k_set_diag, dim3((ld - n + 255) / 256), dim3(256), 0, s, n, ld, ld, A);
    const int nb = ld / BS;
    double *Dinv = scratch;
    double *W = Dinv + BS * BS;
    double *Cold = W + (size_t)BS * ld;
    double *CD = Cold + (size_t)ld * BS;
    for (int p = 0; p < nb; ++p) {
        hipLaunchKernelGGL(k_pivot_inverse, dim3(1), dim3(256), 0, s, p, ld, A, Dinv);
        hipLaunchKernelGGL(k_panels, dim3(nb, 2), dim3(256), 0, s, p, ld, A, Dinv, W, Cold, CD);
        hipLaunchKernelGGL(k_sweep_update, dim3(nb, nb), dim3(256), 0, s, p, ld, A, Dinv, W, Cold, CD);
    }
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_coarse_apply(int nred, const double *Ainv_neg, const int *redToFull, int nnodes, const T *b, T *x,
                         double *tmp_in, double *tmp_out, hipStream_t s) {
    const int ld = (nred + BS - 1) / BS * BS;
    hipLaunchKernelGGL(k_fill_plain<T>, dim3(nblocks(3L * nnodes)), dim3(TPB), 0, s, (size_t)3 * nnodes, x, T(0));
    if (nred == 0) return;
    hipLaunchKernelGGL(k_coarse_gather<T>, dim3(nblocks(nred)), dim3(TPB), 0, s, nred, redToFull, nnodes, b, tmp_in);
    hipLaunchKernelGGL(k_gemv_neg, dim3((nred + 3) / 4), dim3(TPB), 0, s, nred, ld, Ainv_neg, tmp_in, tmp_out);
    hipLaunchKernelGGL(k_coarse_scatter<T>, dim3(nblocks(nred)), dim3(TPB), 0, s, nred, redToFull, nnodes, tmp_out, x);
    VF_HIP(hipGetLastError());
}

template void launch_coarse_apply<double>(int, const double *, const int *, int, const double *, double *, double *,
                                          double *, hipStream_t);
template void launch_coarse_apply<float>(int, const double *, const int *, int, const float *, float *, double *,
                                         double *, hipStream_t);

}  // namespace vf
