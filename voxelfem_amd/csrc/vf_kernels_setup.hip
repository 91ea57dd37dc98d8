// vf_kernels_setup.hip -- per-density-update setup of the multigrid hierarchy on gfx950:
// Galerkin coarsened element matrices (MultigridSolver.hh:704-819) and the assembly of the 27-point
// block rows of the intermediate levels (level 1 straight from the fine moduli). The coarsest level's
// factorisation lives in vf_kernels_chol.hip.
#include "vf_internal.hpp"

#include <algorithm>

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
constexpr int EPB = 16;
// Only the element rows ey in [ylo, yhi) are computed (all rows: [0, ne[1])); the workgroups enumerate
// that band's elements in C order.
__global__ void __launch_bounds__(576) k_galerkin_l2(Grid g2, Grid gf, const double *__restrict__ M2,
                                                     const double *__restrict__ Ef, double *__restrict__ Ke, int ylo,
                                                     int yhi) {
    __shared__ double sE[EPB][64];
    const int e0 = blockIdx.x * EPB;
    const int t = threadIdx.x;
    const int rows = yhi - ylo, nband = g2.ne[0] * rows * g2.ne[2];
    auto band_elem = [&](int b) {
        const int ez = b % g2.ne[2], r = (b / g2.ne[2]) % rows, ex = b / (rows * g2.ne[2]);
        return (ex * g2.ne[1] + ylo + r) * g2.ne[2] + ez;
    };
    for (int w = t; w < EPB * 64; w += 576) {
        const int j = w / 64, k = w % 64;
        double v = 0;
        if (e0 + j < nband) {
            const int e = band_elem(e0 + j);
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
        if (e0 + j < nband) Ke[(size_t)band_elem(e0 + j) * 576 + t] = acc[j];
}

// ---- one more level of Galerkin coarsening: Kc(e) = sum_fi P_fi^T Kf(child fi) P_fi
// (accumulateCoarsenedStiffnessMatrix, MultigridSolver.hh:704-722). phi[fi][fine l][coarse l].
// The eight child matrices are fetched together (eight independent loads per work-item) and staged in LDS,
// so the workgroup waits on global memory once; T = Kf P goes through a double-buffered LDS tile, one
// barrier per child.
// One workgroup per element of the element rows [ylo, yhi).
__global__ void __launch_bounds__(576) k_galerkin_recurse(Grid gc, Grid gf, const double *__restrict__ phi,
                                                          const double *__restrict__ Kf, double *__restrict__ Kc,
                                                          int ylo, int yhi) {
    __shared__ double sK[8][576];
    __shared__ double sT[2][576];
    __shared__ double sphi[512];
    const int rows = yhi - ylo;
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    const int ez = b % gc.ne[2], ey = ylo + (b / gc.ne[2]) % rows, ex = b / (rows * gc.ne[2]);
    const int e = (ex * gc.ne[1] + ey) * gc.ne[2] + ez;
    if (ey >= gc.ymask) {  // uniform over the workgroup
        Kc[(size_t)e * 576 + t] = 0;
        return;
    }
    if (t < 512) sphi[t] = phi[t];
    double v[8];
#pragma unroll
    for (int fi = 0; fi < 8; ++fi) {
        const int fx = 2 * ex + ((fi >> 2) & 1), fy = 2 * ey + ((fi >> 1) & 1), fz = 2 * ez + (fi & 1);
        v[fi] = Kf[(size_t)((fx * gf.ne[1] + fy) * gf.ne[2] + fz) * 576 + t];
    }
#pragma unroll
    for (int fi = 0; fi < 8; ++fi) sK[fi][t] = v[fi];
    __syncthreads();
    const int a = t / 24, bcol = t % 24;
    const int j = bcol / 3, d = bcol % 3, i2 = a / 3, c = a % 3;
    double y = 0;
    for (int fi = 0; fi < 8; ++fi) {
        // T = Kf P : T[a][3j+d] = sum_i Kf[a][3i+d] phi[i][j]
        double s = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) s += sK[fi][a * 24 + 3 * i + d] * sphi[fi * 64 + i * 8 + j];
        sT[fi & 1][t] = s;
        __syncthreads();
        // Kc += P^T T : Kc[3i2+c][b] += sum_i phi[i][i2] T[3i+c][b]
        double r = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) r += sphi[fi * 64 + i * 8 + i2] * sT[fi & 1][(3 * i + c) * 24 + bcol];
        y += r;
    }
    Kc[(size_t)e * 576 + t] = y;
}

// ---- block rows straight from element data, one work-item per node of one colour class.
// Element (i, jj, k) "below" node n = (ix, iy, iz) is the one at (ix - i, iy - jj, iz - k); n is its local
// node ln = 4 i + 2 jj + k, and m = n + o lies in it iff every i + o, jj + o, k + o stays in {0, 1}, with
// lm = 4 (i + ox) + 2 (jj + oy) + (k + oz). The 27 slots split into three groups of nine by ox; in group ox
// only the elements with lx = i + ox in {0, 1} contribute, and for one element and one matrix row 3 ln + a
// the twelve entries 3 lm + c of that group (lx fixed) are contiguous. A work-item holds one group's 81
// FP64 sums in registers, so every element datum it needs is read once per group, and the colour-major
// block layout (BlockSlot) makes each slot entry a contiguous run over the class.
struct ClassNode {
    int ix, iy, iz, j, cnt;
    size_t before;  // nodes of the lower colour classes
};
__device__ inline bool class_node(const Grid &g, int colour, ClassNode &r) {
    const int cx = (colour >> 2) & 1, cy = (colour >> 1) & 1, cz = colour & 1;
    const int mx = (g.nn[0] - cx + 1) / 2, my = (g.nn[1] - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2;
    r.cnt = mx * my * mz;
    r.j = blockIdx.x * TPB + threadIdx.x;
    if (r.j >= r.cnt) return false;
    const int jx = r.j / (my * mz), q = r.j - jx * my * mz, jy = q / mz, jz = q - jy * mz;
    r.ix = 2 * jx + cx;
    r.iy = 2 * jy + cy;
    r.iz = 2 * jz + cz;
    r.before = 0;
    for (int p = 0; p < colour; ++p) {
        const long px = (g.nn[0] - ((p >> 2) & 1) + 1) / 2, py = (g.nn[1] - ((p >> 1) & 1) + 1) / 2,
                   pz = (g.nn[2] - (p & 1) + 1) / 2;
        r.before += (size_t)(px * py * pz);
    }
    return true;
}

// group sums of an intermediate level from its element matrices Ke (576 FP64 per element, row-major 24x24)
// Only the nodes of rows iy in [nylo, nyhi] are assembled.
template <typename OutT>
__global__ void __launch_bounds__(TPB) k_assemble_blocks(Grid g, const double *__restrict__ Ke,
                                                         OutT *__restrict__ blocks, int nylo, int nyhi) {
    ClassNode n;
    if (!class_node(g, blockIdx.y, n)) return;
    if (n.iy < nylo || n.iy > nyhi) return;
    OutT *out = blocks + n.before * 243 + n.j;
    for (int gx = 0; gx < 3; ++gx) {
        const int ox = gx - 1;
        double acc[81];
#pragma unroll
        for (int q = 0; q < 81; ++q) acc[q] = 0;
        for (int i = 0; i < 2; ++i) {
            const int lx = i + ox;
            if (lx < 0 || lx > 1) continue;
            const int ex = n.ix - i;
            const bool xin = (unsigned)ex < (unsigned)g.ne[0];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int ey = n.iy - jj, ez = n.iz - k;
                    if (!xin || (unsigned)ey >= (unsigned)g.ne[1] || (unsigned)ez >= (unsigned)g.ne[2]) continue;
                    const int ln = 4 * i + 2 * jj + k;
                    const double *K = Ke + (size_t)((ex * g.ne[1] + ey) * g.ne[2] + ez) * 576 + 12 * lx;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const double2 *row = reinterpret_cast<const double2 *>(K + (3 * ln + a) * 24);
                        double v[12];
#pragma unroll
                        for (int h = 0; h < 6; ++h) {
                            const double2 w = row[h];
                            v[2 * h] = w.x;
                            v[2 * h + 1] = w.y;
                        }
#pragma unroll
                        for (int ly = 0; ly < 2; ++ly)
#pragma unroll
                            for (int lz = 0; lz < 2; ++lz) {
                                const int s9 = (ly - jj + 1) * 3 + (lz - k + 1);
#pragma unroll
                                for (int c = 0; c < 3; ++c) acc[s9 * 9 + 3 * a + c] += v[6 * ly + 3 * lz + c];
                            }
                    }
                }
        }
#pragma unroll
        for (int q = 0; q < 81; ++q) out[(size_t)(gx * 81 + q) * n.cnt] = (OutT)acc[q];
    }
}

// ---- level-1 block rows straight from the fine moduli: the element datum is sum_fi E_child(fi) *
// Chat_fi[(3 ln + a), (3 lm + c)] over the coarse elements holding both nodes (unmasked: ey < ymask), i.e.
// the assembled form of the level-1 operator that MultigridSolver.hh:483-491 applies on the fly. The
// Chat rows are uniform over the workgroup (scalar loads); only the eight child moduli per element are
// per work-item. Sums in FP64, stored in OutT.
// REFL: the element matrix has the reflection symmetry of a box element with an orthotropic tensor
// (K0[(p ^ e_d, a), (q ^ e_d, c)] = S_d(a) S_d(c) K0[(p, a), (q, c)], S_d(a) = -1 iff a == d), so reflecting
// child 0 gives every child: Chat_fi[(l, a), (m, c)] = S_fi(a) S_fi(c) Chat_0[(l ^ fi, a), (m ^ fi, c)],
// S_fi(a) = -1 iff fi has the bit of axis a set. The scalar working set is then Chat_0 (4.6 KB) instead of
// all eight children (37 KB), which keeps it resident in the scalar cache; the sign is a free source
// negation of the FMA.
template <typename OutT, bool REFL>
__global__ void __launch_bounds__(TPB) k_assemble_l1_blocks(Grid gc, Grid gf, const double *__restrict__ Chat,
                                                            const double *__restrict__ Ef,
                                                            OutT *__restrict__ blocks, int nylo, int nyhi) {
    ClassNode n;
    if (!class_node(gc, blockIdx.y, n)) return;
    if (n.iy < nylo || n.iy > nyhi) return;
    const int eyEnd = min(gc.ne[1], gc.ymask);
    const int fz = gf.ne[2], fyz = gf.ne[1] * gf.ne[2];
    OutT *out = blocks + n.before * 243 + n.j;
    // slot groups (ox, oy): three slots (oz) of nine entries each, 27 FP64 sums in registers
    for (int gxy = 0; gxy < 9; ++gxy) {
        const int ox = gxy / 3 - 1, oy = gxy % 3 - 1;
        double acc[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) acc[q] = 0;
        for (int i = 0; i < 2; ++i) {
            const int lx = i + ox;
            if (lx < 0 || lx > 1) continue;
            for (int jj = 0; jj < 2; ++jj) {
                const int ly = jj + oy;
                if (ly < 0 || ly > 1) continue;
                const int ex = n.ix - i, ey = n.iy - jj;
                const bool xyin = (unsigned)ex < (unsigned)gc.ne[0] && (unsigned)ey < (unsigned)eyEnd;
                // child moduli of the two elements (i, jj, k); an absent element counts as zero moduli
                double E[2][8];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int ez = n.iz - k;
                    const bool in = xyin && (unsigned)ez < (unsigned)gc.ne[2];
                    const size_t fb = in ? ((size_t)(2 * ex) * gf.ne[1] + 2 * ey) * fz + 2 * ez : 0;
#pragma unroll
                    for (int fi = 0; fi < 8; ++fi) {
                        const double w = Ef[fb + ((fi >> 2) & 1) * fyz + ((fi >> 1) & 1) * fz + (fi & 1)];
                        E[k][fi] = in ? w : 0.0;
                    }
                }
                if (REFL) {
                    // Chat_0 row (3 l + a) and column block (lx', ly') with the x and y bits reflected for
                    // the children with fx = 1 and fy = 1: a runtime base per (fx, fy) plus compile-time
                    // offsets; each row serves both elements k (child fi = (fx, fy, k ^ rz))
#pragma unroll
                    for (int fx = 0; fx < 2; ++fx)
#pragma unroll
                        for (int fy = 0; fy < 2; ++fy) {
                            const double *base = Chat + 288 * (i ^ fx) + 144 * (jj ^ fy) + 12 * (lx ^ fx) +
                                                 6 * (ly ^ fy);
#pragma unroll
                            for (int rz = 0; rz < 2; ++rz)
#pragma unroll
                                for (int a = 0; a < 3; ++a) {
                                    const double *row = base + (3 * rz + a) * 24;
#pragma unroll
                                    for (int k = 0; k < 2; ++k) {
                                        const int fzb = k ^ rz, fi = 4 * fx + 2 * fy + fzb;
                                        const bool sa = (fi >> (2 - a)) & 1;
#pragma unroll
                                        for (int lz = 0; lz < 2; ++lz) {
                                            const int s3 = lz - k + 1;
#pragma unroll
                                            for (int c = 0; c < 3; ++c) {
                                                const double e = E[k][fi];
                                                const bool neg = sa != (bool)((fi >> (2 - c)) & 1);
                                                acc[s3 * 9 + 3 * a + c] += (neg ? -e : e) * row[3 * (lz ^ fzb) + c];
                                            }
                                        }
                                    }
                                }
                        }
                } else {
                    const double *base = Chat + 288 * i + 144 * jj + 12 * lx + 6 * ly;
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int fi = 0; fi < 8; ++fi)
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                // row 3 ln + a, columns 3 lm + c (lm = 4 lx + 2 ly + lz) of child fi's matrix
                                const double *row = base + fi * 576 + (3 * k + a) * 24;
#pragma unroll
                                for (int lz = 0; lz < 2; ++lz) {
                                    const int s3 = lz - k + 1;
#pragma unroll
                                    for (int c = 0; c < 3; ++c) acc[s3 * 9 + 3 * a + c] += E[k][fi] * row[3 * lz + c];
                                }
                            }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 27; ++q) out[(size_t)(gxy * 27 + q) * n.cnt] = (OutT)acc[q];
    }
}

}  // namespace

void launch_galerkin_l1_elems(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, double *Ke,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_galerkin_l1, dim3(gc.nelems), dim3(576), 0, s, gc, gf, Chat, Ef, Ke);
    VF_HIP(hipGetLastError());
}

void launch_galerkin_l2_elems(const Grid &g2, const Grid &gf, const double *M2, const double *Ef, double *Ke,
                              hipStream_t s, int ylo, int yhi) {
    if (yhi < 0) yhi = g2.ne[1];
    ylo = std::max(ylo, 0);
    yhi = std::min(yhi, g2.ne[1]);
    if (yhi <= ylo) return;
    const long nband = (long)g2.ne[0] * (yhi - ylo) * g2.ne[2];
    hipLaunchKernelGGL(k_galerkin_l2, dim3((unsigned)((nband + EPB - 1) / EPB)), dim3(576), 0, s, g2, gf, M2, Ef, Ke,
                       ylo, yhi);
    VF_HIP(hipGetLastError());
}

void launch_galerkin_recurse(const Grid &gc, const Grid &gf, const double *phi, const double *Kf, double *Kc,
                             hipStream_t s, int ylo, int yhi) {
    if (yhi < 0) yhi = gc.ne[1];
    ylo = std::max(ylo, 0);
    yhi = std::min(yhi, gc.ne[1]);
    if (yhi <= ylo) return;
    const long nband = (long)gc.ne[0] * (yhi - ylo) * gc.ne[2];
    hipLaunchKernelGGL(k_galerkin_recurse, dim3((unsigned)nband), dim3(576), 0, s, gc, gf, phi, Kf, Kc, ylo, yhi);
    VF_HIP(hipGetLastError());
}

template <typename OutT>
void launch_assemble_blocks(const Grid &g, const double *Ke, OutT *blocks, hipStream_t s, int nylo, int nyhi) {
    if (nyhi < 0) nyhi = g.nn[1] - 1;
    if (nyhi < nylo) return;
    const long maxClass = (long)((g.nn[0] + 1) / 2) * ((g.nn[1] + 1) / 2) * ((g.nn[2] + 1) / 2);
    hipLaunchKernelGGL(k_assemble_blocks<OutT>, dim3(nblocks(maxClass), 8), dim3(TPB), 0, s, g, Ke, blocks, nylo,
                       nyhi);
    VF_HIP(hipGetLastError());
}

template <typename OutT>
void launch_assemble_l1_blocks(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, OutT *blocks,
                               bool reflect, hipStream_t s, int nylo, int nyhi) {
    if (nyhi < 0) nyhi = gc.nn[1] - 1;
    if (nyhi < nylo) return;
    const long maxClass = (long)((gc.nn[0] + 1) / 2) * ((gc.nn[1] + 1) / 2) * ((gc.nn[2] + 1) / 2);
    if (reflect)
        hipLaunchKernelGGL((k_assemble_l1_blocks<OutT, true>), dim3(nblocks(maxClass), 8), dim3(TPB), 0, s, gc, gf,
                           Chat, Ef, blocks, nylo, nyhi);
    else
        hipLaunchKernelGGL((k_assemble_l1_blocks<OutT, false>), dim3(nblocks(maxClass), 8), dim3(TPB), 0, s, gc, gf,
                           Chat, Ef, blocks, nylo, nyhi);
    VF_HIP(hipGetLastError());
}

template void launch_assemble_blocks<double>(const Grid &, const double *, double *, hipStream_t, int, int);
template void launch_assemble_blocks<float>(const Grid &, const double *, float *, hipStream_t, int, int);
template void launch_assemble_l1_blocks<double>(const Grid &, const Grid &, const double *, const double *, double *,
                                                bool, hipStream_t, int, int);
template void launch_assemble_l1_blocks<float>(const Grid &, const Grid &, const double *, const double *, float *,
                                               bool, hipStream_t, int, int);

}  // namespace vf
