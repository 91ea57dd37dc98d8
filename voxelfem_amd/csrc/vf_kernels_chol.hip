// vf_kernels_chol.hip -- direct solve of the coarsest multigrid level on the device (FP64).
//
// The reference factors the reduced coarsest operator with CHOLMOD (TensorProductSimulator.hh:1198-1230)
// after every stiffness update. Here the reduced operator (fixed and detached DoFs removed, natural node
// order, so its half-bandwidth is about 3 (nny nnz + nnz + 1)) is stored as 64 x 64 blocks of its lower
// band and factored by a blocked right-looking band Cholesky A = L L^T:
//   for each block column k: L_kk = chol(A_kk) (one workgroup, LDS), L_ik = A_ik L_kk^-T for the kb blocks
//   below, A_ij -= L_ik L_jk^T for the trailing triangle of the band.
// The work is O(n bw^2) instead of the O(n^3) of a dense factorisation. For the solves, the explicit dense
// lower-triangular X = L^-1 is formed (one workgroup per block column, sequential down the column), so a
// coarse solve x = X^T (X b) is two bandwidth-bound triangular matrix-vector products instead of a chain of
// small sequential triangular solves.
#include "vf_internal.hpp"

namespace vf {
namespace {

constexpr int NB = 64;        // block size
constexpr int BB = NB * NB;   // doubles per block
constexpr int TPB = 256;

// block (I, I-q) of the band, row-major inside the block
__device__ __forceinline__ double *bandblk(double *B, int kb, int I, int q) { return B + ((size_t)I * (kb + 1) + q) * BB; }
__device__ __forceinline__ const double *bandblk(const double *B, int kb, int I, int q) {
    return B + ((size_t)I * (kb + 1) + q) * BB;
}

// ---- assembly of the lower band from the cached coarsest element matrices (one work-item per
//      (row DoF, neighbour slot)); rows/cols of fixed or detached DoFs are absent (red[] < 0).
// Batched forms (blockIdx.z = system): every array argument advances by its z stride per system.
__global__ void __launch_bounds__(TPB) k_band_assemble(Grid g, const double *__restrict__ Ke, const int *__restrict__ red,
                                                       int kb, double *__restrict__ B, size_t zKe, size_t zRed,
                                                       size_t zB) {
    Ke += blockIdx.z * zKe;
    red += blockIdx.z * zRed;
    B += blockIdx.z * zB;
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
    double Bv[3] = {0, 0, 0};
    for (int ex = lo[0]; ex <= hi[0]; ++ex)
        for (int ey = lo[1]; ey <= hi[1]; ++ey)
            for (int ez = lo[2]; ez <= hi[2]; ++ez) {
                const int ln = 4 * (ix - ex) + 2 * (iy - ey) + (iz - ez);
                const int lm = 4 * (ix + off[0] - ex) + 2 * (iy + off[1] - ey) + (iz + off[2] - ez);
                const double *K = Ke + (size_t)((ex * g.ne[1] + ey) * g.ne[2] + ez) * 576;
                for (int c = 0; c < 3; ++c) Bv[c] += K[(3 * ln + a) * 24 + 3 * lm + c];
            }
    for (int c = 0; c < 3; ++c) {
        const int cc = red[3 * m + c];
        if (cc < 0 || cc > rr) continue;  // lower triangle only
        const int I = rr / NB, J = cc / NB;
        bandblk(B, kb, I, I - J)[(rr % NB) * NB + cc % NB] = Bv[c];
    }
}

// identity on the padding rows of the last diagonal block
__global__ void k_band_pad(int n, int nb, int kb, double *B) {
    const int i = n + threadIdx.x;
    if (i < nb * NB) bandblk(B, kb, nb - 1, 0)[(i % NB) * NB + i % NB] = 1.0;
}
// batched: system z holds n[z] unknowns in nb blocks; identity on every row from n[z] on (grid: nb x systems)
__global__ void k_band_pad_b(const int *__restrict__ n, int nb, int kb, double *B, size_t zB) {
    B += blockIdx.y * zB;
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= n[blockIdx.y] && i < nb * NB) bandblk(B, kb, blockIdx.x, 0)[(i % NB) * NB + i % NB] = 1.0;
}

// ---- L_kk = chol(A_kk) and Linv_kk = L_kk^-1 by one workgroup of 256 work-items, four per row.
// Work-item t holds row r = t % 64 at the columns [16 q, 16 q + 16), q = t / 64 (one quarter of the
// columns per wavefront), in registers. The factorisation runs in the unscaled (LDL^T) form
//   A[r][c] -= A[r][j] A[c][j] / A[j][j]   for c > j, r >= c,   j = 0 .. 63,
// with one barrier per step: the wavefront owning column j (q = j / 16, known at compile time in the
// unrolled loop) publishes the column in LDS (double-buffered, so the next step's publication cannot
// overwrite entries still being read) and every work-item then updates its own entries. The pivots
// d_j = A[j][j] give L[r][c] = A[r][c] / sqrt(d_c). The inverse is X = L^-1 = D^-1/2 Y with Y the inverse of
// the unit lower factor L1[r][m] = A[r][m] / d_m, formed right-looking from Y = I:
//   Y[r][c] -= L1[r][m] Y[m][c]   for r > m, c <= m,   m = 0 .. 63,
// again one barrier per step (row m of Y and column m of L1 published in LDS).
// flag != 0 on a non-positive pivot.
constexpr int PTPB = 4 * NB;
constexpr int PQ = NB / 4;  // columns per work-item
// returns false (uniformly) on a non-positive pivot; the block is then left partly factored
__device__ __forceinline__ bool potrf_block(int k, int kb, double *__restrict__ B, double *__restrict__ Linv) {
    __shared__ double sCol[2][NB];
    __shared__ double sRow[2][NB];
    __shared__ double sPiv[NB], sInvPiv[NB], sSqrt[NB], sInvSqrt[NB];
    // q is uniform over the wavefront: the column-range tests below are scalar branches
    const int t = threadIdx.x, r = t % NB, q = __builtin_amdgcn_readfirstlane(t / NB);
    const int c0 = PQ * q;
    double *Dk = bandblk(B, kb, k, 0);
    double a[PQ];
#pragma unroll
    for (int i = 0; i < PQ; ++i) a[i] = Dk[r * NB + c0 + i];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int b = j & 1;
        if (q == j / PQ) {
            if (r >= j) sCol[b][r] = a[j % PQ];
            if (r == j) sPiv[j] = a[j % PQ];
        }
        __syncthreads();
        const double piv = sCol[b][j];
        ok = ok && piv > 0;
        // rows at or above j take f = 0 (a select: their sCol entries are stale); entries right of the
        // diagonal (c > r) are updated too but never used
        const double f = r > j ? sCol[b][r] * (1.0 / (piv > 0 ? piv : 1.0)) : 0.0;
        if (q > j / PQ) {
#pragma unroll
            for (int i = 0; i < PQ; ++i) a[i] -= f * sCol[b][c0 + i];
        } else if (q == j / PQ) {
#pragma unroll
            for (int i = j % PQ + 1; i < PQ; ++i) a[i] -= f * sCol[b][c0 + i];
        }
    }
    if (!ok) return false;  // uniform: every work-item saw the same pivots
    if (t < NB) {
        const double d = sPiv[t], sq = sqrt(d);
        sSqrt[t] = sq;
        sInvSqrt[t] = 1.0 / sq;
        sInvPiv[t] = 1.0 / d;
    }
    __syncthreads();
    double y[PQ];
#pragma unroll
    for (int i = 0; i < PQ; ++i) y[i] = (c0 + i == r) ? 1.0 : 0.0;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int b = m & 1;
        if (r == m) {
#pragma unroll
            for (int i = 0; i < PQ; ++i) sRow[b][c0 + i] = y[i];
        }
        if (q == m / PQ && r > m) sCol[b][r] = a[m % PQ] * sInvPiv[m];
        __syncthreads();
        const double f = r > m ? sCol[b][r] : 0.0;
        if (q < m / PQ) {
#pragma unroll
            for (int i = 0; i < PQ; ++i) y[i] -= f * sRow[b][c0 + i];
        } else if (q == m / PQ) {
#pragma unroll
            for (int i = 0; i <= m % PQ; ++i) y[i] -= f * sRow[b][c0 + i];
        }
    }
    double *Lk = Linv + (size_t)k * BB;
    const double isr = sInvSqrt[r];
#pragma unroll
    for (int i = 0; i < PQ; ++i) {
        const int c = c0 + i;
        Dk[r * NB + c] = c < r ? a[i] * sInvSqrt[c] : (c == r ? sSqrt[r] : 0.0);
        Lk[r * NB + c] = c <= r ? y[i] * isr : 0.0;
    }
    return true;
}
// nbz (batched, optional): system z uses only its first nbz[z] blocks (the rest is padding nothing reads)
__global__ void __launch_bounds__(PTPB) k_potrf_diag(int k, int kb, double *__restrict__ B, double *__restrict__ Linv,
                                                     int *__restrict__ flag, size_t zB, size_t zL,
                                                     const int *__restrict__ nbz) {
    if (nbz && k >= nbz[blockIdx.z]) return;
    B += blockIdx.z * zB;
    Linv += blockIdx.z * zL;
    if (!potrf_block(k, kb, B, Linv) && threadIdx.x == 0) atomicOr(flag + blockIdx.z, 1);
}

// C (+)= alpha * P * Q^T for 64 x 64 row-major blocks, 256 work-items with 4 x 4 register tiles.
// sP and sQ are LDS staging (P, Q copied in); accumulates into acc[4][4] for rows 4*(t/16).., cols 4*(t%16)..
__device__ __forceinline__ void gemm_nt_64(const double (*sP)[NB + 1], const double (*sQ)[NB + 1], int t,
                                           double acc[4][4]) {
    const int r0 = 4 * (t / 16), c0 = 4 * (t % 16);
    for (int m = 0; m < NB; ++m) {
        double p[4], q[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            p[i] = sP[r0 + i][m];
            q[i] = sQ[c0 + i][m];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] += p[i] * q[j];
    }
}

// ---- L_ik = A_ik Linv_kk^T for i = k+1 .. min(k+kb, nb-1)   (grid: one workgroup per block)
__global__ void __launch_bounds__(TPB) k_trsm_panel(int k, int nb, int kb, double *__restrict__ B,
                                                    const double *__restrict__ Linv, size_t zB, size_t zL,
                                                    const int *__restrict__ nbz) {
    if (nbz) nb = min(nb, nbz[blockIdx.z]);
    B += blockIdx.z * zB;
    Linv += blockIdx.z * zL;
    __shared__ double sP[NB][NB + 1];
    __shared__ double sQ[NB][NB + 1];
    const int q = 1 + blockIdx.x, I = k + q;
    if (I >= nb) return;
    const int t = threadIdx.x;
    double *Aik = bandblk(B, kb, I, q);
    const double *Lk = Linv + (size_t)k * BB;
    for (int w = t; w < BB; w += TPB) {
        sP[w / NB][w % NB] = Aik[w];
        sQ[w / NB][w % NB] = Lk[w];
    }
    __syncthreads();
    double acc[4][4] = {};
    gemm_nt_64(sP, sQ, t, acc);
    __syncthreads();
    const int r0 = 4 * (t / 16), c0 = 4 * (t % 16);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Aik[(r0 + i) * NB + c0 + j] = acc[i][j];
}

// ---- A_IJ -= L_Ik L_Jk^T for k < J <= I <= min(k+kb, nb-1)  (grid: one workgroup per pair)
__global__ void __launch_bounds__(TPB) k_band_update(int k, int nb, int kb, double *__restrict__ B, size_t zB,
                                                     const int *__restrict__ nbz) {
    if (nbz) nb = min(nb, nbz[blockIdx.z]);
    B += blockIdx.z * zB;
    __shared__ double sP[NB][NB + 1];
    __shared__ double sQ[NB][NB + 1];
    // decode pair index -> (a, b) with 1 <= b <= a <= m
    int p = blockIdx.x, a = 1;
    while (p >= a) { p -= a; ++a; }
    const int b = p + 1;
    const int I = k + a, J = k + b;
    if (I >= nb) return;
    const int t = threadIdx.x;
    const double *Lik = bandblk(B, kb, I, a), *Ljk = bandblk(B, kb, J, b);
    double *Aij = bandblk(B, kb, I, a - b);
    for (int w = t; w < BB; w += TPB) {
        sP[w / NB][w % NB] = Lik[w];
        sQ[w / NB][w % NB] = Ljk[w];
    }
    __syncthreads();
    double acc[4][4] = {};
    gemm_nt_64(sP, sQ, t, acc);
    const int r0 = 4 * (t / 16), c0 = 4 * (t % 16);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) Aij[(r0 + i) * NB + c0 + j] -= acc[i][j];
}

// ---- X = L^-1 (dense lower triangular, row-major, ld = nb*64), one workgroup per 16-column strip of a
//      block column J: X_JJ = Linv_JJ; X_IJ = -Linv_II * sum_{q=1..min(kb,I-J)} L_{I,I-q} X_{I-q,J}.
// Columns of X are independent, so the strips of one block column run in parallel and each workgroup's
// sequential chain down the column is four times shorter in work than a whole block column's. Both products
// are FP64 MFMA (v_mfma_f64_16x16x4f64): wave w owns rows 16w..16w+15 of the 64 x 16 strip; operand lane
// layout A[i][k] at lane i + 16k, B[k][j] at lane j + 16k, result D[4 v + l / 16][l % 16] in element v of
// lane l (the FP64 form interleaves the four result rows of a lane, unlike the FP32 16x16 forms). The operands come straight from global memory (L2-resident band blocks and earlier strips of X);
// T = sum_q L X goes through LDS once so every wave sees all 64 of its rows as the K dimension of Linv T.
typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int SW = 16;
__global__ void __launch_bounds__(TPB) k_band_inverse(int nb, int kb, const double *__restrict__ B,
                                                      const double *__restrict__ Linv, double *__restrict__ X,
                                                      size_t zB, size_t zL, size_t zX, const int *__restrict__ nbz) {
    const size_t ld = (size_t)nb * NB;  // leading dimension of X (the padded size)
    if (nbz) nb = min(nb, nbz[blockIdx.z]);
    B += blockIdx.z * zB;
    Linv += blockIdx.z * zL;
    X += blockIdx.z * zX;
    __shared__ double sT[NB][SW + 1];
    const int J = blockIdx.x / (NB / SW), cs = (blockIdx.x % (NB / SW)) * SW;  // strip columns J*NB+cs ..
    if (J >= nb) return;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int col0 = J * NB + cs;
    {
        const double *Lj = Linv + (size_t)J * BB;
        for (int e = t; e < NB * SW; e += TPB) {
            const int r = e / SW, c = e % SW;
            X[(size_t)(J * NB + r) * ld + col0 + c] = Lj[r * NB + cs + c];
        }
    }
    __threadfence();
    __syncthreads();
    const int rowA = 16 * w + li;
    for (int I = J + 1; I < nb; ++I) {
        dx4 acc = {0.0, 0.0, 0.0, 0.0};
        const int qmax = min(kb, I - J);
        for (int q = 1; q <= qmax; ++q) {
            const double *Lb = bandblk(B, kb, I, q) + rowA * NB + lk;
            const double *Xb = X + ((size_t)(I - q) * NB + lk) * ld + col0 + li;
            double a[16], b[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                a[s] = Lb[4 * s];
                b[s] = Xb[(size_t)(4 * s) * ld];
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) sT[16 * w + 4 * v + lk][li] = acc[v];
        __syncthreads();
        const double *Li = Linv + (size_t)I * BB + rowA * NB + lk;
        double a[16], b[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            a[s] = Li[4 * s];
            b[s] = sT[4 * s + lk][li];
        }
        dx4 out = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 16; ++s) out = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], out, 0, 0, 0);
        double *Xi = X + ((size_t)I * NB + 16 * w + lk) * ld + col0 + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) Xi[(size_t)(4 * v) * ld] = -out[v];
        // the strip is read back by the other waves in the next kb steps; sT is rewritten next step
        __threadfence();
        __syncthreads();
    }
}

// ---- block row I of X = L^-1 for every block column J <= I at once (batched systems): the same products as
//      one step of k_band_inverse, so X is identical. Row by row, each band block L_{I,I-q} is read by the
//      strips of one launch while it is hot in the caches, instead of once per strip and row as in the
//      column-walking form, whose re-reads of the band dominate when nb is large (grid: (I + 1) * 4 strips,
//      systems; a system with nbz[z] <= I is done).
__global__ void __launch_bounds__(TPB) k_band_inverse_row(int I, int nb, int kb, const double *__restrict__ B,
                                                          const double *__restrict__ Linv, double *__restrict__ X,
                                                          size_t zB, size_t zL, size_t zX,
                                                          const int *__restrict__ nbz) {
    const size_t ld = (size_t)nb * NB;
    if (nbz && I >= nbz[blockIdx.z]) return;
    B += blockIdx.z * zB;
    Linv += blockIdx.z * zL;
    X += blockIdx.z * zX;
    __shared__ double sT[NB][SW + 1];
    const int J = blockIdx.x / (NB / SW), cs = (blockIdx.x % (NB / SW)) * SW;
    const int t = threadIdx.x, w = t >> 6, l = t & 63, li = l & 15, lk = l >> 4;
    const int col0 = J * NB + cs;
    if (J == I) {
        const double *Lj = Linv + (size_t)J * BB;
        for (int e = t; e < NB * SW; e += TPB) {
            const int r = e / SW, c = e % SW;
            X[(size_t)(J * NB + r) * ld + col0 + c] = Lj[r * NB + cs + c];
        }
        return;
    }
    const int rowA = 16 * w + li;
    dx4 acc = {0.0, 0.0, 0.0, 0.0};
    const int qmax = min(kb, I - J);
    for (int q = 1; q <= qmax; ++q) {
        const double *Lb = bandblk(B, kb, I, q) + rowA * NB + lk;
        const double *Xb = X + ((size_t)(I - q) * NB + lk) * ld + col0 + li;
        double a[16], b[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            a[s] = Lb[4 * s];
            b[s] = Xb[(size_t)(4 * s) * ld];
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) sT[16 * w + 4 * v + lk][li] = acc[v];
    __syncthreads();
    const double *Li = Linv + (size_t)I * BB + rowA * NB + lk;
    double a[16], b[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        a[s] = Li[4 * s];
        b[s] = sT[4 * s + lk][li];
    }
    dx4 out = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 16; ++s) out = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], out, 0, 0, 0);
    double *Xi = X + ((size_t)I * NB + 16 * w + lk) * ld + col0 + li;
#pragma unroll
    for (int v = 0; v < 4; ++v) Xi[(size_t)(4 * v) * ld] = -out[v];
}

// ---- y = X b over the lower triangle (one wavefront per row)
__global__ void __launch_bounds__(TPB) k_trmv_lower(int n, int ld, const double *__restrict__ X,
                                                    const double *__restrict__ b, double *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    const double *Xr = X + (size_t)row * ld;
    double s = 0;
    for (int j = lane; j <= row; j += 64) s += Xr[j] * b[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[row] = s;
}

// ---- XT = X^T for the 64 x 64 tiles on and below the diagonal of X (so XT is filled on and above
//      its diagonal; the upper triangle of X is zero)
__global__ void __launch_bounds__(TPB) k_transpose_lower(int ld, const double *__restrict__ X, double *__restrict__ XT,
                                                         size_t zX, const int *__restrict__ nbz) {
    X += blockIdx.z * zX;
    XT += blockIdx.z * zX;
    const int bi = blockIdx.y, bj = blockIdx.x;  // tile row / column of X
    if (bj > bi || (nbz && bi >= nbz[blockIdx.z])) return;
    __shared__ double tl[NB][NB + 1];
    const int t = threadIdx.x;
    for (int w = t; w < NB * NB; w += TPB) {
        const int r = w / NB, c = w % NB;
        tl[r][c] = X[(size_t)(bi * NB + r) * ld + bj * NB + c];
    }
    __syncthreads();
    for (int w = t; w < NB * NB; w += TPB) {
        const int r = w / NB, c = w % NB;
        XT[(size_t)(bj * NB + r) * ld + bi * NB + c] = tl[c][r];
    }
}

// ---- x = XT y over the upper triangle of XT (one wavefront per row, coalesced row reads)
__global__ void __launch_bounds__(TPB) k_trmv_upper(int n, int ld, const double *__restrict__ XT,
                                                    const double *__restrict__ y, double *__restrict__ x) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    const double *Xr = XT + (size_t)row * ld;
    double s = 0;
    for (int j = row + lane; j < n; j += 64) s += Xr[j] * y[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[row] = s;
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_cgather(int nred, const int *__restrict__ redToFull, int nnodes,
                                                 const T *__restrict__ b, double *__restrict__ out) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= nred) return;
    const int f = redToFull[i];
    out[i] = (double)b[(f % 3) * nnodes + f / 3];
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_cscatter(int nred, const int *__restrict__ redToFull, int nnodes,
                                                  const double *__restrict__ in, T *__restrict__ x) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= nred) return;
    const int f = redToFull[i];
    x[(f % 3) * nnodes + f / 3] = (T)in[i];
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_cfill(size_t n, T *x, T v) {
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) x[i] = v;
}

// ---- batched coarse solve over independent systems (blockIdx.y = system z): system z owns n[z] reduced
//      unknowns; its DoF i is component f % 3 of node f / 3 + z * nodeStride of the level vector (N nodes),
//      f = redToFull[z * redStride + i]; its factors X, XT (leading dimension ld) start at z * ld * ld.
// skip (optional): systems with skip[z] == 2 are left out (their x stays zero)
template <typename T>
__global__ void __launch_bounds__(TPB) k_cgather_b(const int *__restrict__ n, const int *__restrict__ redToFull,
                                                   size_t redStride, int N, size_t nodeStride, int ld,
                                                   const T *__restrict__ b, double *__restrict__ out,
                                                   const int *__restrict__ skip) {
    const int z = blockIdx.y, i = blockIdx.x * TPB + threadIdx.x;
    if (i >= ld || (skip && skip[z] == 2)) return;
    double v = 0.0;
    if (i < n[z]) {
        const int f = redToFull[z * redStride + i];
        v = (double)b[(size_t)(f % 3) * N + f / 3 + z * nodeStride];
    }
    out[(size_t)z * ld + i] = v;
}
template <typename T>
__global__ void __launch_bounds__(TPB) k_cscatter_b(const int *__restrict__ n, const int *__restrict__ redToFull,
                                                    size_t redStride, int N, size_t nodeStride, int ld,
                                                    const double *__restrict__ in, T *__restrict__ x,
                                                    const int *__restrict__ skip) {
    const int z = blockIdx.y, i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n[z] || (skip && skip[z] == 2)) return;
    const int f = redToFull[z * redStride + i];
    x[(size_t)(f % 3) * N + f / 3 + z * nodeStride] = (T)in[(size_t)z * ld + i];
}
// y_z = X_z b_z (lower triangle) and x_z = XT_z y_z (upper triangle), one wavefront per row
__global__ void __launch_bounds__(TPB) k_trmv_lower_b(const int *__restrict__ n, int ld, const double *__restrict__ X,
                                                      const double *__restrict__ b, double *__restrict__ y,
                                                      const int *__restrict__ skip) {
    const int z = blockIdx.y, lane = threadIdx.x & 63;
    const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    if (row >= n[z] || (skip && skip[z] == 2)) return;
    const double *Xr = X + (size_t)z * ld * ld + (size_t)row * ld;
    const double *bz = b + (size_t)z * ld;
    double s = 0;
    for (int j = lane; j <= row; j += 64) s += Xr[j] * bz[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[(size_t)z * ld + row] = s;
}
__global__ void __launch_bounds__(TPB) k_trmv_upper_b(const int *__restrict__ n, int ld, const double *__restrict__ XT,
                                                      const double *__restrict__ y, double *__restrict__ x,
                                                      const int *__restrict__ skip) {
    const int z = blockIdx.y, lane = threadIdx.x & 63;
    const int row = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
    const int nz = n[z];
    if (row >= nz || (skip && skip[z] == 2)) return;
    const double *Xr = XT + (size_t)z * ld * ld + (size_t)row * ld;
    const double *yz = y + (size_t)z * ld;
    double s = 0;
    for (int j = row + lane; j < nz; j += 64) s += Xr[j] * yz[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[(size_t)z * ld + row] = s;
}

inline unsigned nblk(long count) { return (unsigned)((count + TPB - 1) / TPB); }

}  // namespace

size_t band_storage(int n, int kb) {
    const int nb = (n + NB - 1) / NB;
    return (size_t)nb * (kb + 1) * BB;
}

void band_cholesky_inverse(const Grid &g, const double *Ke, const int *red, int n, int kb, double *B, double *Linv,
                           double *X, double *XT, int *flag, hipStream_t s) {
    // fill kernels rather than memsets: the sequence may be captured into a HIP graph
    launch_fill<double>(band_storage(n, kb), B, 0.0, s);
    hipLaunchKernelGGL(k_band_assemble, dim3(nblk((long)g.nnodes * 81)), dim3(TPB), 0, s, g, Ke, red, kb, B, 0, 0, 0);
    band_factor_inverse(n, kb, B, Linv, X, XT, flag, s);
}

void band_factor_inverse(int n, int kb, double *B, double *Linv, double *X, double *XT, int *flag, hipStream_t s) {
    const int nb = (n + NB - 1) / NB;
    const size_t ld = (size_t)nb * NB;
    launch_fill<double>(ld * ld, X, 0.0, s);
    launch_fill<int>(1, flag, 0, s);
    if ((int)ld > n) hipLaunchKernelGGL(k_band_pad, dim3(1), dim3(NB), 0, s, n, nb, kb, B);
    for (int k = 0; k < nb; ++k) {
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(PTPB), 0, s, k, kb, B, Linv, flag, 0, 0, nullptr);
        const int m = std::min(kb, nb - 1 - k);
        if (m > 0) {
            hipLaunchKernelGGL(k_trsm_panel, dim3(m), dim3(TPB), 0, s, k, nb, kb, B, Linv, 0, 0, nullptr);
            hipLaunchKernelGGL(k_band_update, dim3(m * (m + 1) / 2), dim3(TPB), 0, s, k, nb, kb, B, 0, nullptr);
        }
    }
    // small systems: one launch whose strips walk down their block columns; large ones (where re-reading the
    // band per strip and row dominates): one launch per block row
    if (nb <= 32)
        hipLaunchKernelGGL(k_band_inverse, dim3(nb * (NB / SW)), dim3(TPB), 0, s, nb, kb, B, Linv, X, 0, 0, 0, nullptr);
    else
        for (int I = 0; I < nb; ++I)
            hipLaunchKernelGGL(k_band_inverse_row, dim3((I + 1) * (NB / SW)), dim3(TPB), 0, s, I, nb, kb, B, Linv, X, 0,
                               0, 0, nullptr);
    hipLaunchKernelGGL(k_transpose_lower, dim3(nb, nb), dim3(TPB), 0, s, (int)ld, X, XT, 0, nullptr);
    VF_HIP(hipGetLastError());
}

void band_cholesky_inverse_batched(const Grid &g, const double *Ke, size_t keStride, const int *red, size_t redStride,
                                   int nsys, const int *n, const int *nbz, int nb, int kb, double *B, double *Linv,
                                   double *X, double *XT, int *flags, hipStream_t s) {
    if (nsys < 1 || nsys > 65535 || nb < 1) throw std::runtime_error("batched band Cholesky: bad system count or size");
    const size_t ld = (size_t)nb * NB, zB = band_storage(nb * NB, kb), zL = (size_t)nb * BB, zX = ld * ld;
    launch_fill<double>(zB * nsys, B, 0.0, s);
    launch_fill<double>(zX * nsys, X, 0.0, s);
    launch_fill<int>(nsys, flags, 0, s);
    hipLaunchKernelGGL(k_band_assemble, dim3(nblk((long)g.nnodes * 81), 1, nsys), dim3(TPB), 0, s, g, Ke, red, kb, B,
                       keStride, redStride, zB);
    hipLaunchKernelGGL(k_band_pad_b, dim3(nb, nsys), dim3(NB), 0, s, n, nb, kb, B, zB);
    for (int k = 0; k < nb; ++k) {
        hipLaunchKernelGGL(k_potrf_diag, dim3(1, 1, nsys), dim3(PTPB), 0, s, k, kb, B, Linv, flags, zB, zL, nbz);
        const int m = std::min(kb, nb - 1 - k);
        if (m > 0) {
            hipLaunchKernelGGL(k_trsm_panel, dim3(m, 1, nsys), dim3(TPB), 0, s, k, nb, kb, B, Linv, zB, zL, nbz);
            hipLaunchKernelGGL(k_band_update, dim3(m * (m + 1) / 2, 1, nsys), dim3(TPB), 0, s, k, nb, kb, B, zB, nbz);
        }
    }
    for (int I = 0; I < nb; ++I)
        hipLaunchKernelGGL(k_band_inverse_row, dim3((I + 1) * (NB / SW), 1, nsys), dim3(TPB), 0, s, I, nb, kb, B, Linv,
                           X, zB, zL, zX, nbz);
    hipLaunchKernelGGL(k_transpose_lower, dim3(nb, nb, nsys), dim3(TPB), 0, s, (int)ld, X, XT, zX, nbz);
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_coarse_chol_apply_batched(int nsys, const int *n, int ld, const double *X, const double *XT,
                                      const int *redToFull, size_t redStride, int nnodes, size_t nodeStride, const T *b,
                                      T *x, double *tmp_in, double *tmp_out, hipStream_t s, const int *skip) {
    hipLaunchKernelGGL(k_cfill<T>, dim3(nblk(3L * nnodes)), dim3(TPB), 0, s, (size_t)3 * nnodes, x, T(0));
    const dim3 gv(nblk(ld), nsys), gr((ld + 3) / 4, nsys);
    hipLaunchKernelGGL(k_cgather_b<T>, gv, dim3(TPB), 0, s, n, redToFull, redStride, nnodes, nodeStride, ld, b, tmp_in,
                       skip);
    hipLaunchKernelGGL(k_trmv_lower_b, gr, dim3(TPB), 0, s, n, ld, X, tmp_in, tmp_out, skip);
    hipLaunchKernelGGL(k_trmv_upper_b, gr, dim3(TPB), 0, s, n, ld, XT, tmp_out, tmp_in, skip);
    hipLaunchKernelGGL(k_cscatter_b<T>, gv, dim3(TPB), 0, s, n, redToFull, redStride, nnodes, nodeStride, ld, tmp_in, x,
                       skip);
    VF_HIP(hipGetLastError());
}
template void launch_coarse_chol_apply_batched<double>(int, const int *, int, const double *, const double *,
                                                       const int *, size_t, int, size_t, const double *, double *,
                                                       double *, double *, hipStream_t, const int *);
template void launch_coarse_chol_apply_batched<float>(int, const int *, int, const double *, const double *,
                                                      const int *, size_t, int, size_t, const float *, float *,
                                                      double *, double *, hipStream_t, const int *);

template <typename T>
void launch_coarse_chol_apply(int nred, const double *X, const double *XT, const int *redToFull, int nnodes,
                              const T *b, T *x, double *tmp_in, double *tmp_out, hipStream_t s, int ncomp) {
    const int ld = (nred + NB - 1) / NB * NB;
    hipLaunchKernelGGL(k_cfill<T>, dim3(nblk((long)ncomp * nnodes)), dim3(TPB), 0, s, (size_t)ncomp * nnodes, x, T(0));
    if (nred == 0) return;
    hipLaunchKernelGGL(k_cgather<T>, dim3(nblk(nred)), dim3(TPB), 0, s, nred, redToFull, nnodes, b, tmp_in);
    hipLaunchKernelGGL(k_trmv_lower, dim3((nred + 3) / 4), dim3(TPB), 0, s, nred, ld, X, tmp_in, tmp_out);
    hipLaunchKernelGGL(k_trmv_upper, dim3((nred + 3) / 4), dim3(TPB), 0, s, nred, ld, XT, tmp_out, tmp_in);
    hipLaunchKernelGGL(k_cscatter<T>, dim3(nblk(nred)), dim3(TPB), 0, s, nred, redToFull, nnodes, tmp_in, x);
    VF_HIP(hipGetLastError());
}

template void launch_coarse_chol_apply<double>(int, const double *, const double *, const int *, int, const double *,
                                               double *, double *, double *, hipStream_t, int);
template void launch_coarse_chol_apply<float>(int, const double *, const double *, const int *, int, const float *,
                                              float *, double *, double *, hipStream_t, int);

}  // namespace vf
