// vf_kernels_l1mfma.hip -- the level-1 operator on the fly from the fine moduli on the matrix cores (gfx950).
//
// Level 1 is the Galerkin coarsening of the fine operator whose coarse element matrix is
// sum_fi E_fi Chat_fi and is never assembled (MultigridSolver.hh:483-491). For element position e of the
// eight around a node (the node is local node e of that element) the 24 products
//   t_e[(fi, a)] = sum_{(m, c)} Chat_fi[(3 e + a), (3 m + c)] u_{e, m, c}
// are a (24 x 24) x (24 x nodes) matrix product whose left factor is a constant. A wavefront computes it for
// a tile of 16 nodes with 2 x 6 MFMA 16x16x4 (8 of the 32 result rows are zero); the fragments of the constant
// factor are staged in LDS in lane order, the node values are the B operand. The (fi, a) row of each result
// element is fixed per lane, so the moduli weighting f[a] = sum_e sum_fi E_{e, fi} t_e[(fi, a)] is a per-lane
// multiply-add, followed by a sum over the four lane groups that hold the same node. The diagonal
// block the smoother needs, M = sum_e sum_fi E_{e, fi} Chat_fi[(3 e + .), (3 e + .)], is split over the lane
// groups by child and summed the same way. FP32 runs the FP32 MFMA (twice the FP64 rate), FP64 the FP64 one.
//
// MFMA 16x16x4 operand layout: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k; result D[i][j] in element v
// of lane j + 16 g with i = 4 g + v (FP32) or i = 4 v + g (FP64).
#include "vf_internal.hpp"

namespace vf {
namespace {

constexpr int TPB = 256;
constexpr int TILE = 16;            // nodes per wavefront tile
constexpr int NFRAG = 8 * 2 * 6 * 64;  // fragments: element, result tile, K step, lane

template <typename T>
struct MfmaOp;
template <>
struct MfmaOp<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // lane group and element that hold result row p (0..15) of a tile
    static __host__ __device__ constexpr int group(int p) { return p / 4; }
    static __host__ __device__ constexpr int elem(int p) { return p % 4; }
};
template <>
struct MfmaOp<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __host__ __device__ constexpr int group(int p) { return p % 4; }
    static __host__ __device__ constexpr int elem(int p) { return p / 4; }
};

__device__ __forceinline__ int eidx(const Grid &g, int ex, int ey, int ez) { return (ex * g.ne[1] + ey) * g.ne[2] + ez; }
// offset of child fi of a coarse element from its first child in the fine element array
__device__ __forceinline__ int child_off(const Grid &gf, int fi) {
    return ((fi >> 2) & 1) * gf.ne[1] * gf.ne[2] + ((fi >> 1) & 1) * gf.ne[2] + (fi & 1);
}

// Result rows are ordered so that lane group g holds, in its eight result slots q = 4 t + v (tile t,
// element v), the rows (fi, a) = (2 g + q / 3, q % 3) for q < 6 and zero rows for q = 6, 7: a lane then needs
// the moduli of its two children 2g, 2g+1 only (the same two the diagonal block uses), and the row of each
// slot is known at compile time.
// sA[((e * 2 + t) * 6 + s) * 64 + l] = A_e[16 t + l % 16][4 s + l / 16] with A_e[(fi, a)][3 m + c] =
// Chat[fi][(3 e + a) * 24 + 3 m + c]; sD[(e * 8 + fi) * 9 + 3 a + c] = the diagonal block.
template <typename T>
__device__ void stage_fragments(const T *__restrict__ Chat, T *sA, T *sD) {
    using Op = MfmaOp<T>;
    for (int idx = threadIdx.x; idx < NFRAG; idx += TPB) {
        const int l = idx & 63, s = (idx >> 6) % 6, t = (idx / 384) & 1, e = idx / 768;
        const int p = l & 15, k = 4 * s + (l >> 4);
        const int q = 4 * t + Op::elem(p);
        T v = T(0);
        if (q < 6) v = Chat[(2 * Op::group(p) + q / 3) * 576 + (3 * e + q % 3) * 24 + k];
        sA[idx] = v;
    }
    if (sD)
        for (int idx = threadIdx.x; idx < 576; idx += TPB) {
            const int q = idx % 9, fe = idx / 9, fi = fe % 8, e = fe / 8;
            sD[idx] = Chat[fi * 576 + (3 * e + q / 3) * 24 + 3 * e + q % 3];
        }
}

// Row of the node (ix, iy, iz) owned by this lane's column (lane % 16); every lane of the wavefront takes part
// (the matrix instructions need all 64), lanes of a padding column work on a valid clamped node. On return
// every lane of the column holds f (and M) of its node.
template <typename T, bool WANT_M>
__device__ __forceinline__ void l1_row_mfma(const Grid &gc, const Grid &gf, const T *sA, const T *sD,
                                            const T *__restrict__ Ef, const T *__restrict__ u, int ix, int iy,
                                            int iz, int lane, T f[3], T M[9]) {
    using Op = MfmaOp<T>;
    using v4 = typename Op::v4;
    const int g = lane >> 4;
    const int N = gc.nnodes, sx = gc.nn[1] * gc.nn[2], sy = gc.nn[2];
    const int n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
    const int off0 = child_off(gf, 2 * g), off1 = child_off(gf, 2 * g + 1);
    T F[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] = T(0);
    if (WANT_M)
#pragma unroll
        for (int q = 0; q < 9; ++q) M[q] = T(0);
    // operands of element e: the six node values of this lane's K slice and the moduli of children 2g, 2g+1;
    // an absent element reads as zero. The node offsets split into a per-lane part (K slice) and a
    // per-element part (uniform), so a load costs one address add.
    int koff[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int kk = 4 * s + g, m = kk / 3, c = kk - 3 * m;
        koff[s] = c * N + ((m >> 2) & 1) * sx + ((m >> 1) & 1) * sy + (m & 1);
    }
    const int fb0 = eidx(gf, 2 * ix, 2 * iy, 2 * iz);
    const bool xl = ix >= 1, xh = ix < gc.ne[0], yl = iy >= 1, yh = iy < gc.ne[1], zl = iz >= 1, zh = iz < gc.ne[2];
    auto fetch = [&](int e, T bv[6], T ev[2]) {
        const int i = (e >> 2) & 1, j = (e >> 1) & 1, k = e & 1;
        const bool ok = (i ? xl : xh) && (j ? yl : yh) && (k ? zl : zh);
        const int nb = ok ? n - (i * sx + j * sy + k) : 0;
#pragma unroll
        for (int s = 0; s < 6; ++s) bv[s] = ok ? u[nb + koff[s]] : T(0);
        const int fb = ok ? fb0 - 2 * (i * gf.ne[1] * gf.ne[2] + j * gf.ne[2] + k) : 0;
        ev[0] = ok ? Ef[fb + off0] : T(0);
        ev[1] = ok ? Ef[fb + off1] : T(0);
    };
    T bv[6], ev[2];
    fetch(0, bv, ev);
    // elements one at a time (a rolled loop, so the register file holds one element's operands), the next
    // element's loads in flight while this one's matrix products run
#pragma unroll 1
    for (int e = 0; e < 8; ++e) {
        T bn[6], en[2];
        if (e < 7) fetch(e + 1, bn, en);
        v4 acc0 = {T(0), T(0), T(0), T(0)}, acc1 = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            acc0 = Op::mac(sA[((e * 2 + 0) * 6 + s) * 64 + lane], bv[s], acc0);
            acc1 = Op::mac(sA[((e * 2 + 1) * 6 + s) * 64 + lane], bv[s], acc1);
        }
        // slots q = 0..3 (tile 0), 4..5 (tile 1): child 2g for q < 3, 2g+1 for 3 <= q < 6
        F[0] += ev[0] * acc0[0];
        F[1] += ev[0] * acc0[1];
        F[2] += ev[0] * acc0[2];
        F[3] += ev[1] * acc0[3];
        F[4] += ev[1] * acc1[0];
        F[5] += ev[1] * acc1[1];
        if (WANT_M) {
            const T *d0 = sD + (e * 8 + 2 * g) * 9;
#pragma unroll
            for (int q = 0; q < 9; ++q) M[q] += ev[0] * d0[q] + ev[1] * d0[9 + q];
        }
        if (e < 7) {
#pragma unroll
            for (int q = 0; q < 6; ++q) bv[q] = bn[q];
            ev[0] = en[0];
            ev[1] = en[1];
        }
    }
    f[0] = F[0] + F[3];
    f[1] = F[1] + F[4];
    f[2] = F[2] + F[5];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        f[a] += __shfl_xor(f[a], 16, 64);
        f[a] += __shfl_xor(f[a], 32, 64);
    }
    if (WANT_M)
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            M[q] += __shfl_xor(M[q], 16, 64);
            M[q] += __shfl_xor(M[q], 32, 64);
        }
}

template <typename T>
__device__ __forceinline__ void store_row(T *y, const T *b, int N, int n, const T f[3], int mode, uint8_t dm) {
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

// y (mode) = level-1 K u over the nodes below the first detached row, tiles of 16 consecutive nodes
template <typename T, int MODE>
__global__ void __launch_bounds__(TPB) k_l1_apply_mfma(Grid gc, Grid gf, const T *__restrict__ Chat,
                                                       const T *__restrict__ Ef, const T *__restrict__ u,
                                                       T *__restrict__ y, const T *__restrict__ b,
                                                       const uint8_t *__restrict__ dmask) {
    __shared__ T sA[NFRAG];
    stage_fragments<T>(Chat, sA, nullptr);
    __syncthreads();
    const int nz = gc.nn[2], slab = gc.ydet * nz, count = gc.nn[0] * slab;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntiles = (count + TILE - 1) / TILE;
    for (int tile = blockIdx.x * (TPB / 64) + w; tile < ntiles; tile += gridDim.x * (TPB / 64)) {
        const int t0 = tile * TILE + (lane & 15);
        const int t = min(t0, count - 1);
        const int ix = t / slab, r = t - ix * slab, iy = r / nz, iz = r - iy * nz;
        T f[3], M[9];
        l1_row_mfma<T, false>(gc, gf, sA, nullptr, Ef, u, ix, iy, iz, lane, f, M);
        if (lane < 16 && t0 < count) {
            const int n = (ix * gc.nn[1] + iy) * nz + iz;
            store_row(y, b, gc.nnodes, n, f, MODE, dmask ? dmask[n] : uint8_t(0));
        }
    }
}

// one colour of the level-1 Gauss-Seidel sweep (MultigridSolver.hh:347-442), tiles of 16 nodes of the class
template <typename T>
__global__ void __launch_bounds__(TPB) k_l1_gs_mfma(Grid gc, Grid gf, const T *__restrict__ Chat,
                                                    const T *__restrict__ Ef, T *__restrict__ u,
                                                    const T *__restrict__ b, const uint8_t *__restrict__ dmask, int cx,
                                                    int cy, int cz, int forward) {
    __shared__ T sA[NFRAG];
    __shared__ T sD[576];
    stage_fragments<T>(Chat, sA, sD);
    __syncthreads();
    const int mx = (gc.nn[0] - cx + 1) / 2, my = (gc.ydet - cy + 1) / 2, mz = (gc.nn[2] - cz + 1) / 2;
    const int cnt = mx * my * mz;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntiles = (cnt + TILE - 1) / TILE;
    const int N = gc.nnodes;
    for (int tile = blockIdx.x * (TPB / 64) + w; tile < ntiles; tile += gridDim.x * (TPB / 64)) {
        const int j0 = tile * TILE + (lane & 15);
        const int jc = min(j0, cnt - 1);
        const int jx = jc / (my * mz), r = jc - jx * my * mz, jy = r / mz, jz = r - jy * mz;
        const int ix = 2 * jx + cx, iy = 2 * jy + cy, iz = 2 * jz + cz;
        T f[3], M[9];
        l1_row_mfma<T, true>(gc, gf, sA, sD, Ef, u, ix, iy, iz, lane, f, M);
        if (lane < 16 && j0 < cnt) {
            const int n = (ix * gc.nn[1] + iy) * gc.nn[2] + iz;
            const uint8_t dm = dmask ? dmask[n] : uint8_t(0);
            if (dm != 7) {
                const T bms[3] = {b[n] - f[0], b[N + n] - f[1], b[2 * N + n] - f[2]};
                T ud[3];
                if (dm) {
                    ud[0] = ud[1] = ud[2] = T(0);
                    for (int q = 0; q < 3; ++q) {
                        const int i = forward ? q : 2 - q;
                        const T fr = (dm & (1 << i)) ? T(0) : T(1);
                        ud[i] = (bms[i] - (M[3 * i] * ud[0] + M[3 * i + 1] * ud[1] + M[3 * i + 2] * ud[2])) *
                                (fr / M[4 * i]);
                    }
                } else {
                    const T c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8],
                            c02 = M[3] * M[7] - M[4] * M[6];
                    const T id = T(1) / (M[0] * c00 + M[1] * c01 + M[2] * c02);
                    ud[0] = (c00 * bms[0] + (M[2] * M[7] - M[1] * M[8]) * bms[1] + (M[1] * M[5] - M[2] * M[4]) * bms[2]) * id;
                    ud[1] = (c01 * bms[0] + (M[0] * M[8] - M[2] * M[6]) * bms[1] + (M[2] * M[3] - M[0] * M[5]) * bms[2]) * id;
                    ud[2] = (c02 * bms[0] + (M[1] * M[6] - M[0] * M[7]) * bms[1] + (M[0] * M[4] - M[1] * M[3]) * bms[2]) * id;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) u[c * N + n] += ud[c];
            }
        }
    }
}

inline unsigned tile_grid(long count, int cus) {
    const long tiles = (count + TILE - 1) / TILE;
    const long wgs = (tiles + TPB / 64 - 1) / (TPB / 64);
    return (unsigned)std::max(1L, std::min(wgs, 8L * cus));
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        VF_HIP(hipGetDevice(&dev));
        VF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return cus;
}

}  // namespace

template <typename T>
void launch_l1_apply_mfma(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, const T *u, T *y, const T *b,
                          const uint8_t *dmask, int mode, hipStream_t s) {
    const long count = (long)gc.nn[0] * gc.ydet * gc.nn[2];
    if (count > 0) {
        const dim3 grid(tile_grid(count, device_cus()));
        switch (mode) {
        case AM_SET: hipLaunchKernelGGL((k_l1_apply_mfma<T, AM_SET>), grid, dim3(TPB), 0, s, gc, gf, Chat, Ef, u, y, b, dmask); break;
        case AM_ADD: hipLaunchKernelGGL((k_l1_apply_mfma<T, AM_ADD>), grid, dim3(TPB), 0, s, gc, gf, Chat, Ef, u, y, b, dmask); break;
        case AM_SUB: hipLaunchKernelGGL((k_l1_apply_mfma<T, AM_SUB>), grid, dim3(TPB), 0, s, gc, gf, Chat, Ef, u, y, b, dmask); break;
        default: hipLaunchKernelGGL((k_l1_apply_mfma<T, AM_RESIDUAL>), grid, dim3(TPB), 0, s, gc, gf, Chat, Ef, u, y, b, dmask); break;
        }
    }
    if (mode == AM_SET && gc.ydet < gc.nn[1]) launch_zero_detached<T>(gc, y, s);  // detached rows of y are zero
    VF_HIP(hipGetLastError());
}

template <typename T>
void launch_l1_gs_mfma(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, T *u, const T *b,
                       const uint8_t *dmask, int color, int forward, hipStream_t s) {
    const int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    const long mx = (gc.nn[0] - cx + 1) / 2, my = (gc.ydet - cy + 1) / 2, mz = (gc.nn[2] - cz + 1) / 2;
    const long count = (mx > 0 && my > 0 && mz > 0) ? mx * my * mz : 0;
    if (!count) return;
    hipLaunchKernelGGL(k_l1_gs_mfma<T>, dim3(tile_grid(count, device_cus())), dim3(TPB), 0, s, gc, gf, Chat, Ef, u, b,
                       dmask, cx, cy, cz, forward);
    VF_HIP(hipGetLastError());
}

template void launch_l1_apply_mfma<double>(const Grid &, const Grid &, const double *, const double *, const double *,
                                           double *, const double *, const uint8_t *, int, hipStream_t);
template void launch_l1_apply_mfma<float>(const Grid &, const Grid &, const float *, const float *, const float *,
                                          float *, const float *, const uint8_t *, int, hipStream_t);
template void launch_l1_gs_mfma<double>(const Grid &, const Grid &, const double *, const double *, double *,
                                        const double *, const uint8_t *, int, int, hipStream_t);
template void launch_l1_gs_mfma<float>(const Grid &, const Grid &, const float *, const float *, float *,
                                       const float *, const uint8_t *, int, int, hipStream_t);

}  // namespace vf
