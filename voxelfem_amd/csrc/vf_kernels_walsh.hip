// vf_kernels_walsh.hip -- element-centric fine K u in the Walsh (Hadamard) basis of the Q1 element.
//
// With W[k][m] = prod_a s_a(m)^{k_a} (s_a(m) = -1 or +1 for the low or high side of axis a; k, m in
// {0,1}^3, index 4x+2y+z), the element stiffness of an axis-aligned box with an orthotropic (in
// particular isotropic) elasticity tensor is K0 = (1/64) W^T Khat W where Khat = W K0 W^T has only 48
// nonzero entries: Khat[(k,d),(k,d)] and Khat[(k,d),(k ^ e_d ^ e_c, c)] for c != d with k_d != k_c
// (the monomials of the element's displacement basis are orthogonal under 2-point Gauss integration).
// One element product y_e = E_e K0 u_e therefore costs two 3-stage butterflies per component (72 + 72
// additions), 48 products with the Khat entries and 24 scalings by E_e, instead of the 576
// multiply-adds of the dense 24 x 24 product (TPSStencils.hh:453-471). The host verifies the sparsity
// on the actual K0 and selects this kernel only where it holds.
//
// Work layout: a workgroup owns a (y,z) tile of 16 x 16 element columns and marches along x. The node
// plane ahead is staged in LDS; each work-item forms its element's 24 outputs, keeps the four of the
// face ahead in registers (they belong to the next node plane), and publishes the four of the current
// plane to LDS, where the owner of each node adds the four (y,z)-neighbouring elements' values in a
// fixed order. The tile's outer element row/column only serves its neighbours: a tile owns 15 x 15
// node columns. Every node value is summed in the same order whatever the tiling, so results are
// deterministic.
#include "vf_internal.hpp"

#include <algorithm>

namespace vf {
namespace {

template <typename T>
using CT = __attribute__((address_space(4))) T;

constexpr int TE = 16;       // element columns per tile side
constexpr int TN = TE - 1;   // node columns owned per tile side
constexpr int RN = TE + 1;   // node columns staged per tile side
constexpr int NTW = TE * TE; // work-items

__device__ __forceinline__ constexpr int bitof(int a) { return 1 << (2 - a); }

// Rows of Khat coupled by its off-diagonal entries: (k, d) with (k ^ e_d ^ e_c, c) for c != d, k_d != k_c. They
// form 14 groups of one to three rows (listed as k * 3 + d), closed under the coupling, so the product can be
// formed one group at a time in place.
struct WalshGroup {
    int n;
    int r[3];
};
__device__ constexpr WalshGroup kWalshGroups[14] = {
    {1, {0, 0, 0}},   {1, {1, 0, 0}},   {1, {2, 0, 0}},    {2, {3, 14, 0}},  {2, {4, 8, 0}},
    {3, {5, 7, 12}},  {2, {6, 13, 0}},  {3, {9, 16, 20}},  {2, {10, 15, 0}}, {2, {11, 18, 0}},
    {2, {17, 19, 0}}, {1, {21, 0, 0}},  {1, {22, 0, 0}},   {1, {23, 0, 0}}};

// y = W^T diag/offdiag(Khat/64) W u for one element, u and y as [8 nodes][3 components], in place
template <typename T>
__device__ __forceinline__ void walsh_apply(const T *__restrict__ C, T Ee, T v[8][3]) {
    // forward butterflies: v[m][d] -> z[k][d]
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int bb = bitof(a);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (m & bb) continue;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T lo = v[m][d], hi = v[m | bb][d];
                v[m][d] = lo + hi;
                v[m | bb][d] = hi - lo;
            }
        }
    }
    // zh[k][d] = E (Khat[(k,d),(k,d)] z[k][d] + sum_{c != d, k_d != k_c} Khat[(k,d),(k^e_d^e_c,c)] z[..][c]) / 64,
    // group by group: a group's rows only read the group's own entries, so they are overwritten after the group
    // (the temporaries are one group, not a second 24-entry array)
#pragma unroll
    for (int gi = 0; gi < 14; ++gi) {
        const WalshGroup grp = kWalshGroups[gi];
        T out[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q >= grp.n) continue;
            const int k = grp.r[q] / 3, d = grp.r[q] % 3;
            // coefficients of row group k (nine values, loaded through an opaque pointer so they are not
            // hoisted out of the element loop and spilled)
            const T *Cg = C + 16 * k;
            asm volatile("" : "+s"(Cg));
            const CT<T> *Ck = (const CT<T> *)Cg;
            T acc = Ck[3 * d + d] * v[k][d];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (c == d) continue;
                const int kd = (k & bitof(d)) ? 1 : 0, kc = (k & bitof(c)) ? 1 : 0;
                if (kd == kc) continue;
                acc += Ck[3 * d + c] * v[k ^ bitof(d) ^ bitof(c)][c];
            }
            out[q] = Ee * acc;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q >= grp.n) continue;
            v[grp.r[q] / 3][grp.r[q] % 3] = out[q];
        }
    }
    // backward butterflies: y[m][d] = sum_k W[k][m] zh[k][d]
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int bb = bitof(a);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k & bb) continue;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T z0 = v[k][d], z1 = v[k | bb][d];
                v[k][d] = z0 - z1;
                v[k | bb][d] = z0 + z1;
            }
        }
    }
}

// dotp (SET mode, optional): per-workgroup partial sums of <u, y> over the owned nodes, in a fixed order.
template <typename T, int MODE>
__global__ void __launch_bounds__(NTW, sizeof(T) == 8 ? 3 : 5) k_ku_walsh(Grid g, const T *__restrict__ C, const T *__restrict__ E,
                                                  const T *__restrict__ u, T *__restrict__ y, const T *__restrict__ b,
                                                  const uint8_t *__restrict__ dmask, int ntZ, int xchunk,
                                                  double *__restrict__ dotp) {
    __shared__ T su[2][3 * RN * RN];  // node planes, double-buffered
    __shared__ T sp[4][3 * NTW];      // per element column: its values at the four corners of the plane
    __shared__ double sred[NTW / 64];
    const int N = g.nnodes, nx = g.nn[0], ny = g.nn[1], nz = g.nn[2];
    const int tiles = gridDim.x;
    const int logical = xcd_remap(blockIdx.x + blockIdx.y * tiles, tiles * gridDim.y);
    const int tile = logical % tiles, chunk = logical / tiles;
    const int tz = tile % ntZ, ty = tile / ntZ;
    const int sy = TN * ty - 1, sz = TN * tz - 1;  // first element (and staged node) of the tile
    const int x0 = chunk * xchunk, x1 = min(x0 + xchunk, nx);
    const int t = threadIdx.x, ly = t / TE, lz = t - ly * TE;
    const int ey = sy + ly, ez = sz + lz;
    const bool elemOk = (unsigned)ey < (unsigned)g.ne[1] && (unsigned)ez < (unsigned)g.ne[2];
    const int Y = ey + 1, Z = ez + 1;  // node owned by this work-item
    const bool owner = ly < TN && lz < TN && Y < g.ydet && Z < nz;

    // Node planes are fetched one step ahead into registers (PF values per work-item) and written to
    // LDS at the start of the next step, so the global-load latency overlaps a whole step of compute.
    constexpr int PLANE = 3 * RN * RN;
    constexpr int PF = (PLANE + NTW - 1) / NTW;
    T pf[PF];
    auto fetch_plane = [&](int ix) {
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int w = t + q * NTW;
            T v = T(0);
            if (w < PLANE) {
                const int c = w / (RN * RN), r = w - c * (RN * RN), ry = r / RN, rz = r - ry * RN;
                const int gy = sy + ry, gz = sz + rz;
                if ((unsigned)ix < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
                    v = u[(size_t)c * N + ((size_t)ix * ny + gy) * nz + gz];
            }
            pf[q] = v;
        }
    };
    auto store_plane = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PF; ++q) {
            const int w = t + q * NTW;
            if (w < PLANE) su[buf][w] = pf[q];
        }
    };
    auto fetch_E = [&](int ex) {
        T v = T(0);
        if (elemOk && (unsigned)ex < (unsigned)g.ne[0]) v = E[((size_t)ex * g.ne[1] + ey) * g.ne[2] + ez];
        return v;
    };
    auto read_face = [&](int buf, T f[4][3]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = q >> 1, k = q & 1;
#pragma unroll
            for (int c = 0; c < 3; ++c) f[q][c] = su[buf][c * RN * RN + (ly + j) * RN + lz + k];
        }
    };

    T carry[4][3];    // contributions of the previous element to the current node plane
    double dacc = 0;  // <u, y> over the owned nodes (SET mode with dotp)
#pragma unroll
    for (int q = 0; q < 4; ++q) carry[q][0] = carry[q][1] = carry[q][2] = T(0);
    fetch_plane(x0 - 1);
    store_plane((x0 + 1) & 1);
    fetch_plane(x0);
    T Enext = fetch_E(x0 - 1);
    for (int ex = x0 - 1; ex < x1; ++ex) {
        // element plane ex: node planes ex (face i = 0) and ex + 1 (face i = 1)
        const int buf = (ex + 1) & 1;
        store_plane(buf);     // plane ex + 1, fetched during the previous step
        fetch_plane(ex + 2);  // consumed by the next step
        const T Ee = Enext;
        Enext = fetch_E(ex + 1);
        // right-hand side and mask of the node finished in this step, requested early
        const int ix = ex;
        const size_t n = ((size_t)ix * ny + Y) * nz + Z;
        uint8_t dm = 0;
        T bv[3] = {0, 0, 0};
        if (owner && ix >= x0) {
            if (dmask) dm = dmask[n];
            if (MODE == AM_RESIDUAL) {
                bv[0] = b[n];
                bv[1] = b[(size_t)N + n];
                bv[2] = b[(size_t)2 * N + n];
            }
        }
        __syncthreads();
        // both node planes of the element are in LDS: plane ex (stored during the previous step) and plane
        // ex + 1; the slot of plane ex is next overwritten after this step's second barrier
        T v[8][3];
        read_face(buf ^ 1, v);       // local nodes 0jk (face i = 0)
        read_face(buf, v + 4);       // local nodes 1jk (face i = 1)
        // u at the owned node (Y, Z) of plane ex: corner (1,1) of the element's face i = 0
        const T un0 = v[3][0], un1 = v[3][1], un2 = v[3][2];
        walsh_apply<T>(C, Ee, v);
        if (ex >= x0) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) sp[q][c * NTW + t] = carry[q][c] + v[q][c];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) carry[q][c] = v[4 + q][c];
        __syncthreads();
        if (ex >= x0 && owner) {
            // corners (j,k) of the elements at (ly+1-j, lz+1-k) that hold node (Y, Z)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                T f = sp[3][c * NTW + ly * TE + lz];
                f += sp[2][c * NTW + ly * TE + lz + 1];
                f += sp[1][c * NTW + (ly + 1) * TE + lz];
                f += sp[0][c * NTW + (ly + 1) * TE + lz + 1];
                T out = MODE == AM_RESIDUAL ? bv[c] - f : f;
                if (dm & (1 << c)) out = T(0);
                y[(size_t)c * N + n] = out;
                if (MODE == AM_SET && dotp) dacc += (double)out * (double)(c == 0 ? un0 : (c == 1 ? un1 : un2));
            }
        }
    }
    if (MODE == AM_SET && dotp) {
        for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o, 64);
        if ((t & 63) == 0) sred[t >> 6] = dacc;
        __syncthreads();
        if (t == 0) {
            double r = 0;
            for (int i = 0; i < NTW / 64; ++i) r += sred[i];
            dotp[logical] = r;  // chunk-major: the partials of one x-chunk are contiguous
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_zero_rows_w(Grid g, T *__restrict__ y) {
    const int nz = g.nn[2], rows = g.nn[1] - g.ydet, slab = rows * nz;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= g.nn[0] * slab) return;
    const int ix = t / slab, r = t - ix * slab, iy = g.ydet + r / nz, iz = r % nz;
    const int n = (ix * g.nn[1] + iy) * nz + iz;
    y[n] = T(0);
    y[g.nnodes + n] = T(0);
    y[(size_t)2 * g.nnodes + n] = T(0);
}

int cu_count_w() {
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

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

static int walsh_xchunk_of(const Grid &g, const void *kern, int alignX) {
    const int tiles = cdiv(g.ydet, TN) * cdiv(g.nn[2], TN);
    const int chunks0 = plane_chunks(tiles, g.nn[0], 4 * cu_count_w(), resident_workgroups(kern, NTW));
    int xchunk = cdiv(g.nn[0], chunks0);
    if (alignX > 0) {  // the largest divisor of alignX not above xchunk: no x-chunk straddles a multiple of alignX
        int best = 1;
        for (int d = 1; d <= alignX && d <= xchunk; ++d)
            if (alignX % d == 0) best = d;
        xchunk = best;
    }
    return xchunk;
}
int walsh_xchunk(const Grid &g, int mode, int alignX) {
    const void *kern = mode == AM_RESIDUAL ? (const void *)k_ku_walsh<double, AM_RESIDUAL> : (const void *)k_ku_walsh<double, AM_SET>;
    return walsh_xchunk_of(g, kern, alignX);
}

template <typename T>
void launch_fine_apply_walsh(const Grid &g, const T *C, const T *E, const T *u, T *y, const T *b,
                             const uint8_t *dmask, int mode, hipStream_t s, double *dotPartials, double *dotOut,
                             int alignX) {
    if (mode != AM_SET && mode != AM_RESIDUAL) throw std::runtime_error("Walsh K u supports the SET and RESIDUAL modes");
    if (dotOut && (mode != AM_SET || !dotPartials)) throw std::runtime_error("the fused <u, K u> needs SET mode and partials");
    if (dotOut && g.ydet <= 0) launch_fill<double>(1, dotOut, 0.0, s);
    if (g.ydet > 0) {
        // tiles cover node columns [TN*t, TN*t + TN); rows beyond ydet are skipped by the owners
        const int ntY = cdiv(g.ydet, TN), ntZ = cdiv(g.nn[2], TN);
        const int tiles = ntY * ntZ;
        const void *kern = mode == AM_RESIDUAL ? (const void *)k_ku_walsh<T, AM_RESIDUAL> : (const void *)k_ku_walsh<T, AM_SET>;
        const int xchunk = walsh_xchunk_of(g, kern, alignX);
        const int chunks = cdiv(g.nn[0], xchunk);
        if (chunks > 65535) throw std::runtime_error("grid too large for the Walsh K u");
        const dim3 grid(tiles, chunks);
        if (mode == AM_RESIDUAL)
            hipLaunchKernelGGL((k_ku_walsh<T, AM_RESIDUAL>), grid, dim3(NTW), 0, s, g, C, E, u, y, b, dmask, ntZ, xchunk,
                               nullptr);
        else
            hipLaunchKernelGGL((k_ku_walsh<T, AM_SET>), grid, dim3(NTW), 0, s, g, C, E, u, y, b, dmask, ntZ, xchunk,
                               dotOut ? dotPartials : nullptr);
        if (dotOut) launch_final_sum(dotPartials, (int)(grid.x * grid.y), dotOut, s);
    }
    if (mode == AM_SET && g.ydet < g.nn[1]) {
        const long cz = (long)g.nn[0] * (g.nn[1] - g.ydet) * g.nn[2];
        hipLaunchKernelGGL(k_zero_rows_w<T>, dim3((unsigned)((cz + 255) / 256)), dim3(256), 0, s, g, y);
    }
    VF_HIP(hipGetLastError());
}

template void launch_fine_apply_walsh<double>(const Grid &, const double *, const double *, const double *, double *,
                                              const double *, const uint8_t *, int, hipStream_t, double *, double *,
                                              int);
template void launch_fine_apply_walsh<float>(const Grid &, const float *, const float *, const float *, float *,
                                             const float *, const uint8_t *, int, hipStream_t, double *, double *, int);

}  // namespace vf
