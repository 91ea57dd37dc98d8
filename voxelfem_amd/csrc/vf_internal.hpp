// vf_internal.hpp -- shared declarations of the HIP implementation (kernels + host orchestration).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace vf {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NanError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define VF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::vf::HipError(std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Geometry of one grid level, passed to kernels by value.
// Nodes with iy < ydet are active ("non-detached"); elements with ey < ymask are unmasked.
struct Grid {
    int ne[3];
    int nn[3];
    int nnodes;
    int nelems;
    int ydet;
    int ymask;
};

inline Grid make_grid(const int64_t ne[3]) {
    Grid g;
    for (int d = 0; d < 3; ++d) {
        g.ne[d] = (int)ne[d];
        g.nn[d] = (int)ne[d] + 1;
    }
    g.nnodes = g.nn[0] * g.nn[1] * g.nn[2];
    g.nelems = g.ne[0] * g.ne[1] * g.ne[2];
    g.ydet = g.nn[1];
    g.ymask = g.ne[1];
    return g;
}

// Workgroups are dispatched to the eight XCDs (each with its own L2) round-robin in the order of their
// linear id. This maps a linear id to a logical one such that consecutive logical workgroups - which
// the kernels give neighbouring tiles sharing halo data - run on the same XCD.
__device__ __forceinline__ int xcd_remap(int linear, int total) {
    const int q = total / 8, r = total % 8;
    const int xcd = linear % 8, slot = linear / 8;
    return xcd * q + (xcd < r ? xcd : r) + slot;
}

// Simple owning device buffer.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { resize(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void resize(size_t count) {
        if (count == n) return;
        release();
        if (count) VF_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    T *data() const { return p; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

// ---------------------------------------------------------------------------------------------
// Kernel launchers (vf_kernels_*.hip). Every launcher enqueues on `s` and does not synchronize.
// T is double or float. Vector fields are SoA planes (3 * g.nnodes).
// ---------------------------------------------------------------------------------------------
enum ApplyMode { AM_SET = 0, AM_ADD = 1, AM_SUB = 2, AM_RESIDUAL = 3 };

// Assembled block rows (levels 2..L-1) are stored colour-major: the nodes of colour
// c = 4(ix%2) + 2(iy%2) + (iz%2) form a class of cnt_c nodes numbered j = ((ix/2) my + iy/2) mz + iz/2
// (m_ = nodes of that parity along each axis), classes follow each other in colour order, and inside a
// class entry k of neighbour slot s sits at (s*9 + k) * cnt_c + j. A smoothing pass over one colour then
// streams contiguous memory.
struct BlockSlot {
    size_t base;  // first entry of the node's colour class
    int cnt;      // nodes in the class
    int j;        // node index inside the class
};
__host__ __device__ inline BlockSlot block_slot(const Grid &g, int ix, int iy, int iz) {
    const int cx = ix & 1, cy = iy & 1, cz = iz & 1;
    const int c = 4 * cx + 2 * cy + cz;
    size_t before = 0;
    for (int q = 0; q < c; ++q) {
        const long mx = (g.nn[0] - ((q >> 2) & 1) + 1) / 2, my = (g.nn[1] - ((q >> 1) & 1) + 1) / 2,
                   mz = (g.nn[2] - (q & 1) + 1) / 2;
        before += (size_t)(mx * my * mz);
    }
    const int my = (g.nn[1] - cy + 1) / 2, mz = (g.nn[2] - cz + 1) / 2, mx = (g.nn[0] - cx + 1) / 2;
    BlockSlot r;
    r.base = before * 243;
    r.cnt = mx * my * mz;
    r.j = ((ix >> 1) * my + (iy >> 1)) * mz + (iz >> 1);
    return r;
}

// fine level (TPSStencils.hh:431-728).  AM_SET: y = K u (detached rows zeroed);
// AM_RESIDUAL: y = b - K u on active rows; dmask != nullptr zeroes constrained output components.
template <typename T>
void launch_fine_apply(const Grid &g, const T *K0, const T *E, const T *u, T *y, const T *b, const uint8_t *dmask,
                       int mode, hipStream_t s);
// level 1 with on-the-fly Galerkin (MultigridSolver.hh:475-499)
// LDS-tiled fine K u (AM_SET or AM_RESIDUAL only); same values as launch_fine_apply.
template <typename T>
void launch_fine_apply_tiled(const Grid &g, const T *K0, const T *E, const T *u, T *y, const T *b,
                             const uint8_t *dmask, int mode, hipStream_t s);
// Workgroups of `threads` work-items of `kernel` (static LDS only) that the current device holds at once.
int resident_workgroups(const void *kernel, int threads);
// Chunk count along x for a launch of `tiles` x chunks workgroups that march over `n` planes: about
// `want` workgroups, never more than `resident` (a workgroup left waiting for a free slot would run its
// whole chunk after the others finish).
int plane_chunks(int tiles, int n, int want, int resident);
// Element-centric fine K u in the Walsh basis (AM_SET or AM_RESIDUAL); C = the 8 x 16 table of the
// nonzero Walsh-basis element stiffness entries divided by 64 (see vf_kernels_walsh.hip).
// With dotOut (SET mode only), *dotOut = <u, y> is also formed from per-workgroup partials (dotPartials
// needs one entry per workgroup, at most tiles x nn[0]); the partials are stored chunk-major (entry
// chunk * tiles + tile, tiles = ceil(ydet / 15) * ceil(nn[2] / 15), chunk c owning node planes
// [c xchunk, (c + 1) xchunk)). alignX > 0: xchunk divides alignX.
template <typename T>
void launch_fine_apply_walsh(const Grid &g, const T *C, const T *E, const T *u, T *y, const T *b,
                             const uint8_t *dmask, int mode, hipStream_t s, double *dotPartials = nullptr,
                             double *dotOut = nullptr, int alignX = 0);
// x-chunk length the Walsh K u uses on g with alignX (as above)
int walsh_xchunk(const Grid &g, int mode, int alignX);
// fixed-order sum of count partials into *out
void launch_final_sum(const double *partials, int count, double *out, hipStream_t s);
// One colour of the fine block Gauss-Seidel sweep on the structured element matrix, register gather
template <typename T>
void launch_fine_gs_gather_f(const Grid &g, const T *F, const T *E, T *u, const T *b, const uint8_t *dmask, int color,
                             int forward, hipStream_t s);
// One x-parity of a fine block Gauss-Seidel sweep: planes of that parity read from uin, neighbour
// planes from unb, result planes written to uout (uout != uin). A forward sweep is parity 0 with
// unb = uin then parity 1 with unb = uout; a backward sweep runs parity 1 first.
template <typename T>
void launch_fine_gs_phase(const Grid &g, const T *K0, const T *E, const T *uin, const T *unb, T *uout, const T *b,
                          const uint8_t *dmask, int parity, int forward, hipStream_t s, int zin = 0);
// level-1 operator on the fly from the fine moduli (reflect: the element matrix is reflection-symmetric;
// the product then runs element by element in the Walsh basis from Cw = W^T Chat_0 W / 64, rows (3 k + a),
// columns (3 k' + c), W[l][k] = prod over the axes a in k of the side of node l along a; otherwise from Chat)
template <typename T>
void launch_l1_apply(const Grid &gc, const Grid &gf, const T *Chat, const T *Cw, const T *Ef, const T *u, T *y,
                     const T *b, const uint8_t *dmask, int mode, bool reflect, hipStream_t s);
// level-1 operator on the fly with the matrix cores (full Chat, any element matrix)
template <typename T>
void launch_l1_apply_mfma(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, const T *u, T *y, const T *b,
                          const uint8_t *dmask, int mode, hipStream_t s);
template <typename T>
void launch_l1_gs_mfma(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, T *u, const T *b,
                       const uint8_t *dmask, int color, int forward, hipStream_t s);
// zero the rows of y at and above the first detached row (ydet)
template <typename T>
void launch_zero_detached(const Grid &g, T *y, hipStream_t s);
// assembled 27-point block rows (levels >= 2)
template <typename T>
void launch_block_apply(const Grid &g, const T *blocks, const T *u, T *y, const T *b, const uint8_t *dmask, int mode,
                        hipStream_t s);
// cached per-element matrices (coarsest level operator; only used by applyK(l) on that level)
template <typename T>
void launch_elem_apply(const Grid &g, const T *Ke, const T *u, T *y, int mode, hipStream_t s);

// one colour of the multicolour block Gauss-Seidel (MultigridSolver.hh:347-442)
template <typename T>
void launch_fine_gs(const Grid &g, const T *K0, const T *E, T *u, const T *b, const uint8_t *dmask, int color,
                    int forward, hipStream_t s);
template <typename T>
void launch_l1_gs(const Grid &gc, const Grid &gf, const T *Chat, const T *Ef, T *u, const T *b, const uint8_t *dmask,
                  int color, int forward, bool reflect, hipStream_t s);
template <typename T>
void launch_block_gs(const Grid &g, const T *blocks, T *u, const T *b, const uint8_t *dmask, int color, int forward,
                     hipStream_t s);

// transfers (MultigridSolver.hh:178-275)
template <typename T>
void launch_restrict(const Grid &gf, const Grid &gc, const T *rf, T *rc, hipStream_t s, T *xzero = nullptr);
template <typename T>
void launch_interpolate(const Grid &gf, const Grid &gc, const T *xc, T *xf, int accumulate, hipStream_t s);

// Dirichlet handling on explicit node lists (TensorProductSimulator.hh:575-593)
template <typename T>
void launch_zero_dirichlet(int count, const int *nodes, const uint8_t *masks, int nnodes, T *u, hipStream_t s);
void launch_enforce_dirichlet(int count, const int *nodes, const uint8_t *masks, int nnodes, const double *vals,
                              double *u, hipStream_t s);

// vector helpers over the active (non-detached) region (+margin rows) of a level
template <typename T>
void launch_fill_active(const Grid &g, int margin, T *x, T value, hipStream_t s);
template <typename T>
void launch_fill(size_t n, T *x, T value, hipStream_t s);
template <typename Tdst, typename Tsrc>
void launch_convert(size_t n, const Tsrc *src, Tdst *dst, hipStream_t s);
template <typename Tdst, typename Tsrc>
void launch_copy_active(const Grid &g, int margin, const Tsrc *src, Tdst *dst, hipStream_t s);

// deterministic reductions; result written to device scalar `out` (double)
// dot over active region: out = sum a*b (3 components)
void launch_dot_active(const Grid &g, const double *a, const double *b, double *partials, double *out, hipStream_t s);
void launch_dot2_active(const Grid &g, const double *a, const double *b, const double *c, double *partials,
                        double *out2, hipStream_t s);  // out2[0] = a.b, out2[1] = c.b
void launch_dot_full(size_t n, const double *a, const double *b, double *partials, double *out, hipStream_t s);
void launch_div_scalar(const double *in, double d, double *out, hipStream_t s);
// constant-stress element load coefficients: w[c][d] = sigma[c][d] * V / (4 h_d)
struct StressCoef {
    double w[3][3];
};
void launch_constant_strain_load(const Grid &g, const double *rho, const StressCoef &sc, double *f, hipStream_t s);
void launch_sample_nodal(const Grid &g, const double lo[3], const double hi[3], const double h[3], const double *u,
                         int ncomp, const double *p, int np, double *out, int *bad, hipStream_t s);
int reduction_partials(const Grid &g);
int reduction_partials_full(size_t n);

// PCG vector kernels (MultigridSolver.hh:1119-1143); scalars live in device memory `sc`
// (layout documented in vf_kernels_vec.hip).
void launch_pcg_direction(const Grid &g, const double *s, double *d, const double *sc, int mode, hipStream_t st);
// fused PCG steps (no host callback): sc[0] = <r,s> (and sc[1] = <Ad,s> when adp is given) with s in the
// cycle's precision; first ? d = s : (x += sc[5] d, d = s + beta d); r -= alpha Ad with <r,r> and an
// optional float copy of r; x += sc[5] d after the last iteration
template <typename TS>
void launch_pcg_dots_s(const Grid &g, const double *r, const TS *s, const double *adp, double *partials, double *out,
                       hipStream_t st);
template <typename TS>
void launch_pcg_direction_x(const Grid &g, const TS *s, double *d, double *x, const double *sc, int mode, int first,
                            hipStream_t st);
void launch_pcg_residual(const Grid &g, double *r, const double *Ad, const double *sc, double *partials, double *rr_out,
                         float *r32, hipStream_t st);
void launch_pcg_finish_x(const Grid &g, double *x, const double *d, const double *sc, hipStream_t st);
void launch_pcg_update(const Grid &g, double *x, double *r, const double *d, const double *Ad, const double *sc,
                       double *partials, double *rr_out, hipStream_t st);
void launch_pcg_scalars(double *sc, int op, hipStream_t st);

// setup-time Galerkin coarsening and assembly (MultigridSolver.hh:704-905)
void launch_galerkin_l1_elems(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, double *Ke,
                              hipStream_t s);
// element matrices of the element rows [ylo, yhi) (default: all rows)
void launch_galerkin_l2_elems(const Grid &g2, const Grid &gf, const double *M2, const double *Ef, double *Ke,
                              hipStream_t s, int ylo = 0, int yhi = -1);
void launch_galerkin_recurse(const Grid &gc, const Grid &gf, const double *phi, const double *Kf, double *Kc,
                             hipStream_t s, int ylo = 0, int yhi = -1);
// block rows (FP64 sums stored as OutT = double or float) of the node rows [nylo, nyhi] (default: all)
template <typename OutT>
void launch_assemble_blocks(const Grid &g, const double *Ke, OutT *blocks, hipStream_t s, int nylo = 0, int nyhi = -1);
// level-1 block rows assembled from the fine moduli and the eight coarsened element matrices
// (reflect: the element matrix is reflection-symmetric, so child 0's matrix gives all eight)
template <typename OutT>
void launch_assemble_l1_blocks(const Grid &gc, const Grid &gf, const double *Chat, const double *Ef, OutT *blocks,
                               bool reflect, hipStream_t s, int nylo = 0, int nyhi = -1);
// coarsest level: band Cholesky of the reduced operator (half-bandwidth kb blocks of 64) and the explicit
// X = L^-1 (dense lower triangular, leading dimension n rounded up to 64) and its transpose XT (filled on
// and above the diagonal); flag != 0 if not positive definite
size_t band_storage(int n, int kb);
void band_cholesky_inverse(const Grid &g, const double *Ke, const int *red, int n, int kb, double *B, double *Linv,
                           double *X, double *XT, int *flag, hipStream_t s);
// the same on a lower band already assembled in B (band_storage(n, kb) doubles, zero outside the band)
void band_factor_inverse(int n, int kb, double *B, double *Linv, double *X, double *XT, int *flag, hipStream_t s);
// x = X^T X b on the reduced unknowns (zero elsewhere); reduced DoF i is component f % 3 of node f / 3,
// f = redToFull[i]; x holds ncomp planes of nnodes values
template <typename T>
void launch_coarse_chol_apply(int nred, const double *X, const double *XT, const int *redToFull, int nnodes, const T *b,
                              T *x, double *tmp_in, double *tmp_out, hipStream_t s, int ncomp = 3);
// the same for nsys independent systems in one launch sequence: system z has n[z] (device) unknowns in
// nbz[z] = ceil(n[z] / 64) of the nb blocks (only those are factored and inverted), its element matrices at
// Ke + z keStride on the sub-grid g, its reduced map at red + z redStride; band storage
// band_storage(nb * 64, kb) per system, Linv nb * 4096, X and XT (nb * 64)^2 (leading dimension nb * 64);
// flags[z] != 0 if system z is not positive definite
void band_cholesky_inverse_batched(const Grid &g, const double *Ke, size_t keStride, const int *red, size_t redStride,
                                   int nsys, const int *n, const int *nbz, int nb, int kb, double *B, double *Linv,
                                   double *X, double *XT, int *flags, hipStream_t s);
// x = X_z^T X_z b over every system's reduced unknowns (DoF i of system z: redToFull[z redStride + i] on the
// nodes shifted by z nodeStride), zero elsewhere; tmp_in / tmp_out hold nsys * ld doubles; skip (optional,
// device): systems z with skip[z] == 2 are left out (zero)
template <typename T>
void launch_coarse_chol_apply_batched(int nsys, const int *n, int ld, const double *X, const double *XT,
                                      const int *redToFull, size_t redStride, int nnodes, size_t nodeStride, const T *b,
                                      T *x, double *tmp_in, double *tmp_out, hipStream_t s, const int *skip = nullptr);

// Stacked layer-by-layer batches: segment k of the stacked grid spans the node planes [k P, (k + 1) P)
// (the last one up to the end); per-segment scalars at sc + 16 k in the PCG layout; state[k]: 0 iterating,
// 1 converged at the last check (deferred x update due), 2 done.
int active_partials(const Grid &g);  // partial sums of one active-row reduction on g
void launch_seg_sum_rows(const Grid &g, const double *partials, int nsets, size_t setStride, int P, int nseg,
                         double *sc, int off, hipStream_t s);
void launch_seg_sum_chunks(const double *partials, int tiles, int xchunk, int nchunks, int P, int nseg, double *sc,
                           int off, hipStream_t s);
void launch_seg_dot(const Grid &g, const double *a, const double *b, double *partials, int P, int nseg, double *sc,
                    int off, hipStream_t s);
template <typename TS>
void launch_seg_dots_s(const Grid &g, const double *r, const TS *sv, const double *adp, double *partials, int P,
                       int nseg, double *sc, hipStream_t s);
void launch_seg_scalars(double *sc, const int *state, int nseg, int op, hipStream_t s);
template <typename TS>
void launch_seg_direction_x(const Grid &g, const TS *sv, double *d, double *x, const double *sc, const int *state,
                            int P, int nseg, int mode, int first, hipStream_t s);
void launch_seg_residual(const Grid &g, double *r, const double *Ad, double *sc, const int *state, int P, int nseg,
                         double *partials, float *r32, hipStream_t s);
void launch_seg_finish_x(const Grid &g, double *x, const double *d, const double *sc, const int *state, int P,
                         int nseg, hipStream_t s);
template <typename T>
void launch_stack_elems(const Grid &gs, const Grid &go, int P, int nseg, const int *top, const T *src, T *dst,
                        hipStream_t s);
void launch_stack_nodes(const Grid &gs, const Grid &go, int P, int nseg, const int *top, const double *src,
                        double *dst, hipStream_t s);
void launch_fold_gradient(const Grid &go, const Grid &gs, int P, int nseg, const int *top, const double *gst,
                          double *g, hipStream_t s);
// mask[n] = 7 for the nodes of level grid g in the void columns (local x > Wl in segments of Pl planes) or at
// or above their segment's first detached row ydet[k]
void launch_stack_mask(const Grid &g, int Pl, int Wl, int nseg, const int *ydet, uint8_t *mask, hipStream_t s);
// v = 0 on every node whose mask is 7
template <typename T>
void launch_zero_held(const Grid &g, const uint8_t *mask, T *v, hipStream_t s);

// elementwise physics
void launch_young(const Grid &g, const double *rho, double *E, float *E32, int law, double E0, double Emin,
                  double gamma, double q, int maskActive, hipStream_t s);
void launch_zero_young_layers(const Grid &g, int y0, int y1, double *E, float *E32, hipStream_t s);
void launch_compliance_gradient(const Grid &g, const double *K0, const double *rho, const double *u, double *out,
                                int law, double E0, double Emin, double gamma, double q, const double *grav, double vol,
                                int accumulate, hipStream_t s);
void launch_energy_density(const Grid &g, const double *K0, const double *E, const double *u, double *out,
                           hipStream_t s);
void launch_self_weight(const Grid &g, const double *rho, const double *grav, double vol, double *f, hipStream_t s);

// filters / optimiser
// y += a x
void launch_axpy_plain(size_t n, double a, const double *x, double *y, hipStream_t s);
void launch_smoothing_filter(const int dims[3], double rmin, int type, const double *in, double *out, hipStream_t s);
// UpsampleFilter / VertexToCellFilter on an nd-dimensional grid of inDims values (adjoint: in holds the output
// grid's gradient, out the input grid's); LangelaarFilter on a planar nx x ny grid (lam: nx ny scratch)
void upsample_dims(int nd, const int64_t *in, int factor, int64_t *out);
void launch_upsample(int nd, const int64_t *inDims, int factor, const double *in, double *out, bool adjoint,
                     hipStream_t s);
void launch_vertex_to_cell(int nd, const int64_t *inDims, const double *in, double *out, bool adjoint, hipStream_t s);
void launch_langelaar(int nx, int ny, const double *in, double *out, double *smax, hipStream_t s);
void launch_langelaar_backprop(int nx, int ny, const double *g, const double *vars, const double *filtered,
                               const double *smax, double *lam, double *out, hipStream_t s);
void launch_projection(double beta, size_t n, const double *in, double *out, hipStream_t s);
void launch_projection_backprop(double beta, size_t n, const double *g, const double *vars, double *out,
                                hipStream_t s);
void launch_oc_update(size_t n, const double *x0, const double *dJ, const double *dc, double lambda, double move,
                      double p, double *out, hipStream_t s);

}  // namespace vf
