/*
 * voxelfem.h -- C ABI of the MI355X-native VoxelFEM solve-and-sensitivity library
 * (libvoxelfem.so, hand-written HIP kernels for gfx950).
 *
 * Every entry point returns VF_OK (0) or a negative status; the message of the last failure on
 * the calling thread is returned by vf_last_error(). No C++ exception crosses this boundary.
 *
 * Data conventions (identical to the reference's, MeshFEM/VoxelFEM):
 *   - grid node (ix,iy,iz) has flat index (ix*nny + iy)*nnz + iz, element (ex,ey,ez) has
 *     (ex*ny + ey)*nz + ez                             (NDVector.hh:256-275)
 *   - element local node l = 4i+2j+k at (ex+i,ey+j,ez+k), element DoF 3l+c
 *                                                       (TensorProductSimulator.hh:254-278)
 *   - a nodal vector field of nn nodes is passed as 3 component planes (SoA, 3*nn doubles,
 *     plane c holds component c): the memory image of the reference's ColMajor VField
 *                                                       (TensorProductSimulator.hh:180)
 *   - the build (layer) direction is axis 1               (TensorProductSimulator.hh:1851)
 * Every vector argument comes with a `where` flag: VF_HOST (0) for host memory, VF_DEVICE (1)
 * for device memory on the library's device. Device results are produced on the handle's
 * stream; host results are complete when the call returns.
 */
#ifndef VOXELFEM_H
#define VOXELFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VF_OK 0
#define VF_ERR_ARG (-1)
#define VF_ERR_RUNTIME (-2)
#define VF_ERR_NAN (-3)
#define VF_ERR_HIP (-4)

#define VF_HOST 0
#define VF_DEVICE 1

/* interpolation laws (TensorProductSimulator.hh:166 InterpolationLaw) */
#define VF_SIMP 0
#define VF_RAMP 1

/* preconditioner precision inside the FP64 PCG */
#define VF_PREC_FP64 0
#define VF_PREC_FP32 1

typedef struct vf_sim vf_sim;  /* TensorProductSimulator<double,1,1,1> (or <double,1,1>, vf_sim_create_2d) */
typedef struct vf_mg vf_mg;    /* MultigridSolver of the same instantiation */

/* ---- library -------------------------------------------------------------------------- */
const char *vf_last_error(void);
/* version string of the library build */
const char *vf_version(void);
/* number of visible devices (0 when no GPU) */
int vf_device_count(int *count);

/* ---- device memory helpers (no reference counterpart: the reference is host-only). kind:
 *      0 host->device, 1 device->host, 2 device->device. Copies are ordered on `stream` and
 *      complete on return when they involve host memory. */
int vf_malloc(void **ptr, size_t bytes);
int vf_free(void *ptr);
int vf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int vf_stream_sync(void *stream);
/* create / destroy a stream of default priority that orders itself with the null stream (work on the
 * null stream, e.g. a framework's default stream, and work on this stream do not overlap) */
int vf_stream_create(void **stream);
int vf_stream_destroy(void *stream);

/* ---- simulator: TensorProductSimulator (TensorProductSimulator.hh:209-285, factory
 *      python_bindings/VoxelFEM.cc:423-430). stream: a hipStream_t or NULL for the null stream. */
int vf_sim_create(vf_sim **out, const int64_t ne[3], const double minCorner[3], const double maxCorner[3], void *stream);
/* the planar instantiation TensorProductSimulator<double,1,1> (python_bindings/VoxelFEM.cc:301-308, factory
 * 198-212): bilinear quads on ne[0] x ne[1] elements, plane-stress isotropic tensor. Every vf_sim_* / vf_mg_* /
 * vf_lbl_* entry point accepts a planar simulator (and a hierarchy built on it) with these differences: nodal
 * fields are SoA 2*nn (x then y plane), node n = ix * (ne[1] + 1) + iy; Dirichlet masks use bits 0..1; K0 is
 * 8x8 (64 doubles); gravity and stretchings use their first two entries (get_stretchings sets h[2] = 0);
 * constant_strain_load takes the 2x2 strain in eps[0..3]; sample_nodal_field takes np x 2 points;
 * level_nodes reports nn[2] = 1; the V-cycle runs in FP64 only (VF_PREC_FP32 is refused), the batched
 * layer-by-layer entry solves the layers one after the other, and the autotuner and kernel benchmarks
 * (3D kernel families) are refused. */
int vf_sim_create_2d(vf_sim **out, const int64_t ne[2], const double minCorner[2], const double maxCorner[2],
                     void *stream);
/* 2 for a planar simulator, 3 otherwise (the Degrees... of the instantiation) */
int vf_sim_dimension(vf_sim *sim, int *dim);
int vf_sim_destroy(vf_sim *sim);
int vf_sim_set_stream(vf_sim *sim, void *stream);
/* ETensor of an isotropic material (setETensor, TensorProductSimulator.hh:333-347) */
int vf_sim_set_material(vf_sim *sim, double young, double poisson);
/* interpolationLaw/E_0/E_min/gamma/q properties (python_bindings/VoxelFEM.cc:115-119) */
int vf_sim_set_law(vf_sim *sim, int law, double E0, double Emin, double gamma, double q);
/* setDensities (TensorProductSimulator.hh:784-789) */
int vf_sim_set_densities(vf_sim *sim, const double *rho, int where);
int vf_sim_get_densities(vf_sim *sim, double *rho, int where);
/* getYoungModulusScaleFactor (TensorProductSimulator.hh:968) */
int vf_sim_get_young(vf_sim *sim, double *E, int where);
/* fullDensityElementStiffnessMatrix (TensorProductSimulator.hh:1077): 24x24 row-major, host */
int vf_sim_get_K0(vf_sim *sim, double *K0);
/* Dirichlet conditions per node (BCBuilder::apply, TensorProductSimulator.hh:511-560):
 * mask[n] bit c set = component c constrained; values: SoA 3*nn or NULL for zero. Host. */
int vf_sim_set_dirichlet(vf_sim *sim, const uint8_t *mask, const double *values);
/* nodal point forces used by buildLoadVector (TensorProductSimulator.hh:1269-1276), SoA, host */
int vf_sim_set_forces(vf_sim *sim, const double *f);
/* gravity property (TensorProductSimulator.hh:1776-1777) */
int vf_sim_set_gravity(vf_sim *sim, const double g[3]);
/* setFabricationMaskHeightByLayer / decrementFabricationMaskHeightByLayer
 * (TensorProductSimulator.hh:311-329) */
int vf_sim_set_mask_layer(vf_sim *sim, int64_t layer);
int vf_sim_get_mask_info(vf_sim *sim, int64_t *firstMaskedElementLayer, int64_t *firstDetachedNodeLayer, double *height);
/* applyK (TensorProductSimulator.hh:1410-1438 -> TPSStencils.hh:431-728): y = K u (SoA) */
int vf_sim_apply_K(vf_sim *sim, const double *u, double *y, int where);
/* buildLoadVector (TensorProductSimulator.hh:1269-1288) */
int vf_sim_build_load(vf_sim *sim, double *f, int where);
/* complianceGradient (TensorProductSimulator.hh:972-1005, 1043-1048) */
int vf_sim_compliance_gradient(vf_sim *sim, const double *u, double *g, int where);
/* accumulateComplianceGradient (TensorProductSimulator.hh:1008-1040): g += dC/drho over unmasked layers */
int vf_sim_accumulate_compliance_gradient(vf_sim *sim, const double *u, double *g, int where);
/* elementEnergyDensity (TensorProductSimulator.hh:1057-1073) */
int vf_sim_energy_density(vf_sim *sim, const double *u, double *out, int where);
/* constantStrainLoad (TensorProductSimulator.hh:1119-1145, element load 84-113): the load of the constant strain
 * eps (3x3 row-major, symmetric) through the material's isotropic tensor, each element weighted by its density
 * (not its interpolated modulus). F: SoA 3*nn. */
int vf_sim_constant_strain_load(vf_sim *sim, const double eps[9], double *F, int where);
/* sampleNodalField (TensorProductSimulator.hh:1161-1179): u holds ncomp nodal planes of nn values, p the np query
 * points (np x 3, row-major), out ncomp planes of np values. Each point is clamped to the domain and located in
 * its element, whose trilinear interpolant is evaluated at the clamped point's coordinates themselves, as the
 * reference does (it passes the point, not its reference coordinates in the element). */
int vf_sim_sample_nodal_field(vf_sim *sim, const double *u, int64_t ncomp, const double *p, int64_t np, double *out,
                              int where);
/* setStretchings / getStretchings (TensorProductSimulator.hh:54-59,819-823,1779): element size per axis; the
 * element stiffness, element volume and element lookup follow it, node positions do not. */
int vf_sim_set_stretchings(vf_sim *sim, const double h[3]);
int vf_sim_get_stretchings(vf_sim *sim, double h[3]);
/* dot product a.b of two nodal fields (3*nn values): masked = 0 over every node (the compliance
 * 1/2 f.u of TopologyOptimizationObjective.hh:41-43 is half of it), masked = 1 over the non-detached
 * nodes only (maskedNodalDotProduct, TensorProductSimulator.hh:445-447) */
int vf_sim_dot(vf_sim *sim, const double *a, const double *b, int masked, int where, double *out);

/* ---- multigrid: MultigridSolver (MultigridSolver.hh:35-121, binding VoxelFEM.cc:155-195) */
int vf_mg_create(vf_mg **out, vf_sim *fine, int64_t numCoarseningLevels);
int vf_mg_destroy(vf_mg *mg);
int vf_mg_num_levels(vf_mg *mg, int64_t *levels);
int vf_mg_level_nodes(vf_mg *mg, int64_t l, int64_t nn[3]);
int vf_mg_level_dirichlet(vf_mg *mg, int64_t l, uint8_t *mask);
/* setSymmetricGaussSeidel (MultigridSolver.hh:123-125) */
int vf_mg_set_symmetric_gs(vf_mg *mg, int symmetric);
/* updateStiffnessMatrices (MultigridSolver.hh:846-905): Galerkin hierarchy + coarse factorization,
 * queued on the solver stream (returns before the device work completes; the factorization is issued
 * from a worker thread onto a second stream). A non-positive-definite coarsest operator is reported by
 * vf_mg_finish_update or by the next coarse solve, whichever comes first, and again by every later solve
 * until the stiffness is updated. */
int vf_mg_update_stiffness(vf_mg *mg);
/* Completes the host side of the last vf_mg_update_stiffness: joins the worker thread, orders the
 * solver's stream after the factorization and reports its failure (the reference's
 * updateStiffnessMatrices reports it on return). A later vf_stream_sync of the solver's stream then
 * covers the whole update. No-op when nothing is pending. */
int vf_mg_finish_update(vf_mg *mg);
/* setFabricationMaskHeightByLayer (MultigridSolver.hh:1022-1028) */
int vf_mg_set_mask_layer(vf_mg *mg, int64_t layer);
/* Partial stiffness updates after layers are removed from the top of the mask (m_partialStiffnessMatrixUpdate,
 * MultigridSolver.hh:907-1017; on by default): the element matrices of the coarse element rows holding the newly
 * masked fine layers are recomputed and the block rows of the nodes they touch reassembled, instead of the whole
 * hierarchy; the coarsest level is refactored. enable = 0 rebuilds the whole hierarchy every time. The count is
 * the number of partial updates made so far. */
int vf_mg_set_partial_updates(vf_mg *mg, int enable);
/* Queue the Galerkin chain of levels >= 2 (and the block rows assembled from it) on a side stream of default
 * priority, so the first cycle's fine-level and level-1 work, which needs only the moduli, overlaps it
 * (enable = 1); 0 (the default: measured no faster on a 384x192x192 cantilever, where the chain takes ~4 ms)
 * queues it on the solver's stream. No reference counterpart (the reference's updateStiffnessMatrices,
 * MultigridSolver.hh:846-905, is synchronous). */
int vf_mg_set_overlap_setup(vf_mg *mg, int enable);
int vf_mg_partial_update_count(vf_mg *mg, int64_t *count);
/* applyK(l, u) (MultigridSolver.hh:471-504) */
int vf_mg_apply_K(vf_mg *mg, int64_t l, const double *u, double *y, int where);
/* computeResidual (MultigridSolver.hh:527-541) */
int vf_mg_residual(vf_mg *mg, int64_t l, const double *u, const double *b, double *r, int where);
/* one multicolour block Gauss-Seidel sweep (smoothingMulticoloredGS, MultigridSolver.hh:452-458), in place */
int vf_mg_smooth(vf_mg *mg, int64_t l, double *u, const double *b, int forward, int where);
/* restriction l -> l+1 (MultigridSolver.hh:214-275) */
int vf_mg_restrict(vf_mg *mg, int64_t l, const double *fine, double *coarse, int where);
/* interpolation / accum_interpolation l+1 -> l (MultigridSolver.hh:178-212) */
int vf_mg_interpolate(vf_mg *mg, int64_t l, const double *coarse, double *fine, int accumulate, int where);
/* coarsest direct solve (TensorProductSimulator::solve, TensorProductSimulator.hh:1198-1230) */
int vf_mg_coarse_solve(vf_mg *mg, const double *b, double *x, int where);
/* debug_get_x / debug_get_b (MultigridSolver.hh:1158-1159): level l's iterate (which = 0) or right-hand side
 * (which = 1) of the last multigrid cycle, 3*nn_l doubles (SoA; an FP32 cycle's buffers are widened). */
int vf_mg_debug_get(vf_mg *mg, int64_t l, int which, double *out, int where);
/* zeroOutDirichletComponents(l, u) (MultigridSolver.hh:507-513), in place */
int vf_mg_zero_dirichlet(vf_mg *mg, int64_t l, double *u, int where);
/* MultigridSolver::solve (MultigridSolver.hh:546-575): numSteps V-cycles (the first one a full
 * multigrid cycle when fullMultigrid) from the initial guess x, in place. zeroDirichlet replaces the
 * Dirichlet values by zero (residual systems). cb(user, i, x) after every cycle (host copy). */
typedef int (*vf_mg_callback)(void *user, int64_t it, const double *x);
int vf_mg_solve(vf_mg *mg, double *x, const double *f, int64_t numSteps, int64_t smoothing, int stiffnessUpdated,
                int zeroDirichlet, int fullMultigrid, vf_mg_callback cb, void *user, int where);
/* applyPreconditionerInv (MultigridSolver.hh:577-580): s = M^-1 r with mgIterations cycles */
int vf_mg_precondition(vf_mg *mg, const double *r, double *s, int64_t mgIterations, int64_t smoothing,
                       int fullMultigrid, int precision, int where);

/* per-iteration callback (it_callback(i, x, r), MultigridSolver.hh:1147-1150); x and r are host
 * arrays of 3*nn doubles (SoA) valid during the call. Return nonzero to abort the solve. */
typedef int (*vf_pcg_callback)(void *user, int64_t it, const double *x, const double *r);

typedef struct vf_pcg_stats {
    int64_t iterations;
    double residual_norm;     /* ||r|| over non-detached nodes at exit */
    double rhs_norm;          /* ||b|| over non-detached nodes */
    double seconds;           /* wall time of the call */
} vf_pcg_stats;

/* preconditionedConjugateGradient (MultigridSolver.hh:1046-1152), x updated in place.
 * precision = VF_PREC_FP64 reproduces the reference (Fletcher-Reeves beta, FP64 V-cycle);
 * VF_PREC_FP32 runs the V-cycle in FP32 (coarse solve FP64) inside an FP64 flexible PCG
 * (Polak-Ribiere beta). Returns VF_ERR_NAN with "NaN encountered" like the reference. */
int vf_mg_pcg(vf_mg *mg, double *x, const double *b, int64_t maxIter, double tol, int64_t mgIterations,
              int64_t smoothing, int fullMultigrid, int precision, int dirichletAlreadySatisfied,
              vf_pcg_callback cb, void *user, int where, vf_pcg_stats *stats);

/* LayerByLayerEvaluator::run with the "zero" initial guess (LayerByLayer.hh:223-300): every layer
 * solve starts from zero, so the solves are independent. objective = 1/2 sum_l f_l.u_l / #layers,
 * grad = sum_l g_l / #layers (LayerByLayer.hh:299-300). perLayer arrays (optional, host) receive
 * f_l.u_l and the PCG iteration counts, top layer first. */
int vf_lbl_run_zero(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                    int64_t smoothing, int fullMultigrid, int precision, double *objective, double *grad,
                    int gradWhere, double *perLayerCompliance, int64_t *perLayerIterations);
/* LayerByLayerEvaluator::run with the reference's initial guesses (LayerByLayer.hh:24-217,223-300):
 * initMethod VF_LBL_INIT_ZERO; VF_LBL_INIT_FD with history 1 (the previous layer's solution, "constant")
 * or 2 (2 u_0 - u_1, "fd"); VF_LBL_INIT_SUBSPACE with history N ("N=N", the reference's default N = 3): the
 * Galerkin projection of the next layer's system onto the last N solutions, its coefficients from the
 * pseudo-inverse of the N x N subspace matrix. uFull (optional, in memory `uFullWhere`): the full-design
 * solution of a previous run; with zeroInit = 0 and uFullValid = 1 the full-design solve starts from it
 * (LayerByLayer.hh:234,263-264), and on return it holds this run's full-design solution (282-283). */
#define VF_LBL_INIT_ZERO 0
#define VF_LBL_INIT_FD 1
#define VF_LBL_INIT_SUBSPACE 2
/* lblCallback(l, compliance, grad_compliance, u) after every layer solve (LayerByLayer.hh:222,277-279): l is the
 * layer height, compliance = f_l.u_l, grad the layer's compliance gradient (numElements host doubles) and u
 * its solution (3*numNodes host doubles, SoA), valid during the call. Return nonzero to abort the run. */
typedef int (*vf_lbl_callback)(void *user, int64_t l, double compliance, const double *grad, const double *u);
/* itcb (optional) is the PCG it_callback of every layer solve (as in vf_mg_pcg); lblcb (optional) the per-layer
 * callback above. */
int vf_lbl_run(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations, int64_t smoothing,
               int fullMultigrid, int precision, int initMethod, int64_t history, int zeroInit, double *uFull,
               int uFullValid, int uFullWhere, double *objective, double *grad, int gradWhere,
               double *perLayerCompliance, int64_t *perLayerIterations, vf_pcg_callback itcb, void *ituser,
               vf_lbl_callback lblcb, void *lbluser);
/* The same evaluation with the layer solves batched (LayerByLayer.hh:223-300 with the zero guess): the
 * partial structures of a batch are laid side by side along x in one stacked grid, separated by one
 * coarsest-level element of void, and solved as one block-diagonal problem: one hierarchy, one Galerkin
 * chain, one V-cycle launch sequence per iteration for the whole batch, with per-structure PCG scalars
 * and stopping tests and a batched block-diagonal coarsest solve. `pipelines` caps the structures per
 * batch (0: as many as fit 2^25 stacked nodes); with useGraphs the steady-state iteration is captured once
 * into a HIP graph and replayed. Each structure's iterates are those of its own solve up to the rounding of
 * the PCG reductions; results are summed in layer order. Needs a multigrid preconditioner and box elements
 * (otherwise the layers are solved one after the other, with the same results). No reference counterpart
 * for the two extra arguments. */
int vf_lbl_run_zero_batched(vf_mg *mg, int64_t layerIncrement, int64_t maxIter, double tol, int64_t mgIterations,
                            int64_t smoothing, int fullMultigrid, int precision, int64_t pipelines, int useGraphs,
                            double *objective, double *grad, int gradWhere, double *perLayerCompliance,
                            int64_t *perLayerIterations);

/* ---- filters and optimiser kernels (TopologyOptimizationFilter.hh, OptimalityCriterion.hh) --- */
/* SmoothingFilter::apply == backprop (TopologyOptimizationFilter.hh:285-400). rmin is the kernel
 * radius of the linear hat: the reference's integer `radius` r corresponds to rmin = r + 1.
 * type 0 = Const, 1 = Linear. dims: element grid (3). */
int vf_smoothing_filter(const int64_t dims[3], double rmin, int type, const double *in, double *out, int where,
                        void *stream);
/* ProjectionFilter::apply / backprop (TopologyOptimizationFilter.hh:199-225) */
/* SmoothingFilter::apply on a planar grid of dims[0] x dims[1] cells (the same filter at N = 2: 2D offsets,
 * Linear weight rmin - |(a, b)|) */
int vf_smoothing_filter_2d(const int64_t dims[2], double rmin, int type, const double *in, double *out, int where,
                           void *stream);
/* UpsampleFilter::apply / backprop (TopologyOptimizationFilter.hh:419-525), degree 1, on an nd-dimensional
 * (nd = 2 or 3) grid of inDims values (each >= 2): out has (inDims[k] - 1) * factor + 1 values along axis k,
 * the multilinear interpolant of the coarse cells; the backprop maps a gradient on that fine grid back to the
 * inDims grid (the transpose). C order, last index fastest. */
int vf_upsample_filter(int64_t nd, const int64_t *inDims, int64_t factor, const double *in, double *out, int where,
                       void *stream);
int vf_upsample_filter_backprop(int64_t nd, const int64_t *inDims, int64_t factor, const double *gOut, double *gIn,
                                int where, void *stream);
/* VertexToCellFilter::apply / backprop (TopologyOptimizationFilter.hh:527-597): the cell values (inDims[k] - 1
 * along each axis) are the means of their 2^nd corner vertices; the backprop is the transpose. */
int vf_vertex_to_cell(int64_t nd, const int64_t *inDims, const double *in, double *out, int where, void *stream);
int vf_vertex_to_cell_backprop(int64_t nd, const int64_t *inDims, const double *gOut, double *gIn, int where,
                               void *stream);
/* LangelaarFilter::apply / backprop (TopologyOptimizationFilter.hh:600-712) on a planar grid dims[0] x dims[1]
 * (y the build direction), device pointers: out = the filtered densities, smax (nx ny) the smooth maxima of the
 * supporting voxels, which backprop reads with the filtered densities of the same apply; scratch is a workspace
 * of nx ny doubles. out and scratch must not alias the inputs or each other (out also serves as workspace). */
int vf_langelaar(const int64_t dims[2], const double *in, double *out, double *smax, void *stream);
int vf_langelaar_backprop(const int64_t dims[2], const double *g, const double *vars, const double *filtered,
                          const double *smax, double *scratch, double *out, void *stream);
int vf_projection(double beta, int64_t n, const double *in, double *out, int where, void *stream);
int vf_projection_backprop(double beta, int64_t n, const double *g, const double *vars, double *out, int where,
                           void *stream);
/* OCOptimizer::step elementwise update (OptimalityCriterion.hh:64-83) */
int vf_oc_update(int64_t n, const double *x0, const double *dJ, const double *dc, double lambda, double move,
                 double p, double *out, int where, void *stream);
/* mean of n values (TotalVolumeConstraint::evaluate, TopologyOptimizationConstraint.hh:30-32) */
int vf_mean(int64_t n, const double *x, int where, void *stream, double *out);
/* Allocation-free variant of vf_mean. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). x must be device memory; out is a host double (outWhere = VF_HOST: the call synchronises
 * `stream`) or a device double (VF_DEVICE: the result is stream-ordered and nothing is synchronised).
 * The partial sums live in a fixed device array of the library (256 KiB per device, held until the
 * process ends), so n is limited to 67,108,864 values; larger n fails with VF_ERR_ARG. Every call on a
 * device shares that array: calls that overlap on one device (other streams, other threads) race. */
int vf_mean_noalloc(int64_t n, const double *x, void *stream, double *out, int outWhere);

/* ---- measurement helpers (no reference counterpart: used by bench.py and the tests) ---------- */
/* Select the fine-level kernels process-wide: 0 (default) = per solver and precision, the fastest
 * family on the solver's grid (timed once, on first use); 1 = register-gather kernels with one launch
 * per colour; 2 = LDS-tiled kernels on the structured element matrix with the two-launch smoother;
 * 3 = element-centric Walsh-basis K u with the LDS-tiled smoother; 4 = Walsh-basis K u with the
 * register-gather smoother on the structured element matrix; 5 = as 3 with level 1 on the fly on the
 * matrix cores. All families compute the same operators; the tests compare them. Variant 1 keeps level 1
 * on the fly on the vector units. */
int vf_set_kernel_variant(int variant);
/* Fine-level kernel families the solver uses in `precision` (running the autotuner if it has not run):
 * choice[0] for K u (0 register-gather, 1 LDS-tiled, 2 Walsh), choice[1] for the smoother
 * (0 register-gather, 1 LDS-tiled two-launch, 2 register-gather on the structured element matrix);
 * ms[] = the autotuner's timings (gather, tiled, Walsh K u; gather, tiled, structured-gather sweep; 0
 * where it did not run). */
int vf_mg_tuning(vf_mg *mg, int precision, int choice[2], double ms[6]);
/* Level-1 operator representation the solver uses in `precision` (running the autotuner if it has not
 * run): choice[0] for the residual, choice[1] for the smoother (0 assembled 27-point block rows, 1 on the
 * fly from the fine moduli on the vector units, 2 on the fly on the matrix cores); ms[] = the autotuner's
 * timings (residual: blocks, vector, matrix; one sweep: blocks, vector, matrix; 0 where it did not run).
 * Implementation detail of this library: the reference always applies level 1 on the fly
 * (MultigridSolver.hh:483-491). */
int vf_mg_tuning_level1(vf_mg *mg, int precision, int choice[2], double ms[6]);
/* Replay the steady-state PCG iteration of vf_mg_pcg from a HIP graph captured on its second iteration
 * (enable = 1) instead of launching its kernels one by one (0, the default). Needs the solver's stream
 * to be a created stream, not the null stream; otherwise the iterations run eagerly. Same results. */
int vf_mg_set_use_graphs(vf_mg *mg, int enable);
/* Time the fine-level K d of every PCG iteration where it runs (enable = 1; resets the totals): HIP events
 * on the solver's stream around each launch, also inside the replayed graph, read after each iteration's
 * host synchronisation. Only the fused iteration (no host callback, mgIterations > 0) is timed.
 * vf_mg_kernel_timing returns the number of timed launches and their summed milliseconds. */
int vf_mg_set_kernel_timing(vf_mg *mg, int enable);
int vf_mg_kernel_timing(vf_mg *mg, int64_t *count, double *total_ms);
/* Average time of one level-`level` operation of the hierarchy over `reps` launches (HIP events on
 * the solver's stream, vectors allocated and filled beforehand): which = 0 K u, 1 residual b - K u,
 * 2 one forward smoothing sweep (all eight colours). */
int vf_bench_mg_kernel(vf_mg *mg, int which, int64_t level, int64_t reps, int precision, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* VOXELFEM_H */
