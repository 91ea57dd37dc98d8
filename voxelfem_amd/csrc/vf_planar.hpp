// vf_planar.hpp -- the planar (2D, degrees [1,1]) instantiation: TensorProductSimulator<double,1,1> and
// MultigridSolver<double,1,1> on the device (FP64), behind the same C entry points as the 3D path.
//
// Layout: node (ix, iy) -> ix * nny + iy, element (ex, ey) -> ex * ney + ey (C order, y fastest); nodal
// fields are two component planes of nnodes doubles; local element node l = 2 i + j for the corner (i, j).
// Every level >= 1 holds its element matrices (64 doubles per element), the Galerkin products of the
// level below; the level operators, smoother and residual gather them per node.
#pragma once
#include <memory>
#include <vector>

#include "vf_internal.hpp"

namespace vf {
namespace planar {

// nodes with iy < ydet are active (non-detached); elements with ey < ymask are unmasked
struct PGrid {
    int ne[2], nn[2];
    int nnodes, nelems;
    int ydet, ymask;
};

struct PSim {
    hipStream_t stream = nullptr;
    int64_t ne[2];
    double lo[2], hi[2], h[2];
    PGrid g;
    double young = 1, poisson = 0;  // ETensor default E = 1, nu = 0 (TensorProductSimulator.hh:2114)
    int law = 0;
    double E0 = 1, Emin = 1e-4, gamma = 3, q = 3;  // TensorProductSimulator.hh:2137-2142
    double grav[2] = {0, 0};
    double K0h[64];
    DBuf<double> K0, rho, E;
    std::vector<uint8_t> hmask;
    std::vector<double> hvals;
    DBuf<uint8_t> dmask;
    DBuf<double> dvals;
    std::vector<double> hforce;
    DBuf<double> force;
    bool hasForce = false;
    int64_t firstMasked, firstDetached;
    double maskHeight;
    uint64_t version = 0, bcVersion = 0;
    DBuf<double> partials, scal, stA, stB, stC;

    PSim(const int64_t ne_[2], const double lo_[2], const double hi_[2], hipStream_t s);
    size_t nn() const { return (size_t)g.nnodes; }
    size_t nel() const { return (size_t)g.nelems; }
    size_t ndof() const { return 2 * (size_t)g.nnodes; }
    void refresh_grid();
    void update_K0();
    void update_E();
    void set_material(double E, double nu);
    void set_stretchings(const double hh[2]);
    void set_law(int law_, double E0_, double Emin_, double gamma_, double q_);
    void set_densities(const double *rho_host_or_dev, bool onDevice);
    void set_dirichlet(const uint8_t *mask, const double *values);
    void set_forces(const double *f);
    void set_mask_height(double H, bool updateYoung);
    void remove_top_layers(int64_t count);
    // device-array operations on the simulator's stream
    void apply_K(const double *u, double *y, bool zeroInit = true, bool negate = false);
    void build_load(double *f);
    void compliance_gradient(const double *u, double *g, bool accumulate);
    void energy_density(const double *u, double *out);
    double dot(const double *a, const double *b, bool masked);
    void zero_dirichlet(double *u);
    void enforce_dirichlet(double *u);
    void zero_detached(double *u);
    // constantStrainLoad for the 2 x 2 strain eps (row-major), with the densities as moduli
    void constant_strain_load(const double eps[4], double *F);
    // sampleNodalField: p holds np points (x, y); out receives ncomp planes of np values; bad != 0 when a
    // point cannot be located
    void sample(const double *u, int ncomp, const double *p, int np, double *out, int *bad);
    double *stage(int which, size_t n);
};

struct PLevel {
    PGrid g;
    double h[2];
    std::vector<uint8_t> hmask;
    DBuf<uint8_t> dmask;
    DBuf<double> Ke;  // levels >= 1: element matrices
    DBuf<double> x, b, r;
};

struct PMG {
    PSim *sim;
    int L;
    std::vector<std::unique_ptr<PLevel>> lv;  // lv[0]: the fine grid (operator from sim->E, sim->K0)
    double Chat[4 * 64], phi[64];
    DBuf<double> dChat, dphi;
    bool symmetricGS = true;
    bool stiffReady = false;
    uint64_t stiffVersion = 0, stiffBc = 0;
    int64_t maskedRowsAtBuild = -1;
    // coarsest level: reduced DoF map (DoF f = 3 n + c, as the shared coarse-solve kernels expect) and the
    // explicit inverse of the band Cholesky factor
    std::vector<int> red, redToFull;
    DBuf<int> dRed, dRedToFull, flag;
    int nred = 0, kb = 1;
    DBuf<double> B, Linv, X, XT, tin, tout;
    bool factored = false;
    DBuf<double> pd, pAd, pr, px, pb, ps;

    PMG(PSim *s, int levels);
    hipStream_t st() const { return sim->stream; }
    void update_chat();
    size_t level_n(int l) const { return 2 * (size_t)lv[l]->g.nnodes; }
    void sync_masks();
    void update_stiffness();
    void ensure_stiffness();
    void factor_coarsest();
    // level operations on device arrays
    void apply_K(int l, const double *u, double *y, bool zeroInit, bool negate);
    void residual(int l, const double *u, const double *b, double *r);
    void smooth(int l, double *u, const double *b, bool forward);
    void restrict_to(int l, const double *fine, double *coarse);
    void interpolate(int l, const double *coarse, double *fine, bool accumulate);
    void coarse_solve(const double *b, double *x);
    void zero_dirichlet(int l, double *u);
    void vcycle(int l, int nu, bool residualSystem);
    void fmg(int l, int nu, bool residualSystem);
    void precondition(const double *r, double *s, int mgIter, int nu, bool fullMG);
    void solve(double *x, const double *f, int numSteps, int nu, bool stiffnessUpdated, bool zeroDirichlet,
               bool fullMG, int (*cb)(void *, int64_t, const double *), void *user);
    int64_t pcg(double *x, const double *b, int64_t maxIter, double tol, int mgIter, int nu, bool fullMG,
                bool dirichletSatisfied, int (*cb)(void *, int64_t, const double *, const double *), void *user,
                double *residualNorm, double *rhsNorm);
};

// SmoothingFilter on a planar grid (dims[0] x dims[1]), TopologyOptimizationFilter.hh:297-400 at N = 2
void launch_smoothing_filter_2d(const int dims[2], double rmin, int type, const double *in, double *out, hipStream_t s);

}  // namespace planar
}  // namespace vf
