// CPU oracle for the planar (2D, degrees [1,1]) instantiation of the reference's hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, for N = 2, the reference's templated algorithm
// (TensorProductSimulator<double,1,1>, MultigridSolver<double,1,1>, LayerByLayerEvaluator, the 2D
// SmoothingFilter); it is loaded by tests/ as the checker and never linked by the product. Every function
// cites the reference file:line it follows; the 3D oracle (vf_oracle.cpp) is the same algorithm at N = 3.
//
// Dimension-specific facts used here:
//  * bilinear quad (4 nodes, local node l = 2 i + j for the corner (i, j)), DoFs component-minor;
//  * node index (ix, iy) -> ix * nny + iy, element (ex, ey) -> ex * ney + ey (C order, last index fastest,
//    NDVector.hh:8-9);
//  * isotropic tensor: lambda = E nu / ((1 + nu)(1 - (N - 1) nu)), mu = E / (2 (1 + nu)), the formula of
//    MeshFEM's ElasticityTensor::setIsotropic (an un-vendored submodule of the reference): at N = 2 it is
//    the plane-stress tensor. Pinned against the closed-form plane-stress Q1 stiffness of the unit square
//    (tests/test_oracle2d_pins.py).
//  * 2^N = 4 colours and 3^N = 9 neighbour blocks; integrated Q1 shape function 1/4 of the element area.
#include <omp.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr long LAYER_NONE = LONG_MAX;
thread_local std::string g_err;
using Vec = std::vector<double>;

// Element stiffness (TensorProductSimulator.hh:67-80 Element_T::Stiffness, strain basis of
// TensorProductPolynomialInterpolant.hh:197-232, 2-point tensor Gauss rule TensorProductQuadrature.hh:134-143,
// exact for Q1).
void element_stiffness(double E, double nu, const double h[2], double *K0) {
    const double lam = E * nu / ((1 + nu) * (1 - nu));
    const double mu = E / (2 * (1 + nu));
    const double gp[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
    std::fill(K0, K0 + 64, 0.0);
    const double w = 0.25 * h[0] * h[1];
    for (int qx = 0; qx < 2; ++qx)
        for (int qy = 0; qy < 2; ++qy) {
            const double p[2] = {gp[qx], gp[qy]};
            double g[4][2];
            for (int l = 0; l < 4; ++l) {
                const int idx[2] = {(l >> 1) & 1, l & 1};
                double v[2], dv[2];
                for (int d = 0; d < 2; ++d) {
                    v[d] = idx[d] ? p[d] : 1 - p[d];
                    dv[d] = (idx[d] ? 1.0 : -1.0) / h[d];
                }
                g[l][0] = dv[0] * v[1];
                g[l][1] = v[0] * dv[1];
            }
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 2; ++c)
                    for (int b = 0; b < 4; ++b)
                        for (int d = 0; d < 2; ++d) {
                            double val = lam * g[a][c] * g[b][d] + mu * g[a][d] * g[b][c];
                            if (c == d) val += mu * (g[a][0] * g[b][0] + g[a][1] * g[b][1]);
                            K0[(2 * a + c) * 8 + 2 * b + d] += w * val;
                        }
        }
}

// Coarse shape functions at the nodes of child fi (MultigridSolver.hh:664-687): phi[fi][fine_local][coarse_local]
void child_interpolation(double phi[4][4][4]) {
    for (int fi = 0; fi < 4; ++fi) {
        const int co[2] = {(fi >> 1) & 1, fi & 1};
        for (int lf = 0; lf < 4; ++lf) {
            const int fo[2] = {(lf >> 1) & 1, lf & 1};
            double x[2];
            for (int d = 0; d < 2; ++d) x[d] = 0.5 * (co[d] + fo[d]);
            for (int lc = 0; lc < 4; ++lc) {
                const int cc[2] = {(lc >> 1) & 1, lc & 1};
                double v = 1;
                for (int d = 0; d < 2; ++d) v *= cc[d] ? x[d] : 1 - x[d];
                phi[fi][lf][lc] = v;
            }
        }
    }
}

// Ke_c += I^T Ke_f I with I[2i+c, 2j+d] = phi(i,j) delta_cd (MultigridSolver.hh:704-722)
void accumulate_coarsened(const double phi[4][4], const double *Kf, double *Kc) {
    double T[64];
    for (int a = 0; a < 8; ++a)
        for (int j = 0; j < 4; ++j)
            for (int d = 0; d < 2; ++d) {
                double s = 0;
                for (int i = 0; i < 4; ++i) s += Kf[a * 8 + 2 * i + d] * phi[i][j];
                T[a * 8 + 2 * j + d] = s;
            }
    for (int i2 = 0; i2 < 4; ++i2)
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 8; ++b) {
                double s = 0;
                for (int i = 0; i < 4; ++i) s += phi[i][i2] * T[(2 * i + c) * 8 + b];
                Kc[(2 * i2 + c) * 8 + b] += s;
            }
}

void inv2(const double M[4], double Mi[4]) {
    const double id = 1.0 / (M[0] * M[3] - M[1] * M[2]);
    Mi[0] = M[3] * id; Mi[1] = -M[1] * id; Mi[2] = -M[2] * id; Mi[3] = M[0] * id;
}

// Simulator state (TensorProductSimulator.hh:209-279 construction, 290-324 masks, 575-593 Dirichlet,
// 2088-2102 Young's moduli) at N = 2.
struct Sim {
    long ne[2], nn[2];
    double lo[2], hi[2], h[2];
    long nnodes, nelems;
    double young = 1, poisson = 0;
    int law = 0;
    double E0 = 1, Emin = 1e-4, gamma = 3, q = 3;
    double grav[2] = {0, 0};
    Vec rho, E, K0;
    std::vector<uint8_t> dmask;
    Vec dval, force;
    std::vector<long> dnodes;
    long firstMasked = LAYER_NONE, firstDetached = LAYER_NONE;
    double maskHeight = std::numeric_limits<double>::infinity();
    Vec Ke;      // per element 64 (coarsest level)
    Vec blocks;  // per node 9 x 4 (assembled block rows, intermediate levels)

    Sim(long nx, long ny, const double *lo_, const double *hi_) {
        ne[0] = nx; ne[1] = ny;
        for (int d = 0; d < 2; ++d) {
            if (ne[d] < 1) throw std::runtime_error("grid must have at least one element per dimension");
            nn[d] = ne[d] + 1; lo[d] = lo_[d]; hi[d] = hi_[d]; h[d] = (hi[d] - lo[d]) / ne[d];
        }
        nnodes = nn[0] * nn[1];
        nelems = ne[0] * ne[1];
        rho.assign(nelems, 0.0);
        dmask.assign(nnodes, 0);
        dval.assign(2 * nnodes, 0.0);
        force.assign(2 * nnodes, 0.0);
        updateK0();
        updateE();
    }
    long node(long ix, long iy) const { return ix * nn[1] + iy; }
    long elem(long ex, long ey) const { return ex * ne[1] + ey; }
    void updateK0() { K0.resize(64); element_stiffness(young, poisson, h, K0.data()); }
    long nondetY() const { return std::min<long>(firstDetached, nn[1]); }
    long unmaskedY() const { return std::min<long>(firstMasked, ne[1]); }
    bool elemMasked(long ey) const { return ey >= firstMasked; }
    double unmaskedE(double r) const {
        if (law == 0) return Emin + std::pow(r, gamma) * (E0 - Emin);
        return Emin + r * (E0 - Emin) / (1 + q * (1 - r));
    }
    void updateE() {
        E.resize(nelems);
        const bool maskActive = maskHeight < hi[1];
        for (long e = 0; e < nelems; ++e) E[e] = (maskActive && elemMasked(e % ne[1])) ? 0.0 : unmaskedE(rho[e]);
    }
    void setDirichlet(const uint8_t *m, const double *v) {
        dnodes.clear();
        for (long n = 0; n < nnodes; ++n) {
            dmask[n] = m[n] & 3;
            if (dmask[n]) dnodes.push_back(n);
        }
        for (long i = 0; i < 2 * nnodes; ++i) dval[i] = v ? v[i] : 0.0;
    }
    // setFabricationMaskHeight (TensorProductSimulator.hh:286-310)
    void setMaskHeight(double hgt, bool updateYoung) {
        if (hgt < 0 || hgt > hi[1]) throw std::runtime_error("Fabrication height has to be in between 0 and the domain height");
        maskHeight = hgt;
        firstMasked = (long)std::ceil(maskHeight / h[1] - 1e-10);
        firstDetached = firstMasked + 1;
        if (updateYoung) updateE();
    }
    void zeroDirichlet(double *u) const {
        for (long n : dnodes)
            for (int c = 0; c < 2; ++c) if (dmask[n] & (1 << c)) u[c * nnodes + n] = 0;
    }
    void enforceDirichlet(double *u) const {
        for (long n : dnodes)
            for (int c = 0; c < 2; ++c) if (dmask[n] & (1 << c)) u[c * nnodes + n] = dval[c * nnodes + n];
    }
    // masked reductions over non-detached nodes (TensorProductSimulator.hh:411-459)
    double maskedDot(const double *a, const double *b) const {
        const long ny = nondetY();
        double s = 0;
        for (long ix = 0; ix < nn[0]; ++ix) {
            double t = 0;
            for (int c = 0; c < 2; ++c)
                for (long iy = 0; iy < ny; ++iy) t += a[c * nnodes + node(ix, iy)] * b[c * nnodes + node(ix, iy)];
            s += t;
        }
        return s;
    }
    template <class F> void maskedVisit(F &&f, long margin = 0) const {
        const long ny = std::min<long>(nondetY() + margin, nn[1]);
        for (long ix = 0; ix < nn[0]; ++ix)
            for (int c = 0; c < 2; ++c)
                for (long iy = 0; iy < ny; ++iy) f(c * nnodes + node(ix, iy));
    }
    // fine-level matrix-free K u (TPSStencils.hh:163-185, generic applyK): rows of detached nodes are
    // skipped (and zeroed when zeroInit)
    void applyK(const double *u, double *y, bool zeroInit, bool negate) const {
        const long ny = nondetY();
        const double *K = K0.data();
        for (long ix = 0; ix < nn[0]; ++ix)
            for (long iy = 0; iy < ny; ++iy) {
                const long n = node(ix, iy);
                double f[2] = {0, 0};
                for (int i = 1; i >= 0; --i)
                    for (int j = 1; j >= 0; --j) {
                        const long ex = ix - i, ey = iy - j;
                        if (ex < 0 || ey < 0 || ex >= ne[0] || ey >= ne[1]) continue;
                        const double Ee = E[elem(ex, ey)];
                        const int ln = 2 * i + j;
                        double acc[2] = {0, 0};
                        for (int m = 0; m < 4; ++m) {
                            const long nm = node(ex + ((m >> 1) & 1), ey + (m & 1));
                            for (int d = 0; d < 2; ++d) {
                                const double um = u[d * nnodes + nm];
                                for (int c = 0; c < 2; ++c) acc[c] += K[(2 * m + d) * 8 + 2 * ln + c] * um;
                            }
                        }
                        for (int c = 0; c < 2; ++c) f[c] += Ee * acc[c];
                    }
                for (int c = 0; c < 2; ++c) {
                    double &out = y[c * nnodes + n];
                    const double base = zeroInit ? 0.0 : out;
                    out = negate ? base - f[c] : base + f[c];
                }
            }
        if (zeroInit)
            for (long ix = 0; ix < nn[0]; ++ix)
                for (long iy = ny; iy < nn[1]; ++iy)
                    for (int c = 0; c < 2; ++c) y[c * nnodes + node(ix, iy)] = 0;
    }
    // buildLoadVector (TensorProductSimulator.hh:1269-1288): integrated Q1 shape function = 1/4
    void buildLoad(double *f) const {
        std::copy(force.begin(), force.end(), f);
        if (grav[0] == 0 && grav[1] == 0) return;
        const double area = h[0] * h[1];
        for (long ex = 0; ex < ne[0]; ++ex)
            for (long ey = 0; ey < ne[1]; ++ey) {
                if (elemMasked(ey)) continue;
                const double w = 0.25 * rho[elem(ex, ey)] * area;
                for (int l = 0; l < 4; ++l) {
                    const long nd = node(ex + ((l >> 1) & 1), ey + (l & 1));
                    for (int c = 0; c < 2; ++c) f[c * nnodes + nd] += grav[c] * w;
                }
            }
    }
    void gatherElem(const double *u, long ex, long ey, double ue[8]) const {
        for (int l = 0; l < 4; ++l) {
            const long nd = node(ex + ((l >> 1) & 1), ey + (l & 1));
            for (int c = 0; c < 2; ++c) ue[2 * l + c] = u[c * nnodes + nd];
        }
    }
    double quadForm(const double ue[8]) const {
        double s = 0;
        for (int a = 0; a < 8; ++a) {
            double r = 0;
            for (int b = 0; b < 8; ++b) r += K0[a * 8 + b] * ue[b];
            s += ue[a] * r;
        }
        return s;
    }
    double dEdrho(double r) const {
        if (law == 0) return gamma * std::pow(r, gamma - 1.0) * (E0 - Emin);
        return (1 + q) * (E0 - Emin) / std::pow(1 + q * (1 - r), 2);
    }
    // computeComplianceGradient / accumulateComplianceGradient (TensorProductSimulator.hh:972-1040)
    void complianceGradient(const double *u, double *g, bool accumulate) const {
        const bool sw = grav[0] != 0 || grav[1] != 0;
        const double area = h[0] * h[1];
        for (long ex = 0; ex < ne[0]; ++ex)
            for (long ey = 0; ey < ne[1]; ++ey) {
                const long e = elem(ex, ey);
                if (elemMasked(ey)) { if (!accumulate) g[e] = 0; continue; }
                double ue[8];
                gatherElem(u, ex, ey, ue);
                double val = -0.5 * dEdrho(rho[e]) * quadForm(ue);
                if (sw)
                    for (int l = 0; l < 4; ++l) val += 0.25 * (grav[0] * ue[2 * l] + grav[1] * ue[2 * l + 1]) * area;
                if (accumulate) g[e] += val; else g[e] = val;
            }
    }
    // elementEnergyDensity (TensorProductSimulator.hh:1057-1073)
    void energyDensity(const double *u, double *out) const {
        for (long ex = 0; ex < ne[0]; ++ex)
            for (long ey = 0; ey < ne[1]; ++ey) {
                double ue[8];
                gatherElem(u, ex, ey, ue);
                out[elem(ex, ey)] = 0.5 * E[elem(ex, ey)] * quadForm(ue);
            }
    }
};

// banded Cholesky of the reduced coarsest operator (in place of CHOLMOD, TensorProductSimulator.hh:1181-1230)
struct BandChol {
    long n = 0, bw = 0;
    Vec L;
    std::vector<long> fullToRed, redToFull;
    double &at(long i, long j) { return L[i * (bw + 1) + (j - i + bw)]; }
    double at(long i, long j) const { return L[i * (bw + 1) + (j - i + bw)]; }
    void factor() {
        for (long j = 0; j < n; ++j) {
            const long i0 = std::max<long>(0, j - bw);
            double s = at(j, j);
            for (long k = i0; k < j; ++k) s -= at(j, k) * at(j, k);
            if (!(s > 0)) throw std::runtime_error("Coarse Cholesky factorization failed (matrix not positive definite)");
            const double d = std::sqrt(s);
            at(j, j) = d;
            const long iend = std::min<long>(n, j + bw + 1);
            for (long i = j + 1; i < iend; ++i) {
                const long k0 = std::max<long>(i0, i - bw);
                double t = at(i, j);
                for (long k = k0; k < j; ++k) t -= at(i, k) * at(j, k);
                at(i, j) = t / d;
            }
        }
    }
    void solve(double *x) const {
        for (long i = 0; i < n; ++i) {
            double s = x[i];
            for (long k = std::max<long>(0, i - bw); k < i; ++k) s -= at(i, k) * x[k];
            x[i] = s / at(i, i);
        }
        for (long i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (long k = i + 1; k <= std::min<long>(n - 1, i + bw); ++k) s -= at(k, i) * x[k];
            x[i] = s / at(i, i);
        }
    }
};

// Multigrid hierarchy and PCG (MultigridSolver.hh) at N = 2.
struct MG {
    Sim *fine;
    std::vector<std::unique_ptr<Sim>> coarse;
    std::vector<Sim *> sims;
    std::vector<Vec> x, b, r;
    Vec d, Ad;
    double Chat[4][64];  // coarsenedFineK0s (MultigridSolver.hh:116-120)
    double phi[4][4][4];
    bool symmetricGS = true;
    std::unique_ptr<BandChol> chol;
    bool factorUpToDate = false;
    long L;

    MG(Sim *f, long levels) : fine(f), L(levels) {
        sims.push_back(f);
        long ne[2] = {f->ne[0], f->ne[1]};
        for (long l = 1; l <= levels; ++l) {
            for (int dd = 0; dd < 2; ++dd) {
                if (ne[dd] % 2) throw std::runtime_error("Grid size currently must be divisible by 2^numCoarseningLevels (nonuniform coarsening not yet implemented)");
                ne[dd] /= 2;
            }
            auto s = std::make_unique<Sim>(ne[0], ne[1], f->lo, f->hi);
            s->young = f->young; s->poisson = f->poisson; s->updateK0();
            // Dirichlet coarsening (MultigridSolver.hh:58-103)
            Sim *fs = sims.back();
            std::vector<uint8_t> cm(s->nnodes, 0);
            for (long n : fs->dnodes) {
                const long iy = n % fs->nn[1], ix = n / fs->nn[1];
                const long id[2] = {ix, iy};
                long c0[2], c1[2];
                for (int dd = 0; dd < 2; ++dd) {
                    if (id[dd] % 2 == 0) c0[dd] = c1[dd] = id[dd] / 2;
                    else { c0[dd] = (id[dd] - 1) / 2; c1[dd] = (id[dd] + 1) / 2; }
                }
                for (long a = c0[0]; a <= c1[0]; ++a)
                    for (long bb = c0[1]; bb <= c1[1]; ++bb) cm[s->node(a, bb)] |= fs->dmask[n];
            }
            s->setDirichlet(cm.data(), nullptr);
            sims.push_back(s.get());
            coarse.push_back(std::move(s));
        }
        for (auto *s : sims) {
            x.emplace_back(2 * s->nnodes, 0.0);
            b.emplace_back(2 * s->nnodes, 0.0);
            r.emplace_back(2 * s->nnodes, 0.0);
        }
        child_interpolation(phi);
        for (int fi = 0; fi < 4; ++fi) {
            std::fill(Chat[fi], Chat[fi] + 64, 0.0);
            accumulate_coarsened(phi[fi], f->K0.data(), Chat[fi]);
        }
    }
    Sim &sim(long l) { return *sims.at(l); }

    void setMaskLayer(long layer) {  // MultigridSolver.hh:1022-1028
        fine->setMaskHeight(fine->h[1] * layer, true);
        for (long l = 1; l <= L; ++l) sims[l]->setMaskHeight(fine->maskHeight, false);
        factorUpToDate = false;
    }

    // Galerkin coarsening (MultigridSolver.hh:744-819, 846-905)
    void coarsenedElement(long l, long ex, long ey, double *out) {
        Sim &c = sim(l);
        std::fill(out, out + 64, 0.0);
        if (c.elemMasked(ey)) return;
        Sim &f = sim(l - 1);
        if (l == 1) {
            for (int fi = 0; fi < 4; ++fi) {
                const double Ef = f.E[f.elem(2 * ex + ((fi >> 1) & 1), 2 * ey + (fi & 1))];
                for (int k = 0; k < 64; ++k) out[k] += Ef * Chat[fi][k];
            }
        } else {
            double child[64];
            for (int fi = 0; fi < 4; ++fi) {
                coarsenedElement(l - 1, 2 * ex + ((fi >> 1) & 1), 2 * ey + (fi & 1), child);
                accumulate_coarsened(phi[fi], child, out);
            }
        }
        if (l >= 2 && l < L) {
            for (int i = 0; i < 4; ++i) {
                const long ni = c.node(ex + ((i >> 1) & 1), ey + (i & 1));
                for (int j = 0; j <= i; ++j) {
                    const int off[2] = {((j >> 1) & 1) - ((i >> 1) & 1), (j & 1) - (i & 1)};
                    const int slot = (off[0] + 1) * 3 + (off[1] + 1);
                    double *B = &c.blocks[(ni * 9 + slot) * 4];
                    for (int a = 0; a < 2; ++a)
                        for (int bb = 0; bb < 2; ++bb) B[2 * a + bb] += out[(2 * i + a) * 8 + 2 * j + bb];
                }
            }
        } else if (l == L) {
            std::copy(out, out + 64, &c.Ke[c.elem(ex, ey) * 64]);
        }
    }
    void updateStiffness() {
        if (L == 0) { factorUpToDate = false; return; }
        for (long l = 2; l < L; ++l) sim(l).blocks.assign(sim(l).nnodes * 9 * 4, 0.0);
        Sim &cs = sim(L);
        cs.Ke.assign(cs.nelems * 64, 0.0);
        for (int color = 0; color < 4; ++color) {
            const int o[2] = {(color >> 1) & 1, color & 1};
            for (long ex = o[0]; ex < cs.ne[0]; ex += 2)
                for (long ey = o[1]; ey < cs.ne[1]; ey += 2) {
                    double tmp[64];
                    coarsenedElement(L, ex, ey, tmp);
                }
        }
        // mirror the lower blocks into the upper ones (full symmetric block storage)
        for (long l = 2; l < L; ++l) {
            Sim &c = sim(l);
            for (long n = 0; n < c.nnodes; ++n) {
                const long iy = n % c.nn[1], ix = n / c.nn[1];
                for (int slot = 5; slot < 9; ++slot) {
                    const long jx = ix + slot / 3 - 1, jy = iy + slot % 3 - 1;
                    if (jx < 0 || jy < 0 || jx >= c.nn[0] || jy >= c.nn[1]) continue;
                    const double *S = &c.blocks[(c.node(jx, jy) * 9 + (8 - slot)) * 4];
                    double *D = &c.blocks[(n * 9 + slot) * 4];
                    for (int a = 0; a < 2; ++a)
                        for (int bb = 0; bb < 2; ++bb) D[2 * a + bb] = S[2 * bb + a];
                }
            }
        }
        factorUpToDate = false;
    }

    // level operators (MultigridSolver.hh:471-504)
    void applyK(long l, const double *u, double *y, bool zeroInit, bool negate) {
        Sim &s = sim(l);
        if (l == 0) { s.applyK(u, y, zeroInit, negate); return; }
        const long nn = s.nnodes;
        if (zeroInit) std::fill(y, y + 2 * nn, 0.0);
        const double sgn = negate ? -1.0 : 1.0;
        if (l >= 2 && s.blocks.empty() && s.Ke.empty()) updateStiffness();
        if (l == 1) {
            Sim &f = sim(0);
            const long ney = s.unmaskedY();
            for (long ex = 0; ex < s.ne[0]; ++ex)
                for (long ey = 0; ey < ney; ++ey) {
                    double ue[8], ke[8] = {0};
                    s.gatherElem(u, ex, ey, ue);
                    for (int fi = 0; fi < 4; ++fi) {
                        const double Ef = f.E[f.elem(2 * ex + ((fi >> 1) & 1), 2 * ey + (fi & 1))];
                        for (int a = 0; a < 8; ++a) {
                            double t = 0;
                            for (int bb = 0; bb < 8; ++bb) t += Chat[fi][a * 8 + bb] * ue[bb];
                            ke[a] += Ef * t;
                        }
                    }
                    for (int m = 0; m < 4; ++m) {
                        const long nd = s.node(ex + ((m >> 1) & 1), ey + (m & 1));
                        for (int c = 0; c < 2; ++c) y[c * nn + nd] += sgn * ke[2 * m + c];
                    }
                }
            return;
        }
        if (!s.blocks.empty()) {
            const long ny = s.nondetY();
            for (long ix = 0; ix < s.nn[0]; ++ix)
                for (long iy = 0; iy < ny; ++iy) {
                    const long n = s.node(ix, iy);
                    double acc[2] = {0, 0};
                    for (int slot = 0; slot < 9; ++slot) {
                        const long jx = ix + slot / 3 - 1, jy = iy + slot % 3 - 1;
                        if (jx < 0 || jy < 0 || jx >= s.nn[0] || jy >= s.nn[1]) continue;
                        const long m = s.node(jx, jy);
                        const double *B = &s.blocks[(n * 9 + slot) * 4];
                        for (int a = 0; a < 2; ++a)
                            for (int bb = 0; bb < 2; ++bb) acc[a] += B[2 * a + bb] * u[bb * nn + m];
                    }
                    for (int c = 0; c < 2; ++c) y[c * nn + n] += sgn * acc[c];
                }
            return;
        }
        for (long ex = 0; ex < s.ne[0]; ++ex)
            for (long ey = 0; ey < s.ne[1]; ++ey) {
                double ue[8];
                s.gatherElem(u, ex, ey, ue);
                const double *K = &s.Ke[s.elem(ex, ey) * 64];
                for (int a = 0; a < 8; ++a) {
                    const long nd = s.node(ex + (((a / 2) >> 1) & 1), ey + ((a / 2) & 1));
                    double t = 0;
                    for (int bb = 0; bb < 8; ++bb) t += K[a * 8 + bb] * ue[bb];
                    y[(a % 2) * nn + nd] += sgn * t;
                }
            }
    }
    void computeResidual(long l, const double *u, const double *bb, double *res) {  // 527-541
        Sim &s = sim(l);
        s.maskedVisit([&](long i) { res[i] = bb[i]; }, 4);
        applyK(l, u, res, false, true);
        s.zeroDirichlet(res);
    }

    // multicolour block Gauss-Seidel (MultigridSolver.hh:347-458)
    void smoothNode(long l, long ix, long iy, double *u, const double *bb, bool forward) {
        Sim &s = sim(l);
        const long n = s.node(ix, iy), nn = s.nnodes;
        const uint8_t dm = s.dmask[n];
        if (dm == 3) return;
        double bms[2] = {bb[n], bb[nn + n]};
        double M[4] = {0, 0, 0, 0};
        if (l == 0 || (l == 1 && L > 1)) {
            for (int i = 1; i >= 0; --i)
                for (int j = 1; j >= 0; --j) {
                    const long ex = ix - i, ey = iy - j;
                    if (ex < 0 || ey < 0 || ex >= s.ne[0] || ey >= s.ne[1]) continue;
                    const int ln = 2 * i + j;
                    double ue[8];
                    s.gatherElem(u, ex, ey, ue);
                    double cols[8][2];
                    if (l == 0) {
                        const double Ee = s.E[s.elem(ex, ey)];
                        for (int a = 0; a < 8; ++a)
                            for (int c = 0; c < 2; ++c) cols[a][c] = Ee * s.K0[a * 8 + 2 * ln + c];
                    } else {
                        Sim &f = sim(0);
                        for (int a = 0; a < 8; ++a) cols[a][0] = cols[a][1] = 0;
                        for (int fi = 0; fi < 4; ++fi) {
                            const double Ef = f.E[f.elem(2 * ex + ((fi >> 1) & 1), 2 * ey + (fi & 1))];
                            for (int a = 0; a < 8; ++a)
                                for (int c = 0; c < 2; ++c) cols[a][c] += Ef * Chat[fi][a * 8 + 2 * ln + c];
                        }
                    }
                    for (int c = 0; c < 2; ++c) {
                        double t = 0;
                        for (int a = 0; a < 8; ++a) t += cols[a][c] * ue[a];
                        bms[c] -= t;
                        for (int c2 = 0; c2 < 2; ++c2) M[2 * c + c2] += cols[2 * ln + c][c2];
                    }
                }
        } else {
            for (int slot = 0; slot < 9; ++slot) {
                const long jx = ix + slot / 3 - 1, jy = iy + slot % 3 - 1;
                if (jx < 0 || jy < 0 || jx >= s.nn[0] || jy >= s.nn[1]) continue;
                const long m = s.node(jx, jy);
                const double *B = &s.blocks[(n * 9 + slot) * 4];
                for (int a = 0; a < 2; ++a)
                    for (int c = 0; c < 2; ++c) bms[a] -= B[2 * a + c] * u[c * nn + m];
                if (slot == 4) std::copy(B, B + 4, M);
            }
        }
        if (dm) {
            double ud[2] = {0, 0};
            for (int t = 0; t < 2; ++t) {
                const int i = forward ? t : 1 - t;
                const double free = (dm & (1 << i)) ? 0.0 : 1.0;
                ud[i] = (bms[i] - (M[2 * i] * ud[0] + M[2 * i + 1] * ud[1])) * (free / M[3 * i]);
            }
            for (int c = 0; c < 2; ++c) u[c * nn + n] += ud[c];
        } else {
            double Mi[4];
            inv2(M, Mi);
            for (int c = 0; c < 2; ++c) u[c * nn + n] += Mi[2 * c] * bms[0] + Mi[2 * c + 1] * bms[1];
        }
    }
    void smooth(long l, double *u, const double *bb, bool forward) {
        Sim &s = sim(l);
        const long ny = s.nondetY();
        for (int t = 0; t < 4; ++t) {
            const int color = forward ? t : 3 - t;
            const int o[2] = {(color >> 1) & 1, color & 1};
            for (long ix = o[0]; ix < s.nn[0]; ix += 2)
                for (long iy = o[1]; iy < ny; iy += 2) smoothNode(l, ix, iy, u, bb, forward);
        }
    }

    // transfers (MultigridSolver.hh:178-275)
    void restriction(long l, const double *rf, double *rc) {
        Sim &f = sim(l), &c = sim(l + 1);
        const long fny = f.nondetY(), cny = c.nondetY();
        for (long ix = 0; ix < c.nn[0]; ++ix)
            for (long iy = 0; iy < cny; ++iy) {
                double v[2] = {0, 0};
                for (long ox = -1; ox <= 1; ++ox) {
                    const long fx = 2 * ix + ox;
                    if (fx < 0 || fx >= f.nn[0]) continue;
                    for (long oy = -1; oy <= 1; ++oy) {
                        const long fy = 2 * iy + oy;
                        if (fy < 0 || fy >= fny) continue;
                        const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0);
                        const long m = f.node(fx, fy);
                        for (int cc = 0; cc < 2; ++cc) v[cc] += w * rf[cc * f.nnodes + m];
                    }
                }
                const long n = c.node(ix, iy);
                for (int cc = 0; cc < 2; ++cc) rc[cc * c.nnodes + n] = v[cc];
            }
        for (long ix = 0; ix < c.nn[0]; ++ix)
            for (long iy = cny; iy < c.nn[1]; ++iy)
                for (int cc = 0; cc < 2; ++cc) rc[cc * c.nnodes + c.node(ix, iy)] = 0;
    }
    void interpolate(long l, const double *xc, double *xf, bool accumulate) {
        Sim &f = sim(l), &c = sim(l + 1);
        const long fny = accumulate ? f.nondetY() : f.nn[1];
        for (long ix = 0; ix < f.nn[0]; ++ix)
            for (long iy = 0; iy < fny; ++iy) {
                const long id[2] = {ix, iy};
                long e[2];
                double w[2][2];
                for (int dd = 0; dd < 2; ++dd) {
                    e[dd] = std::min<long>(id[dd] / 2, c.ne[dd]);
                    const long lf = id[dd] - 2 * e[dd];
                    w[dd][0] = lf ? 0.5 : 1.0;
                    w[dd][1] = lf ? 0.5 : 0.0;
                }
                double v[2] = {0, 0};
                for (int lc = 0; lc < 4; ++lc) {
                    const int a = (lc >> 1) & 1, bb = lc & 1;
                    const double coeff = w[0][a] * w[1][bb];
                    if (coeff == 0) continue;
                    const long m = c.node(e[0] + a, e[1] + bb);
                    for (int k = 0; k < 2; ++k) v[k] += coeff * xc[k * c.nnodes + m];
                }
                const long n = f.node(ix, iy);
                for (int k = 0; k < 2; ++k) {
                    if (accumulate) xf[k * f.nnodes + n] += v[k];
                    else xf[k * f.nnodes + n] = v[k];
                }
            }
    }

    // coarsest direct solve (TensorProductSimulator.hh:1181-1230)
    void factorCoarsest() {
        Sim &s = sim(L);
        const long nn = s.nnodes, nd = 2 * nn;
        auto ch = std::make_unique<BandChol>();
        ch->fullToRed.assign(nd, -1);
        const long ny = s.nondetY();
        for (long n = 0; n < nn; ++n) {
            const long iy = n % s.nn[1];
            for (int c = 0; c < 2; ++c) {
                const bool fixed = (s.dmask[n] & (1 << c)) || iy >= ny;
                if (!fixed) { ch->fullToRed[2 * n + c] = (long)ch->redToFull.size(); ch->redToFull.push_back(2 * n + c); }
            }
        }
        if (L == 0)
            for (long n : s.dnodes)
                for (int c = 0; c < 2; ++c)
                    if ((s.dmask[n] & (1 << c)) && s.dval[c * nn + n] != 0)
                        throw std::runtime_error("Nonzero Dirichlet constraints currently unsupported");
        ch->n = (long)ch->redToFull.size();
        const long fullBw = 2 * (s.nn[1] + 1) + 1;
        ch->bw = std::min<long>(fullBw, std::max<long>(ch->n - 1, 0));
        ch->L.assign(ch->n * (ch->bw + 1), 0.0);
        for (long ex = 0; ex < s.ne[0]; ++ex)
            for (long ey = 0; ey < s.ne[1]; ++ey) {
                double Kloc[64];
                const long e = s.elem(ex, ey);
                if (L == 0) { for (int k = 0; k < 64; ++k) Kloc[k] = s.E[e] * s.K0[k]; }
                else std::copy(&s.Ke[e * 64], &s.Ke[e * 64] + 64, Kloc);
                long dof[8];
                for (int a = 0; a < 8; ++a) {
                    const int ln = a / 2;
                    dof[a] = ch->fullToRed[2 * s.node(ex + ((ln >> 1) & 1), ey + (ln & 1)) + a % 2];
                }
                for (int a = 0; a < 8; ++a) {
                    if (dof[a] < 0) continue;
                    for (int bb = 0; bb < 8; ++bb) {
                        if (dof[bb] < 0 || dof[bb] > dof[a]) continue;
                        ch->at(dof[a], dof[bb]) += Kloc[a * 8 + bb];
                    }
                }
            }
        ch->factor();
        chol = std::move(ch);
        factorUpToDate = true;
    }
    void coarseSolve(const double *bb, double *xx) {
        if (!factorUpToDate) factorCoarsest();
        Sim &s = sim(L);
        const long nn = s.nnodes;
        Vec tmp(chol->n);
        for (long i = 0; i < chol->n; ++i) { const long f = chol->redToFull[i]; tmp[i] = bb[(f % 2) * nn + f / 2]; }
        chol->solve(tmp.data());
        std::fill(xx, xx + 2 * nn, 0.0);
        for (long i = 0; i < chol->n; ++i) { const long f = chol->redToFull[i]; xx[(f % 2) * nn + f / 2] = tmp[i]; }
    }

    // cycles (MultigridSolver.hh:546-658)
    void vcycle(long l, long nu, bool residualSystem) {
        if (l == L) { coarseSolve(b[l].data(), x[l].data()); return; }
        Sim &s = sim(l), &c = sim(l + 1);
        if (residualSystem) s.zeroDirichlet(x[l].data()); else s.enforceDirichlet(x[l].data());
        for (long i = 0; i < nu; ++i) smooth(l, x[l].data(), b[l].data(), true);
        computeResidual(l, x[l].data(), b[l].data(), r[l].data());
        restriction(l, r[l].data(), b[l + 1].data());
        c.maskedVisit([&](long i) { x[l + 1][i] = 0; }, 4);
        vcycle(l + 1, nu, true);
        interpolate(l, x[l + 1].data(), x[l].data(), true);
        for (long i = 0; i < nu; ++i) smooth(l, x[l].data(), b[l].data(), !symmetricGS);
    }
    void fullMultigrid(long l, long nu, bool residualSystem) {
        if (l == L) { coarseSolve(b[l].data(), x[l].data()); return; }
        restriction(l, b[l].data(), b[l + 1].data());
        fullMultigrid(l + 1, nu, residualSystem);
        interpolate(l, x[l + 1].data(), x[l].data(), false);
        vcycle(l, nu, residualSystem);
    }
    // applyPreconditionerInv (MultigridSolver.hh:577-580) -> result in x[0]
    void precondition(const double *res, long mgIter, long nu, bool fmg) {
        if (L > 0 && nu > 0 && sim(L).Ke.empty()) throw std::runtime_error("precondition before update_stiffness");
        Sim &s = sim(0);
        std::fill(x[0].begin(), x[0].end(), 0.0);
        if (b[0].data() != res) std::copy(res, res + 2 * s.nnodes, b[0].begin());
        if (nu == 0) { std::copy(res, res + 2 * s.nnodes, x[0].begin()); return; }
        for (long i = 0; i < mgIter; ++i) {
            if (fmg && i == 0) fullMultigrid(0, nu, true);
            else vcycle(0, nu, true);
        }
    }
    // preconditionedConjugateGradient (MultigridSolver.hh:1046-1152)
    long pcg(double *xo, const double *bb, long maxIter, double tol, long mgIter, long nu, bool fmg,
             bool dirichletSatisfied, double *hist, long histLen) {
        Sim &fs = sim(0);
        const long nd = 2 * fs.nnodes;
        Vec &res = b[0];
        if (L == 0) {
            computeResidual(0, xo, bb, res.data());
            double rr = 0, bn = 0;
            for (long i = 0; i < nd; ++i) { rr += res[i] * res[i]; bn += bb[i] * bb[i]; }
            if (rr < tol * tol * bn) return 0;
            coarseSolve(bb, xo);
            return 1;
        }
        if (!dirichletSatisfied) fs.enforceDirichlet(xo);
        const double bnorm = fs.maskedDot(bb, bb);
        computeResidual(0, xo, bb, res.data());
        double rr = fs.maskedDot(res.data(), res.data());
        if (std::isnan(rr)) throw std::runtime_error("NaN encountered");
        Vec &s = x[0];
        d.assign(nd, 0.0);
        Ad.assign(nd, 0.0);
        double rz = 0;
        bool haveD = false, updated = false;
        long it = 0;
        while (it < maxIter && rr > tol * tol * bnorm) {
            ++it;
            if (mgIter > 0) {
                if (!updated) { updateStiffness(); updated = true; }
                precondition(res.data(), mgIter, nu, fmg);
            } else s = res;
            fs.zeroDirichlet(s.data());
            const double rzOld = rz;
            rz = fs.maskedDot(res.data(), s.data());
            if (haveD) { const double beta = rz / rzOld; fs.maskedVisit([&](long i) { d[i] = s[i] + beta * d[i]; }); }
            else { d = s; haveD = true; }
            applyK(0, d.data(), Ad.data(), true, false);
            fs.zeroDirichlet(Ad.data());
            const double alpha = rz / fs.maskedDot(d.data(), Ad.data());
            fs.maskedVisit([&](long i) { xo[i] += alpha * d[i]; res[i] -= alpha * Ad[i]; });
            rr = fs.maskedDot(res.data(), res.data());
            if (std::isnan(rr)) throw std::runtime_error("NaN encountered at iteration" + std::to_string(it));
            if (hist && it - 1 < histLen) hist[it - 1] = std::sqrt(rr);
        }
        return it;
    }
};

long reflect_index(long i, long s) {  // TopologyOptimizationFilter.hh:343-349
    while (i < 0 || i >= s) {
        if (i >= s) i = 2 * s - i - 1;
        if (i < 0) i = -i - 1;
    }
    return i;
}

// SmoothingFilter::apply at N = 2 (TopologyOptimizationFilter.hh:297-400): normalised reflect-BC convolution
// over the 2D offsets |a|, |b| <= R; rmin = radius + 1 (Linear weight rmin - |(a, b)|, kept where > 0).
void smoothing_filter(const long dims[2], double rmin, int type, const double *in, double *out) {
    const long R = (long)std::ceil(rmin) - 1;
    double W = 0;
    std::vector<std::array<long, 2>> offs;
    std::vector<double> ws;
    for (long a = -R; a <= R; ++a)
        for (long b = -R; b <= R; ++b) {
            const double w = type == 1 ? rmin - std::sqrt(double(a * a + b * b)) : 1.0;
            if (w <= 0) continue;
            offs.push_back({a, b});
            ws.push_back(w);
            W += w;
        }
    for (long i = 0; i < dims[0]; ++i)
        for (long j = 0; j < dims[1]; ++j) {
            double s = 0;
            for (size_t t = 0; t < offs.size(); ++t)
                s += ws[t] * in[reflect_index(i + offs[t][0], dims[0]) * dims[1] + reflect_index(j + offs[t][1], dims[1])];
            out[i * dims[1] + j] = s / W;
        }
}

}  // namespace

#define ORC_TRY(...) try { __VA_ARGS__; return 0; } catch (std::exception &e) { g_err = e.what(); return -1; }

extern "C" {

const char *orc2_last_error() { return g_err.c_str(); }
void orc2_element_stiffness(double E, double nu, const double *h, double *out) { element_stiffness(E, nu, h, out); }

void *orc2_sim_create(long nx, long ny, const double *lo, const double *hi) {
    try { return new Sim(nx, ny, lo, hi); } catch (std::exception &e) { g_err = e.what(); return nullptr; }
}
void orc2_sim_destroy(void *s) { delete (Sim *)s; }
int orc2_sim_set_etensor(void *p, double E, double nu) { ORC_TRY(auto *s = (Sim *)p; s->young = E; s->poisson = nu; s->updateK0()) }
int orc2_sim_set_law(void *p, int law, double E0, double Emin, double gamma, double q) {
    ORC_TRY(auto *s = (Sim *)p; s->law = law; s->E0 = E0; s->Emin = Emin; s->gamma = gamma; s->q = q; s->updateE())
}
int orc2_sim_set_densities(void *p, const double *rho) { ORC_TRY(auto *s = (Sim *)p; std::copy(rho, rho + s->nelems, s->rho.begin()); s->updateE()) }
int orc2_sim_get_E(void *p, double *out) { ORC_TRY(auto *s = (Sim *)p; std::copy(s->E.begin(), s->E.end(), out)) }
// test hook: replace the moduli (e.g. by the device's, which may differ from std::pow's by an ulp)
int orc2_sim_set_E(void *p, const double *E) { ORC_TRY(auto *s = (Sim *)p; std::copy(E, E + s->nelems, s->E.begin())) }
int orc2_sim_get_K0(void *p, double *out) { ORC_TRY(auto *s = (Sim *)p; std::copy(s->K0.begin(), s->K0.end(), out)) }
int orc2_sim_set_dirichlet(void *p, const uint8_t *mask, const double *vals) { ORC_TRY(((Sim *)p)->setDirichlet(mask, vals)) }
int orc2_sim_set_forces(void *p, const double *f) { ORC_TRY(auto *s = (Sim *)p; std::copy(f, f + 2 * s->nnodes, s->force.begin())) }
int orc2_sim_set_gravity(void *p, const double *g) { ORC_TRY(auto *s = (Sim *)p; s->grav[0] = g[0]; s->grav[1] = g[1]) }
int orc2_sim_set_mask_height(void *p, double hgt) { ORC_TRY(((Sim *)p)->setMaskHeight(hgt, true)) }
int orc2_sim_apply_K(void *p, const double *u, double *y) { ORC_TRY(((Sim *)p)->applyK(u, y, true, false)) }
int orc2_sim_build_load(void *p, double *f) { ORC_TRY(((Sim *)p)->buildLoad(f)) }
int orc2_sim_compliance_gradient(void *p, const double *u, double *g) { ORC_TRY(((Sim *)p)->complianceGradient(u, g, false)) }
int orc2_sim_energy_density(void *p, const double *u, double *out) { ORC_TRY(((Sim *)p)->energyDensity(u, out)) }
int orc2_sim_masked_dot(void *p, const double *a, const double *b, double *out) { ORC_TRY(*out = ((Sim *)p)->maskedDot(a, b)) }

void *orc2_mg_create(void *sim, long levels) {
    try { return new MG((Sim *)sim, levels); } catch (std::exception &e) { g_err = e.what(); return nullptr; }
}
void orc2_mg_destroy(void *m) { delete (MG *)m; }
int orc2_mg_set_symmetric(void *m, int s) { ORC_TRY(((MG *)m)->symmetricGS = s != 0) }
int orc2_mg_set_mask_layer(void *m, long layer) { ORC_TRY(((MG *)m)->setMaskLayer(layer)) }
int orc2_mg_update_stiffness(void *m) { ORC_TRY(((MG *)m)->updateStiffness()) }
int orc2_mg_num_nodes(void *m, long l, long *out) { ORC_TRY(*out = ((MG *)m)->sim(l).nnodes) }
int orc2_mg_dirichlet_mask(void *m, long l, uint8_t *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.dmask.begin(), s.dmask.end(), out)) }
int orc2_mg_apply_K(void *m, long l, const double *u, double *y) { ORC_TRY(((MG *)m)->applyK(l, u, y, true, false)) }
int orc2_mg_residual(void *m, long l, const double *u, const double *b, double *r) {
    ORC_TRY(auto *mg = (MG *)m; std::fill(r, r + 2 * mg->sim(l).nnodes, 0.0); mg->computeResidual(l, u, b, r))
}
int orc2_mg_smooth(void *m, long l, double *u, const double *b, int forward) { ORC_TRY(((MG *)m)->smooth(l, u, b, forward != 0)) }
int orc2_mg_restrict(void *m, long l, const double *rf, double *rc) { ORC_TRY(((MG *)m)->restriction(l, rf, rc)) }
int orc2_mg_interpolate(void *m, long l, const double *xc, double *xf, int accumulate) { ORC_TRY(((MG *)m)->interpolate(l, xc, xf, accumulate != 0)) }
int orc2_mg_coarse_solve(void *m, const double *b, double *x) { ORC_TRY(((MG *)m)->coarseSolve(b, x)) }
int orc2_mg_blocks(void *m, long l, double *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.blocks.begin(), s.blocks.end(), out)) }
int orc2_mg_element_matrices(void *m, long l, double *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.Ke.begin(), s.Ke.end(), out)) }
int orc2_mg_chat(void *m, double *out) { ORC_TRY(auto *mg = (MG *)m; for (int fi = 0; fi < 4; ++fi) std::copy(mg->Chat[fi], mg->Chat[fi] + 64, out + 64 * fi)) }
int orc2_mg_precondition(void *m, const double *r, double *s, long mgIter, long nu, int fmg) {
    ORC_TRY(auto *mg = (MG *)m; mg->precondition(r, mgIter, nu, fmg != 0); std::copy(mg->x[0].begin(), mg->x[0].end(), s))
}
int orc2_mg_pcg(void *m, double *x, const double *b, long maxIter, double tol, long mgIter, long nu, int fmg,
                int dirichletSatisfied, double *hist, long histLen, long *iters) {
    ORC_TRY(*iters = ((MG *)m)->pcg(x, b, maxIter, tol, mgIter, nu, fmg != 0, dirichletSatisfied != 0, hist, histLen))
}
int orc2_smoothing_filter(const long *dims, double rmin, int type, const double *in, double *out) {
    ORC_TRY(smoothing_filter(dims, rmin, type, in, out))
}

// LayerByLayerEvaluator::run with the "zero" initial guess (LayerByLayer.hh:223-300)
int orc2_lbl_run_zero(void *m, long layerIncrement, long maxIter, double tol, long mgIter, long nu, int fmg,
                      double *objective, double *grad, double *perLayerCompliance, long *perLayerIters) {
    ORC_TRY({
        auto *mg = (MG *)m;
        Sim &s = mg->sim(0);
        const long numLayers = s.ne[1];
        std::fill(grad, grad + s.nelems, 0.0);
        double total = 0;
        long count = 0;
        Vec f(2 * s.nnodes), u(2 * s.nnodes);
        mg->setMaskLayer(numLayers);
        for (long l = numLayers; l > 0; l -= std::min(layerIncrement, l)) {
            if (l < numLayers) {
                // decrementFabricationMaskHeightByLayer (TensorProductSimulator.hh:311-324)
                s.firstMasked -= layerIncrement;
                s.firstDetached = s.firstMasked + 1;
                s.maskHeight -= layerIncrement * s.h[1];
                for (long e = 0; e < s.nelems; ++e) {
                    const long ey = e % s.ne[1];
                    if (ey >= s.firstMasked && ey < s.firstMasked + layerIncrement) s.E[e] = 0.0;
                }
                for (long lv = 1; lv <= mg->L; ++lv) mg->sim(lv).setMaskHeight(s.maskHeight, false);
            }
            s.buildLoad(f.data());
            std::fill(u.begin(), u.end(), 0.0);
            const long its = mg->pcg(u.data(), f.data(), maxIter, tol, mgIter, nu, fmg != 0, true, nullptr, 0);
            const double comp = s.maskedDot(f.data(), u.data());
            if (perLayerCompliance) perLayerCompliance[count] = comp;
            if (perLayerIters) perLayerIters[count] = its;
            total += comp;
            s.complianceGradient(u.data(), grad, true);
            ++count;
        }
        *objective = 0.5 * total / count;
        for (long e = 0; e < s.nelems; ++e) grad[e] /= count;
    })
}

}  // extern "C"
