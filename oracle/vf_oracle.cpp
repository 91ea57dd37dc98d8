// vf_oracle.cpp -- CPU restatement (FP64, OpenMP) of VoxelFEM's solve-and-sensitivity path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP implementation in
// voxelfem_amd/ and the timed CPU baseline in bench.py. The product never links, loads
// or calls it. Every function cites the reference file:line whose semantics it restates
// (paths relative to the MeshFEM/VoxelFEM repository root).
//
// Conventions restated from the reference:
//   * node (ix,iy,iz) -> (ix*nny + iy)*nnz + iz ; element (ex,ey,ez) -> (ex*ny + ey)*nz + ez
//     (NDVector.hh:256-275, TensorProductSimulator.hh:180,226)
//   * local node l = 4i+2j+k sits at (ex+i, ey+j, ez+k); element DoF order 3l+c
//     (TensorProductSimulator.hh:254-278, TPSStencils.hh:479-488)
//   * nodal vector fields are component planes (SoA): u[c*nn + n]  (VField ColMajor)
//   * build/layer direction is axis 1 (TensorProductSimulator.hh:1851)
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>

namespace {

constexpr long LAYER_NONE = LONG_MAX;
thread_local std::string g_err;

using Vec = std::vector<double>;

// ---------------------------------------------------------------------------------------
// Element stiffness  (TensorProductSimulator.hh:67-80 Element_T::Stiffness with the strain
// basis of TensorProductPolynomialInterpolant.hh:197-232 and the isotropic tensor).
// The integrand eps(phi_a e_c):C:eps(phi_b e_d) = lam d_c phi_a d_d phi_b
//   + mu (delta_cd grad phi_a.grad phi_b + d_d phi_a d_c phi_b) is integrated with the
// tensor 2-point Gauss rule (TensorProductQuadrature.hh:134-143), exact for Q1.
// ---------------------------------------------------------------------------------------
void element_stiffness(double E, double nu, const double h[3], double *K0) {
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = E / (2 * (1 + nu));
    const double gp[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
    std::fill(K0, K0 + 576, 0.0);
    const double vol = h[0] * h[1] * h[2];
    for (int qx = 0; qx < 2; ++qx)
    for (int qy = 0; qy < 2; ++qy)
    for (int qz = 0; qz < 2; ++qz) {
        const double p[3] = {gp[qx], gp[qy], gp[qz]};
        const double w = 0.125 * vol;
        double g[8][3];
        for (int l = 0; l < 8; ++l) {
            const int idx[3] = {(l >> 2) & 1, (l >> 1) & 1, l & 1};
            double v[3], dv[3];
            for (int d = 0; d < 3; ++d) {
                v[d] = idx[d] ? p[d] : 1 - p[d];
                dv[d] = (idx[d] ? 1.0 : -1.0) / h[d];
            }
            g[l][0] = dv[0] * v[1] * v[2];
            g[l][1] = v[0] * dv[1] * v[2];
            g[l][2] = v[0] * v[1] * dv[2];
        }
        for (int a = 0; a < 8; ++a)
        for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 8; ++b)
        for (int d = 0; d < 3; ++d) {
            double val = lam * g[a][c] * g[b][d] + mu * g[a][d] * g[b][c];
            if (c == d) val += mu * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
            K0[(3 * a + c) * 24 + 3 * b + d] += w * val;
        }
    }
}

// Coarse shape-function values on the 8 nodes of child element fi (MultigridSolver.hh:664-687):
// phi[fi][fine_local][coarse_local].
void child_interpolation(double phi[8][8][8]) {
    for (int fi = 0; fi < 8; ++fi) {
        const int co[3] = {(fi >> 2) & 1, (fi >> 1) & 1, fi & 1};
        for (int lf = 0; lf < 8; ++lf) {
            const int fo[3] = {(lf >> 2) & 1, (lf >> 1) & 1, lf & 1};
            double x[3];
            for (int d = 0; d < 3; ++d) x[d] = 0.5 * (co[d] + fo[d]);
            for (int lc = 0; lc < 8; ++lc) {
                const int cc[3] = {(lc >> 2) & 1, (lc >> 1) & 1, lc & 1};
                double v = 1;
                for (int d = 0; d < 3; ++d) v *= cc[d] ? x[d] : 1 - x[d];
                phi[fi][lf][lc] = v;
            }
        }
    }
}

// Ke_c += I^T Ke_f I with I[3i+c, 3j+d] = phi(i,j) delta_cd (MultigridSolver.hh:704-722).
void accumulate_coarsened(const double phi[8][8], const double *Kf, double *Kc) {
    double T[576];  // T = Kf * I
    for (int a = 0; a < 24; ++a)
        for (int j = 0; j < 8; ++j)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int i = 0; i < 8; ++i) s += Kf[a * 24 + 3 * i + d] * phi[i][j];
                T[a * 24 + 3 * j + d] = s;
            }
    for (int i2 = 0; i2 < 8; ++i2)
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 24; ++b) {
                double s = 0;
                for (int i = 0; i < 8; ++i) s += phi[i][i2] * T[(3 * i + c) * 24 + b];
                Kc[(3 * i2 + c) * 24 + b] += s;
            }
}

void inv3(const double M[9], double Mi[9]) {
    const double c00 = M[4] * M[8] - M[5] * M[7];
    const double c01 = M[5] * M[6] - M[3] * M[8];
    const double c02 = M[3] * M[7] - M[4] * M[6];
    const double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
    const double id = 1.0 / det;
    Mi[0] = c00 * id; Mi[1] = (M[2] * M[7] - M[1] * M[8]) * id; Mi[2] = (M[1] * M[5] - M[2] * M[4]) * id;
    Mi[3] = c01 * id; Mi[4] = (M[0] * M[8] - M[2] * M[6]) * id; Mi[5] = (M[2] * M[3] - M[0] * M[5]) * id;
    Mi[6] = c02 * id; Mi[7] = (M[1] * M[6] - M[0] * M[7]) * id; Mi[8] = (M[0] * M[4] - M[1] * M[3]) * id;
}

// ---------------------------------------------------------------------------------------
// Simulator state: TensorProductSimulator.hh:209-279 (construction), 290-324 (masks),
// 575-593 (Dirichlet), 2088-2102 (Young's moduli).
// ---------------------------------------------------------------------------------------
struct Sim {
    long ne[3], nn[3];
    double lo[3], hi[3], h[3];
    long nnodes, nelems;
    double young = 1, poisson = 0;          // ETensor default (E=1, nu=0), TensorProductSimulator.hh:2114
    int law = 0;                            // 0 SIMP, 1 RAMP
    double E0 = 1, Emin = 1e-4, gamma = 3, q = 3;  // TensorProductSimulator.hh:2137-2142
    double grav[3] = {0, 0, 0};
    Vec rho, E, K0;
    std::vector<uint8_t> dmask;             // per-node Dirichlet component bits
    Vec dval;                               // SoA Dirichlet values
    std::vector<long> dnodes;               // nodes with any Dirichlet component
    Vec force;                              // SoA point forces
    long firstMasked = LAYER_NONE, firstDetached = LAYER_NONE;
    double maskHeight = std::numeric_limits<double>::infinity();
    // coarsened operators (filled by the multigrid hierarchy)
    Vec Ke;      // per element 576 (cached element stiffness, coarsest level)
    Vec blocks;  // per node 27 x 9 (assembled block rows, intermediate levels)

    Sim(long nx, long ny, long nz, const double *lo_, const double *hi_) {
        ne[0] = nx; ne[1] = ny; ne[2] = nz;
        for (int d = 0; d < 3; ++d) {
            if (ne[d] < 1) throw std::runtime_error("grid must have at least one element per dimension");
            nn[d] = ne[d] + 1; lo[d] = lo_[d]; hi[d] = hi_[d]; h[d] = (hi[d] - lo[d]) / ne[d];
        }
        nnodes = nn[0] * nn[1] * nn[2];
        nelems = ne[0] * ne[1] * ne[2];
        rho.assign(nelems, 0.0);
        dmask.assign(nnodes, 0);
        dval.assign(3 * nnodes, 0.0);
        force.assign(3 * nnodes, 0.0);
        updateK0();
        updateE();
    }
    long node(long ix, long iy, long iz) const { return (ix * nn[1] + iy) * nn[2] + iz; }
    long elem(long ex, long ey, long ez) const { return (ex * ne[1] + ey) * ne[2] + ez; }
    void updateK0() { K0.resize(576); element_stiffness(young, poisson, h, K0.data()); }
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
        #pragma omp parallel for
        for (long e = 0; e < nelems; ++e) {
            const long ey = (e / ne[2]) % ne[1];
            E[e] = (maskActive && elemMasked(ey)) ? 0.0 : unmaskedE(rho[e]);
        }
    }
    void setDirichlet(const uint8_t *m, const double *v) {
        dnodes.clear();
        for (long n = 0; n < nnodes; ++n) {
            dmask[n] = m[n] & 7;
            if (dmask[n]) dnodes.push_back(n);
        }
        for (long i = 0; i < 3 * nnodes; ++i) dval[i] = v ? v[i] : 0.0;
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
            for (int c = 0; c < 3; ++c) if (dmask[n] & (1 << c)) u[c * nnodes + n] = 0;
    }
    void enforceDirichlet(double *u) const {
        for (long n : dnodes)
            for (int c = 0; c < 3; ++c) if (dmask[n] & (1 << c)) u[c * nnodes + n] = dval[c * nnodes + n];
    }
    // Masked reductions over non-detached nodes (TensorProductSimulator.hh:411-459).
    double maskedDot(const double *a, const double *b) const {
        const long ny = nondetY(), rowLen = ny * nn[2];
        double s = 0;
        #pragma omp parallel for reduction(+ : s)
        for (long ix = 0; ix < nn[0]; ++ix) {
            const long base = ix * nn[1] * nn[2];
            double t = 0;
            for (int c = 0; c < 3; ++c)
                for (long i = 0; i < rowLen; ++i) t += a[c * nnodes + base + i] * b[c * nnodes + base + i];
            s += t;
        }
        return s;
    }
    template <class F> void maskedVisit(F &&f, long margin = 0) const {
        long ny = std::min<long>(nondetY() + margin, nn[1]);
        const long rowLen = ny * nn[2];
        #pragma omp parallel for
        for (long ix = 0; ix < nn[0]; ++ix) {
            const long base = ix * nn[1] * nn[2];
            for (int c = 0; c < 3; ++c)
                for (long i = 0; i < rowLen; ++i) f(c * nnodes + base + i);
        }
    }
    // Fine-level matrix-free K u, node-centric (TPSStencils.hh:431-728):
    // y_n (+/-)= sum over in-grid incident elements E_e K0[3 l_n + c, :] u_e ; rows of
    // detached nodes are skipped (and zeroed when zeroInit).
    void applyK(const double *u, double *y, bool zeroInit, bool negate) const {
        const long ny = nondetY();
        const double *K = K0.data();
        #pragma omp parallel for collapse(2)
        for (long ix = 0; ix < nn[0]; ++ix)
            for (long iy = 0; iy < ny; ++iy)
                for (long iz = 0; iz < nn[2]; ++iz) {
                    const long n = node(ix, iy, iz);
                    double f[3] = {0, 0, 0};
                    for (int i = 1; i >= 0; --i)
                    for (int j = 1; j >= 0; --j)
                    for (int k = 1; k >= 0; --k) {
                        const long ex = ix - i, ey = iy - j, ez = iz - k;
                        if (ex < 0 || ey < 0 || ez < 0 || ex >= ne[0] || ey >= ne[1] || ez >= ne[2]) continue;
                        const double Ee = E[elem(ex, ey, ez)];
                        const int ln = 4 * i + 2 * j + k;
                        double acc[3] = {0, 0, 0};
                        for (int m = 0; m < 8; ++m) {
                            const long nm = node(ex + ((m >> 2) & 1), ey + ((m >> 1) & 1), ez + (m & 1));
                            for (int d = 0; d < 3; ++d) {
                                const double um = u[d * nnodes + nm];
                                for (int c = 0; c < 3; ++c) acc[c] += K[(3 * m + d) * 24 + 3 * ln + c] * um;
                            }
                        }
                        for (int c = 0; c < 3; ++c) f[c] += Ee * acc[c];
                    }
                    for (int c = 0; c < 3; ++c) {
                        double &out = y[c * nnodes + n];
                        const double base = zeroInit ? 0.0 : out;
                        out = negate ? base - f[c] : base + f[c];
                    }
                }
        if (zeroInit && ny < nn[1]) {
            #pragma omp parallel for
            for (long ix = 0; ix < nn[0]; ++ix)
                for (long iy = ny; iy < nn[1]; ++iy)
                    for (long iz = 0; iz < nn[2]; ++iz)
                        for (int c = 0; c < 3; ++c) y[c * nnodes + node(ix, iy, iz)] = 0;
        }
    }
    // buildLoadVector (TensorProductSimulator.hh:1269-1288); integrated Q1 shape function = 1/8.
    void buildLoad(double *f) const {
        std::copy(force.begin(), force.end(), f);
        if (grav[0] == 0 && grav[1] == 0 && grav[2] == 0) return;
        const double vol = h[0] * h[1] * h[2];
        for (long ex = 0; ex < ne[0]; ++ex)
        for (long ey = 0; ey < ne[1]; ++ey) {
            if (elemMasked(ey)) continue;
            for (long ez = 0; ez < ne[2]; ++ez) {
                const double w = 0.125 * rho[elem(ex, ey, ez)] * vol;
                for (int l = 0; l < 8; ++l) {
                    const long nd = node(ex + ((l >> 2) & 1), ey + ((l >> 1) & 1), ez + (l & 1));
                    for (int c = 0; c < 3; ++c) f[c * nnodes + nd] += grav[c] * w;
                }
            }
        }
    }
    void gatherElem(const double *u, long ex, long ey, long ez, double ue[24]) const {
        for (int l = 0; l < 8; ++l) {
            const long nd = node(ex + ((l >> 2) & 1), ey + ((l >> 1) & 1), ez + (l & 1));
            for (int c = 0; c < 3; ++c) ue[3 * l + c] = u[c * nnodes + nd];
        }
    }
    double quadForm(const double ue[24]) const {
        double s = 0;
        for (int a = 0; a < 24; ++a) {
            double r = 0;
            for (int b = 0; b < 24; ++b) r += K0[a * 24 + b] * ue[b];
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
        const bool sw = grav[0] != 0 || grav[1] != 0 || grav[2] != 0;
        const double vol = h[0] * h[1] * h[2];
        #pragma omp parallel for collapse(2)
        for (long ex = 0; ex < ne[0]; ++ex)
            for (long ey = 0; ey < ne[1]; ++ey)
                for (long ez = 0; ez < ne[2]; ++ez) {
                    const long e = elem(ex, ey, ez);
                    if (elemMasked(ey)) { if (!accumulate) g[e] = 0; continue; }
                    double ue[24];
                    gatherElem(u, ex, ey, ez, ue);
                    double val = -0.5 * dEdrho(rho[e]) * quadForm(ue);
                    if (sw) {
                        for (int l = 0; l < 8; ++l)
                            val += 0.125 * (grav[0] * ue[3 * l] + grav[1] * ue[3 * l + 1] + grav[2] * ue[3 * l + 2]) * vol;
                    }
                    if (accumulate) g[e] += val; else g[e] = val;
                }
    }
    // elementEnergyDensity (TensorProductSimulator.hh:1057-1073)
    void energyDensity(const double *u, double *out) const {
        #pragma omp parallel for collapse(2)
        for (long ex = 0; ex < ne[0]; ++ex)
            for (long ey = 0; ey < ne[1]; ++ey)
                for (long ez = 0; ez < ne[2]; ++ez) {
                    double ue[24];
                    gatherElem(u, ex, ey, ez, ue);
                    out[elem(ex, ey, ez)] = 0.5 * E[elem(ex, ey, ez)] * quadForm(ue);
                }
    }
};

// ---------------------------------------------------------------------------------------
// Dense/banded Cholesky of the reduced coarsest operator (replaces CHOLMOD in
// TensorProductSimulator::solve, TensorProductSimulator.hh:1181-1230).
// ---------------------------------------------------------------------------------------
struct BandChol {
    long n = 0, bw = 0;          // half bandwidth (number of sub-diagonals)
    Vec L;                       // row-major band: L[i*(bw+1) + (j - i + bw)], j in [i-bw, i]
    std::vector<long> fullToRed; // -1 for fixed DoFs
    std::vector<long> redToFull;
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
            #pragma omp parallel for if (iend - j > 64)
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

// ---------------------------------------------------------------------------------------
// Multigrid hierarchy + PCG: MultigridSolver.hh (all).
// ---------------------------------------------------------------------------------------
struct MG {
    Sim *fine;
    std::vector<std::unique_ptr<Sim>> coarse;  // levels 1..L
    std::vector<Sim *> sims;
    std::vector<Vec> x, b, r;
    Vec d, Ad;
    double Chat[8][576];  // coarsenedFineK0s (MultigridSolver.hh:116-120)
    double phi[8][8][8];
    bool symmetricGS = true;
    std::unique_ptr<BandChol> chol;
    bool factorUpToDate = false;
    long L;

    MG(Sim *f, long levels) : fine(f), L(levels) {
        sims.push_back(f);
        long ne[3] = {f->ne[0], f->ne[1], f->ne[2]};
        for (long l = 1; l <= levels; ++l) {
            for (int dd = 0; dd < 3; ++dd) {
                if (ne[dd] % 2) throw std::runtime_error("Grid size currently must be divisible by 2^numCoarseningLevels (nonuniform coarsening not yet implemented)");
                ne[dd] /= 2;
            }
            auto s = std::make_unique<Sim>(ne[0], ne[1], ne[2], f->lo, f->hi);
            s->young = f->young; s->poisson = f->poisson; s->updateK0();
            // Dirichlet coarsening (MultigridSolver.hh:58-103): a constrained fine node
            // constrains every coarse node of the containing coarse element that lies on the same
            // vertex/edge/face; along an axis where the fine index is even the coarse index is i/2,
            // where it is odd both neighbours (i-1)/2, (i+1)/2 qualify.
            Sim *fs = sims.back();
            std::vector<uint8_t> cm(s->nnodes, 0);
            for (long n : fs->dnodes) {
                const long iz = n % fs->nn[2], iy = (n / fs->nn[2]) % fs->nn[1], ix = n / (fs->nn[2] * fs->nn[1]);
                const long id[3] = {ix, iy, iz};
                long c0[3], c1[3];
                for (int dd = 0; dd < 3; ++dd) {
                    if (id[dd] % 2 == 0) c0[dd] = c1[dd] = id[dd] / 2;
                    else { c0[dd] = (id[dd] - 1) / 2; c1[dd] = (id[dd] + 1) / 2; }
                }
                for (long a = c0[0]; a <= c1[0]; ++a)
                    for (long bb = c0[1]; bb <= c1[1]; ++bb)
                        for (long c = c0[2]; c <= c1[2]; ++c) cm[s->node(a, bb, c)] |= fs->dmask[n];
            }
            s->setDirichlet(cm.data(), nullptr);
            sims.push_back(s.get());
            coarse.push_back(std::move(s));
        }
        for (auto *s : sims) {
            x.emplace_back(3 * s->nnodes, 0.0);
            b.emplace_back(3 * s->nnodes, 0.0);
            r.emplace_back(3 * s->nnodes, 0.0);
        }
        child_interpolation(phi);
        for (int fi = 0; fi < 8; ++fi) {
            std::fill(Chat[fi], Chat[fi] + 576, 0.0);
            accumulate_coarsened(phi[fi], f->K0.data(), Chat[fi]);
        }
    }
    Sim &sim(long l) { return *sims.at(l); }

    void setMaskLayer(long layer) {  // MultigridSolver.hh:1022-1028
        fine->setMaskHeight(fine->h[1] * layer, true);
        for (long l = 1; l <= L; ++l) sims[l]->setMaskHeight(fine->maskHeight, false);
        factorUpToDate = false;
    }

    // --- Galerkin coarsening (MultigridSolver.hh:744-819, 846-905) -------------------------
    void coarsenedElement(long l, long ex, long ey, long ez, double *out) {
        Sim &c = sim(l);
        std::fill(out, out + 576, 0.0);
        if (c.elemMasked(ey)) return;
        Sim &f = sim(l - 1);
        if (l == 1) {
            for (int fi = 0; fi < 8; ++fi) {
                const double Ef = f.E[f.elem(2 * ex + ((fi >> 2) & 1), 2 * ey + ((fi >> 1) & 1), 2 * ez + (fi & 1))];
                for (int k = 0; k < 576; ++k) out[k] += Ef * Chat[fi][k];
            }
        } else {
            double child[576];
            for (int fi = 0; fi < 8; ++fi) {
                coarsenedElement(l - 1, 2 * ex + ((fi >> 2) & 1), 2 * ey + ((fi >> 1) & 1), 2 * ez + (fi & 1), child);
                accumulate_coarsened(phi[fi], child, out);
            }
        }
        const bool blockLevel = (l >= 2) && (l < L);
        if (blockLevel) {
            // accumulate lower blocks (row node index > col node index) and the diagonal
            for (int i = 0; i < 8; ++i) {
                const long ni = c.node(ex + ((i >> 2) & 1), ey + ((i >> 1) & 1), ez + (i & 1));
                for (int j = 0; j <= i; ++j) {
                    const int off[3] = {((j >> 2) & 1) - ((i >> 2) & 1), ((j >> 1) & 1) - ((i >> 1) & 1), (j & 1) - (i & 1)};
                    const int slot = (off[0] + 1) * 9 + (off[1] + 1) * 3 + (off[2] + 1);
                    double *B = &c.blocks[(ni * 27 + slot) * 9];
                    for (int a = 0; a < 3; ++a)
                        for (int bb = 0; bb < 3; ++bb) B[3 * a + bb] += out[(3 * i + a) * 24 + 3 * j + bb];
                }
            }
        } else if (l == L) {
            std::copy(out, out + 576, &c.Ke[c.elem(ex, ey, ez) * 576]);
        }
    }
    void updateStiffness() {
        if (L == 0) { factorUpToDate = false; return; }
        for (long l = 2; l < L; ++l) sim(l).blocks.assign(sim(l).nnodes * 27 * 9, 0.0);
        Sim &cs = sim(L);
        cs.Ke.assign(cs.nelems * 576, 0.0);
        for (int color = 0; color < 8; ++color) {
            const int o[3] = {(color >> 2) & 1, (color >> 1) & 1, color & 1};
            #pragma omp parallel for collapse(3) schedule(dynamic)
            for (long ex = o[0]; ex < cs.ne[0]; ex += 2)
                for (long ey = o[1]; ey < cs.ne[1]; ey += 2)
                    for (long ez = o[2]; ez < cs.ne[2]; ez += 2) {
                        double tmp[576];
                        coarsenedElement(L, ex, ey, ez, tmp);
                    }
        }
        // mirror lower blocks into the upper ones (BLOCKK full symmetric storage)
        for (long l = 2; l < L; ++l) {
            Sim &c = sim(l);
            #pragma omp parallel for
            for (long n = 0; n < c.nnodes; ++n) {
                const long iz = n % c.nn[2], iy = (n / c.nn[2]) % c.nn[1], ix = n / (c.nn[2] * c.nn[1]);
                for (int slot = 14; slot < 27; ++slot) {
                    const int dx = slot / 9 - 1, dy = (slot / 3) % 3 - 1, dz = slot % 3 - 1;
                    const long jx = ix + dx, jy = iy + dy, jz = iz + dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= c.nn[0] || jy >= c.nn[1] || jz >= c.nn[2]) continue;
                    const long m = c.node(jx, jy, jz);
                    const double *S = &c.blocks[(m * 27 + (26 - slot)) * 9];
                    double *D = &c.blocks[(n * 27 + slot) * 9];
                    for (int a = 0; a < 3; ++a)
                        for (int bb = 0; bb < 3; ++bb) D[3 * a + bb] = S[3 * bb + a];
                }
            }
        }
        factorUpToDate = false;
    }

    // --- level operators (MultigridSolver.hh:471-504) ---------------------------------------
    void applyK(long l, const double *u, double *y, bool zeroInit, bool negate) {
        Sim &s = sim(l);
        if (l == 0) { s.applyK(u, y, zeroInit, negate); return; }
        const long nn = s.nnodes;
        if (zeroInit) std::fill(y, y + 3 * nn, 0.0);
        const double sgn = negate ? -1.0 : 1.0;
        if (l >= 2 && s.blocks.empty() && s.Ke.empty()) updateStiffness();
        if (l == 1) {
            Sim &f = sim(0);
            const long ney = s.unmaskedY();
            for (int color = 0; color < 8; ++color) {
                const int o[3] = {(color >> 2) & 1, (color >> 1) & 1, color & 1};
                #pragma omp parallel for collapse(2)
                for (long ex = o[0]; ex < s.ne[0]; ex += 2)
                    for (long ey = o[1]; ey < ney; ey += 2)
                        for (long ez = o[2]; ez < s.ne[2]; ez += 2) {
                            double ue[24], ke[24] = {0};
                            s.gatherElem(u, ex, ey, ez, ue);
                            for (int fi = 0; fi < 8; ++fi) {
                                const double Ef = f.E[f.elem(2 * ex + ((fi >> 2) & 1), 2 * ey + ((fi >> 1) & 1), 2 * ez + (fi & 1))];
                                for (int a = 0; a < 24; ++a) {
                                    double t = 0;
                                    for (int bb = 0; bb < 24; ++bb) t += Chat[fi][a * 24 + bb] * ue[bb];
                                    ke[a] += Ef * t;
                                }
                            }
                            for (int m = 0; m < 8; ++m) {
                                const long nd = s.node(ex + ((m >> 2) & 1), ey + ((m >> 1) & 1), ez + (m & 1));
                                for (int c = 0; c < 3; ++c) y[c * nn + nd] += sgn * ke[3 * m + c];
                            }
                        }
            }
            return;
        }
        if (!s.blocks.empty()) {
            const long ny = s.nondetY();
            #pragma omp parallel for collapse(2)
            for (long ix = 0; ix < s.nn[0]; ++ix)
                for (long iy = 0; iy < ny; ++iy)
                    for (long iz = 0; iz < s.nn[2]; ++iz) {
                        const long n = s.node(ix, iy, iz);
                        double acc[3] = {0, 0, 0};
                        for (int slot = 0; slot < 27; ++slot) {
                            const long jx = ix + slot / 9 - 1, jy = iy + (slot / 3) % 3 - 1, jz = iz + slot % 3 - 1;
                            if (jx < 0 || jy < 0 || jz < 0 || jx >= s.nn[0] || jy >= s.nn[1] || jz >= s.nn[2]) continue;
                            const long m = s.node(jx, jy, jz);
                            const double *B = &s.blocks[(n * 27 + slot) * 9];
                            for (int a = 0; a < 3; ++a)
                                for (int bb = 0; bb < 3; ++bb) acc[a] += B[3 * a + bb] * u[bb * nn + m];
                        }
                        for (int c = 0; c < 3; ++c) y[c * nn + n] += sgn * acc[c];
                    }
            return;
        }
        // cached per-element matrices (coarsest level): element-coloured scatter
        for (int color = 0; color < 8; ++color) {
            const int o[3] = {(color >> 2) & 1, (color >> 1) & 1, color & 1};
            #pragma omp parallel for collapse(2)
            for (long ex = o[0]; ex < s.ne[0]; ex += 2)
                for (long ey = o[1]; ey < s.ne[1]; ey += 2)
                    for (long ez = o[2]; ez < s.ne[2]; ez += 2) {
                        double ue[24];
                        s.gatherElem(u, ex, ey, ez, ue);
                        const double *K = &s.Ke[s.elem(ex, ey, ez) * 576];
                        for (int a = 0; a < 24; ++a) {
                            const long nd = s.node(ex + (((a / 3) >> 2) & 1), ey + (((a / 3) >> 1) & 1), ez + ((a / 3) & 1));
                            double t = 0;
                            for (int bb = 0; bb < 24; ++bb) t += K[a * 24 + bb] * ue[bb];
                            y[(a % 3) * nn + nd] += sgn * t;
                        }
                    }
        }
    }
    void computeResidual(long l, const double *u, const double *bb, double *res) {  // 527-541
        Sim &s = sim(l);
        s.maskedVisit([&](long i) { res[i] = bb[i]; }, 4);
        applyK(l, u, res, false, true);
        s.zeroDirichlet(res);
    }

    // --- multicolour block Gauss-Seidel (MultigridSolver.hh:347-458) -----------------------
    void smoothNode(long l, long ix, long iy, long iz, double *u, const double *bb, bool forward) {
        Sim &s = sim(l);
        const long n = s.node(ix, iy, iz), nn = s.nnodes;
        const uint8_t dm = s.dmask[n];
        if (dm == 7) return;
        double bms[3] = {bb[n], bb[nn + n], bb[2 * nn + n]};
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (l == 0 || (l == 1 && L > 1)) {
            for (int i = 1; i >= 0; --i)
            for (int j = 1; j >= 0; --j)
            for (int k = 1; k >= 0; --k) {
                const long ex = ix - i, ey = iy - j, ez = iz - k;
                if (ex < 0 || ey < 0 || ez < 0 || ex >= s.ne[0] || ey >= s.ne[1] || ez >= s.ne[2]) continue;
                const int ln = 4 * i + 2 * j + k;
                double ue[24];
                s.gatherElem(u, ex, ey, ez, ue);
                double cols[24][3];  // Ke[:, 3ln:3ln+3]
                if (l == 0) {
                    const double Ee = s.E[s.elem(ex, ey, ez)];
                    for (int a = 0; a < 24; ++a)
                        for (int c = 0; c < 3; ++c) cols[a][c] = Ee * s.K0[a * 24 + 3 * ln + c];
                } else {
                    Sim &f = sim(0);
                    for (int a = 0; a < 24; ++a) for (int c = 0; c < 3; ++c) cols[a][c] = 0;
                    for (int fi = 0; fi < 8; ++fi) {
                        const double Ef = f.E[f.elem(2 * ex + ((fi >> 2) & 1), 2 * ey + ((fi >> 1) & 1), 2 * ez + (fi & 1))];
                        for (int a = 0; a < 24; ++a)
                            for (int c = 0; c < 3; ++c) cols[a][c] += Ef * Chat[fi][a * 24 + 3 * ln + c];
                    }
                }
                for (int c = 0; c < 3; ++c) {
                    double t = 0;
                    for (int a = 0; a < 24; ++a) t += cols[a][c] * ue[a];
                    bms[c] -= t;
                    for (int c2 = 0; c2 < 3; ++c2) M[3 * c + c2] += cols[3 * ln + c][c2];
                }
            }
        } else {
            for (int slot = 0; slot < 27; ++slot) {
                const long jx = ix + slot / 9 - 1, jy = iy + (slot / 3) % 3 - 1, jz = iz + slot % 3 - 1;
                if (jx < 0 || jy < 0 || jz < 0 || jx >= s.nn[0] || jy >= s.nn[1] || jz >= s.nn[2]) continue;
                const long m = s.node(jx, jy, jz);
                const double *B = &s.blocks[(n * 27 + slot) * 9];
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) bms[a] -= B[3 * a + c] * u[c * nn + m];
                if (slot == 13) std::copy(B, B + 9, M);
            }
        }
        if (dm) {
            double ud[3] = {0, 0, 0};
            for (int t = 0; t < 3; ++t) {
                const int i = forward ? t : 2 - t;
                const double free = (dm & (1 << i)) ? 0.0 : 1.0;
                ud[i] = (bms[i] - (M[3 * i] * ud[0] + M[3 * i + 1] * ud[1] + M[3 * i + 2] * ud[2])) * (free / M[4 * i]);
            }
            for (int c = 0; c < 3; ++c) u[c * nn + n] += ud[c];
        } else {
            double Mi[9];
            inv3(M, Mi);
            for (int c = 0; c < 3; ++c) u[c * nn + n] += Mi[3 * c] * bms[0] + Mi[3 * c + 1] * bms[1] + Mi[3 * c + 2] * bms[2];
        }
    }
    void smooth(long l, double *u, const double *bb, bool forward) {
        Sim &s = sim(l);
        const long ny = s.nondetY();
        for (int t = 0; t < 8; ++t) {
            const int color = forward ? t : 7 - t;
            const int o[3] = {(color >> 2) & 1, (color >> 1) & 1, color & 1};
            #pragma omp parallel for collapse(2)
            for (long ix = o[0]; ix < s.nn[0]; ix += 2)
                for (long iy = o[1]; iy < ny; iy += 2)
                    for (long iz = o[2]; iz < s.nn[2]; iz += 2) smoothNode(l, ix, iy, iz, u, bb, forward);
        }
    }

    // --- transfers (MultigridSolver.hh:178-275) ---------------------------------------------
    void restriction(long l, const double *rf, double *rc) {  // fine level l -> l+1
        Sim &f = sim(l), &c = sim(l + 1);
        const long fny = f.nondetY(), cny = c.nondetY();
        #pragma omp parallel for collapse(2)
        for (long ix = 0; ix < c.nn[0]; ++ix)
            for (long iy = 0; iy < cny; ++iy)
                for (long iz = 0; iz < c.nn[2]; ++iz) {
                    double v[3] = {0, 0, 0};
                    for (long ox = -1; ox <= 1; ++ox) {
                        const long fx = 2 * ix + ox;
                        if (fx < 0 || fx >= f.nn[0]) continue;
                        for (long oy = -1; oy <= 1; ++oy) {
                            const long fy = 2 * iy + oy;
                            if (fy < 0 || fy >= fny) continue;
                            for (long oz = -1; oz <= 1; ++oz) {
                                const long fz = 2 * iz + oz;
                                if (fz < 0 || fz >= f.nn[2]) continue;
                                const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0);
                                const long m = f.node(fx, fy, fz);
                                for (int cc = 0; cc < 3; ++cc) v[cc] += w * rf[cc * f.nnodes + m];
                            }
                        }
                    }
                    const long n = c.node(ix, iy, iz);
                    for (int cc = 0; cc < 3; ++cc) rc[cc * c.nnodes + n] = v[cc];
                }
        // zero every detached coarse row (the reference zeroes the first detached layer and
        // never reads the others)
        for (long ix = 0; ix < c.nn[0]; ++ix)
            for (long iy = cny; iy < c.nn[1]; ++iy)
                for (long iz = 0; iz < c.nn[2]; ++iz)
                    for (int cc = 0; cc < 3; ++cc) rc[cc * c.nnodes + c.node(ix, iy, iz)] = 0;
    }
    void interpolate(long l, const double *xc, double *xf, bool accumulate) {  // coarse l+1 -> fine l
        Sim &f = sim(l), &c = sim(l + 1);
        const long fny = accumulate ? f.nondetY() : f.nn[1];
        #pragma omp parallel for collapse(2)
        for (long ix = 0; ix < f.nn[0]; ++ix)
            for (long iy = 0; iy < fny; ++iy)
                for (long iz = 0; iz < f.nn[2]; ++iz) {
                    const long id[3] = {ix, iy, iz};
                    long e[3]; double w[3][2];
                    for (int dd = 0; dd < 3; ++dd) {
                        e[dd] = std::min<long>(id[dd] / 2, c.ne[dd]);
                        const long lf = id[dd] - 2 * e[dd];
                        w[dd][0] = lf ? 0.5 : 1.0;
                        w[dd][1] = lf ? 0.5 : 0.0;
                    }
                    double v[3] = {0, 0, 0};
                    for (int lc = 0; lc < 8; ++lc) {
                        const int a = (lc >> 2) & 1, bb = (lc >> 1) & 1, cc = lc & 1;
                        const double coeff = w[0][a] * w[1][bb] * w[2][cc];
                        if (coeff == 0) continue;
                        const long m = c.node(e[0] + a, e[1] + bb, e[2] + cc);
                        for (int k = 0; k < 3; ++k) v[k] += coeff * xc[k * c.nnodes + m];
                    }
                    const long n = f.node(ix, iy, iz);
                    for (int k = 0; k < 3; ++k) {
                        if (accumulate) xf[k * f.nnodes + n] += v[k];
                        else xf[k * f.nnodes + n] = v[k];
                    }
                }
    }

    // --- coarsest direct solve (TensorProductSimulator.hh:1181-1230) ------------------------
    void factorCoarsest() {
        Sim &s = sim(L);
        const long nn = s.nnodes, nd = 3 * nn;
        auto ch = std::make_unique<BandChol>();
        ch->fullToRed.assign(nd, -1);
        ch->redToFull.clear();
        const long ny = s.nondetY();
        for (long n = 0; n < nn; ++n) {
            const long iy = (n / s.nn[2]) % s.nn[1];
            for (int c = 0; c < 3; ++c) {
                const bool fixed = (s.dmask[n] & (1 << c)) || iy >= ny;
                if (!fixed) { ch->fullToRed[3 * n + c] = (long)ch->redToFull.size(); ch->redToFull.push_back(3 * n + c); }
            }
        }
        if (L == 0) {
            for (long n : s.dnodes)
                for (int c = 0; c < 3; ++c)
                    if ((s.dmask[n] & (1 << c)) && s.dval[c * nn + n] != 0)
                        throw std::runtime_error("Nonzero Dirichlet constraints currently unsupported");
        }
        ch->n = (long)ch->redToFull.size();
        // half bandwidth in reduced indices: bounded by the full-DoF bandwidth
        const long fullBw = 3 * (s.nn[1] * s.nn[2] + s.nn[2] + 1) + 2;
        ch->bw = std::min<long>(fullBw, std::max<long>(ch->n - 1, 0));
        ch->L.assign(ch->n * (ch->bw + 1), 0.0);
        for (long ex = 0; ex < s.ne[0]; ++ex)
            for (long ey = 0; ey < s.ne[1]; ++ey)
                for (long ez = 0; ez < s.ne[2]; ++ez) {
                    double Kloc[576];
                    const long e = s.elem(ex, ey, ez);
                    if (L == 0) { for (int k = 0; k < 576; ++k) Kloc[k] = s.E[e] * s.K0[k]; }
                    else std::copy(&s.Ke[e * 576], &s.Ke[e * 576] + 576, Kloc);
                    long dof[24];
                    for (int a = 0; a < 24; ++a) {
                        const int ln = a / 3;
                        dof[a] = ch->fullToRed[3 * s.node(ex + ((ln >> 2) & 1), ey + ((ln >> 1) & 1), ez + (ln & 1)) + a % 3];
                    }
                    for (int a = 0; a < 24; ++a) {
                        if (dof[a] < 0) continue;
                        for (int bb = 0; bb < 24; ++bb) {
                            if (dof[bb] < 0 || dof[bb] > dof[a]) continue;
                            ch->at(dof[a], dof[bb]) += Kloc[a * 24 + bb];
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
        for (long i = 0; i < chol->n; ++i) { const long f = chol->redToFull[i]; tmp[i] = bb[(f % 3) * nn + f / 3]; }
        chol->solve(tmp.data());
        std::fill(xx, xx + 3 * nn, 0.0);
        for (long i = 0; i < chol->n; ++i) { const long f = chol->redToFull[i]; xx[(f % 3) * nn + f / 3] = tmp[i]; }
    }

    // --- cycles (MultigridSolver.hh:546-658) --------------------------------------------------
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
        // the hierarchy must have been built (updateStiffness), as before the reference's applyPreconditionerInv
        if (L > 0 && nu > 0 && sim(L).Ke.empty()) throw std::runtime_error("precondition before update_stiffness");
        Sim &s = sim(0);
        std::fill(x[0].begin(), x[0].end(), 0.0);
        if (b[0].data() != res) std::copy(res, res + 3 * s.nnodes, b[0].begin());
        if (nu == 0) { std::copy(res, res + 3 * s.nnodes, x[0].begin()); return; }
        for (long i = 0; i < mgIter; ++i) {
            if (fmg && i == 0) fullMultigrid(0, nu, true);
            else vcycle(0, nu, true);
        }
    }

    // preconditionedConjugateGradient (MultigridSolver.hh:1046-1152)
    long pcg(double *xo, const double *bb, long maxIter, double tol, long mgIter, long nu, bool fmg,
             bool dirichletSatisfied, double *hist, long histLen) {
        Sim &fs = sim(0);
        const long nd = 3 * fs.nnodes;
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
        d.assign(nd, 0.0); Ad.assign(nd, 0.0);
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

// ---------------------------------------------------------------------------------------
// Filters (TopologyOptimizationFilter.hh:189-245, 285-400)
// ---------------------------------------------------------------------------------------
long reflect_index(long i, long s) {  // TopologyOptimizationFilter.hh:343-349
    while (i < 0 || i >= s) {
        if (i >= s) i = 2 * s - i - 1;
        if (i < 0) i = -i - 1;
    }
    return i;
}

// Normalised reflect-BC convolution; rmin plays the role of (radius + 1) of the reference
// (its Linear kernel is (r+1) - |d|, clipped to > 0, over |d_i| <= r).  `type` 0 = Const, 1 = Linear.
void smoothing_filter(const long dims[3], double rmin, int type, const double *in, double *out) {
    const long R = (long)std::ceil(rmin) - 1;
    std::vector<std::array<long, 3>> offs;
    std::vector<double> ws;
    double W = 0;
    for (long a = -R; a <= R; ++a)
        for (long b = -R; b <= R; ++b)
            for (long c = -R; c <= R; ++c) {
                const double w = type == 1 ? rmin - std::sqrt(double(a * a + b * b + c * c)) : 1.0;
                if (w <= 0) continue;
                offs.push_back({a, b, c});
                ws.push_back(w);
                W += w;
            }
    // flat offsets of the taps for voxels whose whole stencil lies inside the grid (no reflection):
    // the same terms summed in the same order as the general path
    std::vector<long> flat(offs.size());
    for (size_t t = 0; t < offs.size(); ++t) flat[t] = (offs[t][0] * dims[1] + offs[t][1]) * dims[2] + offs[t][2];
    #pragma omp parallel for collapse(2)
    for (long i = 0; i < dims[0]; ++i)
        for (long j = 0; j < dims[1]; ++j) {
            const bool rowInside = i >= R && i + R < dims[0] && j >= R && j + R < dims[1];
            const long row = (i * dims[1] + j) * dims[2];
            // interior of the row: tap by tap over contiguous k (each output still sums its taps in order)
            const long k0 = R, k1 = dims[2] - R;
            if (rowInside && k1 > k0) {
                std::vector<double> acc(dims[2], 0.0);
                for (size_t t = 0; t < offs.size(); ++t) {
                    const double w = ws[t];
                    const double *src = in + row + flat[t];
                    for (long k = k0; k < k1; ++k) acc[k] += w * src[k];
                }
                for (long k = k0; k < k1; ++k) out[row + k] = acc[k] / W;
            }
            for (long k = 0; k < dims[2]; ++k) {
                const long o = row + k;
                double s = 0;
                if (rowInside && k >= k0 && k < k1) {
                    continue;
                } else {
                    for (size_t t = 0; t < offs.size(); ++t) {
                        const long ii = reflect_index(i + offs[t][0], dims[0]);
                        const long jj = reflect_index(j + offs[t][1], dims[1]);
                        const long kk = reflect_index(k + offs[t][2], dims[2]);
                        s += ws[t] * in[(ii * dims[1] + jj) * dims[2] + kk];
                    }
                }
                out[o] = s / W;
            }
        }
}

}  // namespace

// =========================================================================================
// C ABI for the Python test harness (ctypes).
// =========================================================================================
#define ORC_TRY(...) try { __VA_ARGS__; return 0; } catch (std::exception &e) { g_err = e.what(); return -1; }

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }

void orc_element_stiffness(double E, double nu, const double *h, double *out) { element_stiffness(E, nu, h, out); }

void *orc_sim_create(long nx, long ny, long nz, const double *lo, const double *hi) {
    try { return new Sim(nx, ny, nz, lo, hi); } catch (std::exception &e) { g_err = e.what(); return nullptr; }
}
void orc_sim_destroy(void *s) { delete (Sim *)s; }
int orc_sim_set_etensor(void *p, double E, double nu) { ORC_TRY(auto *s = (Sim *)p; s->young = E; s->poisson = nu; s->updateK0()) }
int orc_sim_set_law(void *p, int law, double E0, double Emin, double gamma, double q) {
    ORC_TRY(auto *s = (Sim *)p; s->law = law; s->E0 = E0; s->Emin = Emin; s->gamma = gamma; s->q = q; s->updateE())
}
int orc_sim_set_densities(void *p, const double *rho) { ORC_TRY(auto *s = (Sim *)p; std::copy(rho, rho + s->nelems, s->rho.begin()); s->updateE()) }
int orc_sim_get_E(void *p, double *out) { ORC_TRY(auto *s = (Sim *)p; std::copy(s->E.begin(), s->E.end(), out)) }
int orc_sim_get_K0(void *p, double *out) { ORC_TRY(auto *s = (Sim *)p; std::copy(s->K0.begin(), s->K0.end(), out)) }
int orc_sim_set_dirichlet(void *p, const uint8_t *mask, const double *vals) { ORC_TRY(((Sim *)p)->setDirichlet(mask, vals)) }
int orc_sim_set_forces(void *p, const double *f) { ORC_TRY(auto *s = (Sim *)p; std::copy(f, f + 3 * s->nnodes, s->force.begin())) }
int orc_sim_set_gravity(void *p, const double *g) { ORC_TRY(auto *s = (Sim *)p; for (int c = 0; c < 3; ++c) s->grav[c] = g[c]) }
int orc_sim_apply_K(void *p, const double *u, double *y) { ORC_TRY(((Sim *)p)->applyK(u, y, true, false)) }
int orc_sim_build_load(void *p, double *f) { ORC_TRY(((Sim *)p)->buildLoad(f)) }
int orc_sim_compliance_gradient(void *p, const double *u, double *g) { ORC_TRY(((Sim *)p)->complianceGradient(u, g, false)) }
int orc_sim_energy_density(void *p, const double *u, double *out) { ORC_TRY(((Sim *)p)->energyDensity(u, out)) }

void *orc_mg_create(void *sim, long levels) {
    try { return new MG((Sim *)sim, levels); } catch (std::exception &e) { g_err = e.what(); return nullptr; }
}
void orc_mg_destroy(void *m) { delete (MG *)m; }
int orc_mg_set_symmetric(void *m, int s) { ORC_TRY(((MG *)m)->symmetricGS = s != 0) }
int orc_mg_set_mask_layer(void *m, long layer) { ORC_TRY(((MG *)m)->setMaskLayer(layer)) }
int orc_mg_update_stiffness(void *m) { ORC_TRY(((MG *)m)->updateStiffness()) }
int orc_mg_num_nodes(void *m, long l, long *out) { ORC_TRY(*out = ((MG *)m)->sim(l).nnodes) }
int orc_mg_dirichlet_mask(void *m, long l, uint8_t *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.dmask.begin(), s.dmask.end(), out)) }
int orc_mg_apply_K(void *m, long l, const double *u, double *y) { ORC_TRY(((MG *)m)->applyK(l, u, y, true, false)) }
int orc_mg_residual(void *m, long l, const double *u, const double *b, double *r) {
    ORC_TRY(auto *mg = (MG *)m; std::fill(r, r + 3 * mg->sim(l).nnodes, 0.0); mg->computeResidual(l, u, b, r))
}
int orc_mg_smooth(void *m, long l, double *u, const double *b, int forward) { ORC_TRY(((MG *)m)->smooth(l, u, b, forward != 0)) }
int orc_mg_restrict(void *m, long l, const double *rf, double *rc) { ORC_TRY(((MG *)m)->restriction(l, rf, rc)) }
int orc_mg_interpolate(void *m, long l, const double *xc, double *xf, int accumulate) { ORC_TRY(((MG *)m)->interpolate(l, xc, xf, accumulate != 0)) }
int orc_mg_coarse_solve(void *m, const double *b, double *x) { ORC_TRY(((MG *)m)->coarseSolve(b, x)) }
int orc_mg_blocks(void *m, long l, double *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.blocks.begin(), s.blocks.end(), out)) }
int orc_mg_element_matrices(void *m, long l, double *out) { ORC_TRY(auto &s = ((MG *)m)->sim(l); std::copy(s.Ke.begin(), s.Ke.end(), out)) }
int orc_mg_chat(void *m, double *out) { ORC_TRY(auto *mg = (MG *)m; for (int fi = 0; fi < 8; ++fi) std::copy(mg->Chat[fi], mg->Chat[fi] + 576, out + 576 * fi)) }
int orc_mg_precondition(void *m, const double *r, double *s, long mgIter, long nu, int fmg) {
    ORC_TRY(auto *mg = (MG *)m; mg->precondition(r, mgIter, nu, fmg != 0); std::copy(mg->x[0].begin(), mg->x[0].end(), s))
}
int orc_mg_pcg(void *m, double *x, const double *b, long maxIter, double tol, long mgIter, long nu, int fmg,
               int dirichletSatisfied, double *hist, long histLen, long *iters) {
    ORC_TRY(*iters = ((MG *)m)->pcg(x, b, maxIter, tol, mgIter, nu, fmg != 0, dirichletSatisfied != 0, hist, histLen))
}

int orc_smoothing_filter(const long *dims, double rmin, int type, const double *in, double *out) {
    ORC_TRY(smoothing_filter(dims, rmin, type, in, out))
}
// ProjectionFilter::apply / backprop (TopologyOptimizationFilter.hh:199-225)
int orc_projection(double beta, long n, const double *in, double *out) {
    ORC_TRY(const double t = std::tanh(0.5 * beta);
            for (long i = 0; i < n; ++i) out[i] = (t + std::tanh(beta * (in[i] - 0.5))) / (2 * t))
}
int orc_projection_backprop(double beta, long n, const double *g, const double *vars, double *out) {
    ORC_TRY(const double scale = 1.0 / (2 * std::tanh(0.5 * beta) / beta);
            for (long i = 0; i < n; ++i) { const double th = std::tanh(beta * (vars[i] - 0.5)); out[i] = g[i] * (1.0 - th * th) * scale; })
}
// OCOptimizer::step's elementwise update (OptimalityCriterion.hh:64-83)
int orc_oc_update(long n, const double *x0, const double *dJ, const double *dc, double lambda, double m, double p, double *out) {
    ORC_TRY(for (long i = 0; i < n; ++i) {
        double r = x0[i] * std::pow(dJ[i] / (dc[i] * lambda), p);
        r = std::min(std::max(std::min(std::max(r, x0[i] - m), x0[i] + m), 0.0), 1.0);
        if (!std::isfinite(r)) r = x0[i];
        out[i] = r;
    })
}

// LayerByLayerEvaluator::run with the "zero" initial guess (LayerByLayer.hh:223-300):
// returns J = 0.5 * sum_l f_l.u_l / #layers and grad = sum_l g_l / #layers.
int orc_lbl_run_zero(void *m, long layerIncrement, long maxIter, double tol, long mgIter, long nu, int fmg,
                     double *objective, double *grad, double *perLayerCompliance, long *perLayerIters) {
    ORC_TRY({
        auto *mg = (MG *)m;
        Sim &s = mg->sim(0);
        const long numLayers = s.ne[1];
        std::fill(grad, grad + s.nelems, 0.0);
        double total = 0;
        long count = 0;
        Vec f(3 * s.nnodes), u(3 * s.nnodes);
        mg->setMaskLayer(numLayers);
        for (long l = numLayers; l > 0; l -= std::min(layerIncrement, l)) {
            if (l < numLayers) {
                // decrementFabricationMaskHeightByLayer (TensorProductSimulator.hh:311-324)
                s.firstMasked -= layerIncrement; s.firstDetached = s.firstMasked + 1;
                s.maskHeight -= layerIncrement * s.h[1];
                for (long e = 0; e < s.nelems; ++e) {
                    const long ey = (e / s.ne[2]) % s.ne[1];
                    if (ey >= s.firstMasked && ey < s.firstMasked + layerIncrement) s.E[e] = 0.0;
                }
                for (long lv = 1; lv <= mg->L; ++lv) mg->sim(lv).setMaskHeight(s.maskHeight, false);
            }
            s.buildLoad(f.data());
            std::fill(u.begin(), u.end(), 0.0);
            long its = mg->pcg(u.data(), f.data(), maxIter, tol, mgIter, nu, fmg != 0, true, nullptr, 0);
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
