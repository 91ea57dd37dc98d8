// pyoptimizer.cpp -- pybind11 module `pyOptimizer`: the Method of Moving Asymptotes of the reference
// (MethodOfMovingAsymptotes.hh, binding python_bindings/Optimizer.cc:11-22), re-implemented as host C++.
//
// MMA updates the design on the host between two objective/gradient evaluations that the caller supplies as
// Python callbacks (the layer-by-layer drivers evaluate them on the device through pyVoxelFEM and hand back
// numpy arrays). Its work is O((m+1) n) per subproblem Newton step; the per-variable loops run in parallel
// (OpenMP) and every reduction is summed over fixed 8192-element blocks in block order, so results do not
// depend on the thread count.
//
// Problem (MethodOfMovingAsymptotes.hh:22-27):
//   min_x f_0(x) + a_0 z + sum_i (c_i y_i + d_i y_i^2 / 2)
//   s.t.  f_i(x) - a_i z - y_i <= 0 (i = 1..m),  xmin <= x <= xmax,  y >= 0,  z >= 0
// with a_0 = 1, a = 0, c = 1000, d = 1 (34-36).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <functional>
#include <limits>
#include <stdexcept>
#include <vector>

namespace py = pybind11;

namespace {

using Vec = std::vector<double>;
constexpr size_t BLK = 8192;

template <class F>
double bsum(size_t n, F &&f) {
    const size_t nb = (n + BLK - 1) / BLK;
    Vec part(nb);
#pragma omp parallel for schedule(static)
    for (long b = 0; b < (long)nb; ++b) {
        double s = 0;
        const size_t e = std::min(n, (size_t)(b + 1) * BLK);
        for (size_t j = (size_t)b * BLK; j < e; ++j) s += f(j);
        part[b] = s;
    }
    double s = 0;
    for (double v : part) s += v;
    return s;
}

template <class F, class R>
double breduce(size_t n, double init, F &&f, R &&r) {
    const size_t nb = (n + BLK - 1) / BLK;
    Vec part(nb, init);
#pragma omp parallel for schedule(static)
    for (long b = 0; b < (long)nb; ++b) {
        double s = init;
        const size_t e = std::min(n, (size_t)(b + 1) * BLK);
        for (size_t j = (size_t)b * BLK; j < e; ++j) s = r(s, f(j));
        part[b] = s;
    }
    double s = init;
    for (double v : part) s = r(s, v);
    return s;
}

template <class F>
void pfor(size_t n, F &&f) {
#pragma omp parallel for schedule(static)
    for (long j = 0; j < (long)n; ++j) f((size_t)j);
}

// dense solve with partial pivoting of the (m+1) x (m+1) system of the subproblem's Newton step (the
// reference uses a column-pivoted QR; both solve the same nonsingular system)
Vec dense_solve(std::vector<Vec> A, Vec b) {
    const size_t k = b.size();
    for (size_t c = 0; c < k; ++c) {
        size_t piv = c;
        for (size_t r = c + 1; r < k; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        if (A[piv][c] == 0) throw std::runtime_error("MMA: singular subproblem system");
        std::swap(A[c], A[piv]);
        std::swap(b[c], b[piv]);
        for (size_t r = c + 1; r < k; ++r) {
            const double f = A[r][c] / A[c][c];
            if (f == 0) continue;
            for (size_t q = c; q < k; ++q) A[r][q] -= f * A[c][q];
            b[r] -= f * b[c];
        }
    }
    Vec x(k);
    for (size_t c = k; c-- > 0;) {
        double s = b[c];
        for (size_t q = c + 1; q < k; ++q) s -= A[c][q] * x[q];
        x[c] = s / A[c][c];
    }
    return x;
}

class MMA {
public:
    using F = std::function<py::array_t<double>(py::array_t<double>)>;

    MMA(int numVars, int numConstr, const Vec &xmin, const Vec &xmax, py::function f, py::function df)
        : n((size_t)numVars), m((size_t)numConstr), xmin(xmin), xmax(xmax), f(std::move(f)), df(std::move(df)) {
        if (numVars <= 0 || numConstr < 0) throw std::runtime_error("MMA: need numVars > 0 and numConstr >= 0");
        if (xmin.size() != n || xmax.size() != n) throw std::runtime_error("MMA: xmin/xmax size mismatch");
        xdiff.resize(n);
        for (size_t j = 0; j < n; ++j) xdiff[j] = xmax[j] - xmin[j];
        a.assign(m, 0.0);
        c.assign(m, 1000.0);
        d.assign(m, 1.0);
    }

    void enableGCMMA(bool on) { gcmma = on; }
    void setInitialVar(const Vec &x0) {
        if (x0.size() != n) throw std::runtime_error("MMA: initial variable size mismatch");
        push(x0);
    }
    Vec getOptimalVar() const {
        if (hist.empty()) throw std::runtime_error("Must specify an initial value");
        return hist.front();
    }

    // step (MethodOfMovingAsymptotes.hh:63-163)
    void step() {
        if (hist.empty()) throw std::runtime_error("Must specify an initial value");
        ++outer;
        const Vec &x = hist[0];
        if (l.size() != n) { l.assign(n, 0.0); u.assign(n, 0.0); }
        if (outer <= 2) {
            pfor(n, [&](size_t j) {
                l[j] = x[j] - ASYINIT * xdiff[j];
                u[j] = x[j] + ASYINIT * xdiff[j];
            });
        } else {
            const Vec &x1 = hist[1], &x2 = hist[2];
            pfor(n, [&](size_t j) {
                const double zzz = (x[j] - x1[j]) * (x1[j] - x2[j]);
                const double g = zzz > 0 ? 1.2 : 0.7;
                l[j] = std::max(std::min(x[j] - g * (x1[j] - l[j]), x[j] - 0.01 * xdiff[j]), x[j] - 10 * xdiff[j]);
                u[j] = std::min(std::max(x[j] + g * (u[j] - x1[j]), x[j] + 0.01 * xdiff[j]), x[j] + 10 * xdiff[j]);
            });
        }
        alpha.resize(n);
        beta.resize(n);
        pfor(n, [&](size_t j) {
            alpha[j] = std::max(std::max(xmin[j], l[j] + ALBEFA * (x[j] - l[j])), x[j] - MOVE * xdiff[j]);
            beta[j] = std::min(std::min(xmax[j], u[j] - ALBEFA * (u[j] - x[j])), x[j] + MOVE * xdiff[j]);
        });
        fcur = call_f(x);
        call_df(x);  // fills dplus / dminus
        Vec umx(n), xml(n);
        pfor(n, [&](size_t j) { umx[j] = u[j] - x[j]; xml[j] = x[j] - l[j]; });
        if (!gcmma) {
            Vec rho(m + 1, RAA0);
            make_pq(umx, xml, rho);
            const Vec g = g_all(umx, xml);
            r.resize(m + 1);
            for (size_t i = 0; i <= m; ++i) r[i] = fcur[i] - g[i];
            push(subsolve());
        } else {
            Vec rho(m + 1), xin, diff;
            for (int inner = 0;; ++inner) {
                if (inner == 0) {
                    for (size_t i = 0; i <= m; ++i) {
                        const double *dp = &dplus[i * n], *dm = &dminus[i * n];
                        rho[i] = std::max(bsum(n, [&](size_t j) { return 0.1 / (double)n * (dp[j] + dm[j]) * xdiff[j]; }),
                                          1e-6);
                    }
                } else {
                    // next_rho (MethodOfMovingAsymptotes.hh:143-163), distances at the last inner solution
                    const double dd = bsum(n, [&](size_t j) {
                        const double t = xin[j] - x[j];
                        return (u[j] - l[j]) * t * t / ((u[j] - xin[j]) * (xin[j] - l[j]) * xdiff[j]);
                    });
                    for (size_t i = 0; i <= m; ++i) {
                        const double delta = diff[i] / dd;
                        rho[i] = delta < 0 ? rho[i] : std::min(1.1 * (rho[i] + delta), 10 * rho[i]);
                    }
                }
                make_pq(umx, xml, rho);
                const Vec g = g_all(umx, xml);
                r.resize(m + 1);
                for (size_t i = 0; i <= m; ++i) r[i] = fcur[i] - g[i];
                xin = subsolve();
                // isFeasible (212-215): f(x_inner) below its approximation in every row
                const Vec fin = call_f(xin);
                Vec ui(n), xi(n);
                pfor(n, [&](size_t j) { ui[j] = u[j] - xin[j]; xi[j] = xin[j] - l[j]; });
                const Vec gi = g_all(ui, xi);
                diff.resize(m + 1);
                double mx = -std::numeric_limits<double>::infinity();
                for (size_t i = 0; i <= m; ++i) {
                    diff[i] = fin[i] - (gi[i] + r[i]);
                    mx = std::max(mx, diff[i]);
                }
                if (mx < 0) break;
            }
            push(xin);
        }
    }

private:
    static constexpr double RAA0 = 1e-5, ALBEFA = 0.1, MOVE = 0.5, ASYINIT = 0.5;
    const size_t n, m;
    const Vec xmin, xmax;
    Vec xdiff;
    py::function f, df;
    Vec a, c, d;
    const double a0 = 1.0;
    bool gcmma = false;
    int outer = 0;
    std::deque<Vec> hist;  // x^k, x^(k-1), x^(k-2)
    Vec l, u, alpha, beta, fcur, r;
    Vec dplus, dminus, p, q;  // (m+1) x n, row-major

    void push(const Vec &x) {
        hist.push_front(x);
        if (hist.size() > 3) hist.pop_back();
    }
    Vec call_f(const Vec &x) {
        py::array_t<double> xa((py::ssize_t)n);
        std::copy(x.begin(), x.end(), xa.mutable_data());
        auto res = py::array_t<double, py::array::forcecast | py::array::c_style>(f(xa));
        if ((size_t)res.size() != m + 1) throw std::runtime_error("MMA: f must return numConstr + 1 values");
        return Vec(res.data(), res.data() + m + 1);
    }
    void call_df(const Vec &x) {
        py::array_t<double> xa((py::ssize_t)n);
        std::copy(x.begin(), x.end(), xa.mutable_data());
        auto res = py::array_t<double, py::array::forcecast | py::array::c_style>(df(xa));
        if ((size_t)res.size() != (m + 1) * n) throw std::runtime_error("MMA: df_dx must return (numConstr + 1) x numVars values");
        const double *g = res.data();
        dplus.resize((m + 1) * n);
        dminus.resize((m + 1) * n);
        pfor((m + 1) * n, [&](size_t k) {
            dplus[k] = std::max(g[k], 0.0);
            dminus[k] = std::max(-g[k], 0.0);
        });
    }
    void make_pq(const Vec &umx, const Vec &xml, const Vec &rho) {
        p.resize((m + 1) * n);
        q.resize((m + 1) * n);
        for (size_t i = 0; i <= m; ++i)
            pfor(n, [&](size_t j) {
                const size_t k = i * n + j;
                const double rd = rho[i] / xdiff[j];
                p[k] = umx[j] * umx[j] * (1.001 * dplus[k] + 0.001 * dminus[k] + rd);
                q[k] = xml[j] * xml[j] * (0.001 * dplus[k] + 1.001 * dminus[k] + rd);
            });
    }
    // g_i = sum_j p_ij / (u_j - x_j) + q_ij / (x_j - l_j), i = 0..m (sub_g_eval, 166-189)
    Vec g_all(const Vec &umx, const Vec &xml) const {
        Vec g(m + 1);
        for (size_t i = 0; i <= m; ++i) {
            const double *pi = &p[i * n], *qi = &q[i * n];
            g[i] = bsum(n, [&](size_t j) { return pi[j] / umx[j] + qi[j] / xml[j]; });
        }
        return g;
    }

    // ---- the subproblem's primal-dual interior-point Newton method (Subproblem, 240-465)
    struct State {
        Vec x, xi, eta, y, lam, s, mu;
        double z = 1, zeta = 1;
    };
    struct Derived {
        Vec umx, xml, plam, qlam, dpsi, gvec;
    };
    Derived derive(const State &st) const {
        Derived D;
        D.umx.resize(n); D.xml.resize(n); D.plam.resize(n); D.qlam.resize(n); D.dpsi.resize(n);
        pfor(n, [&](size_t j) {
            D.umx[j] = u[j] - st.x[j];
            D.xml[j] = st.x[j] - l[j];
            double pl = p[j], ql = q[j];
            for (size_t i = 0; i < m; ++i) {
                pl += p[(i + 1) * n + j] * st.lam[i];
                ql += q[(i + 1) * n + j] * st.lam[i];
            }
            D.plam[j] = pl;
            D.qlam[j] = ql;
            D.dpsi[j] = pl / (D.umx[j] * D.umx[j]) - ql / (D.xml[j] * D.xml[j]);
        });
        D.gvec.resize(m);
        for (size_t i = 0; i < m; ++i) {
            const double *pi = &p[(i + 1) * n], *qi = &q[(i + 1) * n];
            D.gvec[i] = bsum(n, [&](size_t j) { return pi[j] / D.umx[j] + qi[j] / D.xml[j]; });
        }
        return D;
    }
    // the equations 5.9 (a)-(i) of the subproblem's KKT system: squared norm (squared_residual) or max norm
    // (KKT_inf_norm)
    template <bool SQ>
    double kkt(const State &st, const Derived &D, double eps) const {
        auto acc = [](double s, double v) { return SQ ? s + v * v : std::max(s, std::fabs(v)); };
        double res = breduce(n, 0.0, [&](size_t j) {
            const double ea = D.dpsi[j] - st.xi[j] + st.eta[j];
            const double ee = st.xi[j] * (st.x[j] - alpha[j]) - eps;
            const double ef = st.eta[j] * (beta[j] - st.x[j]) - eps;
            return SQ ? ea * ea + ee * ee + ef * ef : std::max(std::fabs(ea), std::max(std::fabs(ee), std::fabs(ef)));
        }, [](double s, double v) { return SQ ? s + v : std::max(s, v); });
        double lama = 0;
        for (size_t i = 0; i < m; ++i) lama += st.lam[i] * a[i];
        for (size_t i = 0; i < m; ++i) res = acc(res, c[i] + d[i] * st.y[i] - st.lam[i] - st.mu[i]);
        res = acc(res, a0 - st.zeta - lama);
        for (size_t i = 0; i < m; ++i) res = acc(res, D.gvec[i] - a[i] * st.z - st.y[i] + st.s[i] + r[i + 1]);
        for (size_t i = 0; i < m; ++i) res = acc(res, st.mu[i] * st.y[i] - eps);
        res = acc(res, st.zeta * st.z - eps);
        for (size_t i = 0; i < m; ++i) res = acc(res, st.lam[i] * st.s[i] - eps);
        return res;
    }
    static void axpy(State &st, double t, const State &dl, size_t n, size_t m) {
        pfor(n, [&](size_t j) {
            st.x[j] += t * dl.x[j];
            st.xi[j] += t * dl.xi[j];
            st.eta[j] += t * dl.eta[j];
        });
        for (size_t i = 0; i < m; ++i) {
            st.y[i] += t * dl.y[i];
            st.lam[i] += t * dl.lam[i];
            st.mu[i] += t * dl.mu[i];
            st.s[i] += t * dl.s[i];
        }
        st.z += t * dl.z;
        st.zeta += t * dl.zeta;
    }

    Vec subsolve() {
        State st;
        st.x.resize(n); st.xi.resize(n); st.eta.resize(n);
        pfor(n, [&](size_t j) {
            st.x[j] = 0.5 * (alpha[j] + beta[j]);
            st.xi[j] = std::max(1.0 / (st.x[j] - alpha[j]), 1.0);
            st.eta[j] = std::max(1.0 / (beta[j] - st.x[j]), 1.0);
        });
        st.y.assign(m, 1.0);
        st.lam.assign(m, 1.0);
        st.s.assign(m, 1.0);
        st.mu.resize(m);
        for (size_t i = 0; i < m; ++i) st.mu[i] = std::max(c[i] / 2, 1.0);
        Derived D = derive(st);
        State dl;
        dl.x.resize(n); dl.xi.resize(n); dl.eta.resize(n);
        dl.y.resize(m); dl.lam.resize(m); dl.mu.resize(m); dl.s.resize(m);
        Vec Dx(n), dxv(n), G(m * n);
        double eps = 1;
        while (eps > 1e-7) {
            double res_old = 0;
            for (int it = 0; it < 10; ++it) {
                // Newton direction (solve_for_newton_direction, 311-374)
                pfor(n, [&](size_t j) {
                    const double xa = st.x[j] - alpha[j], bx = beta[j] - st.x[j];
                    dxv[j] = D.dpsi[j] - eps / xa + eps / bx;
                    Dx[j] = 2 * D.plam[j] / (D.umx[j] * D.umx[j] * D.umx[j]) +
                            2 * D.qlam[j] / (D.xml[j] * D.xml[j] * D.xml[j]) + st.xi[j] / xa + st.eta[j] / bx;
                    for (size_t i = 0; i < m; ++i)
                        G[i * n + j] = p[(i + 1) * n + j] / (D.umx[j] * D.umx[j]) - q[(i + 1) * n + j] / (D.xml[j] * D.xml[j]);
                });
                Vec Dy(m), dy(m);
                for (size_t i = 0; i < m; ++i) {
                    Dy[i] = d[i] + st.mu[i] / st.y[i];
                    dy[i] = c[i] + d[i] * st.y[i] - st.lam[i] - eps / st.y[i];
                }
                std::vector<Vec> M(m + 1, Vec(m + 1, 0.0));
                for (size_t ci = 0; ci < m; ++ci)
                    for (size_t cj = ci; cj < m; ++cj) {
                        const double *gi = &G[ci * n], *gj = &G[cj * n];
                        M[ci][cj] = bsum(n, [&](size_t j) { return gi[j] * (gj[j] / Dx[j]); });
                    }
                for (size_t i = 0; i < m; ++i) {
                    M[i][i] += st.s[i] / st.lam[i] + 1.0 / Dy[i];
                    M[i][m] = a[i];
                }
                M[m][m] = -st.zeta / st.z;
                for (size_t i = 0; i <= m; ++i)
                    for (size_t k = 0; k < i; ++k) M[i][k] = M[k][i];
                Vec rhs(m + 1);
                double lama = 0;
                for (size_t i = 0; i < m; ++i) lama += st.lam[i] * a[i];
                for (size_t i = 0; i < m; ++i) {
                    const double *gi = &G[i * n];
                    rhs[i] = D.gvec[i] - a[i] * st.z - st.y[i] + r[i + 1] + eps / st.lam[i] + dy[i] / Dy[i] -
                             bsum(n, [&](size_t j) { return gi[j] * (dxv[j] / Dx[j]); });
                }
                rhs[m] = a0 - lama - eps / st.z;
                const Vec sol = dense_solve(M, rhs);
                for (size_t i = 0; i < m; ++i) dl.lam[i] = sol[i];
                dl.z = sol[m];
                pfor(n, [&](size_t j) {
                    double gl = 0;
                    for (size_t i = 0; i < m; ++i) gl += G[i * n + j] * dl.lam[i];
                    const double xa = st.x[j] - alpha[j], bx = beta[j] - st.x[j];
                    dl.x[j] = -(dxv[j] + gl) / Dx[j];
                    dl.xi[j] = -st.xi[j] + eps / xa - st.xi[j] * dl.x[j] / xa;
                    dl.eta[j] = -st.eta[j] + eps / bx + st.eta[j] * dl.x[j] / bx;
                });
                for (size_t i = 0; i < m; ++i) {
                    dl.y[i] = dl.lam[i] / Dy[i] - dy[i] / Dy[i];
                    dl.mu[i] = (eps - st.mu[i] * dl.y[i]) / st.y[i] - st.mu[i];
                    dl.s[i] = (eps - st.s[i] * dl.lam[i]) / st.lam[i] - st.s[i];
                }
                dl.zeta = (eps - st.zeta * dl.z) / st.z - st.zeta;
                if (it == 0) res_old = kkt<true>(st, D, eps);
                // step keeping every variable positive (step_satisfy_KKT, 402-417), then backtracking
                double mc = breduce(n, std::numeric_limits<double>::infinity(), [&](size_t j) {
                    return std::min(std::min(dl.x[j] / (st.x[j] - alpha[j]), dl.x[j] / (st.x[j] - beta[j])),
                                    std::min(dl.xi[j] / st.xi[j], dl.eta[j] / st.eta[j]));
                }, [](double s, double v) { return std::min(s, v); });
                for (size_t i = 0; i < m; ++i)
                    mc = std::min({mc, dl.y[i] / st.y[i], dl.s[i] / st.s[i], dl.mu[i] / st.mu[i], dl.lam[i] / st.lam[i]});
                mc = std::min({mc, dl.z / st.z, dl.zeta / st.zeta});
                double t = 1 / std::max(-1.01 * mc, 1.0);
                axpy(st, t, dl, n, m);
                D = derive(st);
                double res = kkt<true>(st, D, eps);
                while (res > res_old) {
                    t /= 2;
                    axpy(st, -t, dl, n, m);
                    D = derive(st);
                    res = kkt<true>(st, D, eps);
                }
                res_old = res;
                if (kkt<false>(st, D, eps) <= 0.9 * eps) break;
            }
            eps *= 0.1;
        }
        return st.x;
    }
};

}  // namespace

PYBIND11_MODULE(pyOptimizer, m) {
    m.doc() = "Method of Moving Asymptotes (MethodOfMovingAsymptotes.hh), host-side optimiser of the layer-by-layer "
              "drivers";
    py::class_<MMA>(m, "MMA")
        .def(py::init<int, int, const Vec &, const Vec &, py::function, py::function>(), py::arg("numVars"),
             py::arg("numConstr"), py::arg("xmin"), py::arg("xmax"), py::arg("f"), py::arg("df_dx"))
        .def("setInitialVar", &MMA::setInitialVar, py::arg("x0"))
        .def("step", &MMA::step)
        .def("enableGCMMA", &MMA::enableGCMMA, py::arg("enable"))
        .def("getOptimalVar", [](const MMA &o) {
            const Vec x = o.getOptimalVar();
            py::array_t<double> out((py::ssize_t)x.size());
            std::copy(x.begin(), x.end(), out.mutable_data());
            return out;
        });
}
