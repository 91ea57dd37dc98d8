"""Method of Moving Asymptotes, restated in numpy from MethodOfMovingAsymptotes.hh (Svanberg's MMA with the
reference's constants and its primal-dual interior-point subproblem solver).

TEST INFRASTRUCTURE ONLY: the checker of the product's MMA (pyOptimizer.MMA), never shipped or measured.

Problem (MethodOfMovingAsymptotes.hh:22-27):
    min_x  f_0(x) + a_0 z + sum_i (c_i y_i + d_i y_i^2 / 2)
    s.t.   f_i(x) - a_i z - y_i <= 0,  xmin <= x <= xmax,  y >= 0,  z >= 0
with a_0 = 1, a = 0, c = 1000, d = 1 (34-36). f(x) returns [f_0, f_1..f_m], df_dx(x) the (m+1) x n gradient.
"""
import numpy as np

RAA0, ALBEFA, MOVE, ASYINIT = 1e-5, 0.1, 0.5, 0.5   # MethodOfMovingAsymptotes.hh:218
GAMMA = (0.7, 1.2, 1.0)                             # 217


class MMA:
    def __init__(self, n, m, xmin, xmax, f, df_dx):
        self.n, self.m = n, m
        self.xmin, self.xmax = np.asarray(xmin, float), np.asarray(xmax, float)
        self.xdiff = self.xmax - self.xmin
        self.f, self.df = f, df_dx
        self.a0 = 1.0
        self.a, self.c, self.d = np.zeros(m), np.full(m, 1000.0), np.ones(m)
        self.hist = []          # x^k, x^{k-1}, x^{k-2} (front first)
        self.outer = 0
        self.gcmma = False
        self.l = self.u = None

    def setInitialVar(self, x0):
        self._push(np.array(x0, float))

    def enableGCMMA(self, on):
        self.gcmma = bool(on)

    def _push(self, x):
        self.hist.insert(0, np.array(x, float))
        del self.hist[3:]

    def x(self):
        return self.hist[0]

    # ---- one outer iteration (step, 63-163)
    def step(self):
        if not self.hist:
            raise RuntimeError("Must specify an initial value")
        self.outer += 1
        x = self.x()
        if self.outer <= 2:
            self.l = x - ASYINIT * self.xdiff
            self.u = x + ASYINIT * self.xdiff
        else:
            x1, x2 = self.hist[1], self.hist[2]
            zzz = (x - x1) * (x1 - x2)
            g = np.where(zzz > 0, GAMMA[1], GAMMA[0])
            self.l = np.maximum(np.minimum(x - g * (x1 - self.l), x - 0.01 * self.xdiff), x - 10 * self.xdiff)
            self.u = np.minimum(np.maximum(x + g * (self.u - x1), x + 0.01 * self.xdiff), x + 10 * self.xdiff)
        self.alpha = np.maximum(np.maximum(self.xmin, self.l + ALBEFA * (x - self.l)), x - MOVE * self.xdiff)
        self.beta = np.minimum(np.minimum(self.xmax, self.u - ALBEFA * (self.u - x)), x + MOVE * self.xdiff)
        self.fcur = np.asarray(self.f(x), float)
        dfx = np.asarray(self.df(x), float).reshape(self.m + 1, self.n)
        self.dplus, self.dminus = np.maximum(dfx, 0), np.maximum(-dfx, 0)
        umx, xml = self.u - x, x - self.l
        if not self.gcmma:
            rho = np.full(self.m + 1, RAA0)
            self._pq(umx ** 2, xml ** 2, rho)
            self.r = self.fcur - self._g_all(umx, xml)
            self._push(self._subsolve())
        else:
            umx_old, xml_old = umx.copy(), xml.copy()
            inner = 0
            self.rho = None
            xin = None
            while True:
                if inner == 0:
                    rho = np.maximum(0.1 / self.n * ((self.dplus + self.dminus) @ self.xdiff), 1e-6)
                else:
                    # the asymptote distances at the last inner solution (next_rho, 154-161)
                    dd = np.sum((self.u - self.l) * (xin - x) ** 2 / ((self.u - xin) * (xin - self.l) * self.xdiff))
                    delta = self.diff / dd
                    rho = np.where(delta < 0, self.rho, np.minimum(1.1 * (self.rho + delta), 10 * self.rho))
                self.rho = rho
                self._pq(umx_old ** 2, xml_old ** 2, rho)
                self.r = self.fcur - self._g_all(umx_old, xml_old)
                xin = self._subsolve()
                inner += 1
                umx_i, xml_i = self.u - xin, xin - self.l
                self.diff = np.asarray(self.f(xin), float) - (self._g_all(umx_i, xml_i) + self.r)
                if self.diff.max() < 0:
                    break
            self._push(xin)

    def _pq(self, umx2, xml2, rho):
        rd = rho[:, None] / self.xdiff[None, :]
        self.p = umx2[None, :] * (1.001 * self.dplus + 0.001 * self.dminus + rd)
        self.q = xml2[None, :] * (0.001 * self.dplus + 1.001 * self.dminus + rd)

    def _g_all(self, umx, xml):
        return self.p @ (1.0 / umx) + self.q @ (1.0 / xml)

    # ---- the subproblem's primal-dual Newton method (Subproblem, 240-465)
    def _subsolve(self):
        m, n = self.m, self.n
        al, be = self.alpha, self.beta
        x = 0.5 * (al + be)
        xi = np.maximum(1.0 / (x - al), 1.0)
        eta = np.maximum(1.0 / (be - x), 1.0)
        y, z, zeta = np.ones(m), 1.0, 1.0
        lam, s = np.ones(m), np.ones(m)
        mu = np.maximum(self.c / 2, 1.0)
        st = dict(x=x, xi=xi, eta=eta, y=y, z=z, zeta=zeta, lam=lam, s=s, mu=mu)

        def derived(st):
            umx, xml = self.u - st["x"], st["x"] - self.l
            plam = self.p[0] + self.p[1:].T @ st["lam"]
            qlam = self.q[0] + self.q[1:].T @ st["lam"]
            dpsi = plam / umx ** 2 - qlam / xml ** 2
            gvec = self.p[1:] @ (1.0 / umx) + self.q[1:] @ (1.0 / xml)
            return umx, xml, plam, qlam, dpsi, gvec

        def residual_parts(st, eps, dpsi, gvec):
            return [dpsi - st["xi"] + st["eta"],
                    self.c + self.d * st["y"] - st["lam"] - st["mu"],
                    np.array([self.a0 - st["zeta"] - st["lam"] @ self.a]),
                    gvec - self.a * st["z"] - st["y"] + st["s"] + self.r[1:],
                    st["xi"] * (st["x"] - al) - eps,
                    st["eta"] * (be - st["x"]) - eps,
                    st["mu"] * st["y"] - eps,
                    np.array([st["zeta"] * st["z"] - eps]),
                    st["lam"] * st["s"] - eps]

        def sqres(st, eps):
            _, _, _, _, dpsi, gvec = derived(st)
            return sum(float(np.sum(v * v)) for v in residual_parts(st, eps, dpsi, gvec))

        def kkt_inf(st, eps):
            _, _, _, _, dpsi, gvec = derived(st)
            return max(float(np.max(np.abs(v))) for v in residual_parts(st, eps, dpsi, gvec))

        eps = 1.0
        while eps > 1e-7:
            res_old = None
            for i in range(10):
                umx, xml, plam, qlam, dpsi, gvec = derived(st)
                x, xi, eta, y, z, zeta, lam, s, mu = (st[k] for k in ("x", "xi", "eta", "y", "z", "zeta", "lam", "s", "mu"))
                dx = dpsi - eps / (x - al) + eps / (be - x)
                Dx = 2 * plam / umx ** 3 + 2 * qlam / xml ** 3 + xi / (x - al) + eta / (be - x)
                G = self.p[1:] / umx ** 2 - self.q[1:] / xml ** 2
                Dy = self.d + mu / y
                dy = self.c + self.d * y - lam - eps / y
                M = np.zeros((m + 1, m + 1))
                M[:m, :m] = (G / Dx) @ G.T
                M[:m, :m] += np.diag(s / lam + 1.0 / Dy)
                M[:m, m] = self.a
                M[m, :m] = self.a
                M[m, m] = -zeta / z
                rhs = np.zeros(m + 1)
                rhs[:m] = gvec - self.a * z - y + self.r[1:] + eps / lam + dy / Dy - G @ (dx / Dx)
                rhs[m] = self.a0 - lam @ self.a - eps / z
                sol = np.linalg.solve(M, rhs)
                dlam, dz = sol[:m], sol[m]
                ddx = -(dx + G.T @ dlam) / Dx
                dxi = -xi + eps / (x - al) - xi * ddx / (x - al)
                deta = -eta + eps / (be - x) + eta * ddx / (be - x)
                ddy = dlam / Dy - dy / Dy
                dmu = (eps - mu * ddy) / y - mu
                dzeta = (eps - zeta * dz) / z - zeta
                ds = (eps - s * dlam) / lam - s
                delta = dict(x=ddx, xi=dxi, eta=deta, y=ddy, z=dz, zeta=dzeta, lam=dlam, s=ds, mu=dmu)
                if i == 0:
                    res_old = sqres(st, eps)
                mc = min(np.min(ddx / (x - al)), np.min(ddx / (x - be)), np.min(dxi / xi), np.min(deta / eta),
                         np.min(ddy / y), dz / z, np.min(ds / s), np.min(dmu / mu), dzeta / zeta, np.min(dlam / lam))
                step = 1.0 / max(-1.01 * mc, 1.0)
                for k in st:
                    st[k] = st[k] + step * delta[k]
                res = sqres(st, eps)
                while res > res_old:
                    step /= 2
                    for k in st:
                        st[k] = st[k] - step * delta[k]
                    res = sqres(st, eps)
                res_old = res
                if kkt_inf(st, eps) <= 0.9 * eps:
                    break
            eps *= 0.1
        self.sub_state = st
        return st["x"]
