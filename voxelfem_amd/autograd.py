"""Differentiable compliance: design variables -> smoothing filter -> Heaviside projection -> SIMP ->
MG-PCG solve -> compliance 1/2 f.u, as a ``torch.autograd.Function`` on device tensors.

The forward pass is the path of TopologyOptimizationProblem::setVars + evaluateObjective
(TopologyOptimizationProblem.hh:41-75); the backward pass is evaluateObjectiveGradient
(TopologyOptimizationProblem.hh:78-84): the compliance sensitivity with respect to the physical
densities (TensorProductSimulator.hh:972-1005; for compliance the adjoint state is u itself, so no
second solve is needed), then the projection and smoothing adjoints (TopologyOptimizationFilter.hh:
162-175, 212-225, 306-310). The smoothed field S(x) is recomputed in backward instead of stored.

All data stays in HBM: the C ABI is called with device pointers of float64 torch tensors.
"""
import ctypes

import torch

from . import capi
from . import pyVoxelFEM as pv


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class CompliancePipeline:
    """Owns the simulator state of one compliance evaluation chain.

    tps: a pyVoxelFEM TensorProductSimulator with material, law and boundary conditions set;
    levels: multigrid coarsening levels; rmin: smoothing radius (linear hat, w = rmin - |d|; the
    reference's integer radius r is rmin = r + 1; 0 disables the filter); beta: projection steepness
    (0 disables the projection); precision: "fp64" or "fp32" V-cycle inside the FP64 PCG.
    """

    def __init__(self, tps, levels, rmin=0.0, filter_type=1, beta=0.0, tol=1e-8, max_iter=1000,
                 mg_iterations=1, smoothing=1, full_multigrid=False, precision="fp64", device="cuda:0"):
        self.tps = tps
        self.mg = tps.multigridSolver(levels)
        self.sim_h = tps._handle()
        self.mg_h = self.mg._handle()
        self.dims = tuple(int(n) for n in tps.NbElementsPerDimension)
        self.nel = int(tps.numElements())
        self.nn = int(tps.numNodes())
        self.rmin = float(rmin)
        self.filter_type = int(filter_type)
        self.beta = float(beta)
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.mg_iterations = int(mg_iterations)
        self.smoothing = int(smoothing)
        self.full_multigrid = bool(full_multigrid)
        self.precision = capi.VF_PREC_FP32 if precision == "fp32" else capi.VF_PREC_FP64
        self.device = torch.device(device)
        self.u = torch.zeros(3 * self.nn, dtype=torch.float64, device=self.device)
        self.f = torch.zeros_like(self.u)
        self.last_stats = None
        self._dims64 = (ctypes.c_int64 * 3)(*self.dims)

    # -- filters on device -------------------------------------------------------------------
    def smooth(self, x):
        if self.rmin <= 0:
            return x
        out = torch.empty_like(x)
        capi.check(capi.lib().vf_smoothing_filter(self._dims64, self.rmin, self.filter_type, _ptr(x), _ptr(out),
                                                  capi.VF_DEVICE, None))
        return out

    def project(self, s):
        if self.beta <= 0:
            return s
        out = torch.empty_like(s)
        capi.check(capi.lib().vf_projection(self.beta, self.nel, _ptr(s), _ptr(out), capi.VF_DEVICE, None))
        return out

    def project_backprop(self, g, s):
        if self.beta <= 0:
            return g
        out = torch.empty_like(g)
        capi.check(capi.lib().vf_projection_backprop(self.beta, self.nel, _ptr(g), _ptr(s), _ptr(out),
                                                     capi.VF_DEVICE, None))
        return out

    def physical(self, x):
        return self.project(self.smooth(x))

    # -- state solve ---------------------------------------------------------------------------
    def solve(self, rho, zero_init=True):
        L = capi.lib()
        capi.check(L.vf_sim_set_densities(self.sim_h, _ptr(rho), capi.VF_DEVICE))
        capi.check(L.vf_sim_build_load(self.sim_h, _ptr(self.f), capi.VF_DEVICE))
        if zero_init:
            self.u.zero_()
        torch.cuda.synchronize(self.device)
        self.last_stats = capi.pcg(self.mg_h, self.u.data_ptr(), self.f.data_ptr(), self.max_iter, self.tol,
                                   self.mg_iterations, self.smoothing, self.full_multigrid, self.precision)
        c = ctypes.c_double()
        capi.check(L.vf_sim_dot(self.sim_h, _ptr(self.f), _ptr(self.u), 0, capi.VF_DEVICE, ctypes.byref(c)))
        return 0.5 * c.value

    def compliance_gradient(self, rho, u):
        """dc/drho at densities rho for the equilibrium displacement u (both device tensors)."""
        g = torch.empty(self.nel, dtype=torch.float64, device=self.device)
        L = capi.lib()
        capi.check(L.vf_sim_set_densities(self.sim_h, _ptr(rho), capi.VF_DEVICE))
        capi.check(L.vf_sim_compliance_gradient(self.sim_h, _ptr(u), _ptr(g), capi.VF_DEVICE))
        return g

    def __call__(self, x):
        return ComplianceFunction.apply(x, self)


class ComplianceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, pipe):
        x = x.detach().to(device=pipe.device, dtype=torch.float64).contiguous()
        torch.cuda.synchronize(pipe.device)
        rho = pipe.physical(x)
        c = pipe.solve(rho)
        ctx.pipe = pipe
        # the displacement is the adjoint state; the filtered fields are recomputed in backward
        ctx.save_for_backward(x, pipe.u.clone())
        return torch.tensor(c, dtype=torch.float64, device=pipe.device)

    @staticmethod
    def backward(ctx, grad_c):
        x, u = ctx.saved_tensors
        pipe = ctx.pipe
        torch.cuda.synchronize(pipe.device)
        s = pipe.smooth(x)                         # recomputed, not stored
        rho = pipe.project(s)
        g = pipe.compliance_gradient(rho, u)       # dc/drho (SIMP adjoint + self-weight term)
        g = pipe.project_backprop(g, s)            # through the Heaviside projection
        g = pipe.smooth(g)                         # the reflect-BC smoothing filter is self-adjoint
        torch.cuda.synchronize(pipe.device)
        return grad_c * g, None


def compliance(x, pipe):
    """c(x) with gradient support: ``compliance(x, pipe).backward()`` fills ``x.grad``."""
    return ComplianceFunction.apply(x, pipe)


__all__ = ["CompliancePipeline", "ComplianceFunction", "compliance", "pv"]
