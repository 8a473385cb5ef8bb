"""Differentiable packing: gradients of smooth packing observables through libdem's adjoint
form, a torch.autograd.Function over it, finite-difference checks and calibration.

Parameters per instance, theta = (mu, log E, e, mu_r): sliding friction (particle-particle
and particle-wall), log of Young's modulus in Pa, normal restitution (mapped to the dashpot
coefficient of both laws by the head-on collision ODE) and the rolling-friction arm in units
of d.  The library returns adjoints of the contact-law constants (E*, mu, gamma_n, gamma_t,
mu_r of the pair law, then of the wall law); `MaterialMap.chain` maps them to theta:
E* is proportional to E in both laws, so d/dlogE = E* x adjoint; gamma_n = gamma_t =
gamma(e), whose derivative is taken by a central difference of the ODE-defined map.

Observables (dem_obs): phi = windowed solid fraction of an inner region, CN = windowed mean
soft coordination number; both are smooth functions of the final positions.  The forward
runs `nsteps` steps keeping a checkpoint every `every` steps; the backward recomputes each
segment with tapes and runs the reverse kernels over it.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (OPT_ADJ_ROLL_REG, OPT_ROLL_FLOOR, Context, Material, Obs, Workload,
               damping_for_restitution, reduced_material)

PARAMS = ("mu", "logE", "e", "mu_r")


class MaterialMap:
    """theta = (mu, log E, e, mu_r) -> reduced-unit Material, and the chain rule back."""

    def __init__(self, nu: float = 0.29, rho: float = 2500.0, diam: float = 1e-3,
                 e_step: float = 1e-4):
        self.nu, self.rho, self.diam, self.e_step = nu, rho, diam, e_step

    def material(self, theta) -> Material:
        mu, loge, e, mur = (float(v) for v in theta)
        return reduced_material(E=math.exp(loge), nu=self.nu, rho=self.rho, diam=self.diam,
                                mu=mu, mu_w=mu, e=e, mu_r=mur)

    def dgamma_de(self, e: float) -> float:
        h = self.e_step * e
        return (damping_for_restitution(e + h) - damping_for_restitution(e - h)) / (2.0 * h)

    def chain(self, theta, tb) -> np.ndarray:
        """tb[10] = adjoints of (E*, mu, gamma_n, gamma_t, mu_r) pair law, then wall law."""
        m = self.material(theta)
        g_mu = tb[1] + tb[6]
        g_loge = tb[0] * m.p.estarp + tb[5] * m.w.estarw
        g_e = (tb[2] + tb[3] + tb[7] + tb[8]) * self.dgamma_de(float(theta[2]))
        g_mur = tb[4] + tb[9]
        return np.array([g_mu, g_loge, g_e, g_mur])


class PackingProblem:
    """One context over a workload: fixed initial state, `nsteps` steps, checkpoints every
    `every` steps, observables on `obs`.  theta has one row per instance.

    `roll_floor` (DEM_OPT_ROLL_FLOOR) floors the rolling moment's |w_rel| in the forward at k
    times the one-step chatter scale: the moment is viscous there instead of flipping sign
    every step about rolling rest, the forward is a Lipschitz function of its inputs and the
    adjoint is its exact reverse mode.  With roll_floor = 0 the forward is the reference's
    moment and `roll_reg` regularises the reverse pass only (0: exact reverse mode, unstable
    over long windows of resting contact; 1: |w_rel| floored at the chatter scale in the
    derivative)."""

    def __init__(self, w: Workload, nsteps: int, every: int, obs: Obs,
                 mmap: MaterialMap | None = None, roll_reg: float = 1.0,
                 roll_floor: float = 0.0):
        self.w = w
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.obs = obs
        self.mmap = mmap or MaterialMap()
        self.ctx = Context(w.instances)
        self.ctx.set_option(OPT_ADJ_ROLL_REG, roll_reg)
        self.ctx.set_option(OPT_ROLL_FLOOR, roll_floor)
        self.roll_floor = float(roll_floor)
        self.ninst = len(w.instances)
        self.theta = None
        self.timing = {}

    def _map(self, f, rows):
        # the restitution inversion (an ODE solve per call) runs in the C library, which
        # releases the GIL: instances map in parallel
        if len(rows) == 1:
            return [f(rows[0])]
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(min(16, len(rows))) as ex:
            return list(ex.map(f, rows))

    def _set(self, theta):
        theta = np.asarray(theta, dtype=np.float64).reshape(self.ninst, 4)
        mats = self._map(self.mmap.material, list(theta))
        for k, m in enumerate(mats):
            self.ctx.set_material(k, m)
        self.ctx.load(self.w)
        return theta

    def forward(self, theta):
        """Checkpointed forward: (phi, CN) per instance."""
        self.theta = self._set(theta)
        t0 = time.perf_counter()
        phi, cn = self.ctx.adj_forward(self.nsteps, self.every, self.obs)
        self.timing["forward_s"] = time.perf_counter() - t0
        return phi, cn

    def backward(self, gphi, gcn) -> np.ndarray:
        """d(sum gphi phi + gcn CN)/d theta, one row per instance (after forward)."""
        t0 = time.perf_counter()
        tb = self.ctx.adj_backward(gphi, gcn)
        self.timing["backward_s"] = time.perf_counter() - t0
        self.theta_bar = tb
        return np.stack(self._map(lambda k: self.mmap.chain(self.theta[k], tb[k]),
                                  list(range(self.ninst))))

    def evaluate(self, theta):
        """Plain forward (dem_step, no checkpoints) and the observables: the FD reference."""
        self._set(theta)
        t0 = time.perf_counter()
        self.ctx.step(self.nsteps)
        self.ctx.sync()
        self.timing["plain_s"] = time.perf_counter() - t0
        return self.ctx.observe(self.obs)


def _torch():
    import torch
    return torch


def packing_function():
    """The torch.autograd.Function class (built on first use so that importing this module
    does not import torch)."""
    torch = _torch()

    class PackingObservables(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem):
            phi, cn = problem.forward(theta.detach().cpu().numpy())
            ctx.problem = problem
            return (torch.from_numpy(phi).to(theta.dtype), torch.from_numpy(cn).to(theta.dtype))

        @staticmethod
        def backward(ctx, gphi, gcn):
            p = ctx.problem
            zeros = np.zeros(p.ninst)
            gp = gphi.detach().cpu().numpy() if gphi is not None else zeros
            gc = gcn.detach().cpu().numpy() if gcn is not None else zeros
            g = p.backward(gp, gc)
            return torch.from_numpy(g).reshape(p.ninst, 4), None

    return PackingObservables


def packing_observables(theta, problem: PackingProblem):
    """(phi, CN) per instance as torch tensors differentiable with respect to theta."""
    return packing_function().apply(theta, problem)


def fd_check(problem: PackingProblem, theta, rel: float = 1e-4, which=range(4)) -> dict:
    """Adjoint gradients of phi and CN (instance 0) against central differences in each
    parameter of `which`, with relative step `rel` (for log E: absolute step `rel`, i.e. a
    relative step in E)."""
    theta = np.asarray(theta, dtype=np.float64).reshape(problem.ninst, 4)
    phi, cn = problem.forward(theta)
    g_phi = problem.backward(np.eye(1, problem.ninst, 0)[0], np.zeros(problem.ninst))[0]
    g_cn = problem.backward(np.zeros(problem.ninst), np.eye(1, problem.ninst, 0)[0])[0]
    fd_phi = np.full(4, np.nan)
    fd_cn = np.full(4, np.nan)
    for j in which:
        h = rel if j == 1 or theta[0, j] == 0 else rel * abs(theta[0, j])
        tp, tm = theta.copy(), theta.copy()
        tp[0, j] += h
        tm[0, j] -= h
        pp, cp = problem.evaluate(tp)
        pm, cm = problem.evaluate(tm)
        fd_phi[j] = (pp[0] - pm[0]) / (2 * h)
        fd_cn[j] = (cp[0] - cm[0]) / (2 * h)
    return {"phi": float(phi[0]), "cn": float(cn[0]), "adj_phi": g_phi, "adj_cn": g_cn,
            "fd_phi": fd_phi, "fd_cn": fd_cn}


BOUNDS = np.array([[0.05, 1.0], [math.log(1e6), math.log(1e9)], [0.1, 0.95], [0.0, 0.05]])


def calibrate(problem: PackingProblem, target_phi: float, target_cn: float, theta0,
              iters: int = 10, lr=(0.05, 0.2, 0.05, 0.002), scale=(0.01, 0.5),
              progress=None, decay: float = 1.0) -> dict:
    """Gradient-based calibration of theta (one row per instance, each its own start) to the
    target observables: Adam on the scaled squared misfit with per-parameter step sizes `lr`,
    multiplied by `decay` after every iteration, projected onto BOUNDS after each step.  Every
    iteration is one batched forward + backward through the autograd function;
    `progress(record)` is called after each."""
    torch = _torch()
    step = torch.tensor(lr, dtype=torch.float64)
    # Adam's steps are about lr in the optimised variable whatever the gradient's scale, so
    # optimise z = theta / step with unit learning rate: each parameter moves about its
    # own step size per iteration.
    z = (torch.tensor(np.asarray(theta0, dtype=np.float64).reshape(problem.ninst, 4)) / step)
    z.requires_grad_(True)
    opt = torch.optim.Adam([z], lr=1.0)
    lo = torch.tensor(BOUNDS[:, 0])
    hi = torch.tensor(BOUNDS[:, 1])
    hist = []
    for it in range(iters):
        opt.zero_grad()
        theta = z * step
        phi, cn = packing_observables(theta, problem)
        loss_k = ((phi - target_phi) / scale[0]) ** 2 + ((cn - target_cn) / scale[1]) ** 2
        loss = loss_k.sum()
        loss.backward()
        hist.append({"iter": it, "loss": float(loss.detach()), "loss_per_instance": loss_k.detach().numpy().copy(),
                     "theta": theta.detach().numpy().copy(), "phi": phi.detach().numpy().copy(),
                     "cn": cn.detach().numpy().copy(), **problem.timing})
        if progress:
            progress(hist[-1])
        opt.step()
        for g in opt.param_groups:
            g["lr"] *= decay
        with torch.no_grad():
            z.copy_(torch.max(torch.min(z * step, hi), lo) / step)
    return {"history": hist, "theta": (z * step).detach().numpy()}
