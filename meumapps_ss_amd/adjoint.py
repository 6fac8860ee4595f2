"""Derivatives of a run of semi-implicit steps: the hand-written adjoint of the
C ABI (``mm_adj_*``) driven with checkpoints, and a ``torch.autograd.Function``
over it.

A run starts after the zeroth-order elastic solve (step = elastep0 + 1, zero
displacement, i.e. a restart with u = 0, Multi_MEUMAPPS.f:470-481) and takes
``nsteps`` regular steps (Multi_MEUMAPPS.f:548-1106).  The forward pass keeps a
device checkpoint of the state every ``every`` steps; the backward pass walks
the segments last to first, recomputes the states inside a segment from its
checkpoint, and for every step (last to first) re-runs it with the tape on and
pulls the adjoint state back through it.

Scalar parameters reach the kernels through ``mm_params`` (built by a
user-supplied ``build(theta) -> MMParams``); the adjoint returns dL/d of the
raw parameter arrays, and the chain to each scalar uses the derivative of the
build, taken by central differences of the (linear) set-up.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .params import MM_FP64, MMAdjGrads, MMDiag, MMParams
from .sim import Simulation, soa_to_aos

RAW = ("grad_coeff", "wellph", "cmobs", "b_mob", "trans_ppt")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def raw_vector(p) -> Dict[str, np.ndarray]:
    """The differentiable parameter arrays of an MMParams (or MMAdjGrads)."""
    out = {}
    for k in RAW:
        v = getattr(p, k)
        out[k] = np.array(v if hasattr(v, "__len__") else [v], dtype=np.float64)
    return out


def build_jacobian(build: Callable[[Dict[str, float]], MMParams], theta: Dict[str, float],
                   rel: float = 1e-3) -> Dict[str, Dict[str, np.ndarray]]:
    """d raw / d theta_name for every scalar by central differences of the
    set-up (exact for the linear maps of kappa, mobilities and eigenstrain)."""
    jac = {}
    for name, t in theta.items():
        h = rel * max(1.0, abs(t))
        up = dict(theta, **{name: t + h})
        dn = dict(theta, **{name: t - h})
        a, b = raw_vector(build(up)), raw_vector(build(dn))
        jac[name] = {k: (a[k] - b[k]) / (2 * h) for k in RAW}
    return jac


class AdjointRun:
    """One FP64 context with the adjoint enabled."""

    def __init__(self, params: MMParams):
        import torch  # device memory for the checkpoints

        self.torch = torch
        self.lib = _lib.load()
        self.p = params
        self.sim = Simulation(params, MM_FP64)
        _lib.check(self.lib.mm_adj_enable(self.sim.ctx), "mm_adj_enable")
        nb = ctypes.c_int64()
        _lib.check(self.lib.mm_state_bytes(self.sim.ctx, ctypes.byref(nb)), "mm_state_bytes")
        self.state_bytes = nb.value
        self.V, self.E = params.vartot, params.numel
        self.shape = (params.nz, params.ny, params.nx)

    def close(self):
        self.sim.close()

    # -- checkpoints in device memory owned by torch
    def save(self):
        t = self.torch.empty(self.state_bytes, dtype=self.torch.uint8, device="cuda")
        self.torch.cuda.synchronize()
        _lib.check(self.lib.mm_state_save(self.sim.ctx, ctypes.c_void_p(t.data_ptr())), "mm_state_save")
        return t

    def load(self, t):
        self.torch.cuda.synchronize()
        _lib.check(self.lib.mm_state_load(self.sim.ctx, ctypes.c_void_p(t.data_ptr())), "mm_state_load")

    # -- branch tape (mm_branch_tape): record / count / replay the phi > 0
    #    tests of the f_diff correction and the clamp, one byte per point and
    #    variant per step, in a device buffer owned by torch
    def new_branch_tape(self, nsteps: int):
        nb = ctypes.c_int64()
        _lib.check(self.lib.mm_branch_bytes(self.sim.ctx, ctypes.byref(nb)), "mm_branch_bytes")
        return self.torch.zeros(nsteps * nb.value, dtype=self.torch.uint8, device="cuda")

    def set_branch_tape(self, mode: int, buf=None, nsteps: int = 0):
        ptr = None if buf is None else ctypes.c_void_p(buf.data_ptr())
        self.torch.cuda.synchronize()
        _lib.check(self.lib.mm_branch_tape(self.sim.ctx, mode, ptr, nsteps), "mm_branch_tape")

    def branch_flips(self) -> int:
        n = ctypes.c_int64()
        _lib.check(self.lib.mm_branch_flips(self.sim.ctx, ctypes.byref(n)), "mm_branch_flips")
        return n.value

    def start(self, phi0: np.ndarray, cons0: np.ndarray):
        self.sim.upload(phi=phi0, cons=cons0, u=np.zeros((3, *self.shape)), gradu=np.zeros((9, *self.shape)))
        self.sim.set_step(self.p.elastep0 + 1)

    def forward(self, phi0, cons0, nsteps: int, every: int):
        """Run nsteps from (phi0, cons0); returns (phi_final, checkpoints, rows)."""
        self.start(phi0, cons0)
        ckpts, rows = [], []
        for s in range(nsteps):
            if s % every == 0:
                ckpts.append(self.save())
            rows += self.sim.step(1)
        return self.sim.download()["phi"], ckpts, rows

    def backward(self, lam_phi: np.ndarray, ckpts, nsteps: int, every: int,
                 lam_cons: Optional[np.ndarray] = None):
        """Adjoint of the run: returns (dL/dphi0, dL/dcons0, MMAdjGrads)."""
        n3 = int(np.prod(self.shape))
        lp = np.ascontiguousarray(lam_phi, dtype=np.float64).reshape(self.V * n3)
        lc = (np.zeros(self.E * n3) if lam_cons is None
              else np.ascontiguousarray(lam_cons, dtype=np.float64).reshape(self.E * n3))
        ctx = self.sim.ctx
        _lib.check(self.lib.mm_adj_zero(ctx), "mm_adj_zero")
        _lib.check(self.lib.mm_adj_set(ctx, _dp(lp), _dp(lc), None), "mm_adj_set")
        g = MMAdjGrads()
        nseg = len(ckpts)
        for seg in reversed(range(nseg)):
            s0 = seg * every
            s1 = min(s0 + every, nsteps)
            states = [ckpts[seg]]
            self.load(ckpts[seg])
            for _ in range(s0, s1 - 1):
                _lib.check(self.lib.mm_step(ctx, 1, None), "mm_step")
                states.append(self.save())
            for s in reversed(range(s0, s1)):
                self.load(states[s - s0])
                _lib.check(self.lib.mm_adj_tape_step(ctx, None), "mm_adj_tape_step")
                _lib.check(self.lib.mm_adj_step(ctx, ctypes.byref(g)), "mm_adj_step")
            del states
        out_p = np.empty(self.V * n3)
        out_c = np.empty(self.E * n3)
        _lib.check(self.lib.mm_adj_get(ctx, _dp(out_p), _dp(out_c), None), "mm_adj_get")
        return out_p.reshape(self.V, *self.shape), out_c.reshape(self.E, *self.shape), g


def chain(g: MMAdjGrads, jac: Dict[str, Dict[str, np.ndarray]]) -> Dict[str, float]:
    """dL/dtheta from the raw-array derivatives and the build Jacobian."""
    rv = raw_vector(g)
    return {name: float(sum(np.dot(rv[k], jk[k]) for k in RAW)) for name, jk in jac.items()}


@dataclass
class Problem:
    """A differentiable run: the parameter build, the fixed composition
    field, the step count and the checkpoint interval."""

    build: Callable[[Dict[str, float]], MMParams]
    names: Sequence[str]
    cons0: np.ndarray
    nsteps: int
    every: int = 20
    runs: List[AdjointRun] = field(default_factory=list)


def _make_function():
    import torch

    class SpectralSteps(torch.autograd.Function):
        """phi_final = run(phi0, theta): nsteps semi-implicit steps on one GPU,
        backward by the hand-written adjoint with checkpoint recomputation."""

        @staticmethod
        def forward(ctx, phi0, theta, problem: Problem):
            th = dict(zip(problem.names, theta.detach().cpu().double().tolist()))
            run = AdjointRun(problem.build(th))
            phiT, ckpts, _ = run.forward(phi0.detach().cpu().double().numpy(), problem.cons0, problem.nsteps,
                                         problem.every)
            ctx.run, ctx.ckpts, ctx.problem, ctx.th = run, ckpts, problem, th
            ctx.dev = phi0.device
            return torch.from_numpy(phiT).to(phi0.device)

        @staticmethod
        def backward(ctx, grad_out):
            pr = ctx.problem
            lam0, _, g = ctx.run.backward(grad_out.detach().cpu().double().numpy(), ctx.ckpts, pr.nsteps, pr.every)
            d = chain(g, build_jacobian(pr.build, ctx.th))
            ctx.run.close()
            ctx.ckpts = None
            dth = torch.tensor([d[n] for n in pr.names], dtype=torch.float64)
            return torch.from_numpy(lam0).to(ctx.dev), dth, None

    return SpectralSteps


_FN = None


def spectral_steps(phi0, theta, problem: Problem):
    """Differentiable run (see SpectralSteps)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(phi0, theta, problem)
