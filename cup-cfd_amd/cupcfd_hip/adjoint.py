"""Reverse mode of the gradient + scalar-flux chain on the GPU, for torch.

``ScalarFlux`` is a ``torch.autograd.Function``: its forward runs
GradientPhiGaussDolfyn (``nGradient`` sweeps) and FluxScalarDolfynFaceLoop
through the resident C ABI, from Su = 0, and returns (Su, RFace); its
backward calls ``cupcfd_dev_scalar_adjoint_f64`` (one sweep, all-DEFAULT
mesh) or ``cupcfd_dev_scalar_adjoint_full_f64`` (any mesh, any number of
sweeps) for the cotangents of phi and MassFlux.
``march`` and ``march_adjoint`` run explicit pseudo-time steps
phi <- phi + dtau * Su(phi, mf) / vol and their reverse with checkpoints:
the states at every ``every``-th step are kept, the states inside a segment
are recomputed from its checkpoint when the reverse sweep reaches it.

Tensors are FP64 CUDA tensors in the device mesh's internal numbering (cell
arrays (C,), face arrays (F,), boundary arrays in mesh numbering), as for
``cupcfd_hip.fvm.resident``.
"""
from __future__ import annotations

import torch

from .fvm import resident as R


class Chain:
    """What the chain needs besides (phi, mf): the mesh, phiB, VisEff and the scalar loop's parameters.
    The inputs the chain does not differentiate (Den, TE, CpB, visEffB, PhiB) are fixed tensors."""

    def __init__(self, mesh, phiB, visEff, params, den=None, te=None, cpB=None, visEffB=None, nGradient=1):
        dev = visEff.device
        self.nGradient = nGradient
        # the one-sweep adjoint of all-DEFAULT meshes needs no forward gradient
        self.plain = nGradient == 1 and bool((mesh.host.a["regType"] == 0).all())
        C = visEff.shape[0]
        B = phiB.shape[0]
        self.mesh, self.phiB, self.visEff, self.params = mesh, phiB, visEff, dict(params)
        one = lambda n: torch.ones(n, dtype=torch.float64, device=dev)  # noqa: E731
        self.den = den if den is not None else one(C)
        self.te = te if te is not None else one(C)
        self.cpB = cpB if cpB is not None else one(B)
        self.visEffB = visEffB if visEffB is not None else one(B)
        self.C, self.B = C, B

    def gradient(self, phi):
        g = torch.empty(3 * self.C, dtype=torch.float64, device=phi.device)
        go = torch.empty_like(g)
        R.gradient(self.mesh, self.nGradient, phi, self.phiB, g, go)
        return g

    def forward(self, phi, mf):
        """(Su, RFace) from Su = 0."""
        dev = phi.device
        g = self.gradient(phi)
        su = torch.zeros(self.C, dtype=torch.float64, device=dev)
        au = torch.zeros(self.C, dtype=torch.float64, device=dev)
        rface = torch.zeros(2 * mf.shape[0], dtype=torch.float64, device=dev)
        pB = self.phiB.clone()
        R.scalar(self.mesh, phi, pB, self.visEff, au, su, self.den, mf, self.te, self.cpB, self.visEffB, rface, g,
                 self.params)
        return su, rface

    def adjoint(self, phi, mf, subar, rbar):
        """(phibar, mfbar) = J^T (subar, rbar)."""
        phibar = torch.empty_like(phi)
        mfbar = torch.empty_like(mf)
        if self.plain:
            R.scalar_adjoint(self.mesh, phi, self.visEff, mf, subar.contiguous(), rbar.contiguous(), self.params,
                             phibar, mfbar)
        else:
            R.scalar_adjoint_full(self.mesh, self.nGradient, phi, self.visEff, mf, self.gradient(phi), self.te,
                                  self.den, self.cpB, self.visEffB, subar.contiguous(), rbar.contiguous(),
                                  self.params, phibar, mfbar)
        return phibar, mfbar


class ScalarFlux(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, mf, chain):
        ctx.chain = chain
        ctx.save_for_backward(phi, mf)
        return chain.forward(phi.detach(), mf.detach())

    @staticmethod
    def backward(ctx, subar, rbar):
        phi, mf = ctx.saved_tensors
        z = lambda t, n: t if t is not None else torch.zeros(n, dtype=torch.float64, device=phi.device)  # noqa
        phibar, mfbar = ctx.chain.adjoint(phi, mf, z(subar, phi.shape[0]), z(rbar, 2 * mf.shape[0]))
        return phibar, mfbar, None


def scalar_flux(phi, mf, chain):
    """(Su, RFace) with autograd through the GPU adjoint."""
    return ScalarFlux.apply(phi, mf, chain)


def _step(chain, phi, mf, dtau, vol):
    su, _ = chain.forward(phi, mf)
    return phi + dtau * su / vol


def march(chain, phi0, mf, steps, dtau, vol):
    """phi after ``steps`` explicit pseudo-time steps."""
    phi = phi0
    for _ in range(steps):
        phi = _step(chain, phi, mf, dtau, vol)
    return phi


def march_adjoint(chain, phi0, mf, steps, dtau, vol, phibar_final, every=10):
    """Cotangents (phibar0, mfbar) of the state after ``steps`` steps, given its cotangent ``phibar_final``.
    States are checkpointed every ``every`` steps (every >= steps keeps them all) and recomputed per segment."""
    ckpt = {0: phi0}
    phi = phi0
    for n in range(steps):
        phi = _step(chain, phi, mf, dtau, vol)
        if (n + 1) % every == 0 and n + 1 < steps:
            ckpt[n + 1] = phi
    pbar = phibar_final
    mbar = torch.zeros_like(mf)
    for s0 in sorted(ckpt, reverse=True):
        s1 = min(s0 + every, steps)
        seg = [ckpt[s0]]
        for _ in range(s0, s1 - 1):
            seg.append(_step(chain, seg[-1], mf, dtau, vol))
        for n in range(s1 - 1, s0 - 1, -1):
            # phi_{n+1} = phi_n + dtau * Su(phi_n, mf) / vol
            sub = dtau * pbar / vol
            pb, mb = chain.adjoint(seg[n - s0], mf, sub, torch.zeros(2 * mf.shape[0], dtype=torch.float64,
                                                                       device=mf.device))
            pbar = pbar + pb
            mbar = mbar + mb
    return pbar, mbar
