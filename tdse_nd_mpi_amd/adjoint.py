"""Differentiable propagation: a torch.autograd.Function over the libqnd C ABI.

`Propagate.apply(E, softcore, spec)` runs the length-gauge split-operator
propagation of `spec` on the GPU (libqnd, checkpointed every `spec.every`
steps) under the field E(t_n) (x-polarized, held over the stages of a step)
and returns the dipole acceleration a(t_n) [nsteps, dim] and the final norm.
Its backward runs the adjoint passes (qnd_adjoint) and returns dL/dE and
dL/d(soft-core length); the initial state is held fixed.  The tensors live on
the host (float64): the propagation state never leaves the GPU.
"""
import numpy as np
import torch

from . import Grid


class Spec:
    """A propagation problem: grid, potential, integrator, mask, initial state."""

    def __init__(self, nr, dr, potential, integrator, dt, psi0, precision="c64", mask=0.0, every=64,
                 slot=0):
        self.nr, self.dr, self.potential = tuple(nr), dr, potential
        self.integrator, self.dt, self.psi0 = integrator, dt, psi0
        self.precision, self.mask, self.every, self.slot = precision, mask, every, slot
        self.grid = None
        self.softcore = None
        self.runs = 0          # forward runs made on the grid; the last one is what qnd_adjoint differentiates

    def ensure(self, softcore):
        if self.grid is None:
            self.grid = Grid(self.nr, self.dr, self.precision, n_slots=1)
            self.grid.set_integrator(self.integrator, self.dt)
            self.grid.set_mask(self.mask)
        if self.softcore != softcore:
            self.grid.set_potential(self.potential, softcore=softcore)
            self.softcore = softcore
        return self.grid

    def close(self):
        if self.grid is not None:
            self.grid.close()
            self.grid = None


def field_tables(E, s, dim):
    """Stage and observation tables for a per-step x-polarized field E_n."""
    nt = len(E)
    Et = np.zeros((nt, s, dim))
    Et[:, :, 0] = E[:, None]
    Eo = np.zeros((nt, dim))
    Eo[:, 0] = E
    return Et, np.zeros_like(Et), Eo, np.zeros_like(Eo)


def forward_numpy(spec, E, softcore):
    """Forward run only: (a [nsteps, dim], norms [nsteps], final norm)."""
    g = spec.ensure(float(softcore))
    g.set_field(*field_tables(np.asarray(E, dtype=np.float64), g.s, g.dim))
    g.upload(spec.slot, spec.psi0)
    out = g.forward_checkpointed(spec.slot, len(E), spec.every)
    spec.runs += 1
    return out


class Propagate(torch.autograd.Function):
    """The library keeps the checkpoints of one forward run per grid.  Each
    application records which run it was; a backward whose run is no longer
    the grid's last (another application ran in between) runs its own
    forward again before the adjoint, so every gradient belongs to its own
    field and soft-core length."""

    @staticmethod
    def forward(ctx, E, softcore, spec):
        En = E.detach().cpu().numpy().astype(np.float64)
        a, _, nf = forward_numpy(spec, En, float(softcore))
        ctx.spec = spec
        ctx.run = spec.runs
        ctx.E = En.copy()
        ctx.softcore = float(softcore)
        return torch.from_numpy(a.copy()), torch.tensor(nf, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_a, grad_norm):
        if ctx.spec.runs != ctx.run:
            forward_numpy(ctx.spec, ctx.E, ctx.softcore)
            ctx.run = ctx.spec.runs
        g = ctx.spec.grid
        ga = np.zeros((len(ctx.E), g.dim)) if grad_a is None else grad_a.detach().cpu().numpy()
        gn = 0.0 if grad_norm is None else float(grad_norm)
        dEt, dEo, ds = g.adjoint(ga, gn)
        dE = dEt[:, :, 0].sum(axis=1) + dEo[:, 0]
        return torch.from_numpy(dE), torch.tensor(ds, dtype=torch.float64), None


def harmonic_loss(a, norm_final, dt, omega0, orders=(11, 21), width=0.5):
    """Power of the dipole-acceleration spectrum in windows of +-width*omega0
    around the odd/even harmonics q*omega0, q in [orders], plus the norm lost
    to the absorber: a differentiable torch expression of a(t_n)."""
    nt = a.shape[0]
    w = torch.from_numpy(np.hanning(nt))
    spec = torch.fft.rfft(a[:, 0] * w)
    freqs = 2 * np.pi * np.fft.rfftfreq(nt, dt)
    mask = np.zeros(len(freqs))
    for qn in range(orders[0], orders[1] + 1):
        mask[np.abs(freqs - qn * omega0) <= width * omega0] = 1.0
    power = (spec.real ** 2 + spec.imag ** 2) * torch.from_numpy(mask)
    return power.sum() * dt * dt + (1.0 - norm_final)
