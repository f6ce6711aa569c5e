"""Derivatives of the linewidths and the RTA thermal conductivity with respect to the
parameters of the Morse pair potentials that generate the force constants (inputs.py).

Both orders of force constants depend on the parameters, so a change of a parameter moves
the dispersions (frequencies, eigenvectors, group velocities, the adaptive widths) as well
as Phi3.  Each parameter is one forward-mode direction of the library (tc3_jvp): the
derivatives of the force-constant lists come analytically from the generator
(inputs.morse_ifc_derivatives), and the library carries them through first-order
perturbation of the dynamical matrices, the |Psi|^2 chain and the smeared energy
conservation.  `MorseLinewidths` and `MorseKappaRTA` wrap them as torch.autograd.Functions of
the parameter vector; their backward contracts the cotangent with the Jacobian, one library
call per parameter.
"""
import os
import tempfile

import numpy as np
import torch

from . import inputs
from .engine import Engine
from .io import FC2_UNIT, FC3_UNIT


def _list_arrays(fc2, fc3):
    """The matrices of morse_ifcs lists, in file order (atom by atom), in Hartree/Bohr^n."""
    m2 = np.array([e[-1] for per in fc2 for e in per]) * FC2_UNIT
    m3 = np.array([e[-1] for per in fc3 for e in per]) * FC3_UNIT
    return m2, m3


class MorseModel:
    """A workload family at Morse parameters p: the engine of its input files, run through
    the harmonic step and the scattering pass (FP64), and the derivatives with respect to p."""

    def __init__(self, name, qgrid, temperatures, params=None, seed=0, device=0, integrationtype=2,
                 sigma=1.0, classical=False):
        p0, self.labels, _, _ = inputs.morse_parameters(name)
        self.name, self.seed = name, seed
        self.params = np.array(p0 if params is None else params, dtype=np.float64)
        self.temperatures = np.atleast_1d(np.asarray(temperatures, np.float64))
        d = tempfile.mkdtemp()
        inputs.make_config(name, d, seed=seed, params=self.params)
        self.engine = Engine.from_directory(d, qgrid, device=device)
        for f in os.listdir(d):
            os.remove(os.path.join(d, f))
        os.rmdir(d)
        self.engine.dispersions()
        self.engine.scattering_rates(self.temperatures, precision="fp64", integrationtype=integrationtype,
                                     sigma=sigma, classical=classical)

    def linewidths(self):
        return np.stack([self.engine.linewidths(i)[0] for i in range(len(self.temperatures))])

    def kappa_sma(self):
        ks, _, _ = self.engine.kappa(maxiter=0, offdiag=False)
        return ks

    def jacobian(self):
        """d/dp_k of the linewidths [k, nT, nirr, nb], kappa_sma [k, nT, 3, 3], the frequencies
        [k, nirr, nb] and the group velocities [k, nirr, nb, 3]."""
        out = {"linewidth": [], "kappa_sma": [], "omega": [], "velocity": []}
        for k in range(len(self.params)):
            d2, d3 = inputs.morse_ifc_derivatives(self.name, self.params, k, seed=self.seed)
            m2, m3 = _list_arrays(d2, d3)
            j = self.engine.jvp(m2, m3)
            for key in out:
                out[key].append(j[key])
        return {key: np.stack(v) for key, v in out.items()}


class MorseLinewidths(torch.autograd.Function):
    """linewidths [nT, nirr, nb] (Hartree) of a family as a function of its Morse parameters."""

    @staticmethod
    def forward(ctx, p, name, qgrid, temperatures, seed=0):
        model = MorseModel(name, qgrid, temperatures, p.detach().cpu().numpy(), seed=seed)
        ctx.jac = torch.from_numpy(model.jacobian()["linewidth"]).to(p.device, p.dtype)
        return torch.from_numpy(model.linewidths()).to(p.device, p.dtype)

    @staticmethod
    def backward(ctx, grad):
        return (ctx.jac * grad.unsqueeze(0)).flatten(1).sum(1), None, None, None, None


class MorseKappaRTA(torch.autograd.Function):
    """single-mode kappa [nT, 3, 3] (atomic units, symmetrised) as a function of the Morse
    parameters."""

    @staticmethod
    def forward(ctx, p, name, qgrid, temperatures, seed=0):
        model = MorseModel(name, qgrid, temperatures, p.detach().cpu().numpy(), seed=seed)
        ctx.jac = torch.from_numpy(model.jacobian()["kappa_sma"]).to(p.device, p.dtype)
        return torch.from_numpy(model.kappa_sma().copy()).to(p.device, p.dtype)

    @staticmethod
    def backward(ctx, grad):
        return (ctx.jac * grad.unsqueeze(0)).flatten(1).sum(1), None, None, None, None
