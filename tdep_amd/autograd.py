"""Differentiable three-phonon linewidths and RTA thermal conductivity.

`ThreePhononLinewidths` and `KappaRTA` are torch.autograd.Functions of the third-order
force constants m [ntrip, 3, 3, 3] (the tensors of infile.forceconstant_thirdorder in
Ha/Bohr^3, in the order of the engine's fc3 list).  Their forward runs the GPU scattering
pass of libtc3 (compute_threephonon_scattering, scattering_threephonon.f90:2-217, and
get_kappa, kappa.f90:28-75); their backward calls the library's hand-written adjoint
(tc3_adjoint_gamma / tc3_adjoint_kappa_rta), which recomputes |Psi|^2 instead of keeping
it.  The harmonic part (force constants of second order, q-mesh, eigenvectors) is fixed by
the engine and is not differentiated.

Gradients with respect to a parametrisation of Phi3, such as irreducible coefficients x
with m = B x (type_forcemap_returntensors.f90:294-335 builds such a B), follow by composing
in torch: `ThreePhononLinewidths.apply((B @ x).reshape(-1, 3, 3, 3), engine, T)`.

The tensors cross the boundary on the host in float64; the library moves them to HBM.
"""
import numpy as np
import torch

from .engine import Engine  # noqa: F401  (the type the functions expect)


def _forward(engine, m, temperatures, precision, integrationtype, sigma, classical):
    fc3 = dict(engine.fc3)
    fc3["m"] = m.detach().to("cpu", torch.float64).numpy().copy()
    engine.set_fc3(fc3)
    engine.scattering_rates(temperatures, precision=precision, integrationtype=integrationtype,
                            sigma=sigma, classical=classical)
    token = object()
    engine._adjoint_token = token
    return token


class ThreePhononLinewidths(torch.autograd.Function):
    """linewidths [nT, nirr, nb] (Hartree) = f(m); backward gives dL/dm."""

    @staticmethod
    def forward(ctx, m, engine, temperatures, precision="fp64", integrationtype=2, sigma=1.0,
                classical=False):
        opts = (np.atleast_1d(np.asarray(temperatures, np.float64)), precision, integrationtype, sigma,
                classical)
        ctx.token = _forward(engine, m, *opts)
        ctx.engine, ctx.opts = engine, opts
        ctx.save_for_backward(m)
        lw = np.stack([engine.linewidths(i)[0] for i in range(len(opts[0]))])
        return torch.from_numpy(lw).to(m.device, m.dtype)

    @staticmethod
    def backward(ctx, grad):
        (m,) = ctx.saved_tensors
        eng = ctx.engine
        if getattr(eng, "_adjoint_token", None) is not ctx.token:
            # the engine ran another forward since: restore this one's state
            ctx.token = _forward(eng, m, *ctx.opts)
        g = eng.adjoint_gamma(grad.detach().to("cpu", torch.float64).numpy(), precision=ctx.opts[1])
        return torch.from_numpy(g).to(m.device, m.dtype), None, None, None, None, None, None


class KappaRTA(torch.autograd.Function):
    """single-mode kappa [nT, 3, 3] (atomic units, symmetrised) = f(m); backward gives dL/dm."""

    @staticmethod
    def forward(ctx, m, engine, temperatures, precision="fp64", integrationtype=2, sigma=1.0,
                classical=False):
        opts = (np.atleast_1d(np.asarray(temperatures, np.float64)), precision, integrationtype, sigma,
                classical)
        ctx.token = _forward(engine, m, *opts)
        ctx.engine, ctx.opts = engine, opts
        ctx.save_for_backward(m)
        ks, _, _ = engine.kappa(maxiter=0, offdiag=False)
        return torch.from_numpy(ks.copy()).to(m.device, m.dtype)

    @staticmethod
    def backward(ctx, grad):
        (m,) = ctx.saved_tensors
        eng = ctx.engine
        if getattr(eng, "_adjoint_token", None) is not ctx.token:
            ctx.token = _forward(eng, m, *ctx.opts)
        g = eng.adjoint_kappa_rta(grad.detach().to("cpu", torch.float64).numpy(), precision=ctx.opts[1])
        return torch.from_numpy(g).to(m.device, m.dtype), None, None, None, None, None, None


def linewidths(engine, m, temperatures, **kw):
    return ThreePhononLinewidths.apply(m, engine, temperatures, kw.get("precision", "fp64"),
                                       kw.get("integrationtype", 2), kw.get("sigma", 1.0),
                                       kw.get("classical", False))


def kappa_rta(engine, m, temperatures, **kw):
    return KappaRTA.apply(m, engine, temperatures, kw.get("precision", "fp64"),
                          kw.get("integrationtype", 2), kw.get("sigma", 1.0), kw.get("classical", False))
