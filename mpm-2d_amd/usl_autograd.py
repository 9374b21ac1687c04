"""torch.autograd bridge for an n-step USL run (mpm_hip_adjoint_*).

    adj = pympm.Adjoint(cfg, particles, checkpoint_every=20)
    outs = UslRun.apply(adj, nsteps, m, params, x0, y0, vx0, vy0, v0,
                        sxx0, sxy0, syy0)

`params` is a float64 tensor [7] = (E, nu, mu_s, mu_2, I_0, grains_d,
grains_rho) of the g_local_mu2 law (materialsrc/g_local_mu2.c:30-41,
139-140); the outputs are the final x, y, vx, vy, v, Txx, Txy, Tyy, Tzz,
gammap (OUTPUTS), float64 tensors on the device of x0.  The backward pass
is the library's reverse sweep; gradients flow to `params` and to the
initial positions, velocities, volumes and in-plane stresses.  The mass is
a constant.  Tensors cross the boundary through host memory: the run's
state lives in the library's own device buffers.
"""
import numpy as np
import torch

import pympm

OUTPUTS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap"]


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class UslRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, adj, nsteps, m, params, x0, y0, vx0, vy0, v0, sxx0, sxy0, syy0):
        adj.set_material(**dict(zip(pympm.PARAMS, _host(params).tolist())))
        adj.set_particles(dict(m=_host(m), v=_host(v0), x=_host(x0), y=_host(y0),
                               vx=_host(vx0), vy=_host(vy0), sxx=_host(sxx0),
                               sxy=_host(sxy0), syy=_host(syy0)))
        adj.forward(int(nsteps))
        ctx.adj = adj
        ctx.device = x0.device
        return tuple(torch.from_numpy(adj.get(k)).to(x0.device) for k in OUTPUTS)

    @staticmethod
    def backward(ctx, *grads):
        seed = {k: _host(g) for k, g in zip(OUTPUTS, grads) if g is not None}
        gp, gi = ctx.adj.backward(seed)
        dev = ctx.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        gparams = torch.tensor([gp[k] for k in pympm.PARAMS], dtype=torch.float64, device=dev)
        return (None, None, None, gparams, t(gi["x"]), t(gi["y"]), t(gi["vx"]), t(gi["vy"]),
                t(gi["v"]), t(gi["txx"]), t(gi["txy"]), t(gi["tyy"]))
