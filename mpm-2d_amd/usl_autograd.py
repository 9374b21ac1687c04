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
a constant.

When x0 is on the GPU the particle state never leaves it: the initial
state, the outputs, the seeds and the gradients are copied device to device
between torch's tensors and the library's buffers (the seven constants
cross as host scalars).  torch's stream is synchronised before each call
into the library, and each call returns after its copies are complete.  CPU
tensors go through the host entry points.

Device pointers are shared only when torch and libmpm_hip.so run on one HIP
runtime.  torch's wheel carries its own libamdhip64.so with the soname of
the system one, so the process loads a single runtime when torch is
imported before the library is loaded (the library then binds to torch's);
loaded the other way round, torch's own copy loads beside it and
torch.cuda does not come up.  The bridge checks that the two resolve
hipMalloc to one function and raises otherwise.

One object keeps one forward run's checkpoints: a second UslRun.apply on the
same Adjoint before the first one's backward replaces them, so that
backward raises instead of returning the second run's gradients.
"""
import ctypes
import os

import numpy as np
import torch

import pympm

OUTPUTS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap"]
INPUTS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy"]   # differentiable initial state


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


def _dev(t, device):
    return t.detach().to(device=device, dtype=torch.float64).contiguous()


_shared = None


def _check_shared_runtime():
    """Raises unless torch and the library use one HIP runtime."""
    global _shared
    if _shared is None:
        ours = ctypes.cast(pympm.load().hipMalloc, ctypes.c_void_p).value
        path = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        theirs = ctypes.cast(ctypes.CDLL(path).hipMalloc, ctypes.c_void_p).value \
            if os.path.exists(path) else ours
        _shared = ours == theirs
    if not _shared:
        raise RuntimeError("usl_autograd: torch and libmpm_hip.so run on different HIP "
                           "runtimes; import torch before the library is first loaded")


class UslRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, adj, nsteps, m, params, x0, y0, vx0, vy0, v0, sxx0, sxy0, syy0):
        adj.set_material(**dict(zip(pympm.PARAMS, _host(params).tolist())))
        device = x0.device
        if device.type == "cuda":
            _check_shared_runtime()
            ins = [_dev(t, device) for t in (m, v0, x0, y0, vx0, vy0, sxx0, sxy0, syy0)]
            torch.cuda.current_stream(device).synchronize()
            adj.set_particles_device([t.data_ptr() for t in ins])
            adj.forward(int(nsteps))
            outs = tuple(torch.empty(adj.np, dtype=torch.float64, device=device)
                         for _ in OUTPUTS)
            for k, o in zip(OUTPUTS, outs):
                adj.get_device(k, o.data_ptr())
        else:
            adj.set_particles(dict(m=_host(m), v=_host(v0), x=_host(x0), y=_host(y0),
                                   vx=_host(vx0), vy=_host(vy0), sxx=_host(sxx0),
                                   sxy=_host(sxy0), syy=_host(syy0)))
            adj.forward(int(nsteps))
            outs = tuple(torch.from_numpy(adj.get(k)) for k in OUTPUTS)
        ctx.adj = adj
        ctx.generation = adj.generation
        ctx.device = device
        return outs

    @staticmethod
    def backward(ctx, *grads):
        adj = ctx.adj
        if adj.generation != ctx.generation:
            raise RuntimeError("UslRun.backward: the Adjoint was run forward again after this "
                               "node's forward; it keeps one run's checkpoints (use one "
                               "Adjoint per live forward run)")
        dev = ctx.device
        if dev.type == "cuda":
            seeds = {k: _dev(g, dev) for k, g in zip(OUTPUTS, grads) if g is not None}
            gi = {k: torch.empty(adj.np, dtype=torch.float64, device=dev) for k in INPUTS}
            torch.cuda.current_stream(dev).synchronize()
            gp = adj.backward_device({k: t.data_ptr() for k, t in seeds.items()},
                                     {k: t.data_ptr() for k, t in gi.items()})
        else:
            seed = {k: _host(g) for k, g in zip(OUTPUTS, grads) if g is not None}
            gp, gh = adj.backward(seed)
            gi = {k: torch.from_numpy(np.ascontiguousarray(gh[k])) for k in INPUTS}
        ctx.generation = -1      # the sweep consumed the run
        gparams = torch.tensor([gp[k] for k in pympm.PARAMS], dtype=torch.float64, device=dev)
        return (None, None, None, gparams) + tuple(gi[k] for k in INPUTS)
