"""The differentiable step: a witness current profile, its objective, a gradient, and an optimiser.

The witness is a flat-top beam whose macro-particle weights follow a piecewise-linear profile of S values
(deck key <beam>.pwl_weights, quarz_amd.configs.c4_deck).  The objective is the witness's relative rms energy
spread after one time step,

    J(v) = gamma_rms / gamma_mean,

with the weighted moments of the reference's beam log (src/Beam.cpp:1157-1192).  One evaluation is one full
step: beam deposit, the plasma sweep, the beam push.

Gradient, two ways:

  - "adjoint" (the default): the hand-written adjoint of the step (qz_sim_spread_grad): the sweep checkpointed
    every 64 slices, the reverse xi sweep recomputing each segment from its checkpoint, the adjoints of the
    beam kick, the beam deposit and the objective.  One forward and one reverse sweep per gradient.
  - "fd": central finite differences of the FP64 step, dJ/dv_s = (J(v + eps e_s) - J(v - eps e_s)) / (2 eps),
    the 2S + 1 evaluations run as one batched step (one persistent sweep launch, one workgroup per run): the
    check of the adjoint.

`spread` puts either behind torch.autograd, and `optimise` runs Adam on the profile with it.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Simulation, step_batched


def relative_spread(diag):
    """gamma_rms / gamma_mean of a beam_diag() dict."""
    return diag["gamma_rms"] / diag["gamma_mean"]


class SpreadObjective:
    """J(v) for many profiles v at once.  Holds `nruns` copies of the run made from `deck`, each with a
    snapshot of its freshly sampled beams: every evaluation rewinds a copy, sets its profile and steps it."""

    def __init__(self, deck, beam=1, nseg=16, nruns=None, overrides=None, method="adjoint", every=64):
        if method not in ("adjoint", "fd"):
            raise ValueError("method is 'adjoint' or 'fd'")
        self.beam, self.nseg, self.method, self.every = beam, nseg, method, every
        self.sims = [Simulation(deck, overrides) for _ in range(nruns or (1 if method == "adjoint" else 2 * nseg + 1))]
        for s in self.sims:
            s.snapshot()
        self.evaluations = 0
        self.last_batch_ms = 0.0
        # the per-run host work (profile weights, beam sums) of different runs touches different streams and
        # buffers; the library calls release the GIL, so a few threads run them side by side
        self._pool = ThreadPoolExecutor(max_workers=min(8, len(self.sims)))

    def close(self):
        self._pool.shutdown()
        for s in self.sims:
            s.close()
        self.sims = []

    def evaluate(self, profiles):
        """J of every profile (a sequence of S-vectors), as one batched step."""
        profiles = [np.asarray(v, dtype=np.float64) for v in profiles]
        if not 1 <= len(profiles) <= len(self.sims):
            raise ValueError(f"1..{len(self.sims)} profiles per evaluation")
        sims = self.sims[:len(profiles)]
        for v in profiles:
            if v.shape != (self.nseg,):
                raise ValueError(f"a profile has {self.nseg} values")

        def prepare(sv):
            sv[0].rewind()
            sv[0].set_beam_profile(self.beam, sv[1])

        list(self._pool.map(prepare, zip(sims, profiles)))
        self.last_batch_ms = step_batched(sims)
        self.evaluations += len(sims)
        return np.array(list(self._pool.map(lambda s: relative_spread(s.beam_diag(self.beam)), sims)))

    def value_and_grad(self, v, eps=1e-6, method=None):
        """J(v) and dJ/dv: by the adjoint (one step and its reverse sweep) or by central finite differences (one
        batched step of 2S + 1 runs)."""
        v = np.asarray(v, dtype=np.float64)
        if (method or self.method) == "adjoint":
            s = self.sims[0]
            s.rewind()
            s.set_beam_profile(self.beam, v)
            J, g, ms = s.spread_grad(self.beam, self.every)
            self.last_adjoint_ms = ms
            self.evaluations += 1
            return J, g
        if len(self.sims) < 2 * self.nseg + 1:
            raise ValueError("the gradient needs 2 S + 1 runs")
        prof = [v]
        for s in range(self.nseg):
            e = np.zeros(self.nseg)
            e[s] = eps
            prof += [v + e, v - e]
        J = self.evaluate(prof)
        return J[0], (J[1::2] - J[2::2]) / (2 * eps)


_FN = None


def _torch_function():
    global _FN
    if _FN is not None:
        return _FN
    import torch

    class SpreadFunction(torch.autograd.Function):
        """J(v) as a torch op; its backward is the gradient made with the forward (adjoint or finite
        differences, the objective's method)."""

        @staticmethod
        def forward(ctx, v, objective, eps):
            J, g = objective.value_and_grad(v.detach().cpu().double().numpy(), eps)
            ctx.save_for_backward(torch.from_numpy(g).to(v))
            return v.new_tensor(J)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None

    _FN = SpreadFunction
    return _FN


def spread(v, objective, eps=1e-6):
    """torch: J(v) for a 1-D double tensor v of S profile values, differentiable in v."""
    return _torch_function().apply(v, objective, eps)


def optimise(objective, v0, iters=50, lr=0.02, eps=1e-6, floor=0.05, log=None):
    """Adam on the profile values (kept >= floor).  Returns (v, history of J)."""
    import torch
    v = torch.tensor(np.asarray(v0, dtype=np.float64), requires_grad=True)
    opt = torch.optim.Adam([v], lr=lr)
    hist = []
    for it in range(iters):
        opt.zero_grad()
        J = spread(v, objective, eps)
        J.backward()
        opt.step()
        with torch.no_grad():
            v.clamp_(min=floor)
        hist.append(J.item())
        if log:
            log(it, hist[-1])
    return v.detach().numpy().copy(), hist
