"""Oracle: the reference's training semantics around the network's gradient.
TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

* ``clipnorm_adamw``: one step of OptimiserChain(ClipNorm(clip),
  AdamW(lr)) (train, src/training.jl:50), applied leaf by leaf as
  Optimisers.jl (0.4, the reference's compat bound, Project.toml) does: each
  parameter array's gradient is scaled by min(1, clip / |g|_2); AdamW then
  keeps m, v per element with the bias corrections 1 - beta^t of a shared
  step count t and moves w by lr (mhat / (sqrt(vhat) + eps) + lambda w).
* ``lr_schedule``: linear warm-up then cosine decay (src/training.jl:64-70).
* ``e_post``: the a-posteriori error of reduce! (src/les.jl:94-97).

Pinned by closed forms (tests/test_optim_oracle.py): Adam's first step is
lr * g / (|g| + eps) per element whatever the gradient's scale; two clipped
steps along one direction equal two steps with the unit-norm gradient, while
the unclipped pair does not; the clip acts per array; with a zero gradient the
decay alone gives w (1 - lr lambda); the schedule's end points and midpoint;
e_post's values at u = u_ref, u = 0 and u = 2 u_ref.
"""
import math

import numpy as np


class AdamWState:
    def __init__(self, nparams):
        self.m = np.zeros(nparams)
        self.v = np.zeros(nparams)
        self.t = 0


def clipnorm_adamw(w, g, state: AdamWState, bounds, lr, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0, clip=0.0):
    """w, g flat float64 arrays; bounds: array boundaries [b0=0, b1, ..., len].
    Returns the new w (state updated in place)."""
    g = np.array(g, np.float64)
    if clip > 0:
        for a in range(len(bounds) - 1):
            s = slice(bounds[a], bounds[a + 1])
            nrm = np.linalg.norm(g[s])
            if nrm > clip:
                g[s] *= clip / nrm
    state.t += 1
    state.m = b1 * state.m + (1 - b1) * g
    state.v = b2 * state.v + (1 - b2) * g * g
    mhat = state.m / (1 - b1**state.t)
    vhat = state.v / (1 - b2**state.t)
    return w - lr * (mhat / (np.sqrt(vhat) + eps) + weight_decay * w)


def lr_schedule(step, total, lr, warmup_frac=0.05):
    """learning rate of step 1..total"""
    nwarm = max(1, round(warmup_frac * total))
    if step < nwarm:
        return lr * step / nwarm
    return lr * (1 + math.cos(math.pi * (step - nwarm) / max(1, total - nwarm))) / 2


def e_post(u, u_ref, trunc):
    """||u - trunc(u_ref)|| / ||trunc(u_ref)|| over the stored half spectra
    (plain 2-norm of the stacked arrays); ``trunc`` is the 2/3 truncation"""
    r = np.stack([trunc(c) for c in u_ref])
    return float(np.linalg.norm((np.asarray(u) - r).ravel()) / np.linalg.norm(r.ravel()))
