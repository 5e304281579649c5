"""Training the learned closure on data the GPU DNS generates.

A-priori (the reference's route): march the DNS, form the filtered pairs
(ū, τ) at each save time and filter width (create_data, src/data.jl:110-220;
sfs!, :625-642), normalise them into network batches (create_dataloader,
:705-757) and fit them: ``reference_train`` has the reference's semantics
(relative loss, ClipNorm + AdamW, warm-up + cosine; train,
src/training.jl:3-9,45-109), ``train`` exposes each choice.

A-posteriori (no reference counterpart): the loss of an LES rollout against
the filtered-DNS trajectory, differentiated through the hand-written adjoint
of the RK4 LES step (sc_les_rk4_vjp), whose learned-closure stages push the
cotangent through the network and accumulate its parameter gradient.

Scales: the network sees ū / s_in and predicts τ / s_out with s_in the
velocity r.m.s. and s_out = Δ²⟨|∇ū|²⟩ of the training snapshots at that Δ
(the reference normalises pointwise by |∇ū| because its nets are pointwise;
this net sees a 9³ neighbourhood, so global scales keep the closure a fixed
function of ū).  The extra input channel is the standardised log Re_Δ
(append_redelta, src/data.jl:689-703; compute_redelta_norm,
src/training.jl:179-193).
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import solver as S
from .closures import LearnedClosure, les_rk4, les_rk4_vjp
from .filtering import sfs
from .grid import Grid


@dataclass
class Pair:
    ubar: torch.Tensor  # truncated filtered velocity, spectral, LES grid
    tau: torch.Tensor   # trace-free sub-filter stress, spectral, LES grid
    delta: float
    redelta: float


def filter_reynolds(ubar, g: Grid, visc, delta):
    """filter_reynolds (src/solver.jl:768-774): Re_Δ = Δ² √⟨|∇ū|²⟩ / ν with
    ⟨|∇ū|²⟩ = Σ_ij getenergy(2/3 (i k_j ū_i)): the gradient is truncated to
    the 2/3 box on the device, whatever ū holds."""
    return S.filter_reynolds(ubar, g, visc, delta)


def create_data(u, g_dns: Grid, g_les: Grid, visc, filters, nsnap, dt_save, cfl=0.3, shells=None):
    """March the DNS ``u`` in place (CFL RK4; band forcing as set on g_dns)
    and, at nsnap save times dt_save apart (the first is the current state),
    form the pairs for every filter factor (Δ = Δf · l / n_les).  With
    ``shells`` (energy_shells output) the shell clamp runs after each step
    (maintain_shell_energy!, src/solver.jl:1146-1152).  Returns {Δf: [Pair]}."""
    pairs = {f: [] for f in filters}
    for i in range(nsnap):
        if i > 0:
            t = 0.0
            while t < dt_save * (1 - 1e-12):
                t += S.rk4_cfl(u, g_dns, visc, cfl, dt_save - t)
                if shells is not None:
                    S.maintain_shell_energy(u, shells, g_dns)
        for f in filters:
            delta = f * g_les.l / g_les.n
            ubar, tau = sfs(u, g_dns, g_les, delta)
            pairs[f].append(Pair(ubar, tau, delta, filter_reynolds(ubar, g_les, visc, delta)))
    return pairs


def save_pairs(path, pairs, g_les: Grid, delta_factor, visc):
    """The pairs of one filter width as the reference's fieldsfile
    (src/data.jl:189-191,207-211): inputs = [map(Array, ū)], outputs =
    [map(Array, τ)] in the grid's precision and the reference's RFFT array
    layout, redelta, Δ, Δ_factor, visc (written atomically)."""
    from . import jld2

    if not pairs:
        raise ValueError("no pairs to save")
    deltas = {p.delta for p in pairs}
    if len(deltas) != 1:
        raise ValueError("a fields file holds one filter width")
    keys_v = jld2.VECTOR_KEYS[:g_les.D]
    keys_t = jld2.TENSOR_KEYS if g_les.D == 3 else ("xx", "yy", "xy")
    inputs = [jld2.field_dict(g_les.download_spectral(p.ubar), keys_v) for p in pairs]
    outputs = [jld2.field_dict(g_les.download_spectral(p.tau), keys_t) for p in pairs]
    jld2.save_fields(path, inputs, outputs, [p.redelta for p in pairs], pairs[0].delta, delta_factor, visc)


def load_pairs(path, g_les: Grid):
    """A fieldsfile (the reference's or save_pairs') as device Pairs on g_les"""
    from . import jld2

    inputs, outputs, redelta, delta = jld2.load(path, "inputs", "outputs", "redelta", "Δ")
    keys_v = jld2.VECTOR_KEYS[:g_les.D]
    keys_t = jld2.TENSOR_KEYS if g_les.D == 3 else ("xx", "yy", "xy")
    shape = g_les.ref_spec_shape(1)[1:]
    pairs = []
    for ub, tau, rd in zip(inputs, outputs, np.atleast_1d(redelta)):
        if tuple(ub[keys_v[0]].shape) != shape:
            raise ValueError(f"{path}: arrays of shape {ub[keys_v[0]].shape}, the grid needs {shape}")
        pairs.append(Pair(g_les.upload_spectral(np.stack([ub[k] for k in keys_v])),
                          g_les.upload_spectral(np.stack([tau[k] for k in keys_t])), float(delta), float(rd)))
    return pairs


def velocity_rms(u, g: Grid):
    return math.sqrt(S.spectraldot(u, u, g).real / g.D)


@dataclass
class Normalization:
    s_in: dict = field(default_factory=dict)   # Δ -> velocity scale
    s_out: dict = field(default_factory=dict)  # Δ -> stress scale Δ²⟨|∇ū|²⟩
    mu: float = 0.0                            # log Re_Δ standardisation
    sigma: float = 1.0

    @staticmethod
    def of(pairs, g: Grid):
        """from a list of Pairs (any mix of Δ)"""
        nrm = Normalization()
        for d in sorted({p.delta for p in pairs}):
            ps = [p for p in pairs if p.delta == d]
            nrm.s_in[d] = float(np.mean([velocity_rms(p.ubar, g) for p in ps]))
            nrm.s_out[d] = d**2 * float(np.mean([S.get_dissipation(p.ubar, 1.0, g) for p in ps]))
        logre = np.log([p.redelta for p in pairs])
        nrm.mu = float(np.mean(logre))
        sd = float(np.std(logre))
        nrm.sigma = sd if sd > 1e-3 else 1.0
        return nrm

    def z(self, redelta):
        return (math.log(redelta) - self.mu) / self.sigma

    def closure(self, net, g: Grid, delta, redelta):
        """the trained net as an LES model at filter width Δ"""
        return LearnedClosure(net, g, self.s_in[delta], self.s_out[delta], self.z(redelta))


def pack_pairs(net, g: Grid, pairs, nrm: Normalization, x=None, y=None):
    """Network batch (x bf16 (B, n³, cin_p), target fp32 (B, n³, 6)) of the
    pairs: x = [C2R(ū) / s_in, z(Re_Δ)], y = C2R(τ) / s_out."""
    B = len(pairs)
    x = net.new_input(B) if x is None else x
    y = torch.empty((B, net.nvox, net.cout), dtype=torch.float32, device=net.device) if y is None else y
    v = g.spacevectorfield()
    t = g.spacetensorfield()
    for b, p in enumerate(pairs):
        S.to_phys(v, p.ubar, g)
        S.to_phys(t, p.tau, g)
        net.pack_input(v, scale=1.0 / nrm.s_in[p.delta], extra=nrm.z(p.redelta), x=x[b:b + 1])
        net.pack_target(t, scale=1.0 / nrm.s_out[p.delta], out=y[b:b + 1])
    return x, y


def lr_schedule(step, total, lr, warmup_frac=0.05):
    """learning rate of step 1..total (train, src/training.jl:64-70): linear
    warm-up over nwarm = max(1, round(warmup_frac * total)) steps, then
    cosine decay to zero at step total"""
    nwarm = max(1, round(warmup_frac * total))
    if step < nwarm:
        return lr * step / nwarm
    return lr * (1 + math.cos(math.pi * (step - nwarm) / max(1, total - nwarm))) / 2


def train(net, x, y, nsteps, lr=1e-3, log_every=0, loss="mse", clip=0.0, schedule="constant", warmup_frac=0.05,
          weight_decay=0.0, batchsize=None):
    """Fixed-budget training of the network on (x, y); returns the loss
    history.  ``loss``: "mse" or the reference's "relative"; ``clip``: per-array
    ClipNorm before AdamW (0: none); ``schedule``: "constant" or the
    reference's "cosine" (lr_schedule).  ``batchsize`` cycles through
    contiguous minibatches of the samples in order (None: the whole batch every
    step).  The reference's settings (train, src/training.jl:45-109;
    default_schedule, src/experiment.jl:23-36) are ``reference_train``."""
    if schedule not in ("constant", "cosine"):
        raise ValueError("schedule must be 'constant' or 'cosine'")
    B = x.shape[0]
    bs = B if batchsize is None else int(batchsize)
    if not 1 <= bs <= B:
        raise ValueError("batchsize out of range")
    nb = (B + bs - 1) // bs
    losses = []
    for it in range(nsteps):
        eta = lr if schedule == "constant" else lr_schedule(it + 1, nsteps, lr, warmup_frac)
        j = it % nb
        xb, yb = (x, y) if nb == 1 else (x[j * bs:(j + 1) * bs], y[j * bs:(j + 1) * bs])
        losses.append(net.train_step(xb, yb, eta, loss=loss, clip=clip, weight_decay=weight_decay))
        if log_every and (it % log_every == 0 or it == nsteps - 1):
            print(f"train step {it}: lr {eta:.3e} loss {losses[-1]:.4e}", flush=True)
    return losses


def reference_train(net, x, y, nsteps, lr=1e-3, batchsize=None, log_every=0):
    """train with the reference's semantics: relative loss, ClipNorm(1.0) then
    AdamW(lr), 5 % linear warm-up then cosine decay (src/training.jl:3-9,
    45-70; src/experiment.jl:27-30)"""
    return train(net, x, y, nsteps, lr, log_every, loss="relative", clip=1.0, schedule="cosine", warmup_frac=0.05,
                 batchsize=batchsize)


def apriori_metrics(tau_model, tau_ref):
    """relerr and crosscor of compute_sfs_stats (src/analysis.jl:116-121) on
    physical stresses (torch tensors)"""
    a = tau_ref.double().flatten()
    b = tau_model.double().flatten()
    relerr = float(torch.linalg.norm(b - a) / torch.linalg.norm(a))
    aa, bb = a - a.mean(), b - b.mean()
    crosscor = float(torch.dot(aa, bb) / torch.sqrt(torch.dot(aa, aa) * torch.dot(bb, bb)))
    return relerr, crosscor


def e_post(u, u_ref, g: Grid):
    """a-posteriori error ‖u − ū‖ / ‖ū‖ of an LES state against the filtered
    DNS, ū 2/3-truncated first, with the plain norm of the stored half
    spectra (reduce!, src/les.jl:94-97)"""
    r = u_ref.clone()
    S.twothirds(r, g)
    e = g.vectorfield()
    S.lincomb(e, u, 1.0, r, -1.0, g)
    return math.sqrt(S.norm2(e, g) / S.norm2(r, g))


# ------------------------------------------------------------ a-posteriori
def rollout_loss(u0, refs, dt, g: Grid, visc, model, states=None):
    """L = Σ_t ⟨u_t − r_t, u_t − r_t⟩ / (2 nt) over nt = len(refs) fixed-dt
    RK4 steps of les! from u0 (RFFT-weighted inner product).  ``states``
    (list) receives u_0..u_nt for the backward pass."""
    u = u0.clone()
    if states is not None:
        states.append(u0)
    e = g.vectorfield()
    L = 0.0
    for r in refs:
        les_rk4(u, dt, g, model, visc)
        if states is not None:
            states.append(u.clone())
        S.lincomb(e, u, 1.0, r, -1.0, g)
        L += S.spectraldot(e, e, g).real
    return L / (2 * len(refs))


def rollout_gradient(u0, refs, dt, g: Grid, visc, model):
    """(L, dL/du0) of rollout_loss; with a LearnedClosure its network's
    parameter gradients accumulate (zero them first).  Checkpoints every step
    and recomputes the stages inside each step's transpose."""
    states = []
    L = rollout_loss(u0, refs, dt, g, visc, model, states)
    nt = len(refs)
    ubar = g.vectorfield()
    out = g.vectorfield()
    for t in range(nt, 0, -1):
        S.lincomb(ubar, ubar, 1.0, states[t], 1.0 / nt, g)
        S.lincomb(ubar, ubar, 1.0, refs[t - 1], -1.0 / nt, g)
        les_rk4_vjp(states[t - 1], ubar, dt, g, model, visc, out=out)
        ubar, out = out, ubar
    return L, ubar


def aposteriori_step(net, u0, refs, dt, g: Grid, visc, model, lr=1e-3, clip=0.0):
    """one a-posteriori gradient step: zero grads, rollout + adjoint, then
    Adam (with per-array ClipNorm when clip > 0)"""
    net.zero_grad()
    L, _ = rollout_gradient(u0, refs, dt, g, visc, model)
    net.adamw(lr, clip=clip)
    return L


def regrid(u, g_from: Grid, g_to: Grid):
    """a spectral field moved between grids of equal n and different precision"""
    return g_to.upload_spectral(g_from.download_spectral(u))


def timed(fn, *args, reps=1, **kw):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn(*args, **kw)
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) / reps
