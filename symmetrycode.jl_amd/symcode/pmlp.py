"""Host-side wrapper of the pointwise learned closures (sc_pmlp_* in
include/symcode.h): the reference's `mlp` ("conv" arch, src/nets.jl:247-298)
with its training pairs (create_dataloader, src/data.jl:705-741), its
solver-facing wrapper (fullchain, src/nets.jl:311-334) and its group average
(symmetrize_pointwise, :350-389, the "convsym" model); the tensor-basis
network (tbnn_net / tbnn, :416-583, with create_dataloader_tbnn); and the
group-equivariant network (equivariant_net, :102-238, "equi").

    x = [∇ū / (|∇ū| + eps), z]         z = standardised log Re_Δ (optional)
    y = symm(W4ᵀ gelu(W3ᵀ gelu(W2ᵀ gelu(W1ᵀ x + b1) + b2) + b3))
    τ = Δ² |∇ū|² y

Training runs in FP32 on voxel matrices [nvox][channels] (one row per grid
point); the closure widens the parameters and runs in FP64, as the reference
does with `ps |> f64` before inference.  torch supplies device buffers only;
every operation runs in libsymcode_hip.so.
"""
import ctypes
import math

import numpy as np
import torch

from ._lib import check_pmlp, lib
from .closures import Closure, gradient_phys
from .grid import Grid, _p
from . import solver as S

# hidden widths of the reference's capacity tiers for this architecture
# (default_tiers, src/experiment.jl:50-57)
TIERS = {"p120": (3, 5, 6), "p400": (6, 10, 14), "p1200": (10, 18, 32), "p3000": (22, 36, 44),
         "p8000": (44, 64, 64), "p16000": (64, 96, 96)}

# hidden widths of every architecture's tiers ("conv" = the mlp, "equi" =
# equivariant_net in channels per group copy, "tbnn")
ARCH_TIERS = {
    "conv": TIERS,
    "equi": {"p120": (1, 1, 1), "p400": (2, 2, 2), "p1200": (2, 3, 5), "p3000": (4, 5, 8), "p8000": (4, 8, 16)},
    "tbnn": {"p120": (3, 5, 6), "p400": (6, 10, 14), "p1200": (10, 18, 32), "p3000": (22, 36, 44),
             "p8000": (46, 64, 64)},
}

# |G| of the cube's symmetry group: `same_as_equi` multiplies the widths by it
GROUP_ORDER = {2: 8, 3: 48}


def _vp(t):
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class _PointwiseNet:
    """common device-side state of the pointwise networks (sc_pmlp_*)"""
    _create = "sc_pmlp_create"

    def __init__(self, D, widths, use_redelta, seed, nvox_max, device):
        if len(widths) != 3:
            raise ValueError("the network has three hidden layers")
        w = [int(c) for c in widths]
        net = ctypes.c_void_p()
        rc = getattr(lib(), self._create)(D, 1 if use_redelta else 0, w[0], w[1], w[2], seed, nvox_max,
                                          ctypes.byref(net))
        if rc != 0:
            if rc == 1:
                raise ValueError(f"{self._create}: D in {{2, 3}}, widths in [1, 1024], nvox_max * width < 2^31")
            raise RuntimeError(f"{self._create} failed: " + lib().sc_pmlp_last_error(None).decode())
        self.net = net
        info = (ctypes.c_int64 * 12)()
        check_pmlp(lib().sc_pmlp_info(net, info), net)
        (self.D, self.nin, c1, c2, c3, self.nout, self.nparams, self.nvox_max, self.kind,
         self.row, self.nreg, self.nnet) = list(info)
        self.widths = (c1, c2, c3)
        self.use_redelta = bool(use_redelta)
        self.device = torch.device("cuda", device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check_pmlp(lib().sc_pmlp_set_stream(net, ctypes.c_void_p(stream)), net)

    def __del__(self):
        try:
            if getattr(self, "net", None):
                lib().sc_pmlp_destroy(self.net)
                self.net = None
        except Exception:
            pass

    # --------------------------------------------------------- parameters
    def net_outputs(self):
        return self.nout

    def layer_shapes(self):
        """[(c_in, c_out)] of the four layers"""
        ch = [self.nin, *self.widths, self.net_outputs()]
        return [(ch[i], ch[i + 1]) for i in range(4)]

    def get_params(self):
        h = np.empty(self.nparams, np.float32)
        check_pmlp(lib().sc_pmlp_get_params(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)
        return h

    def set_params(self, h):
        h = np.ascontiguousarray(h, np.float32)
        if h.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        check_pmlp(lib().sc_pmlp_set_params(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)

    def get_grads(self):
        h = np.empty(self.nparams, np.float32)
        check_pmlp(lib().sc_pmlp_get_grads(self.net, h.ctypes.data_as(ctypes.c_void_p)), self.net)
        return h

    def zero_grad(self):
        check_pmlp(lib().sc_pmlp_zero_grad(self.net), self.net)

    def unflatten(self, h):
        """{'W1': (c_in, c_out), 'b1': (c_out,), ..., 'W4'} views of a flat vector"""
        out, o = {}, 0
        for l, (ci, co) in enumerate(self.layer_shapes()):
            out[f"W{l + 1}"] = h[o:o + ci * co].reshape(co, ci).T
            o += ci * co
            if l < 3:
                out[f"b{l + 1}"] = h[o:o + co]
                o += co
        return out

    # ------------------------------------------------------------ compute
    def features(self, GG, delta, tau=None, z=0.0):
        """create_dataloader's normalised pair from the physical gradient GG
        (D² components) and, when given, the physical stress tau (packed):
        returns x (np, row) [and y (np, nout)], FP32"""
        np_ = GG[0].numel()
        fp64 = GG.dtype == torch.float64
        if tau is not None and (tau.dtype != GG.dtype or tau.shape[0] != self.nout):
            raise ValueError("tau must be the packed stress in GG's dtype")
        if GG.shape[0] != self.D * self.D:
            raise ValueError("GG must hold D^2 components")
        x = torch.empty((np_, self.row), dtype=torch.float32, device=self.device)
        y = torch.empty((np_, self.nout), dtype=torch.float32, device=self.device) if tau is not None else None
        check_pmlp(lib().sc_pmlp_features(self.net, _vp(GG), _vp(tau) if tau is not None else None, int(fp64), np_,
                                          float(delta), float(z), _vp(x), _vp(y) if y is not None else None),
                   self.net)
        return x if y is None else (x, y)

    def _flat(self, t, c):
        t = t.reshape(-1, c)
        if t.dtype != torch.float32:
            raise ValueError("voxel matrices are FP32")
        return t

    def forward(self, x, y=None):
        x = self._flat(x, self.row)
        nv = x.shape[0]
        y = torch.empty((nv, self.nout), dtype=torch.float32, device=self.device) if y is None else y
        check_pmlp(lib().sc_pmlp_forward(self.net, _vp(x), _vp(y), nv), self.net)
        return y

    def backward(self, x, dy):
        """accumulate the parameter gradients of the last forward (on x)"""
        x = self._flat(x, self.row)
        dy = self._flat(dy, self.nout)
        check_pmlp(lib().sc_pmlp_backward(self.net, _vp(x), _vp(dy), x.shape[0]), self.net)

    def loss(self, y, target, kind="relative", dy=None):
        """(loss, dL/dy): "mse" = mean squared error, "relative" = the
        reference's Σ(ŷ - y)² / (Σy² + eps(Float32)) (src/training.jl:3-9)"""
        k = {"mse": 0, "relative": 1}[kind]
        y = self._flat(y, self.nout)
        target = self._flat(target, self.nout)
        dy = torch.empty_like(y) if dy is None else dy
        out = ctypes.c_double()
        check_pmlp(lib().sc_pmlp_loss(self.net, _vp(y), _vp(target), _vp(dy), y.shape[0], k, ctypes.byref(out)),
                   self.net)
        return out.value, dy

    def adamw(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, clip=0.0):
        """per-array ClipNorm(clip) (clip <= 0: none), then AdamW"""
        check_pmlp(lib().sc_pmlp_adamw(self.net, lr, beta1, beta2, eps, weight_decay, clip), self.net)

    def train_step(self, x, target, lr=1e-3, loss="relative", clip=0.0, weight_decay=0.0):
        """zero grad, forward, loss, backward, optimiser step; returns the loss"""
        self.zero_grad()
        y = self.forward(x)
        l, dy = self.loss(y, target, loss)
        self.backward(x, dy)
        self.adamw(lr, weight_decay=weight_decay, clip=clip)
        return l


class PointwiseMLP(_PointwiseNet):
    """mlp(setup, nchan; same_as_equi, use_redelta): three GELU hidden layers
    of widths ``widths`` (times |G| with ``same_as_equi``), a bias-free output
    layer of D == 3 ? 6 : 3 channels and the trace removal.  Parameters are
    glorot-uniform (zero biases) from ``seed``; the flat FP32 vector follows
    Lux's order [W1, b1, W2, b2, W3, b3, W4] with each W (c_in, c_out)
    column-major.  ``nvox_max`` bounds the voxels of one training call."""

    def __init__(self, D, widths=TIERS["p3000"], use_redelta=False, same_as_equi=False, seed=0,
                 nvox_max=1 << 21, device=0):
        w = [int(c) * (GROUP_ORDER.get(D, 1) if same_as_equi else 1) for c in widths]
        super().__init__(D, w, use_redelta, seed, nvox_max, device)


class TBNN(_PointwiseNet):
    """tbnn_net(setup, nchan; use_redelta) (src/nets.jl:535-556): three GELU
    hidden layers on the tb_kernel! invariants [+ z], a bias-free output layer
    of the basis coefficients (7 in 3D, 2 in 2D), contracted with the rows'
    deviatoric basis tensors.  Input rows: [invariants, z, basis (component
    fastest)], ``row`` wide."""
    _create = "sc_tbnn_create"

    def __init__(self, D, widths=TIERS["p3000"], use_redelta=False, seed=0, nvox_max=1 << 21, device=0):
        super().__init__(D, widths, use_redelta, seed, nvox_max, device)
        self.nbasis = 7 if D == 3 else 2

    def net_outputs(self):
        return self.nbasis


class EquivariantNet(_PointwiseNet):
    """equivariant_net(setup, nchan; use_redelta) (src/nets.jl:102-238): the
    pointwise network over |G| = 48 (2D: 8) copies of each hidden channel,
    its weights synthesised from compact trainables (get_weight_synthesis,
    src/symmetry.jl:156-197) and the sink's D² outputs made symmetric and
    trace-free; exactly equivariant under the cube's symmetries.  The
    parameter vector is the compact one (``compact_shapes``)."""
    _create = "sc_equi_create"

    def __init__(self, D, widths=(4, 5, 8), use_redelta=False, seed=0, nvox_max=1 << 18, device=0):
        super().__init__(D, widths, use_redelta, seed, nvox_max, device)

    def compact_shapes(self):
        D, (c1, c2, c3), nreg, nten = self.D, self.widths, self.nreg, self.D * self.D
        sh = [("lift.weight", (nten, c1)), ("lift.bias", (c1,))]
        if self.use_redelta:
            sh.append(("lift.weight_re", (1, c1)))
        return sh + [("mid_1.weight", (nreg, c2, c1)), ("mid_1.bias", (c2,)), ("mid_2.weight", (nreg, c3, c2)),
                     ("mid_2.bias", (c3,)), ("sink.weight", (nten, c3))]

    def layer_shapes(self):
        raise NotImplementedError("the trainables are the compact arrays: compact_shapes()")


def redelta_norm(pairs):
    """compute_redelta_norm (src/training.jl:180-186): (μ, σ) of log Re_Δ over
    the pairs, σ the corrected standard deviation"""
    logre = np.log([p.redelta for p in pairs])
    if logre.size < 2:
        raise ValueError("the standardisation needs at least two snapshots")
    return float(np.mean(logre)), float(np.std(logre, ddof=1))


def pointwise_pairs(net, pairs, g: Grid, norm=None):
    """create_dataloader (src/data.jl:705-741; for a TBNN
    create_dataloader_tbnn, :759-805) over Pairs (ubar, tau spectral
    on the LES grid, their Δ and Re_Δ): per snapshot x = ∇ū / (|∇ū| + eps)
    [+ z(Re_Δ)], y = τ / (Δ² |∇ū|² + eps) with ∇ū and τ 2/3-truncated in
    physical space; returns (X, Y) stacked over the snapshots, FP32.  ``norm``
    = (μ, σ) from redelta_norm, needed when the network has the Re_Δ input."""
    if net.use_redelta and norm is None:
        raise ValueError("a network with the Re_Delta input needs norm = redelta_norm(pairs)")
    np_ = g.phys_elems
    X = torch.empty((len(pairs) * np_, net.row), dtype=torch.float32, device=net.device)
    Y = torch.empty((len(pairs) * np_, net.nout), dtype=torch.float32, device=net.device)
    t = g.spacetensorfield()
    for i, p in enumerate(pairs):
        GG = gradient_phys(p.ubar, g)
        th = p.tau.clone()
        S.twothirds(th, g)
        S.to_phys(t, th, g)
        z = (math.log(p.redelta) - norm[0]) / norm[1] if net.use_redelta else 0.0
        x, y = net.features(GG, p.delta, t, z)
        X[i * np_:(i + 1) * np_] = x
        Y[i * np_:(i + 1) * np_] = y
    return X, Y


class PointwiseClosure(Closure):
    """fullchain(setup, net, ...; use_redelta, redelta_norm) as closure kind 7
    of the LES solver: τ = Δ² |∇ū|² net(∇ū / (|∇ū| + eps) [, z]) with the live
    field's z = (log(Δ² √⟨|∇ū|²⟩ / ν) - μ) / σ; for a TBNN, tbnn(...)'s
    τ = Δ² |∇ū|² Σ_b w_b B_b.  ``symmetrize`` wraps an mlp in
    symmetrize_pointwise (the "convsym" model).  It stays attached to the
    grid until another model is attached."""
    kind = 7

    def __init__(self, net, g: Grid, delta, visc=0.0, norm=None, symmetrize=False):
        super().__init__(g)
        if net.use_redelta and (norm is None or visc <= 0):
            raise ValueError("a network with the Re_Delta input needs visc > 0 and norm = (mu, sigma)")
        if symmetrize and net.kind != 0:
            raise ValueError("the group average applies to the mlp")
        self.net, self.delta, self.visc = net, float(delta), float(visc)
        self.mu, self.sigma = (float(norm[0]), float(norm[1])) if norm is not None else (0.0, 1.0)
        self.symmetrize = bool(symmetrize)
        self.attach()

    def attach(self):
        S._c(lib().sc_set_pmlp_closure(self.g.ctx, self.net.net, self.delta, self.visc, self.mu, self.sigma), self.g)
        if self.symmetrize:
            S._c(lib().sc_set_pmlp_symmetrize(self.g.ctx, 1), self.g)

    def __call__(self, u, GG=None):
        self.attach()
        tau = self.g.spacetensorfield()
        S._c(lib().sc_closure_pmlp(self.g.ctx, _p(tau), _p(u)), self.g)
        return tau


def create_pointwise_closure(net, g, delta, visc=0.0, norm=None, symmetrize=False):
    return PointwiseClosure(net, g, delta, visc, norm, symmetrize)


def symmetrize_pointwise(model: PointwiseClosure):
    """symmetrize_pointwise (src/nets.jl:350-389) of an mlp closure: the same
    parameters, exactly equivariant under the cube's symmetries"""
    return PointwiseClosure(model.net, model.g, model.delta, model.visc, (model.mu, model.sigma), symmetrize=True)
