"""Oracle: the pointwise MLP learned closure.  TEST INFRASTRUCTURE ONLY -- see
oracle/__init__.py.

* ``features``: create_dataloader's normalised pair (src/data.jl:729-741),
  x = G / (sqrt(A2) + eps) [+ z], y = tau / (Delta^2 A2 + eps), A2 = sum G^2.
* ``forward`` / ``backward``: mlp (src/nets.jl:247-298) -- PointwiseConv
  layers (src/nets.jl:37-81: y = sigma(W^T x + b) per voxel, W (c_in, c_out))
  with NNlib's tanh-form gelu on the three hidden layers, a bias-free output
  layer and the `symm` trace removal -- and the hand-written gradient of a
  loss L(y) with respect to every parameter array.
* ``relative_loss``: create_loss (src/training.jl:3-9).
* ``fullchain``: the solver-facing model (src/nets.jl:311-334) with the live
  standardised log Re_Delta input (append_redelta, src/data.jl:689-694).
* ``flat`` / ``unflat``: Lux's parameter order [W1, b1, W2, b2, W3, b3, W4],
  each W column-major (c_in fastest).

Pinned (tests/test_pmlp_oracle.py): forward and every gradient against
torch.autograd of the same network built from torch.nn.functional's
gelu(approximate="tanh"); the trace removal is idempotent and leaves a
trace-free output; fullchain is homogeneous of degree two in G, tau(aG) =
a^2 tau(G), without the Re_Delta input, and with it the input z moves by
log(a) / sigma; the features have unit row norm up to eps and reproduce tau.
"""
import math

import numpy as np

C0 = math.sqrt(2.0 / math.pi)


def gelu(z):
    """NNlib's gelu: z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))"""
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + 0.044715 * z**3)))


def gelu_grad(z):
    t = np.tanh(C0 * (z + 0.044715 * z**3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C0 * (1.0 + 3 * 0.044715 * z * z)


def tracefree(y, D):
    """the `symm` layer on rows [nv][nt]: xx, yy[, zz] lose their mean"""
    y = np.array(y, np.float64)
    k = 3 if D == 3 else 2
    y[:, :k] -= y[:, :k].mean(axis=1, keepdims=True)
    return y


def layer_shapes(D, widths, use_redelta):
    nin = D * D + (1 if use_redelta else 0)
    nt = 6 if D == 3 else 3
    ch = [nin, *widths, nt]
    return [(ch[i], ch[i + 1]) for i in range(4)]


def unflat(h, shapes):
    """flat Lux-order vector -> {'W1': (c_in, c_out), 'b1', ..., 'W4'}"""
    out, o = {}, 0
    for l, (ci, co) in enumerate(shapes):
        out[f"W{l + 1}"] = np.asarray(h[o:o + ci * co], np.float64).reshape(co, ci).T.copy()
        o += ci * co
        if l < 3:
            out[f"b{l + 1}"] = np.asarray(h[o:o + co], np.float64).copy()
            o += co
    return out


def flat(p, shapes):
    parts = []
    for l in range(4):
        parts.append(p[f"W{l + 1}"].T.ravel())
        if l < 3:
            parts.append(p[f"b{l + 1}"])
    return np.concatenate(parts)


def array_bounds(shapes):
    """boundaries of the seven parameter arrays in the flat vector"""
    b, o = [0], 0
    for l, (ci, co) in enumerate(shapes):
        o += ci * co
        b.append(o)
        if l < 3:
            o += co
            b.append(o)
    return b


def forward(p, x, D):
    """x [nv][nin] -> (y [nv][nt], cache)"""
    a = np.asarray(x, np.float64)
    zs, acts = [], [a]
    for l in range(1, 4):
        z = a @ p[f"W{l}"] + p[f"b{l}"]
        a = gelu(z)
        zs.append(z)
        acts.append(a)
    y = tracefree(a @ p["W4"], D)
    return y, (zs, acts)


def backward(p, cache, dy, D):
    """gradients {'W1', 'b1', ..., 'W4'} of sum(dy * y) for the cached forward"""
    zs, acts = cache
    d = tracefree(dy, D)  # the trace removal is symmetric
    g = {"W4": acts[3].T @ d}
    d = d @ p["W4"].T
    for l in range(3, 0, -1):
        d = d * gelu_grad(zs[l - 1])
        g[f"W{l}"] = acts[l - 1].T @ d
        g[f"b{l}"] = d.sum(axis=0)
        d = d @ p[f"W{l}"].T
    return g


def relative_loss(yhat, y):
    """sum (yhat - y)^2 / (sum y^2 + eps(Float32)) and its gradient"""
    den = float(np.sum(np.asarray(y, np.float64) ** 2)) + float(np.finfo(np.float32).eps)
    e = np.asarray(yhat, np.float64) - y
    return float(np.sum(e * e)) / den, 2 * e / den


def mse_loss(yhat, y):
    e = np.asarray(yhat, np.float64) - y
    return float(np.mean(e * e)), 2 * e / e.size


def a2_of(GG):
    """pointwise |G|^2 = sum_c G_c^2 of the component-major gradient"""
    a2 = np.zeros(GG.shape[1:])
    for c in range(GG.shape[0]):
        a2 = a2 + np.asarray(GG[c], np.float64) ** 2
    return a2


def features(GG, tau, delta, z=None, eps=np.finfo(np.float64).eps):
    """(x [np][nin], y [np][nt]) in float64: x = G / (sqrt(A2) + eps) [, z],
    y = tau / (Delta^2 A2 + eps); GG, tau component-major physical fields"""
    a2 = a2_of(GG).ravel()
    G = np.stack([np.asarray(g, np.float64).ravel() for g in GG], axis=1)
    x = G / (np.sqrt(a2) + eps)[:, None]
    if z is not None:
        x = np.concatenate([x, np.full((x.shape[0], 1), z)], axis=1)
    y = None
    if tau is not None:
        t = np.stack([np.asarray(c, np.float64).ravel() for c in tau], axis=1)
        y = t / (delta**2 * a2 + eps)[:, None]
    return x, y


def redelta_z(GG, delta, visc, mu, sigma):
    """z = (log(Delta^2 sqrt(mean A2) / visc) - mu) / sigma of the field"""
    re = delta**2 * math.sqrt(float(a2_of(GG).mean())) / visc
    return (math.log(re) - mu) / sigma


def fullchain(p, GG, delta, D, eps=np.finfo(np.float64).eps, redelta=None):
    """tau (component-major, GG's grid shape) = Delta^2 A2 net(x); redelta =
    (visc, mu, sigma) for a network with the Re_Delta input"""
    z = None if redelta is None else redelta_z(GG, delta, *redelta)
    x, _ = features(GG, None, delta, z, eps)
    y, _ = forward(p, x, D)
    s = delta**2 * a2_of(GG).ravel()
    tau = y * s[:, None]
    return np.stack([tau[:, c].reshape(GG.shape[1:]) for c in range(tau.shape[1])])
