"""Oracle: les! with the learned convolutional closure (closure kind 4) and
the transpose of its right-hand side and RK4 / Wray3 steps, including the
parameter gradient.  TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

Closure (solver-facing wrapper of the network, after fullchain,
src/nets.jl:311-334, with global scales):
    v   = C2R(2/3 u)                      (physical velocity, [z, y, x])
    x   = [v / s_in, extra]               (channels-last, voxel x + n (y + n z))
    tau = s_out * net(x)[:, :6]           (xx, yy, zz, xy, yz, zx)
    les!(u) = tensordivergence(stress(u) + to_spec(tau)) (+ band forcing)
(src/les.jl:4-33).

Transpose (RFFT-weighted inner product in spectral space = physical mean):
with g_ij = C2R(i k_j w_i), the stress cotangent is tau_bar = (g_xx, g_yy,
g_zz, g_xy + g_yx, g_yz + g_zy, g_zx + g_xz) under the physical mean; under
plain sums over voxels (the network's own backward) y_bar = s_out tau_bar /
n^3, and the velocity part of J^T w is T to_spec(n^3 x_bar[:, :3] / s_in).

Pinned by central finite differences in FP64 (tests/test_les_cnn_oracle.py):
of the rollout loss along random parameter directions and along a random
initial-condition direction.
"""
import numpy as np
import torch

from . import adjoint as A
from . import cnn as OC
from . import spectral as S

PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0))


class Net:
    """the network's parameters and the closure's scales"""

    def __init__(self, Ws, bs, cin, s_in, s_out, extra=0.0, emulate=True, dtype=torch.float32, tracefree=False):
        self.Ws, self.bs, self.cin = Ws, bs, cin
        self.cin_p = Ws[0].shape[2]
        self.s_in, self.s_out, self.extra = s_in, s_out, extra
        self.emulate, self.dtype, self.tracefree = emulate, dtype, tracefree

    def perturbed(self, dW, db, eps):
        Ws = [W + eps * d for W, d in zip(self.Ws, dW)]
        bs = [b + eps * d for b, d in zip(self.bs, db)]
        return Net(Ws, bs, self.cin, self.s_in, self.s_out, self.extra, self.emulate, self.dtype, self.tracefree)


def _input(u, g: S.Grid, net: Net):
    n = g.n
    v = [S.to_phys(S.twothirds(u[i], g), g) for i in range(3)]
    x = np.zeros((1, n**3, net.cin_p))
    for m in range(3):
        x[0, :, m] = v[m].ravel() / net.s_in
    if net.cin > 3:
        x[0, :, 3] = net.extra
    return x


def cnn_tau(u, g: S.Grid, net: Net):
    """tau (6, n, n, n) and the network's stored forward state"""
    n = g.n
    y, store = OC.forward(net.Ws, net.bs, torch.as_tensor(_input(u, g, net)), n, net.emulate, net.dtype,
                          tracefree_out=net.tracefree)
    y = y.double().numpy()
    tau = np.stack([net.s_out * y[0, :, c].reshape(n, n, n) for c in range(6)])
    return tau, store


def les_rhs(u, g: S.Grid, visc, net: Net, forcing=None):
    """les! with the learned closure (+ band forcing when given)"""
    tau, _ = cnn_tau(u, g, net)
    sig = S.stress(u, visc, g) + np.stack([S.to_spec(tau[i], g) for i in range(6)])
    du = S.tensordivergence(sig, g)
    if forcing is not None:
        du = du + S.band_forcing(u, g, *forcing)
    return du.astype(u.dtype)


def les_rhs_vjp(u, w, g: S.Grid, visc, net: Net, forcing=None, grads=None):
    """J(u)^T w; the parameter gradient (list of weight grads, list of bias
    grads) is added to ``grads`` when given"""
    n = g.n
    out = A.rhs_vjp(u, w, g, visc, forcing)
    _, store = cnn_tau(u, g, net)
    k = g.kfull
    gr = [[S.to_phys(1j * k[j] * w[i], g) for j in range(3)] for i in range(3)]
    tbar = [gr[i][j] if i == j else gr[i][j] + gr[j][i] for (i, j) in PAIRS]
    cout_p = net.Ws[3].shape[1]
    dy = np.zeros((1, n**3, cout_p))
    for c in range(6):
        dy[0, :, c] = net.s_out * tbar[c].ravel() / n**3
    gW, gb, dx = OC.backward(store, torch.as_tensor(dy), n, net.emulate)
    dx = dx.double().numpy()
    for m in range(3):
        a = n**3 * dx[0, :, m].reshape(n, n, n) / net.s_in
        out[m] = out[m] + S.twothirds(S.to_spec(a, g), g)
    if grads is not None:
        for l in range(4):
            grads[0][l] += gW[l].double().numpy()
        for l in range(3):
            grads[1][l] += gb[l].double().numpy()
    return out.astype(u.dtype)


def zero_grads(net: Net):
    return ([np.zeros(W.shape) for W in net.Ws], [np.zeros(np.shape(b)) for b in net.bs])


def rk4(u, dt, g: S.Grid, visc, net: Net, forcing=None):
    return S.rk4(lambda v: les_rhs(v, g, visc, net, forcing), u, dt, g)


def rk4_vjp(u0, ubar, dt, g: S.Grid, visc, net: Net, forcing=None, grads=None):
    return A.rk4_vjp(u0, ubar, dt, g, visc, forcing,
                     rhs=lambda v: les_rhs(v, g, visc, net, forcing),
                     vjp=lambda v, w: les_rhs_vjp(v, w, g, visc, net, forcing, grads))


def wray3_vjp(u0, ubar, dt, g: S.Grid, visc, net: Net, forcing=None, grads=None):
    return A.wray3_vjp(u0, ubar, dt, g, visc, forcing,
                       rhs=lambda v: les_rhs(v, g, visc, net, forcing),
                       vjp=lambda v, w: les_rhs_vjp(v, w, g, visc, net, forcing, grads))


def rollout_loss(u0, refs, dt, g: S.Grid, visc, net: Net, forcing=None):
    """L = sum_t <u_t - r_t, u_t - r_t> / (2 nt) over nt = len(refs) RK4 steps
    (RFFT-weighted inner product); returns (L, states u_0..u_nt)"""
    us = [u0]
    L = 0.0
    for r in refs:
        us.append(rk4(us[-1], dt, g, visc, net, forcing))
        e = us[-1] - r
        L += sum(S.spectraldot(e[i], e[i], g).real for i in range(3))
    return L / (2 * len(refs)), us


def rollout_gradient(u0, refs, dt, g: S.Grid, visc, net: Net, forcing=None):
    """(L, d L / d u0, parameter gradient) of rollout_loss"""
    L, us = rollout_loss(u0, refs, dt, g, visc, net, forcing)
    nt = len(refs)
    grads = zero_grads(net)
    ubar = np.zeros_like(u0)
    for t in range(nt, 0, -1):
        ubar = ubar + (us[t] - refs[t - 1]) / nt
        ubar = rk4_vjp(us[t - 1], ubar, dt, g, visc, net, forcing, grads)
    return L, ubar, grads
