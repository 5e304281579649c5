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
* ``tb_features`` / ``tbnn_forward`` / ``tbnn_backward`` / ``tbnn_model``:
  tb_kernel!'s invariants and deviatoric basis (src/nets.jl:416-500),
  tbnn_net and the tbnn closure (:535-583).
* ``symmetrize``: symmetrize_pointwise (src/nets.jl:350-389) over
  group_stuff's elements (src/symmetry.jl:42-67).
* ``weight_synthesis`` / ``equi_project`` / ``equi_net`` /
  ``equi_fullchain``: equivariant_net (src/nets.jl:102-238) with
  get_weight_synthesis (src/symmetry.jl:156-197), in torch float64 so that
  the trainables' gradients come from autograd.

Pinned (tests/test_pmlp_oracle.py): forward and every gradient against
torch.autograd of the same network built from torch.nn.functional's
gelu(approximate="tanh"); the trace removal is idempotent and leaves a
trace-free output; fullchain is homogeneous of degree two in G, tau(aG) =
a^2 tau(G), without the Re_Delta input, and with it the input z moves by
log(a) / sigma; the features have unit row norm up to eps and reproduce tau.
The TBNN's invariants are unchanged and its basis tensors rotate with any
orthogonal Q (A -> Q A Q^T), are symmetric and trace-free, and vanish as the
closed forms say for pure strain and pure rotation; its forward and gradients
equal torch.autograd's.  The group tables form a group of 48 (8) distinct
signed permutation matrices closed under products; the symmetrized model is
exactly equivariant under every element and leaves an equivariant model
unchanged.  The synthesised equivariant network is exactly equivariant under
every element for random trainables; in 2D the synthesis columns are
orthonormal fixed points of the brute-force Reynolds projectors of
get_weight_projectors (src/symmetry.jl:92-127).
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


def layer_shapes(D, widths, use_redelta, tbnn=False):
    nt = 6 if D == 3 else 3
    if tbnn:
        nin, nout = (5 if D == 3 else 2) + (1 if use_redelta else 0), (7 if D == 3 else 2)
    else:
        nin, nout = D * D + (1 if use_redelta else 0), nt
    ch = [nin, *widths, nout]
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


def mlp_core(p, x):
    """the four PointwiseConv layers without the output head"""
    a = np.asarray(x, np.float64)
    zs, acts = [], [a]
    for l in range(1, 4):
        z = a @ p[f"W{l}"] + p[f"b{l}"]
        a = gelu(z)
        zs.append(z)
        acts.append(a)
    return a @ p["W4"], (zs, acts)


def core_backward(p, cache, d):
    """parameter gradients for the cotangent d of mlp_core's output"""
    zs, acts = cache
    g = {"W4": acts[3].T @ d}
    d = d @ p["W4"].T
    for l in range(3, 0, -1):
        d = d * gelu_grad(zs[l - 1])
        g[f"W{l}"] = acts[l - 1].T @ d
        g[f"b{l}"] = d.sum(axis=0)
        d = d @ p[f"W{l}"].T
    return g


def forward(p, x, D):
    """x [nv][nin] -> (y [nv][nt], cache)"""
    raw, cache = mlp_core(p, x)
    return tracefree(raw, D), cache


def backward(p, cache, dy, D):
    """gradients {'W1', 'b1', ..., 'W4'} of sum(dy * y) for the cached forward"""
    return core_backward(p, cache, tracefree(dy, D))  # the trace removal is symmetric


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


# ------------------------------------------------------------------ TBNN
def grad_matrices(GG, D):
    """(np, D, D) matrices A[i][j] = G_ij of the component-major gradient
    (nonsym order: G_ij at component i + D j)"""
    G = np.stack([np.asarray(g, np.float64).ravel() for g in GG], axis=1)
    A = np.empty((G.shape[0], D, D))
    for i in range(D):
        for j in range(D):
            A[:, i, j] = G[:, i + D * j]
    return A, (G * G).sum(axis=1)


def _dev(B):
    D = B.shape[-1]
    return B - np.trace(B, axis1=1, axis2=2)[:, None, None] / D * np.eye(D)


def _pack(B, D):
    if D == 3:
        return np.stack([B[:, 0, 0], B[:, 1, 1], B[:, 2, 2], B[:, 0, 1], B[:, 1, 2], B[:, 2, 0]], axis=1)
    return np.stack([B[:, 0, 0], B[:, 1, 1], B[:, 0, 1]], axis=1)


def tb_features(GG, D, eps=np.finfo(np.float64).eps):
    """tb_kernel! (src/nets.jl:455-500): (invariants [np][ni], basis
    [np][nb][nt] packed, |G|^2 [np]) of A = G / (|G| + eps)"""
    A, a2 = grad_matrices(GG, D)
    A = A / (np.sqrt(a2) + eps)[:, None, None]
    At = np.swapaxes(A, 1, 2)
    S, R = (A + At) / 2, (A - At) / 2
    tr = lambda M: np.trace(M, axis1=1, axis2=2)
    SS, RR, SR, RS = S @ S, R @ R, S @ R, R @ S
    if D == 3:
        inv = [tr(SS), tr(RR), tr(SS @ S), tr(SR @ R), tr(SS @ R @ R)]
        bas = [S, SS, RR, SR - RS, RS @ R, RS @ S - SS @ R, RS @ R @ R - RR @ S @ R]
    else:
        inv = [tr(SS), tr(RR)]
        bas = [S, SR - RS]
    return np.stack(inv, axis=1), np.stack([_pack(_dev(B), D) for B in bas], axis=1), a2


def tb_rows(GG, D, z=None, eps=np.finfo(np.float64).eps):
    """create_dataloader_tbnn's input rows [invariants, z, basis (component
    fastest)]"""
    inv, bas, _ = tb_features(GG, D, eps)
    cols = [inv] + ([np.full((inv.shape[0], 1), z)] if z is not None else []) + [bas.reshape(bas.shape[0], -1)]
    return np.concatenate(cols, axis=1)


def tbnn_forward(p, x, D, nin):
    """rows x -> (m = sum_b w_b B_b [nv][nt], cache); w = mlp_core(x[:, :nin])"""
    nt, nb = (6, 7) if D == 3 else (3, 2)
    w, cache = mlp_core(p, x[:, :nin])
    B = np.asarray(x[:, nin:], np.float64).reshape(-1, nb, nt)
    return np.einsum("vb,vbt->vt", w, B), (cache, B)


def tbnn_backward(p, cache, dm):
    core, B = cache
    return core_backward(p, core, np.einsum("vt,vbt->vb", dm, B))


def tbnn_model(p, GG, delta, D, eps=np.finfo(np.float64).eps, redelta=None):
    """tbnn(net, ps, st, Δ, g; use_redelta, redelta_norm, visc) (src/nets.jl:558-583)"""
    z = None if redelta is None else redelta_z(GG, delta, *redelta)
    x = tb_rows(GG, D, z, eps)
    nin = (5 if D == 3 else 2) + (0 if z is None else 1)
    m, _ = tbnn_forward(p, x, D, nin)
    m = m * (delta**2 * a2_of(GG).ravel())[:, None]
    return np.stack([m[:, c].reshape(GG.shape[1:]) for c in range(m.shape[1])])


# ------------------------------------------------- the cube's symmetry group
def group_elements(D):
    """group_stuff (src/symmetry.jl:42-67): [(permutation, signs)], the
    permutation index fastest, 0-based axes"""
    if D == 3:
        perms = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0), (1, 0, 2), (0, 2, 1)]
        signs = [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1), (-1, -1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, -1)]
    else:
        perms = [(0, 1), (1, 0)]
        signs = [(1, 1), (-1, 1), (1, -1), (-1, -1)]
    return [(p, s) for s in signs for p in perms]


def element_matrix(p, s):
    """R with (R A R^T)_ab = s_a s_b A_{p_a p_b}: R_{a, p_a} = s_a"""
    D = len(p)
    R = np.zeros((D, D))
    for a in range(D):
        R[a, p[a]] = s[a]
    return R


def symmetrize(model, GG, D):
    """symmetrize_pointwise (src/nets.jl:350-389): 1/|G| sum_g R_g^T
    model(R_g A R_g^T) R_g for a pointwise model(GG) -> packed tau"""
    names = [(i, j) for j in range(D) for i in range(D)]  # nonsym order: i + D j
    comps = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)] if D == 3 else [(0, 0), (1, 1), (0, 1)]
    pidx = {}
    for t, (a, b) in enumerate(comps):
        pidx[(a, b)] = pidx[(b, a)] = t
    acc = None
    els = group_elements(D)
    for p, s in els:
        Arot = np.stack([s[a] * s[b] * GG[p[a] + D * p[b]] for (a, b) in names])
        y = model(Arot)
        q = [0] * D
        for a in range(D):
            q[p[a]] = a
        qs = [s[q[a]] for a in range(D)]
        term = np.stack([qs[a] * qs[b] * y[pidx[(q[a], q[b])]] for (a, b) in comps])
        acc = term if acc is None else acc + term
    return acc / len(els)


# ------------------------------------------------------ equivariant_net
# Written with torch (CPU, float64) so that the compact gradients come from
# autograd; Julia's column-major reshape is _reshape_f.
def _reshape_f(t, shape):
    """column-major reshape (Julia's reshape) of a torch tensor"""
    nd = t.dim()
    t = t.permute(*reversed(range(nd))).reshape(tuple(reversed(shape)))
    return t.permute(*reversed(range(len(shape))))


def cayley_tables(D):
    """group_stuff (src/symmetry.jl:42-67): Cayley table and inverse indices"""
    mats = [element_matrix(p, s) for p, s in group_elements(D)]
    n = len(mats)
    key = {m.astype(int).tobytes(): e for e, m in enumerate(mats)}
    cayley = np.array([[key[(mats[a] @ mats[b]).astype(int).tobytes()] for b in range(n)] for a in range(n)])
    unit = key[np.eye(D).astype(int).tobytes()]
    inverse = np.array([int(np.nonzero(cayley[a] == unit)[0][0]) for a in range(n)])
    return cayley, inverse


def weight_synthesis(D):
    """get_weight_synthesis (src/symmetry.jl:156-197) with the entries of the
    synthesis matrices rounded to FP32, the precision the reference stores
    them in; returns (s_lift, s_mid, s_sink) as float64"""
    els = group_elements(D)
    cayley, inverse = cayley_tables(D)
    nreg, nten = len(els), D * D
    alpha = float(np.float32(1 / math.sqrt(nreg)))
    Q = np.zeros((nreg, nten, nten))
    for h, (p, s) in enumerate(els):
        for j in range(D):
            for i in range(D):
                for y in range(D):
                    for x in range(D):
                        Q[h, x + y * D, i + j * D] = s[x] * s[y] * (p[x] == i) * (p[y] == j)
    relidx = np.array([[cayley[inverse[n], m] for n in range(nreg)] for m in range(nreg)])
    s_mid = np.zeros((nreg * nreg, nreg))
    for n in range(nreg):
        for m in range(nreg):
            s_mid[m + n * nreg, relidx[m, n]] = alpha
    s_lift = np.zeros((nreg * nten, nten))
    s_sink = np.zeros((nten * nreg, nten))
    for h in range(nreg):
        for mu in range(nten):
            for nu in range(nten):
                s_lift[h + mu * nreg, nu] = alpha * Q[h, mu, nu]
                s_sink[mu + h * nten, nu] = alpha * Q[h, mu, nu]
    return s_lift, s_mid, s_sink


def equi_shapes(D, widths, use_redelta):
    """the compact trainable arrays in Lux's order: lift (weight, bias[,
    weight_re]), mid_1, mid_2 (weight, bias), sink"""
    nreg, nten = (48 if D == 3 else 8), D * D
    c1, c2, c3 = widths
    sh = [("lift.weight", (nten, c1)), ("lift.bias", (c1,))]
    if use_redelta:
        sh.append(("lift.weight_re", (1, c1)))
    sh += [("mid_1.weight", (nreg, c2, c1)), ("mid_1.bias", (c2,)),
           ("mid_2.weight", (nreg, c3, c2)), ("mid_2.bias", (c3,)), ("sink.weight", (nten, c3))]
    return sh


def equi_unflat(h, D, widths, use_redelta):
    """flat trainables -> {name: torch float64 tensor (column-major shapes)}"""
    out, o = {}, 0
    for name, shp in equi_shapes(D, widths, use_redelta):
        n = int(np.prod(shp))
        import torch

        out[name] = _reshape_f(torch.tensor(np.asarray(h[o:o + n], np.float64)), shp)
        o += n
    return out


def equi_project(ps, D, syn):
    """project (src/nets.jl:130-170): the network weights (c_in, c_out) and
    biases of the four layers from the trainables"""
    import torch

    s_lift, s_mid, s_sink = (torch.tensor(a) for a in syn)
    nreg, nten = (48 if D == 3 else 8), D * D
    layers = []
    w, b = ps["lift.weight"], ps["lift.bias"]
    c_out = w.shape[1]
    w = _reshape_f(s_lift @ w, (nreg, nten, c_out)).permute(1, 0, 2)
    w = _reshape_f(w, (nten, nreg * c_out))
    if "lift.weight_re" in ps:
        wre = _reshape_f(ps["lift.weight_re"].repeat(nreg, 1), (1, nreg * c_out))
        w = torch.cat([w, wre], dim=0)
    layers.append((w, _reshape_f(b.reshape(1, -1).repeat(nreg, 1), (nreg * c_out,))))
    for i in (1, 2):
        w, b = ps[f"mid_{i}.weight"], ps[f"mid_{i}.bias"]
        _, c_out, c_in = w.shape
        w = s_mid @ _reshape_f(w, (nreg, c_out * c_in))
        w = _reshape_f(w, (nreg, nreg, c_out, c_in)).permute(1, 3, 0, 2)
        layers.append((_reshape_f(w, (nreg * c_in, nreg * c_out)),
                       _reshape_f(b.reshape(1, -1).repeat(nreg, 1), (nreg * c_out,))))
    w = ps["sink.weight"]
    c_in = w.shape[1]
    w = _reshape_f(s_sink @ w, (nten, nreg, c_in)).permute(1, 2, 0)
    layers.append((_reshape_f(w, (nreg * c_in, nten)), None))
    return layers


def equi_net(layers, x, D):
    """the projected network on rows x [nv][nin] with the symmetrising `symm`
    layer (src/nets.jl:196-215): packed [nv][nt]"""
    import torch

    a = x
    for w, b in layers[:3]:
        a = torch.nn.functional.gelu(a @ w + b, approximate="tanh")
    s = a @ layers[3][0]
    if D == 3:
        xx, yy, zz = s[:, 0], s[:, 4], s[:, 8]
        xy, yz, zx = (s[:, 1] + s[:, 3]) / 2, (s[:, 5] + s[:, 7]) / 2, (s[:, 2] + s[:, 6]) / 2
        t = (xx + yy + zz) / 3
        return torch.stack([xx - t, yy - t, zz - t, xy, yz, zx], dim=1)
    xx, yy, xy = s[:, 0], s[:, 3], (s[:, 1] + s[:, 2]) / 2
    t = (xx + yy) / 2
    return torch.stack([xx - t, yy - t, xy], dim=1)


def equi_fullchain(h, D, widths, GG, delta, eps=np.finfo(np.float64).eps, redelta=None):
    """fullchain of equivariant_net at inference (parameters widened to
    FP64): packed tau, component-major"""
    import torch

    syn = weight_synthesis(D)
    layers = equi_project(equi_unflat(h, D, widths, redelta is not None), D, syn)
    z = None if redelta is None else redelta_z(GG, delta, *redelta)
    x, _ = features(GG, None, delta, z, eps)
    y = equi_net(layers, torch.tensor(x), D).numpy()
    s = delta**2 * a2_of(GG).ravel()
    tau = y * s[:, None]
    return np.stack([tau[:, c].reshape(GG.shape[1:]) for c in range(tau.shape[1])])
