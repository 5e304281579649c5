"""Oracle: adjoint (vector-Jacobian product) of the DNS right-hand side and
of the Wray3 / RK4 steps.  TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

The reference has no adjoint (its training is a-priori only), so these
functions have no reference counterpart; they are pinned by central finite
differences of the oracle RHS (exact for the quadratic RHS), by a
dot-product test, and by an independent torch autograd restatement of the
RHS (tests/test_adjoint_oracle.py).

Inner product: the RFFT-weighted real inner product of spectraldot
(src/solver.jl:708), <a, b> = Re sum_k w_k conj(a_k) b_k, which equals the
physical mean of a(x) b(x).  Under it to_phys and to_spec are mutual
transposes, the 2/3 mask, the projection and the viscous operator are
self-adjoint, and the transpose of multiplying by -i k_j is +i k_j.

RHS: f(u)_i = sum_j (-i k_j) S[v_i v_j] + V u + f_band(u), v = C2R(T u).
Transpose applied to w:
    g_ij = C2R(i k_j w_i)                       (physical gradient of w)
    a_m  = sum_j g_mj v_j + sum_i g_im v_i
    J^T w = T S[a] + V w + f_band^T w
with S = to_spec, T the 2/3 mask, V = -nu (|k|^2 + k k^T), and for
f_band = A B u, A = eps / <u, B u>:  f_band^T w = A B w - (2A/<u,Bu>) <Bu, w> B u.
"""
import numpy as np

from . import spectral as S


def rhs_vjp(u, w, g: S.Grid, visc, forcing=None):
    """J(u)^T w for f = convectiondiffusion (+ band forcing)."""
    D = g.D
    v = [S.to_phys(S.twothirds(u[i], g), g) for i in range(D)]
    k = g.kfull
    gr = [[S.to_phys(1j * k[j] * w[i], g) for j in range(D)] for i in range(D)]
    out = []
    for m in range(D):
        a = sum(gr[m][j] * v[j] for j in range(D)) + sum(gr[i][m] * v[i] for i in range(D))
        out.append(S.twothirds(S.to_spec(a, g), g))
    out = np.stack(out)
    # viscous part (self-adjoint): -nu (k^2 w + k (k.w))
    k2 = sum(k[i] * k[i] for i in range(D))
    kw = sum(k[i] * w[i] for i in range(D))
    out = out + np.stack([-visc * (k2 * w[i] + k[i] * kw) for i in range(D)])
    if forcing is not None:
        eps, kband = forcing
        k2i = g.k2int()
        band = (k2i > 0) & (k2i <= kband * kband)
        wt = g.rfft_weight()
        Bu = np.where(band[None], u, 0)
        Sb = float(np.sum(wt * np.sum(np.abs(Bu) ** 2, axis=0)))
        A = eps / Sb
        c = sum(S.spectraldot(Bu[i], w[i], g) for i in range(D)).real
        out = out + A * np.where(band[None], w, 0) - (2 * A / Sb) * c * Bu
    return out.astype(u.dtype)


def rk4_vjp(u0, ubar, dt, g: S.Grid, visc, forcing=None, rhs=None, vjp=None):
    """d(rk4 step)/du0 transposed, applied to ubar (stages recomputed).
    ``rhs(v)`` / ``vjp(v, w)`` replace the DNS right-hand side and its
    transpose (the LES with a learned closure, oracle/les_cnn.py)."""
    f = rhs or (lambda v: S.convectiondiffusion(v, g, visc, forcing=forcing))
    P = lambda v: S.project(v, g)
    k1 = f(u0)
    u1 = P(u0 + (dt / 2) * k1)
    k2 = f(u1)
    u2 = P(u0 + (dt / 2) * k2)
    k3 = f(u2)
    u3 = P(u0 + dt * k3)
    J = vjp or (lambda v, ww: rhs_vjp(v, ww, g, visc, forcing))
    s = P(ubar)
    t3 = P(J(u3, (dt / 6) * s))
    t2 = P(J(u2, (dt / 3) * s + dt * t3))
    t1 = P(J(u1, (dt / 3) * s + (dt / 2) * t2))
    return s + t3 + t2 + t1 + J(u0, (dt / 6) * s + (dt / 2) * t1)


def wray3_vjp(u0, ubar, dt, g: S.Grid, visc, forcing=None, rhs=None, vjp=None):
    """d(wray3 step)/du0 transposed, applied to ubar.

    Forward (src/solver.jl:968-1012): k1 = f(u0); u1 = P(u0 + a1 h k1);
    s1 = u0 + b1 h k1; k2 = f(u1); u2 = P(s1 + a2 h k2); s2 = s1 (b2 = 0);
    k3 = f(u2); u3 = P(s2 + a3 h k3)."""
    a1, a2, a3 = S.WRAY3_A
    b1, _ = S.WRAY3_B
    f = rhs or (lambda v: S.convectiondiffusion(v, g, visc, forcing=forcing))
    P = lambda v: S.project(v, g)
    k1 = f(u0)
    u1 = P(u0 + a1 * dt * k1)
    s1 = u0 + b1 * dt * k1
    k2 = f(u1)
    u2 = P(s1 + a2 * dt * k2)
    J = vjp or (lambda v, ww: rhs_vjp(v, ww, g, visc, forcing))
    # u3 = P(s1 + a3 h k3), k3 = f(u2)
    q3 = P(ubar)                      # cotangent of the pre-projection sum
    s1bar = q3
    u2bar = J(u2, a3 * dt * q3)
    q2 = P(u2bar)                     # u2 = P(s1 + a2 h k2)
    s1bar = s1bar + q2
    u1bar = J(u1, a2 * dt * q2)
    q1 = P(u1bar)                     # u1 = P(u0 + a1 h k1)
    # s1 = u0 + b1 h k1
    k1bar = a1 * dt * q1 + b1 * dt * s1bar
    return q1 + s1bar + J(u0, k1bar)
