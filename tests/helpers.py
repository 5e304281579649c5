"""Shared helpers of the GPU parity tests."""
import numpy as np

from oracle import spectral as S


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def rand_phys(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def oracle_solenoidal(g, seed, trunc=False):
    u = np.stack([S.to_spec(rand_phys(g.phys_shape, seed + i), g) for i in range(g.D)])
    u = S.project(u, g)
    if trunc:
        u = np.stack([S.twothirds(u[i], g) for i in range(g.D)])
    return u


def print_case(*args):
    print("CASE", *args, flush=True)
