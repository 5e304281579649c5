"""Oracle: counter-based normal stream shared with the HIP initial-condition
kernel.  TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference draws its random initial phases with Julia's ``Xoshiro`` and
``randn!`` (src/solver.jl:590), a stream that cannot be reproduced outside
Julia.  Both this oracle and the HIP kernel instead draw the normals of mode
``idx`` (0-based linear index in the reference's RFFT layout) of component
``c`` from a splitmix64 counter stream followed by Box-Muller, so the two
sides see the same initial field bit-for-bit up to the last ulp of log/cos.

    key   = mix64(seed * G + c + 1)
    b_j   = mix64(key + (2 idx + j + 1) * G)       j = 0, 1
    u1    = ((b_0 >> 11) + 1) * 2^-53               in (0, 1]
    u2    = (b_1 >> 11) * 2^-53                     in [0, 1)
    r     = sqrt(-2 log u1),  θ = 2π u2
    value = (r cos θ + i r sin θ) * sqrt(1/2)

with ``G = 0x9E3779B97F4A7C15`` and mix64 the splitmix64 finaliser.
"""
import numpy as np

G = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def complex_normals(seed: int, ncomp: int, shape):
    """Complex normals for components 0..ncomp-1 on an array of ``shape``
    (C-order, which equals the reference's linear index order)."""
    size = int(np.prod(shape))
    idx = np.arange(size, dtype=np.uint64)
    out = np.empty((ncomp, size), np.complex128)
    with np.errstate(over="ignore"):
        for c in range(ncomp):
            key = mix64(np.uint64(seed) * G + np.uint64(c + 1))
            b0 = mix64(key + (np.uint64(2) * idx + np.uint64(1)) * G)
            b1 = mix64(key + (np.uint64(2) * idx + np.uint64(2)) * G)
            u1 = ((b0 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0**-53
            u2 = (b1 >> np.uint64(11)).astype(np.float64) * 2.0**-53
            r = np.sqrt(-2.0 * np.log(u1))
            th = 2.0 * np.pi * u2
            out[c] = (r * np.cos(th) + 1j * (r * np.sin(th))) * np.sqrt(0.5)
    return out.reshape((ncomp,) + tuple(shape))
