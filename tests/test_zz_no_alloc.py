"""sc_energy_noalloc (the entry point that allocates nothing) against
sc_energy, through the C ABI: two calls in a row on one GPU, the second on a
smaller field, and no change of the device's free memory across a call on a
context that never ran a reduction before."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import print_case

pytestmark = pytest.mark.gpu


def energy_pair(sc, g, u):
    L = sc.lib()
    a, b = ctypes.c_double(), ctypes.c_double()
    ncomp = int(u.shape[0])
    assert L.sc_energy_noalloc(g.ctx, ctypes.c_void_p(u.data_ptr()), ncomp, ctypes.byref(a)) == 0, \
        L.sc_last_error(g.ctx)
    assert L.sc_energy(g.ctx, ctypes.c_void_p(u.data_ptr()), ncomp, ctypes.byref(b)) == 0
    return a.value, b.value


def test_energy_noalloc_matches_energy():
    import symcode as sc

    print_case("energy_noalloc 512 fp64 then 64 fp32")
    g1 = sc.Grid(3, 512, dtype=torch.float64)  # 264 x 512 x 512 entries: 5 chunks
    u1 = sc.randomfield(sc.peak_profile, g1, seed=3, totalenergy=0.7, kpeak=4)
    a, b = energy_pair(sc, g1, u1)
    print(f"512: noalloc {a:.17g} energy {b:.17g}")
    assert a > 0 and abs(a - b) <= 1e-14 * b
    del u1, g1
    g2 = sc.Grid(3, 64, dtype=torch.float32)
    u2 = sc.randomfield(sc.peak_profile, g2, seed=4, totalenergy=0.3, kpeak=4)
    free0 = torch.cuda.mem_get_info()[0]
    a2 = ctypes.c_double()
    L = sc.lib()
    assert L.sc_energy_noalloc(g2.ctx, ctypes.c_void_p(u2.data_ptr()), 3, ctypes.byref(a2)) == 0
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0, (free0, free1)
    a, b = energy_pair(sc, g2, u2)
    print(f"64 fp32: noalloc {a:.17g} energy {b:.17g}")
    assert a == a2.value
    assert a > 0 and abs(a - b) <= 1e-14 * b
    # one component, and the argument checks
    a1 = ctypes.c_double()
    assert L.sc_energy_noalloc(g2.ctx, ctypes.c_void_p(u2.data_ptr()), 1, ctypes.byref(a1)) == 0
    e1 = sc.energy(u2[0:1].contiguous(), g2)
    assert abs(a1.value - e1) <= 1e-14 * e1
    assert L.sc_energy_noalloc(g2.ctx, None, 3, ctypes.byref(a1)) == 1
    assert L.sc_energy_noalloc(g2.ctx, ctypes.c_void_p(u2.data_ptr()), 0, ctypes.byref(a1)) == 1
