"""GPU-generated (ubar, tau) pairs through the reference's fieldsfile schema
(src/data.jl:189-191,207-211): the GPU DNS marches, sfs forms the pairs at
two filter widths, save_pairs writes the JLD2 artefact, and the file reads
back bit for bit (here and through h5py, where installed) and uploads back
into equal device Pairs.  The first snapshot's pair is also compared with the
oracle's sfs of the same field (FP32 against the FP64 oracle: 2e-5 for ubar,
1e-4 for tau, a difference of two filtered products)."""
import os
import subprocess

import numpy as np
import pytest
import torch

from helpers import print_case, rel
from oracle import filtering as OF
from oracle import spectral as S

pytestmark = pytest.mark.gpu

CONDA_PY = "/opt/conda/bin/python3.9"


def test_pairs_to_fieldsfile_and_back(tmp_path):
    import symcode as sc
    from symcode import jld2
    from symcode import training as T

    print_case("jld2 fields 32 -> 16 fp32")
    visc = 2e-3
    g_dns = sc.Grid(3, 32, dtype=torch.float32)
    g_les = sc.Grid(3, 16, dtype=torch.float32)
    u = sc.randomfield(sc.peak_profile, g_dns, seed=5, totalenergy=0.2, kpeak=3)
    u0 = g_dns.download_spectral(u).astype(np.complex128)
    pairs = T.create_data(u, g_dns, g_les, visc, [2.0, 3.0], 3, 0.02, cfl=0.3)
    for f, ps in pairs.items():
        p = str(tmp_path / f"fields_delta={f}.jld2")
        T.save_pairs(p, ps, g_les, f, visc)
        assert jld2.keys(p) == ["inputs", "outputs", "redelta", "Δ", "Δ_factor", "visc"]
        ins, outs, rd, d, df, v = jld2.load(p, "inputs", "outputs", "redelta", "Δ", "Δ_factor", "visc")
        assert (d, df, v) == (ps[0].delta, f, visc) and np.array_equal(rd, [q.redelta for q in ps])
        for q, a, b in zip(ps, ins, outs):
            ub, tau = g_les.download_spectral(q.ubar), g_les.download_spectral(q.tau)
            assert list(a) == ["x", "y", "z"] and list(b) == ["xx", "yy", "zz", "xy", "yz", "zx"]
            assert a["x"].dtype == np.complex64 and a["x"].shape == (16, 16, 9)  # Julia (9, 16, 16)
            assert all(np.array_equal(a[k], ub[i]) for i, k in enumerate(a))
            assert all(np.array_equal(b[k], tau[i]) for i, k in enumerate(b))
        back = T.load_pairs(p, g_les)
        assert len(back) == len(ps)
        for q, r in zip(ps, back):
            assert torch.equal(q.ubar, r.ubar) and torch.equal(q.tau, r.tau)
            assert (q.delta, q.redelta) == (r.delta, r.redelta)
        if os.path.exists(CONDA_PY):
            np.save(str(tmp_path / "ux.npy"), g_les.download_spectral(ps[-1].ubar)[0])
            script = ("import sys, h5py, numpy as np\n"
                      "f = h5py.File(sys.argv[1], 'r')\n"
                      "a = f[f['inputs'][-1]['x']][()]\n"
                      "ref = np.load(sys.argv[2])\n"
                      "assert np.array_equal(a['re'], ref.real) and np.array_equal(a['im'], ref.imag)\n"
                      "print('ok')\n")
            r = subprocess.run([CONDA_PY, "-c", script, p, str(tmp_path / "ux.npy")], capture_output=True,
                               text=True, timeout=120)
            if "No module named 'h5py'" not in r.stderr:
                assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]
    # the first snapshot (the initial field) against the oracle's sfs
    go_d, go_l = S.Grid(3, 32, 2 * np.pi, np.float64), S.Grid(3, 16, 2 * np.pi, np.float64)
    for f, ps in pairs.items():
        ob, ot = OF.sfs(u0, go_d, go_l, f * 2 * np.pi / 16)
        assert rel(g_les.download_spectral(ps[0].ubar), ob) < 2e-5
        assert rel(g_les.download_spectral(ps[0].tau), ot) < 1e-4


def test_load_pairs_rejects_wrong_grid(tmp_path):
    import symcode as sc
    from symcode import training as T

    print_case("jld2 wrong grid")
    g = sc.Grid(3, 16, dtype=torch.float32)
    ub = g.upload_spectral(np.zeros(g.ref_spec_shape(3), np.complex64))
    tau = g.upload_spectral(np.zeros(g.ref_spec_shape(6), np.complex64))
    p = str(tmp_path / "f.jld2")
    T.save_pairs(p, [T.Pair(ub, tau, 0.5, 10.0)], g, 1.0, 1e-3)
    with pytest.raises(ValueError):
        T.load_pairs(p, sc.Grid(3, 32, dtype=torch.float32))
