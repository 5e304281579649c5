"""The reference's JLD2 artefacts (src/data.jl:83,200-216; src/solver.jl:76-86)
written and read by symcode.jld2.  The HDF5 layer is pinned against the HDF5
library where the image has one (HDF5 1.10.6 with h5py and h5dump under
/opt/conda): h5py reads every value of our files, our reader reads files
written by h5py, h5dump parses our files' metadata (it checks the object
header checksums).  Those tests skip where that installation is absent."""
import json
import os
import subprocess

import numpy as np
import pytest

from symcode import jld2

CONDA_PY = "/opt/conda/bin/python3.9"
H5DUMP = "/opt/conda/bin/h5dump"


def _has_h5py():
    if not os.path.exists(CONDA_PY):
        return False
    r = subprocess.run([CONDA_PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


needs_h5py = pytest.mark.skipif(not _has_h5py(), reason="no h5py installation to check against")


def _h5py(script, *args):
    r = subprocess.run([CONDA_PY, "-c", script, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


def _field(rng, shape, keys, dtype=np.complex64):
    return {k: (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype) for k in keys}


def _stats(rng):
    return {k: float(rng.standard_normal()) for k in jld2.STAT_KEYS}


def test_lookup3_published_values():
    """lookup3.c's own self-test values for hashlittle"""
    assert jld2.lookup3(b"") == 0xDEADBEEF
    assert jld2.lookup3(b"Four score and seven years ago") == 0x17770551
    assert jld2.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def _fields_file(tmp_path, n=8, nsnap=3, seed=0):
    rng = np.random.default_rng(seed)
    shape = (n, n, n // 2 + 1)  # Julia's RFFT array (n/2 + 1, n, n)
    inputs = [_field(rng, shape, jld2.VECTOR_KEYS) for _ in range(nsnap)]
    outputs = [_field(rng, shape, jld2.TENSOR_KEYS) for _ in range(nsnap)]
    redelta = rng.random(nsnap)
    p = str(tmp_path / "fields.jld2")
    jld2.save_fields(p, inputs, outputs, redelta, 0.0982, 2.0, 2.5e-4)
    return p, inputs, outputs, redelta


def test_fields_roundtrip(tmp_path):
    p, inputs, outputs, redelta = _fields_file(tmp_path)
    assert jld2.keys(p) == ["inputs", "outputs", "redelta", "Δ", "Δ_factor", "visc"]
    ins, outs, rd, d = jld2.load(p, "inputs", "outputs", "redelta", "Δ")
    assert d == 0.0982 and jld2.load(p, "visc") == 2.5e-4 and jld2.load(p, "Δ_factor") == 2.0
    assert np.array_equal(rd, redelta)
    for a, b in zip(ins + outs, inputs + outputs):
        assert list(a) == list(b)
        for k in b:
            assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k])


def test_dns_and_meta_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    u = _field(rng, (6, 6, 4), jld2.VECTOR_KEYS, np.complex128)
    st = [_stats(rng) for _ in range(4)]
    p = str(tmp_path / "dns.jld2")
    jld2.save_dns(p, u, [0.0, 0.1, 0.2, 0.3], st, 12.5)
    u2, times, st2, wt = jld2.load(p, "u", "times", "statistics", "walltime")
    assert all(np.array_equal(u2[k], u[k]) for k in u) and wt == 12.5
    assert np.array_equal(times, [0.0, 0.1, 0.2, 0.3])
    assert st2 == st
    spectra = [rng.random(5), rng.random(5)]
    pm = str(tmp_path / "dns_meta.jld2")
    jld2.save_dns_meta(pm, [0.0, 0.5], spectra, st[:2], 0.7, 3.0, [0.0], st[2:3])
    sp, sd, tint, sw = jld2.load(pm, "spectra_dns", "statistics_dns", "t_int", "statistics_warmup")
    assert all(np.array_equal(a, b) for a, b in zip(sp, spectra)) and sd == st[:2] and tint == 0.7 and sw == st[2:3]
    pl = str(tmp_path / "les_meta.jld2")
    jld2.save_les_meta(pl, spectra, 0.25)
    sl, rm = jld2.load(pl, "spectra_les", "redelta_mean")
    assert all(np.array_equal(a, b) for a, b in zip(sl, spectra)) and rm == 0.25


def test_atomic_write_leaves_no_partial_file(tmp_path, monkeypatch):
    """jldsave_atomic semantics: an interrupted write keeps the old file and
    leaves no temporary behind"""
    p = str(tmp_path / "x.jld2")
    jld2.save(p, a=np.arange(3.0))
    before = open(p, "rb").read()

    def boom(*a, **k):
        raise KeyboardInterrupt

    monkeypatch.setattr(os, "replace", boom)
    with pytest.raises(KeyboardInterrupt):
        jld2.save(p, a=np.arange(5.0))
    monkeypatch.undo()
    assert open(p, "rb").read() == before
    assert sorted(os.listdir(tmp_path)) == ["x.jld2"]


def test_unsupported_values_raise(tmp_path):
    with pytest.raises(TypeError):
        jld2.save(str(tmp_path / "y.jld2"), a={"x": "text"})
    with pytest.raises(KeyError):
        p = str(tmp_path / "z.jld2")
        jld2.save(p, a=1.0)
        jld2.load(p, "b")


def test_artefact_paths():
    """src/experiment.jl:188-216 path strings (Julia prints 2.5e-4 as 0.00025)"""
    assert jld2.dnsdir("/r", "snellius", 2.5e-4, 100, 810) == "/r/snellius_visc=0.00025_seed=100_n=810"
    assert jld2.datadir("/r", "snellius", 1.0e-4, 200, 810, 3.5).endswith("snellius_visc=0.0001_seed=200_n=810/delta=3.5")
    assert jld2._jl_float(1.0e-5) == "1.0e-5" and jld2._jl_float(2.0) == "2.0" and jld2._jl_float(1.5e-4) == "0.00015"


READ_ALL = r"""
import sys, json, h5py, numpy as np
out = {}
with h5py.File(sys.argv[1], "r") as f:
    ub = f.userblock_size
    def digest(a):
        a = np.asarray(a)
        if a.dtype.names:
            return {"fields": list(a.dtype.names), "shape": list(a.shape)}
        return {"dtype": a.dtype.str, "shape": list(a.shape),
                "sum": complex(a.astype(np.complex128).sum()).__repr__() if a.size else "0",
                "first": repr(a.reshape(-1)[0].item()) if a.size else None}
    for k in f:
        d = f[k]
        v = d[()]
        ent = {"top": digest(v)}
        if v.dtype.names and all(v.dtype[n] == h5py.ref_dtype for n in v.dtype.names):
            ent["deref"] = [[digest(f[e[n]][()]) for n in v.dtype.names] for e in np.atleast_1d(v)]
        elif v.dtype.names:
            ent["values"] = [[repr(float(e[n])) for n in v.dtype.names] for e in np.atleast_1d(v)]
        elif v.dtype == h5py.ref_dtype:
            ent["deref"] = [digest(f[e][()]) for e in np.atleast_1d(v)]
        out[k] = ent
print(json.dumps({"userblock": ub, "data": out}))
"""


def _digest(a):
    a = np.asarray(a)
    return {"dtype": a.dtype.str, "shape": list(a.shape),
            "sum": repr(complex(a.astype(np.complex128).sum())) if a.size else "0",
            "first": repr(a.reshape(-1)[0].item()) if a.size else None}


@needs_h5py
def test_h5py_reads_our_fields_file(tmp_path):
    p, inputs, outputs, redelta = _fields_file(tmp_path, n=8, nsnap=2, seed=4)
    got = json.loads(_h5py(READ_ALL, p))
    assert got["userblock"] == 512
    d = got["data"]
    assert set(d) == {"inputs", "outputs", "redelta", "Δ", "Δ_factor", "visc"}
    # complex64 arrays come back as {re, im} compounds: compare the component sums
    for name, fields, keys_ in (("inputs", inputs, jld2.VECTOR_KEYS), ("outputs", outputs, jld2.TENSOR_KEYS)):
        assert d[name]["top"] == {"fields": list(keys_), "shape": [len(fields)]}
        for e, f in zip(d[name]["deref"], fields):
            for ent, k in zip(e, keys_):
                assert ent["fields"] == ["re", "im"] and ent["shape"] == list(f[k].shape)
    assert d["redelta"]["top"] == _digest(redelta)
    assert d["Δ"]["top"]["first"] == repr(0.0982) and d["visc"]["top"]["first"] == repr(2.5e-4)


CHECK_VALUES = r"""
import sys, h5py, numpy as np
f = h5py.File(sys.argv[1], "r")
ref = np.load(sys.argv[2])
for snap in range(int(sys.argv[3])):
    e = f["inputs"][snap]
    for k in ("x", "y", "z"):
        a = f[e[k]][()]
        c = a["re"] + 1j * a["im"]
        assert np.array_equal(c, ref[f"in{snap}{k}"]), (snap, k)
st = f["stats"][()]
assert list(st.dtype.names) == sys.argv[4].split(","), st.dtype.names
assert np.array_equal(np.array([list(r) for r in st]), ref["stats"])
sp = [f[r][()] for r in f["spectra"][()]]
assert all(np.array_equal(a, ref[f"sp{i}"]) for i, a in enumerate(sp))
print("ok")
"""


@needs_h5py
def test_h5py_values_match(tmp_path):
    rng = np.random.default_rng(5)
    shape = (4, 4, 3)
    inputs = [_field(rng, shape, jld2.VECTOR_KEYS) for _ in range(2)]
    st = [_stats(rng) for _ in range(3)]
    spectra = [rng.random(4), rng.random(7)]
    p = str(tmp_path / "v.jld2")
    jld2.save(p, inputs=inputs, stats=[jld2.statistics_record(s) for s in st], spectra=spectra)
    ref = {f"in{i}{k}": f[k] for i, f in enumerate(inputs) for k in f}
    ref["stats"] = np.array([[s[k] for k in jld2.STAT_KEYS] for s in st])
    ref.update({f"sp{i}": s for i, s in enumerate(spectra)})
    np.savez(str(tmp_path / "ref.npz"), **ref)
    assert _h5py(CHECK_VALUES, p, str(tmp_path / "ref.npz"), "2", ",".join(jld2.STAT_KEYS)).strip() == "ok"


WRITE_H5PY = r"""
import sys, h5py, numpy as np
r = np.random.default_rng(7)
with h5py.File(sys.argv[1], "w", libver="latest", userblock_size=512) as f:
    f["a"] = np.arange(24, dtype="<f8").reshape(2, 3, 4)
    f["b"] = (np.arange(5) - 2.5j).astype(np.complex64)
    f["s"] = 3.25
    f["i"] = np.int64(-7)
    rec = np.zeros(3, dtype=[("e", "<f8"), ("n", "<i8")]); rec["e"] = [1.5, 2.5, 3.5]; rec["n"] = [1, 2, 3]
    f["rec"] = rec
    # the referenced arrays in a subgroup: the root keeps at most 8 links,
    # HDF5's threshold for its compact link storage
    g = f.create_group("g")
    x = g.create_dataset("x", data=np.linspace(0, 1, 6).reshape(2, 3))
    y = g.create_dataset("y", data=np.linspace(1, 2, 6).reshape(2, 3))
    refs = f.create_dataset("refs", (2,), dtype=h5py.ref_dtype); refs[0] = x.ref; refs[1] = y.ref
    nt = np.zeros((), dtype=[("x", h5py.ref_dtype), ("y", h5py.ref_dtype)]); nt["x"] = x.ref; nt["y"] = y.ref
    f["nt"] = nt
    f["a"].attrs["note"] = "attribute messages are skipped"
"""


@needs_h5py
def test_reader_reads_h5py_files(tmp_path):
    p = str(tmp_path / "h.h5")
    _h5py(WRITE_H5PY, p)
    v = jld2.load(p)
    assert np.array_equal(v["a"], np.arange(24.0).reshape(2, 3, 4))
    assert v["b"].dtype == np.complex64 and np.array_equal(v["b"], (np.arange(5) - 2.5j).astype(np.complex64))
    assert v["s"] == 3.25 and v["i"] == -7
    assert v["rec"] == [{"e": 1.5, "n": 1}, {"e": 2.5, "n": 2}, {"e": 3.5, "n": 3}]
    assert np.array_equal(v["refs"][0], np.linspace(0, 1, 6).reshape(2, 3))
    assert np.array_equal(v["refs"][1], np.linspace(1, 2, 6).reshape(2, 3))
    assert np.array_equal(v["nt"]["x"], v["refs"][0]) and np.array_equal(v["nt"]["y"], v["refs"][1])
    assert sorted(v) == ["a", "b", "g", "i", "nt", "rec", "refs", "s"]
    assert np.array_equal(v["g"]["x"], v["refs"][0]) and np.array_equal(jld2.load(p, "g/y"), v["refs"][1])


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_h5dump_parses_our_files(tmp_path):
    p, _, _, _ = _fields_file(tmp_path, n=4, nsnap=2, seed=6)
    r = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for name in ("inputs", "outputs", "redelta", "visc"):
        assert f'DATASET "{name}"' in r.stdout
    assert "H5T_STD_REF_OBJECT" in r.stdout
    # values through the library: the scalar visc
    r = subprocess.run([H5DUMP, "-d", "/visc", p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "0.00025" in r.stdout


WRITE_COMMITTED = r"""
import sys, h5py, numpy as np
with h5py.File(sys.argv[1], "w", libver="latest", userblock_size=512) as f:
    f["ComplexF32"] = np.dtype([("re", "<f4"), ("im", "<f4")])  # a committed (named) datatype
    a = np.zeros(6, dtype=f["ComplexF32"].dtype)
    a["re"] = np.arange(6); a["im"] = -np.arange(6)
    f.create_dataset("z", data=a, dtype=f["ComplexF32"])
    rec = np.dtype([("e", "<f8"), ("n", "<i8")])
    f["Stats"] = rec
    r = np.zeros(2, dtype=rec); r["e"] = [0.5, 1.5]; r["n"] = [3, 4]
    f.create_dataset("stats", data=r, dtype=f["Stats"])
"""


@needs_h5py
def test_reader_follows_committed_datatypes(tmp_path):
    """datasets whose datatype is a committed (named) one, as JLD2 writes the
    element types of Julia structs: the datatype message is shared"""
    p = str(tmp_path / "c.h5")
    _h5py(WRITE_COMMITTED, p)
    z = jld2.load(p, "z")
    assert z.dtype == np.complex64 and np.array_equal(z, np.arange(6) - 1j * np.arange(6))
    assert jld2.load(p, "stats") == [{"e": 0.5, "n": 3}, {"e": 1.5, "n": 4}]
    assert sorted(jld2.load(p)) == ["stats", "z"]  # the committed datatypes are not data
