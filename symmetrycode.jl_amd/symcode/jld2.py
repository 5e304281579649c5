"""JLD2 artefact files of the reference, written and read without an HDF5
library.

The reference saves its data-generation artefacts with JLD2
(`jldsave_atomic`, src/solver.jl:76-86): the warm-up DNS state
(`dnsfile`, src/data.jl:83), the DNS-side metadata (`dnsmetafile`,
src/data.jl:200-204), the filtered (ubar, tau) pairs per filter width
(`fieldsfile`, src/data.jl:207-211) and the LES-side metadata
(`lesmetafile`, src/data.jl:212-215); its analysis and training read them
back with `load(file, names...)` (src/analysis.jl:46, src/data.jl:721,
src/les.jl:63).  The paths are those of src/experiment.jl:184-216.

A JLD2 file is an HDF5 file behind a 512-byte user block holding JLD2's
header line.  This module writes and reads the subset of HDF5 those
artefacts need, from the HDF5 file-format specification (version 3.0):

* superblock version 2 at byte 512 (base address 512; every other address
  is relative to it, the end-of-file address is absolute);
* version-2 object headers (``OHDR``, one chunk, Jenkins lookup3 checksum;
  continuation chunks ``OCHK`` and shared messages, e.g. committed
  datatypes, are followed on read);
* groups in compact form: Link Info, Group Info and hard-link messages;
* datasets with Dataspace (v2), Datatype, Fill Value (v3) and contiguous
  Data Layout (v3) messages; compact layouts are read as well;
* datatypes: little-endian IEEE floats and integers, compound (written as
  version 3), object references.

Julia values map as JLD2 stores them:

* ``Array{T,N}`` of size (n1, .., nN) <-> a dataset with dimensions
  (nN, .., n1), i.e. a C-order numpy array of shape (nN, .., n1) holding the
  Julia array's memory.  A spectral field downloaded from the GPU grid
  (``Grid.download_spectral``: [z, y, x]) is the reference's RFFT array
  (n/2 + 1, n, n) as is.
* ``Complex{T}`` <-> compound {re, im} (Julia's field names); compounds
  named {r, i} (h5py's convention) also read as complex.
* a NamedTuple of arrays (a vector or tensor field after ``map(Array, .)``)
  <-> a compound whose members are object references to the arrays'
  datasets; a NamedTuple of numbers (``turbulence_statistics``) <-> a
  compound of those numbers.  Both read back as dicts in field order.
* ``Vector`` of such values <-> a one-dimensional dataset of them; a
  ``Vector`` of arrays (``spectra_dns``) <-> a dataset of object references.
* scalars <-> scalar dataspaces.

Not written: JLD2's committed Julia type descriptions (the ``_types`` group
and ``julia_type`` attributes).  The HDF5 layer is checked against the HDF5
library (h5py and h5dump read these files; this reader reads h5py's); how a
Julia JLD2 reader rebuilds NamedTuple and Complex element types from plain
compounds cannot be checked in this environment (no Julia).
"""
import os
import struct
import tempfile

import numpy as np

JLD2_HEADER = b"HDF5-based Julia Data Format, version 0.2.0\x00 (written by symcode)\x00"
USERBLOCK = 512
UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"


# --------------------------------------------------------------- checksum
def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 ``hashlittle`` (public domain, 2006) over bytes:
    the checksum of HDF5's version-2 metadata structures."""
    M = 0xFFFFFFFF

    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & M

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    i = 0
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], "little")) & M
        b = (b + int.from_bytes(data[i + 4:i + 8], "little")) & M
        c = (c + int.from_bytes(data[i + 8:i + 12], "little")) & M
        a = (a - c) & M; a ^= rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 4); b = (b + a) & M
        i += 12
        n -= 12
    if n == 0:
        return c
    tail = bytes(data[i:i + n]) + b"\x00" * (12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & M
    b = (b + int.from_bytes(tail[4:8], "little")) & M
    c = (c + int.from_bytes(tail[8:12], "little")) & M
    c ^= b; c = (c - rot(b, 14)) & M
    a ^= c; a = (a - rot(c, 11)) & M
    b ^= a; b = (b - rot(a, 25)) & M
    c ^= b; c = (c - rot(b, 16)) & M
    a ^= c; a = (a - rot(c, 4)) & M
    b ^= a; b = (b - rot(a, 14)) & M
    c ^= b; c = (c - rot(b, 24)) & M
    return c


# --------------------------------------------------------------- datatypes
# A datatype is described by a small tuple tree:
#   ("float", size) | ("int", size, signed) | ("ref",)
#   ("compound", size, [(name, offset, dtype), ...])
def _dt_of_numpy(dt):
    dt = np.dtype(dt)
    if dt.kind == "f":
        return ("float", dt.itemsize)
    if dt.kind in "iu":
        return ("int", dt.itemsize, dt.kind == "i")
    if dt.kind == "b":
        return ("int", 1, False)
    if dt.kind == "c":
        h = dt.itemsize // 2
        return ("compound", dt.itemsize, [("re", 0, ("float", h)), ("im", h, ("float", h))])
    raise TypeError(f"no HDF5 datatype for numpy {dt}")


def _encode_dt(t):
    if t[0] == "float":
        size = t[1]
        if size == 4:
            sign, eloc, esz, msz, bias = 31, 23, 8, 23, 127
        elif size == 8:
            sign, eloc, esz, msz, bias = 63, 52, 11, 52, 1023
        else:
            raise TypeError("float size")
        # class 1 version 1; bit field: little endian, no padding, implied
        # mantissa msb (normalization 2); sign bit position
        return (bytes([0x11, 0x20, sign, 0]) + struct.pack("<I", size)
                + struct.pack("<HHBBBBI", 0, 8 * size, eloc, esz, 0, msz, bias))
    if t[0] == "int":
        size, signed = t[1], t[2]
        return bytes([0x10, 0x08 if signed else 0x00, 0, 0]) + struct.pack("<I", size) + struct.pack("<HH", 0, 8 * size)
    if t[0] == "ref":
        return bytes([0x17, 0, 0, 0]) + struct.pack("<I", 8)  # object reference
    if t[0] == "compound":
        size, members = t[1], t[2]
        ob = 1 if size < 1 << 8 else 2 if size < 1 << 16 else 3 if size < 1 << 24 else 4
        out = bytes([0x36]) + struct.pack("<H", len(members)) + b"\x00" + struct.pack("<I", size)
        for name, off, mt in members:
            out += name.encode() + b"\x00" + off.to_bytes(ob, "little") + _encode_dt(mt)
        return out
    raise TypeError(t)


def _decode_dt(buf, p):
    """-> (dtype tuple, position after the message)"""
    buf = bytes(buf)
    cv = buf[p]
    cls, ver = cv & 0x0F, cv >> 4
    bf = buf[p + 1] | buf[p + 2] << 8 | buf[p + 3] << 16
    size = struct.unpack_from("<I", buf, p + 4)[0]
    q = p + 8
    if cls == 0:  # fixed point
        if bf & 1:
            raise ValueError("big-endian integers are not supported")
        return ("int", size, bool(bf & 0x08)), q + 4
    if cls == 1:  # floating point
        if bf & 1 or bf & 0x40:
            raise ValueError("non-little-endian floats are not supported")
        if size not in (2, 4, 8):
            raise ValueError("float size")
        return ("float", size), q + 12
    if cls == 3:  # string: fixed length
        return ("string", size), q
    if cls == 7:  # reference
        if bf & 0x0F != 0:
            raise ValueError("only object references are supported")
        return ("ref",), q
    if cls == 6:  # compound
        nmem = bf & 0xFFFF
        members = []
        for _ in range(nmem):
            e = buf.index(b"\x00", q)
            name = bytes(buf[q:e]).decode()
            if ver < 3:
                q = q + ((e - q) // 8 + 1) * 8  # name padded to a multiple of 8
                off = struct.unpack_from("<I", buf, q)[0]
                q += 4
                if ver == 1:
                    q += 1 + 3 + 4 + 4 + 16  # dimensionality, reserved, permutation, reserved, sizes
            else:
                q = e + 1
                ob = 1 if size < 1 << 8 else 2 if size < 1 << 16 else 3 if size < 1 << 24 else 4
                off = int.from_bytes(buf[q:q + ob], "little")
                q += ob
            mt, q = _decode_dt(buf, q)
            members.append((name, off, mt))
        return ("compound", size, members), q
    raise ValueError(f"unsupported HDF5 datatype class {cls}")


def _np_of_dt(t):
    if t[0] == "float":
        return np.dtype(f"<f{t[1]}")
    if t[0] == "int":
        return np.dtype(f"<{'i' if t[2] else 'u'}{t[1]}")
    if t[0] == "ref":
        return np.dtype("<u8")
    if t[0] == "string":
        return np.dtype(f"S{t[1]}")
    if t[0] == "compound":
        return np.dtype({"names": [m[0] for m in t[2]], "formats": [_np_of_dt(m[2]) for m in t[2]],
                         "offsets": [m[1] for m in t[2]], "itemsize": t[1]})
    raise TypeError(t)


def _is_complex(t):
    if t[0] != "compound" or len(t[2]) != 2:
        return False
    (n0, o0, t0), (n1, o1, t1) = t[2]
    return ((n0, n1) in (("re", "im"), ("r", "i")) and t0 == t1 and t0[0] == "float" and o0 == 0
            and o1 == t0[1] and t[1] == 2 * t0[1])


# --------------------------------------------------------------- writer
class _Writer:
    """Objects are appended in order; addresses are relative to the base
    (the superblock, at byte USERBLOCK)."""

    def __init__(self):
        self.buf = bytearray()
        self.pos = 48  # the superblock occupies relative bytes 0..47

    def alloc(self, data):
        addr = self.pos
        self.buf += data
        self.pos += len(data)
        return addr

    @staticmethod
    def _ohdr(messages):
        body = b"".join(struct.pack("<BHB", mtype, len(m), 0) + m for mtype, m in messages)
        n = len(body)
        if n < 1 << 8:
            flag, sz = 0, n.to_bytes(1, "little")
        elif n < 1 << 16:
            flag, sz = 1, n.to_bytes(2, "little")
        elif n < 1 << 32:
            flag, sz = 2, n.to_bytes(4, "little")
        else:
            flag, sz = 3, n.to_bytes(8, "little")
        head = b"OHDR" + bytes([2, flag]) + sz + body
        return head + struct.pack("<I", lookup3(head))

    @staticmethod
    def _dataspace(shape):
        if shape is None:
            return bytes([2, 0, 0, 0])  # scalar
        return bytes([2, len(shape), 0, 1]) + b"".join(struct.pack("<Q", d) for d in shape)

    def dataset(self, raw, dt, shape):
        """raw bytes of the elements (C order); shape None for a scalar"""
        nbytes = len(raw)
        data_addr = self.alloc(raw) if nbytes else UNDEF
        msgs = [
            (0x0001, self._dataspace(shape)),
            (0x0003, _encode_dt(dt)),
            # fill value v3: allocation time late, fill written if set, none defined
            (0x0005, bytes([3, 0x02 | (2 << 2)])),
            # data layout v3, contiguous
            (0x0008, bytes([3, 1]) + struct.pack("<QQ", data_addr, nbytes)),
        ]
        return self.alloc(self._ohdr(msgs))

    def group(self, links):
        msgs = [(0x0002, bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)),  # link info, compact
                (0x000A, bytes([0, 0]))]  # group info
        for name, addr in links:
            nb = name.encode()
            if len(nb) >= 256:
                raise ValueError("link names are limited to 255 bytes")
            # link v1, hard link; a non-ASCII name carries the UTF-8 charset field
            head = bytes([1, 0, len(nb)]) if nb.isascii() else bytes([1, 0x10, 1, len(nb)])
            msgs.append((0x0006, head + nb + struct.pack("<Q", addr)))
        return self.alloc(self._ohdr(msgs))

    def finish(self, root):
        eof = USERBLOCK + self.pos
        sb = SIGNATURE + bytes([2, 8, 8, 0]) + struct.pack("<QQQQ", USERBLOCK, UNDEF, eof, root)
        sb += struct.pack("<I", lookup3(sb))
        ub = JLD2_HEADER + b"\x00" * (USERBLOCK - len(JLD2_HEADER))
        return ub + sb + bytes(self.buf)


def _is_field(v):
    return isinstance(v, dict) and all(isinstance(x, np.ndarray) and x.ndim > 0 for x in v.values())


def _is_record(v):
    return isinstance(v, dict) and all(np.isscalar(x) or (isinstance(x, np.ndarray) and x.ndim == 0)
                                       for x in v.values())


def _write_value(w, v):
    """-> object header address of a dataset holding v"""
    if isinstance(v, (float, int, np.floating, np.integer, bool, np.bool_)) or (
            isinstance(v, np.ndarray) and v.ndim == 0):
        a = np.asarray(v)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        if a.dtype.kind == "i" and isinstance(v, int):
            a = a.astype(np.int64)
        return w.dataset(a.tobytes(), _dt_of_numpy(a.dtype), None)
    if isinstance(v, np.ndarray):
        a = np.ascontiguousarray(v)
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        return w.dataset(a.tobytes(), _dt_of_numpy(a.dtype), a.shape)
    if _is_field(v):  # NamedTuple of arrays: one compound of references
        return w.dataset(*_field_elements(w, [v]), None)
    if _is_record(v):
        raw, dt = _record_elements([v])
        return w.dataset(raw, dt, None)
    if isinstance(v, (list, tuple)):
        if len(v) == 0:
            return w.dataset(b"", ("float", 8), (0,))
        if all(_is_field(e) for e in v):
            raw, dt = _field_elements(w, v)
            return w.dataset(raw, dt, (len(v),))
        if all(_is_record(e) for e in v):
            raw, dt = _record_elements(v)
            return w.dataset(raw, dt, (len(v),))
        if all(isinstance(e, np.ndarray) and e.ndim > 0 for e in v):
            refs = np.array([_write_value(w, e) for e in v], np.uint64)
            return w.dataset(refs.tobytes(), ("ref",), (len(v),))
        a = np.asarray(v)
        if a.dtype.kind in "fiuc":
            return _write_value(w, a)
    raise TypeError(f"cannot store {type(v).__name__} in a JLD2 artefact")


def _field_elements(w, fields):
    names = list(fields[0].keys())
    for f in fields:
        if list(f.keys()) != names:
            raise ValueError("the fields of a vector must have the same components")
    refs = np.array([[_write_value(w, f[k]) for k in names] for f in fields], np.uint64)
    dt = ("compound", 8 * len(names), [(k, 8 * i, ("ref",)) for i, k in enumerate(names)])
    return refs.tobytes(), dt


def _record_elements(recs):
    names = list(recs[0].keys())
    for r in recs:
        if list(r.keys()) != names:
            raise ValueError("the records of a vector must have the same fields")
    cols = [np.asarray([np.asarray(r[k]) for r in recs]) for k in names]
    cols = [c.astype(np.int64) if c.dtype.kind in "iu" else c.astype(np.float64) if c.dtype.kind == "f" else c
            for c in cols]
    ndt = np.dtype({"names": names, "formats": [c.dtype for c in cols]})
    a = np.empty(len(recs), ndt)
    for k, c in zip(names, cols):
        a[k] = c
    members = [(k, ndt.fields[k][1], _dt_of_numpy(ndt.fields[k][0])) for k in names]
    return a.tobytes(), ("compound", ndt.itemsize, members)


def save(path, **values):
    """jldsave_atomic(path; values...) (src/solver.jl:76-86): the file is
    written under a temporary name in the same directory and renamed over
    ``path``, so an interrupted write leaves no partial file.  Values:
    numbers, numpy arrays, dicts of arrays (NamedTuples of fields), dicts of
    numbers (NamedTuples of statistics) and lists of any of these."""
    w = _Writer()
    links = [(name, _write_value(w, v)) for name, v in values.items()]
    root = w.group(links)
    blob = w.finish(root)
    d = os.path.dirname(os.path.abspath(path))
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".tmp_", suffix=".jld2")
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(blob)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise


# --------------------------------------------------------------- reader
class _Reader:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = memoryview(f.read())
        b = self.buf
        sb = None
        for off in (0, 512, 1024, 2048, 4096):  # the superblock sits at 0 or a power of two >= 512
            if bytes(b[off:off + 8]) == SIGNATURE:
                sb = off
                break
        if sb is None:
            raise ValueError(f"{path}: not an HDF5 file")
        ver = b[sb + 8]
        if ver not in (2, 3):
            raise ValueError(f"{path}: superblock version {ver} is not supported (2 and 3 are)")
        if b[sb + 9] != 8 or b[sb + 10] != 8:
            raise ValueError("only 8-byte offsets and lengths are supported")
        if lookup3(b[sb:sb + 44]) != struct.unpack_from("<I", b, sb + 44)[0]:
            raise ValueError(f"{path}: superblock checksum mismatch")
        self.base, _, _, self.root = struct.unpack_from("<QQQQ", b, sb + 12)

    def at(self, addr):
        return self.base + addr

    def messages(self, addr):
        """(type, body) of an object header, continuation chunks followed"""
        b = self.buf
        p = self.at(addr)
        if bytes(b[p:p + 4]) != b"OHDR":
            raise ValueError("only version-2 object headers are supported")
        if b[p + 4] != 2:
            raise ValueError("object header version")
        flags = b[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16  # times
        if flags & 0x10:
            q += 4  # attribute phase change values
        nsz = 1 << (flags & 3)
        size = int.from_bytes(b[q:q + nsz], "little")
        q += nsz
        out = []
        chunks = [(q, size, p)]
        while chunks:
            start, size, head = chunks.pop(0)
            end = start + size
            if lookup3(b[head:end]) != struct.unpack_from("<I", b, end)[0]:
                raise ValueError("object header checksum mismatch")
            q = start
            while q + 4 <= end:
                mtype, msize, mflags = struct.unpack_from("<BHB", b, q)
                q += 4
                if flags & 0x04:
                    q += 2  # creation order
                body = b[q:q + msize]
                q += msize
                if mtype == 0x0010:  # continuation
                    caddr, clen = struct.unpack_from("<QQ", body, 0)
                    cp = self.at(caddr)
                    if bytes(b[cp:cp + 4]) != b"OCHK":
                        raise ValueError("continuation chunk signature")
                    chunks.append((cp + 4, clen - 8, cp))
                elif mtype != 0:
                    out.append((mtype, body) if not mflags & 0x02 else (mtype, self.shared(mtype, body)))
        return out

    def shared(self, mtype, body):
        """The body of a shared message (flag bit 1): a committed datatype
        (JLD2 commits the datatypes of Julia structs such as Complex) or a
        message kept in another object header; returns that header's message
        of the same type."""
        ver = body[0]
        if ver == 1:
            addr = struct.unpack_from("<Q", body, 8)[0]
        elif ver == 2 or (ver == 3 and body[1] in (2, 3)):
            addr = struct.unpack_from("<Q", body, 2)[0]
        else:
            raise ValueError("shared messages in the shared-message heap are not supported")
        for t, b in self.messages(addr):
            if t == mtype:
                return b
        raise ValueError("shared message target holds no message of its type")

    def links(self, addr):
        res = []
        for mtype, body in self.messages(addr):
            if mtype == 0x0002:  # link info
                fh = struct.unpack_from("<Q", body, 2 + (8 if body[1] & 1 else 0))[0]
                if fh != UNDEF:
                    raise ValueError("groups with dense link storage are not supported")
            if mtype == 0x0006:
                fl = body[1]
                q = 2
                ltype = 0
                if fl & 0x08:
                    ltype = body[q]
                    q += 1
                if fl & 0x04:
                    q += 8
                if fl & 0x10:
                    q += 1
                nl = 1 << (fl & 3)
                ln = int.from_bytes(body[q:q + nl], "little")
                q += nl
                name = bytes(body[q:q + ln]).decode()
                q += ln
                if ltype != 0:
                    continue  # soft / external links are not followed
                res.append((name, struct.unpack_from("<Q", body, q)[0]))
        return res

    def dataset(self, addr):
        shape, dt, layout = None, None, None
        for mtype, body in self.messages(addr):
            if mtype == 0x0001:
                ver, nd, fl = body[0], body[1], body[2]
                if ver == 1:
                    dims = struct.unpack_from(f"<{nd}Q", body, 8)
                    shape = tuple(dims)
                else:
                    stype = body[3]
                    if stype == 2:
                        shape = (0,)
                    else:
                        shape = tuple(struct.unpack_from(f"<{nd}Q", body, 4)) if stype == 1 else ()
            elif mtype == 0x0003:
                dt, _ = _decode_dt(body, 0)
            elif mtype == 0x0008:
                if body[0] not in (3, 4):  # v4 encodes compact / contiguous as v3 does
                    raise ValueError("data layout version")
                cls = body[1]
                if cls == 0:
                    n = struct.unpack_from("<H", body, 2)[0]
                    layout = ("compact", bytes(body[4:4 + n]))
                elif cls == 1:
                    layout = ("contiguous",) + struct.unpack_from("<QQ", body, 2)
                else:
                    raise ValueError("chunked datasets are not supported")
            elif mtype == 0x0006:
                return ("group", addr)
        if dt is None:
            return ("group", addr)
        if layout is None:  # a committed (named) datatype
            return ("datatype", dt)
        nd = _np_of_dt(dt)
        count = int(np.prod(shape)) if shape else 1
        if layout[0] == "compact":
            raw = layout[1]
        elif layout[1] == UNDEF:
            raw = b"\x00" * (count * nd.itemsize)
        else:
            p = self.at(layout[1])
            raw = self.buf[p:p + count * nd.itemsize]
        a = np.frombuffer(raw, nd, count).reshape(shape if shape else ())
        return self.convert(a, dt)

    def convert(self, a, dt):
        """numpy view of the stored elements -> value (a scalar dataspace gives
        a number)"""
        if dt[0] == "float" or dt[0] == "int" or dt[0] == "string":
            return a[()] if a.ndim == 0 else a.copy()
        if _is_complex(dt):
            h = dt[2][0][2][1]
            c = np.empty(a.shape, np.complex64 if h == 4 else np.complex128)
            c.real = a[dt[2][0][0]]
            c.imag = a[dt[2][1][0]]
            return c[()] if a.ndim == 0 else c
        if dt[0] == "ref":
            out = [self.load_ref(int(r)) for r in a.reshape(-1)]
            return out if a.ndim else out[0]
        if dt[0] == "compound":
            if all(m[2][0] == "ref" for m in dt[2]):  # NamedTuple of arrays
                recs = [{m[0]: self.load_ref(int(e[m[0]])) for m in dt[2]} for e in a.reshape(-1)]
            else:
                recs = [{m[0]: self.convert(np.asarray(e[m[0]]), m[2]) for m in dt[2]} for e in a.reshape(-1)]
            return recs if a.ndim else recs[0]
        raise TypeError(dt)

    def load_ref(self, addr):
        v = self.dataset(addr)
        if isinstance(v, tuple) and v and v[0] == "group":
            raise ValueError("references to groups are not supported")
        return v

    def value(self, addr):
        """a dataset's value, a group as a dict of its members, a committed
        datatype as ("datatype", description)"""
        v = self.dataset(addr)
        if isinstance(v, tuple) and v and v[0] == "group":
            return {k: self.value(a) for k, a in self.links(addr)}
        return v

    def lookup(self, path, name):
        addr = self.root
        for part in [s for s in name.split("/") if s]:
            sub = dict(self.links(addr))
            if part not in sub:
                raise KeyError(f"{path}: no {name}")
            addr = sub[part]
        return addr


def load(path, *names):
    """load(file, names...) of JLD2: one value, or a tuple of values in the
    order of ``names`` ("group/name" paths allowed); with no names a dict of
    every top-level value.  Arrays come back as numpy arrays in the file's
    (Julia-reversed) index order, NamedTuples and groups as dicts, vectors as
    lists (vectors of plain numbers as numpy arrays)."""
    r = _Reader(path)
    if not names:  # JLD2's _types group (committed Julia types) is not data
        out = {k: r.value(a) for k, a in r.links(r.root) if k != "_types"}
        return {k: v for k, v in out.items() if not (isinstance(v, tuple) and v and v[0] == "datatype")}
    vals = tuple(r.value(r.lookup(path, n)) for n in names)
    return vals[0] if len(vals) == 1 else vals


def keys(path):
    r = _Reader(path)
    return [k for k, _ in r.links(r.root)]


# --------------------------------------------------------------- reference artefacts
STAT_KEYS = ("e", "uavg", "diss", "l_int", "l_tay", "l_kol", "t_int", "t_tay", "t_kol", "Re_int", "Re_tay", "kmax_eta")
VECTOR_KEYS = ("x", "y", "z")
TENSOR_KEYS = ("xx", "yy", "zz", "xy", "yz", "zx")


def field_dict(a, keys_):
    """[comp, ...] numpy array (Grid.download_spectral) -> NamedTuple-like dict"""
    if a.shape[0] != len(keys_):
        raise ValueError(f"expected {len(keys_)} components, got {a.shape[0]}")
    return {k: np.ascontiguousarray(a[i]) for i, k in enumerate(keys_)}


def statistics_record(stats):
    """turbulence_statistics NamedTuple (src/solver.jl:836-851) from a dict"""
    return {k: float(stats[k]) for k in STAT_KEYS}


def save_dns(path, u, times, statistics, walltime):
    """dnsfile (src/data.jl:83): the warm-up DNS state and its series"""
    save(path, u=u, times=np.asarray(times, np.float64), statistics=[statistics_record(s) for s in statistics],
         walltime=float(walltime))


def save_dns_meta(path, times, spectra_dns, statistics_dns, t_int, walltime, times_warmup, statistics_warmup):
    """dnsmetafile (src/data.jl:200-204)"""
    save(path, times=np.asarray(times, np.float64), spectra_dns=[np.asarray(s, np.float64) for s in spectra_dns],
         statistics_dns=[statistics_record(s) for s in statistics_dns], t_int=float(t_int), walltime=float(walltime),
         times_warmup=np.asarray(times_warmup, np.float64),
         statistics_warmup=[statistics_record(s) for s in statistics_warmup])


def save_fields(path, inputs, outputs, redelta, delta, delta_factor, visc):
    """fieldsfile (src/data.jl:207-211): the (ubar, tau) pairs of one filter
    width; inputs / outputs are lists of {x, y, z} / {xx, yy, zz, xy, yz, zx}
    dicts of spectral arrays"""
    save(path, inputs=list(inputs), outputs=list(outputs), redelta=np.asarray(redelta, np.float64),
         **{"Δ": float(delta), "Δ_factor": float(delta_factor), "visc": float(visc)})


def save_les_meta(path, spectra_les, redelta_mean):
    """lesmetafile (src/data.jl:212-215)"""
    save(path, spectra_les=[np.asarray(s, np.float64) for s in spectra_les], redelta_mean=float(redelta_mean))


def dnsdir(rootdir, name, visc, seed, n_dns):
    """src/experiment.jl:188-189 (Julia's string form of a Float64 visc)"""
    return os.path.join(rootdir, f"{name}_visc={_jl_float(visc)}_seed={seed}_n={n_dns}")


def datadir(rootdir, name, visc, seed, n_dns, delta_factor):
    """src/experiment.jl:192"""
    return os.path.join(dnsdir(rootdir, name, visc, seed, n_dns), f"delta={_jl_float(delta_factor)}")


def _jl_float(x):
    """Julia's shortest round-trip print of a Float64 (0.00015, 1.0e-5, 2.0):
    Python's repr with Julia's exponent form, the same for 1e-5 <= |x| < 1e5
    (the viscosities and filter widths of the reference's paths)"""
    x = float(x)
    r = repr(x)
    if "e" in r:
        m, e = r.split("e")
        if "." not in m:
            m += ".0"
        return f"{m}e{int(e)}"
    return r
