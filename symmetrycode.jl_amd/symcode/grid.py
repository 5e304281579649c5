"""Grid and field allocation: the host-side mirror of the reference's
``Grid{D}`` and field allocators (src/solver.jl:131-347).

Fields are torch tensors on the GPU in the library's device layout
(``include/symcode.h``): spectral ``(ncomp, nz, n, nhp)`` complex, physical
``(ncomp, nz, n, n)`` real.  torch only provides the memory; every operation
runs in libsymcode_hip.so.
"""
import ctypes

import numpy as np
import torch

from ._lib import check, lib


class Grid:
    """``Grid{D}(; l, n)`` (src/solver.jl:131-151) bound to one GPU context.

    ``dtype`` is ``torch.float64`` (FP64 path) or ``torch.float32``."""

    def __init__(self, D, n, l=2 * np.pi, dtype=torch.float64, device=0):
        if n % 2 != 0:
            raise AssertionError("Only even number of grid points supported.")
        self.D, self.n, self.l = D, int(n), float(l)
        self.dtype = dtype
        self.cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
        self.prec = 64 if dtype == torch.float64 else 32
        self.device = torch.device("cuda", device)
        ctx = ctypes.c_void_p()
        check(lib().sc_create(D, self.n, self.l, self.prec, device, ctypes.byref(ctx)), None)
        self.ctx = ctx
        info = (ctypes.c_int64 * 8)()
        check(lib().sc_info(ctx, info), ctx)
        _, _, self.nh, self.nhp, self.nz, self.spec_elems, self.phys_elems, _ = list(info)
        # share torch's stream so allocations and kernels are ordered
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().sc_set_stream(ctx, ctypes.c_void_p(stream)), ctx)

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                lib().sc_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    # ------------------------------------------------------------ metadata
    dim = property(lambda self: self.D)
    tensordim = property(lambda self: 3 if self.D == 2 else 6)
    spacing = property(lambda self: self.l / self.n)
    volume = property(lambda self: (self.l / self.n) ** self.D)

    def ndrange(self):
        """RFFT array size in the reference's (x, y[, z]) order."""
        return (self.nh,) + (self.n,) * (self.D - 1)

    def space_ndrange(self):
        return (self.n,) * self.D

    # ---------------------------------------------------------- allocators
    def spec(self, ncomp):
        return torch.zeros((ncomp, self.nz, self.n, self.nhp), dtype=self.cdtype, device=self.device)

    def phys(self, ncomp):
        return torch.zeros((ncomp, self.nz, self.n, self.n), dtype=self.dtype, device=self.device)

    def scalarfield(self):
        return self.spec(1)

    def vectorfield(self):
        return self.spec(self.D)

    def tensorfield(self):
        return self.spec(self.tensordim)

    def tensorfield_nonsym(self):
        return self.spec(self.D * self.D)

    def spacescalarfield(self):
        return self.phys(1)

    def spacevectorfield(self):
        return self.phys(self.D)

    def spacetensorfield(self):
        return self.phys(self.tensordim)

    def spacetensorfield_nonsym(self):
        return self.phys(self.D * self.D)

    # ----------------------------------------------------- host conversion
    def ref_spec_shape(self, ncomp):
        """numpy shape of ncomp spectral scalars in the reference memory order."""
        return (ncomp,) + (self.n,) * (self.D - 1) + (self.nh,)

    def upload_spectral(self, a):
        """numpy ``[comp, (z,) y, x]`` (the reference layout) -> device field."""
        a = np.ascontiguousarray(a, dtype=np.complex128 if self.prec == 64 else np.complex64)
        if a.ndim == self.D:
            a = a[None]
        t = self.spec(a.shape[0])
        check(lib().sc_upload_spectral(self.ctx, _p(t), a.ctypes.data_as(ctypes.c_void_p), a.shape[0]), self.ctx)
        return t

    def download_spectral(self, t):
        ncomp = t.shape[0]
        a = np.empty(self.ref_spec_shape(ncomp), np.complex128 if self.prec == 64 else np.complex64)
        check(lib().sc_download_spectral(self.ctx, a.ctypes.data_as(ctypes.c_void_p), _p(t), ncomp), self.ctx)
        return a

    def upload_physical(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64 if self.prec == 64 else np.float32)
        if a.ndim == self.D:
            a = a[None]
        t = self.phys(a.shape[0])
        check(lib().sc_upload_physical(self.ctx, _p(t), a.ctypes.data_as(ctypes.c_void_p), a.shape[0]), self.ctx)
        return t

    def download_physical(self, t):
        ncomp = t.shape[0]
        a = np.empty((ncomp,) + (self.n,) * self.D, np.float64 if self.prec == 64 else np.float32)
        check(lib().sc_download_physical(self.ctx, a.ctypes.data_as(ctypes.c_void_p), _p(t), ncomp), self.ctx)
        return a


def _p(t):
    """device pointer of a torch tensor (must be contiguous)."""
    if not t.is_contiguous():
        raise ValueError("field must be contiguous")
    if t.device.type != "cuda":
        raise ValueError("field must live on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def ncomp_of(t):
    return int(t.shape[0])
