"""symcode: MI355X-native pseudo-spectral Navier-Stokes hot path with the
call signatures of agdestein/SymmetryCode.jl (see include/symcode.h).

Every operation runs in libsymcode_hip.so (hand-written HIP for gfx950);
importing raises if the library has not been built.
"""
from ._lib import LIBPATH, SymcodeError, lib  # noqa: F401
from .grid import Grid  # noqa: F401
from . import solver, filtering, closures, autograd, cnn, pmlp  # noqa: F401
from .solver import *  # noqa: F401,F403
from .filtering import cutoff, gaussianfilter, sfs, inverse_cutoff  # noqa: F401
from .closures import (create_clark, create_smagorinsky, create_dynamic_smagorinsky, create_verstappen,  # noqa: F401
                       create_bardina,
                       create_learned_closure, les_rhs, les_wray3, les_wray3_cfl, les_rk4, les_rk4_cfl,
                       les_rhs_vjp, les_rk4_vjp, les_wray3_vjp, solve_les, gradient_phys)
from .pmlp import PointwiseMLP, TBNN, EquivariantNet, PointwiseClosure, create_pointwise_closure, symmetrize_pointwise  # noqa: F401

lib()  # fail loudly at import if the HIP library is missing
