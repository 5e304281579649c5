"""ctypes binding of libsymcode_hip.so (the C ABI in include/symcode.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SYMCODE_LIB selects an alternative build (e.g. a tuning variant)
LIBPATH = os.environ.get("SYMCODE_LIB") or os.path.join(_HERE, "libsymcode_hip.so")

_lib = None

c_int, c_double, c_void_p, c_int64, c_uint64 = (
    ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64)
P = c_void_p
DP = ctypes.POINTER(c_double)

# name -> argtypes (all return int status except sc_last_error)
SIGNATURES = {
    "sc_create": [c_int, c_int, c_double, c_int, c_int, ctypes.POINTER(P)],
    "sc_destroy": [P],
    "sc_info": [P, ctypes.POINTER(c_int64)],
    "sc_set_stream": [P, P],
    "sc_set_sync": [P, c_int],
    "sc_sync": [P],
    "sc_profile": [P, c_int],
    "sc_profile_report": [P, ctypes.c_char_p, c_int],
    "sc_upload_spectral": [P, P, P, c_int],
    "sc_download_spectral": [P, P, P, c_int],
    "sc_upload_physical": [P, P, P, c_int],
    "sc_download_physical": [P, P, P, c_int],
    "sc_to_phys": [P, P, P, c_int],
    "sc_to_spec": [P, P, P, c_int],
    "sc_irfft": [P, P, P, c_int],
    "sc_rfft": [P, P, P, c_int],
    "sc_twothirds": [P, P, c_int],
    "sc_project": [P, P],
    "sc_symmetrize": [P, P, c_int],
    "sc_viscosity": [P, P, P, c_double],
    "sc_tensordivergence": [P, P, P],
    "sc_vectordivergence": [P, P, P],
    "sc_vectorgradient": [P, P, P],
    "sc_strainrate": [P, P, P],
    "sc_gaussianfilter": [P, P, c_int, c_double],
    "sc_cutoff": [P, P, P, P, c_int],
    "sc_inverse_cutoff": [P, P, P, P, c_int],
    "sc_axpy": [P, P, P, c_double, c_int],
    "sc_make_tracefree_spectral": [P, P],
    "sc_make_tracefree_physical": [P, P],
    "sc_nonlinearity": [P, P, P],
    "sc_stress": [P, P, P, c_double],
    "sc_convectiondiffusion": [P, P, P, c_double],
    "sc_wray3": [P, P, c_double, c_double],
    "sc_rk4": [P, P, c_double, c_double],
    "sc_propose_timestep": [P, P, c_double, DP],
    "sc_rk4_cfl": [P, P, c_double, c_double, c_double, DP],
    "sc_wray3_cfl": [P, P, c_double, c_double, c_double, DP],
    "sc_rk4_steps": [P, P, c_int, c_double, c_double, c_double, DP],
    "sc_les_wray3_cfl": [P, P, c_double, c_double, c_double, c_int, c_double, c_double, DP],
    "sc_set_band_forcing": [P, c_double, c_int],
    "sc_energy": [P, P, c_int, DP],
    "sc_energy_noalloc": [P, P, c_int, DP],
    "sc_spectraldot": [P, P, P, c_int, DP, DP],
    "sc_dissipation": [P, P, c_double, DP],
    "sc_filter_reynolds": [P, P, c_double, c_double, DP],
    "sc_palinstrophy": [P, P, c_double, DP],
    "sc_spectrum": [P, P, c_int, c_int, DP],
    "sc_shell_energies": [P, P, c_int, c_int, DP],
    "sc_shell_scale": [P, P, c_int, c_int, DP],
    "sc_maintain_shell_energy": [P, P, c_int, DP],
    "sc_randn": [P, P, c_int, c_uint64],
    "sc_randomfield": [P, P, c_uint64, c_int, DP],
    "sc_taylorgreen": [P, P, c_double, c_int],
    "sc_gradient_phys": [P, P, P],
    "sc_closure_clark": [P, P, P, c_double],
    "sc_closure_smagorinsky": [P, P, P, c_double, c_double],
    "sc_closure_dynsmag": [P, P, P, c_double, DP],
    "sc_closure_verstappen": [P, P, P, c_double, c_double],
    "sc_sfs_dissipation": [P, P, P, P],
    "sc_shell_transfer": [P, P, P, c_int, DP],
    "sc_moments": [P, P, ctypes.c_long, DP],
    "sc_sample_mean": [P, P, ctypes.c_long, DP],
    "sc_apriori_metrics": [P, P, P, ctypes.c_long, DP],
    "sc_closure_bardina": [P, P, P, c_double, c_double],
    "sc_les_rhs": [P, P, P, c_double, c_int, c_double, c_double],
    "sc_les_wray3": [P, P, c_double, c_double, c_int, c_double, c_double],
    "sc_rhs_vjp": [P, P, P, P, c_double],
    "sc_rk4_vjp": [P, P, P, P, c_double, c_double],
    "sc_wray3_vjp": [P, P, P, P, c_double, c_double],
    "sc_cnn_create": [c_int, c_int, c_int, c_int, c_int, c_uint64, ctypes.POINTER(P)],
    "sc_cnn_destroy": [P],
    "sc_cnn_info": [P, ctypes.POINTER(c_int64)],
    "sc_cnn_set_stream": [P, P],
    "sc_cnn_get_params": [P, P],
    "sc_cnn_set_params": [P, P],
    "sc_cnn_get_grads": [P, P],
    "sc_cnn_zero_grad": [P],
    "sc_cnn_pack_input": [P, P, c_int, c_int, ctypes.c_long, ctypes.c_long, c_int, c_double, c_int, c_double, P],
    "sc_cnn_forward": [P, P, P, c_int],
    "sc_cnn_mse": [P, P, P, P, c_int, DP],
    "sc_cnn_backward": [P, P, P, P, c_int],
    "sc_cnn_adam": [P, c_double, c_double, c_double, c_double],
    "sc_norm2": [P, P, c_int, DP],
    "sc_cnn_loss": [P, P, P, P, c_int, c_int, DP],
    "sc_cnn_adamw": [P, c_double, c_double, c_double, c_double, c_double, c_double],
    "sc_cnn_set_tracefree": [P, c_int],
    "sc_cnn_profile": [P, c_int],
    "sc_cnn_profile_report": [P, ctypes.c_char_p, c_int],
    "sc_cnn_pack_target": [P, P, c_int, ctypes.c_long, ctypes.c_long, c_int, c_double, P],
    "sc_set_cnn_closure": [P, P, c_double, c_double, c_double],
    "sc_closure_cnn": [P, P, P],
    "sc_les_rk4": [P, P, c_double, c_double, c_int, c_double, c_double],
    "sc_les_rk4_cfl": [P, P, c_double, c_double, c_double, c_int, c_double, c_double, DP],
    "sc_les_rhs_vjp": [P, P, P, P, c_double, c_int],
    "sc_les_rk4_vjp": [P, P, P, P, c_double, c_double, c_int],
    "sc_les_wray3_vjp": [P, P, P, P, c_double, c_double, c_int],
    "sc_lincomb": [P, P, P, c_double, P, c_double, c_int],
    "sc_pmlp_create": [c_int, c_int, c_int, c_int, c_int, c_uint64, ctypes.c_long, ctypes.POINTER(P)],
    "sc_tbnn_create": [c_int, c_int, c_int, c_int, c_int, c_uint64, ctypes.c_long, ctypes.POINTER(P)],
    "sc_equi_create": [c_int, c_int, c_int, c_int, c_int, c_uint64, ctypes.c_long, ctypes.POINTER(P)],
    "sc_pmlp_destroy": [P],
    "sc_pmlp_info": [P, ctypes.POINTER(c_int64)],
    "sc_pmlp_set_stream": [P, P],
    "sc_pmlp_get_params": [P, P],
    "sc_pmlp_set_params": [P, P],
    "sc_pmlp_get_grads": [P, P],
    "sc_pmlp_zero_grad": [P],
    "sc_pmlp_features": [P, P, P, c_int, ctypes.c_long, c_double, c_double, P, P],
    "sc_pmlp_forward": [P, P, P, ctypes.c_long],
    "sc_pmlp_backward": [P, P, P, ctypes.c_long],
    "sc_pmlp_loss": [P, P, P, P, ctypes.c_long, c_int, DP],
    "sc_pmlp_adamw": [P, c_double, c_double, c_double, c_double, c_double, c_double],
    "sc_set_pmlp_closure": [P, P, c_double, c_double, c_double, c_double],
    "sc_closure_pmlp": [P, P, P],
    "sc_set_pmlp_symmetrize": [P, c_int],
}


class SymcodeError(RuntimeError):
    pass


def lib():
    """Load libsymcode_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with `python symmetrycode.jl_amd/build.py` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = c_int
        L.sc_last_error.argtypes = [P]
        L.sc_last_error.restype = ctypes.c_char_p
        L.sc_cnn_last_error.argtypes = [P]
        L.sc_cnn_last_error.restype = ctypes.c_char_p
        L.sc_pmlp_last_error.argtypes = [P]
        L.sc_pmlp_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check_cnn(rc, net):
    if rc != 0:
        msg = lib().sc_cnn_last_error(net)
        raise SymcodeError(f"symcode cnn error {rc}: {msg.decode() if msg else ''}")


def check_pmlp(rc, net):
    if rc != 0:
        msg = lib().sc_pmlp_last_error(net)
        raise SymcodeError(f"symcode pmlp error {rc}: {msg.decode() if msg else ''}")


def check(rc, ctx=None):
    if rc != 0:
        msg = lib().sc_last_error(ctx)
        raise SymcodeError(f"symcode error {rc}: {msg.decode() if msg else ''}")
