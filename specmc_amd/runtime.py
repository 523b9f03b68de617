"""ctypes binding of ``libspecmc_hip.so`` (the C ABI of include/specmc_hip.h).

There is no CPU fallback: importing the engine without the built library, or
stepping a model without a HIP device, raises ``RuntimeError``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libspecmc_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
L = ctypes.c_long
LL = ctypes.c_longlong
D = ctypes.c_double

# name -> argtypes, in the order of include/specmc_hip.h
SIGNATURES = {
    "smx_version": [],
    "smx_last_error": [],
    "smx_gemm": [I, I, I, I, D, D, P, L, L, L, L, P, L, L, L, L, P, L, L, L, L,
                 I, I, P, I, P, I, I, P],
    "smx_gemm_ex": [I, I, I, I, I, D, D, P, L, L, L, L, P, L, L, L, L, P, L, L, L, L,
                    I, I, P, I, P, I, I, P],
    "smx_gemm_tune": [I, I, I, I, I, I, I, I],
    "smx_gemm_tune_clear": [],
    "smx_lu_factor": [I, I, I, P, P, I, P, P],
    "smx_lu_solve": [I, I, I, P, P, P, P, I, D, P, P, I, I, P, P],
    "smx_invert": [I, I, I, P, I, P, P],
    "smx_clock_start": [I, I],
    "smx_clock_stop": [I, P, P, P],
    "smx_set_gemm_mfma": [I],
    "smx_set_xprod_mfma": [I],
    "smx_set_gj_rows": [I],
    "smx_set_gj_reg": [I],
    "smx_set_gj_mfma": [I],
    "smx_set_inv_pairs": [I],
    "smx_set_lud_fuse": [I],
    "smx_dense_lu": [I, I, I, P, L, P, P, P],
    "smx_dense_lu_solve": [I, I, I, P, L, P, P, P, P, P],
    "smx_dense_lu_perm": [I, I, P, P, P, P],
    "smx_dense_lu_solve_mw_bytes": [I, I, I],
    "smx_dense_lu_solve_mw": [I, I, I, P, L, P, P, P, P, P, L, P],
    "smx_eta_drift": [I, I, L, P, P, P, P, P],
    "smx_rownorm": [I, I, L, L, P, P, P, P],
    "smx_wdot": [I, I, L, P, L, P, L, P, P, L, P, P],
    "smx_inv_apply": [I, I, I, P, P, P, I, D, P, P, I, I, P, P],
    "smx_tau_assemble": [I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P],
    "smx_cheb_mult": [I, I, I, P, P, P, P],
    "smx_schur_blocks": [I, I, I, P, P, P, P, P, D, P, P, P],
    "smx_schur_tau": [I, I, I, P, P, P, P],
    "smx_eta_law": [I, I, I, I, P, P, P, P, P, P, P],
    "smx_eta_guard_profile": [I, I, I, I, P, P, P, I, P, P, P],
    "smx_xprod": [I, I, I, I, I, I, I, P, P, P, L, P, P, P, P, P, P, P],
    "smx_schur_apply_work": [I, I, I],
    "smx_set_schur_compact": [I],
    "smx_schur_apply": [I, I, I, I, I, I, P, P, L, P, P, P, P, P, P, P, P, P, P, P, LL, P, P],
    "smx_schur_apply_static_bytes": [],
    "smx_schur_apply_static": [I, I, I, I, I, I, P, P, L, P, P, P, P, P, P, P, P, P, P, P, P],
    "smx_advect_prod": [I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P],
    "smx_axpby": [I, L, I, D, P, P, D, P, P, P, P, P],
    "smx_rowscale": [I, I, I, I, P, P, D, P, I, P, P, P],
    "smx_tau_rows": [I, I, I, I, P, P, P, P, P, P],
    "smx_stokes_combine": [I, I, I, I, P, P, P, P, P, P, P, P, P],
    "smx_theta_combine": [I, I, I, I, P, P, P, P, P, P, P, P, P, P],
    "smx_theta_rhs": [I, I, I, I, P, P, P, P, P, P, P, P, P, P, P],
    "smx_phi_grid": [I, I, L, P, P, P, P, P, P, P],
    "smx_diag": [I, I, I, I, P, P, P, P, L, P, P, P, P, P, P],
    "smx_maxabs": [I, I, L, P, P, P, P],
    "smx_advect_prod_general": [I, I, I, I, I] + [P] * 18,
    "smx_phi_grid_general": [I, I, L, P, P, P, P, P, P, P, P, P],
    "smx_diag_general": [I, I, I, I, P, P, P, P, P, P, P, L, P, P, P, P, P, P],
    "smx_maxabs_pair": [I, I, L, P, P, P, P, P],
    "smx_fg_init": [I, I, L, I, P, P, P, P, P, P, D, D, I, P, P],
    "smx_fg_cycle_end": [I, I, L, I, P, P, P, P, P, P, P, P, P, P],
    "smx_fg_post": [I, I, L, I, P, P, P, P, P, P, P, P, P, P, P, I, P],
    "smx_fg_count": [I, P, P, P],
    "smx_cfl_dt": [I, P, P, P, P, P],
}

# entry points that return a value rather than a status
_VALUE_RETURNING = ("smx_schur_apply_work", "smx_schur_apply_static_bytes",
                    "smx_dense_lu_solve_mw_bytes")

_lib = None


class _Checked:
    """Calls an entry point and raises on a nonzero status."""

    def __init__(self, lib, name):
        self.fn = getattr(lib, name)
        self.name = name
        self.nargs = len(SIGNATURES[name])
        self.err = lib.smx_last_error

    def __call__(self, *args):
        # ctypes passes surplus arguments unconverted (as C int), which
        # would truncate pointers: insist on the declared count
        if len(args) != self.nargs:
            raise TypeError("%s takes %d arguments, got %d" % (self.name, self.nargs, len(args)))
        rc = self.fn(*args)
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s"
                               % (self.name, rc, self.err().decode()))
        return rc


class _Lib:
    def __init__(self, cdll):
        self.cdll = cdll
        for name in SIGNATURES:
            if name in ("smx_version", "smx_last_error"):
                continue
            if name in _VALUE_RETURNING:
                setattr(self, name[4:], getattr(cdll, name))
                continue
            setattr(self, name[4:], _Checked(cdll, name))

    def version(self):
        return self.cdll.smx_version()


def load():
    """-> the bound library (built in-tree by ``specmc_amd.build_ext``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "the HIP library %s is not built; run "
                "`python -m specmc_amd.build_ext` (or __graft_entry__.build())"
                % LIB_PATH)
        cdll = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(cdll, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        cdll.smx_last_error.restype = ctypes.c_char_p
        cdll.smx_schur_apply_work.restype = ctypes.c_longlong
        cdll.smx_schur_apply_static_bytes.restype = ctypes.c_longlong
        cdll.smx_dense_lu_solve_mw_bytes.restype = ctypes.c_long
        _lib = _Lib(cdll)
    return _lib


def require_device():
    """Raise unless a HIP device is usable through torch."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("specmc_amd runs on a HIP device (MI355X); none is "
                           "visible -- there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())
