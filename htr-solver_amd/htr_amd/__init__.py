"""htr_amd -- Python host side of the MI355X HTR right-hand-side library.

Loads libhtr_amd.so (built from htr-solver_amd/csrc for gfx950) and exposes
the C ABI of include/htr_amd.h through ctypes.  Device memory and streams come
from PyTorch (plumbing only); every computation runs in the library's HIP
kernels.  There is no CPU fallback: if the library or a GPU is missing, the
calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.environ.get("HTR_AMD_LIBRARY") or os.path.join(PKG, "libhtr_amd.so")

HTR_EULER_HYBRID, HTR_EULER_TENO, HTR_EULER_TENOA, HTR_EULER_TENOLAD, HTR_EULER_KG = range(5)
HTR_CHEM_FP64, HTR_CHEM_FP32_IR = 0, 1
HTR_CHEM_JAC_ANALYTIC = 0x10   # or-ed into the precision: analytic Rosenbrock Jacobian
HTR_BC_DIRICHLET, HTR_BC_ADIABATIC_WALL, HTR_BC_ISOTHERMAL_WALL, HTR_BC_NSCBC_INFLOW, HTR_BC_NSCBC_OUTFLOW = 0, 1, 2, 3, 4
ERRORS = {-1: "HTR_E_ARG", -2: "HTR_E_UNSUPPORTED", -3: "HTR_E_DEVICE", -4: "HTR_E_FIELD", -5: "HTR_E_NUMERIC"}


class HtrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class View(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("sx", C.c_int64), ("sy", C.c_int64), ("sz", C.c_int64), ("sc", C.c_int64)]


class Rect(C.Structure):
    _fields_ = [("lo", C.c_int64 * 3), ("hi", C.c_int64 * 3)]


class Grid(C.Structure):
    _fields_ = [("bounds", Rect), ("exec", Rect)]


class Fields(C.Structure):
    _fields_ = [("nType", View * 3), ("m_e", View * 3), ("m_d", View * 3), ("m_s", View * 3),
                ("rho", View), ("mu", View), ("lam", View), ("Di", View), ("SoS", View),
                ("pressure", View), ("temperature", View), ("MassFracs", View), ("MolarFracs", View),
                ("velocity", View), ("Conserved", View), ("Conserved_old", View), ("Conserved_t", View),
                ("Conserved_t_old", View), ("shockSensor", View * 3), ("DucrosSensor", View),
                ("MolarFracs_profile", View), ("velocity_profile", View), ("temperature_profile", View),
                ("dudtBoundary", View), ("dTdtBoundary", View)]


class Diag(C.Structure):
    _fields_ = [("decisions", View * 3), ("nsteps", View), ("flags", View)]


EXPORTS = [
    "htr_init", "htr_last_error", "htr_version",
    "htr_mix_create_constprop", "htr_mix_create_airmix", "htr_mix_create_species", "htr_mix_add_reaction",
    "htr_mix_finalize", "htr_mix_nspec", "htr_mix_refs", "htr_mix_destroy",
    "htr_update_properties", "htr_update_conserved", "htr_update_primitive",
    "htr_ducros_sensor", "htr_shock_sensor", "htr_euler_flux", "htr_diffusion_flux",
    "htr_chem_source", "htr_chem_update", "htr_max_spectral_radius",
    "htr_rk_update", "htr_rk_init", "htr_full_rhs",
    "htr_profile", "htr_profile_report", "htr_chem_launch_config", "htr_stencil_table",
    "htr_workspace_reserve", "htr_rk_stage", "htr_full_rhs_cfl", "htr_set_bc",
    "htr_max_spectral_radius_noalloc", "htr_nscbc_inflow", "htr_nscbc_outflow", "htr_nscbc_farfield",
    "htr_max_mach", "htr_rhs_overlap",
]

_lib = None


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HtrError(-3, "libhtr_amd.so not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, d, i, f = C.c_void_p, C.c_double, C.c_int, C.c_float
    G, F, D = C.POINTER(Grid), C.POINTER(Fields), C.POINTER(Diag)
    ip, fp, dp = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double)
    sig = {
        "htr_init": (i, [i]), "htr_last_error": (C.c_char_p, []), "htr_version": (C.c_char_p, []),
        "htr_profile": (i, [i]), "htr_profile_report": (i, [C.c_char_p, C.c_size_t]),
        "htr_chem_launch_config": (i, [i, i]),
        "htr_rhs_overlap": (i, [i]),
        "htr_stencil_table": (i, [i, i, i, dp]),
        "htr_workspace_reserve": (i, [G, F, vp, vp]),
        "htr_rk_stage": (i, [G, F, i, d, d, vp, vp, D, vp]),
        "htr_full_rhs_cfl": (i, [G, F, i, d, d, d, d, vp, i, vp, vp]),
        "htr_set_bc": (i, [i, i, G, F, d, vp, vp]),
        "htr_nscbc_inflow": (i, [i, G, F, vp, vp]),
        "htr_nscbc_outflow": (i, [i, G, F, d, d, d, vp, vp]),
        "htr_max_mach": (i, [i, G, F, dp, vp]),
        "htr_nscbc_farfield": (i, [i, G, F, d, d, d, vp, vp]),
        "htr_mix_create_constprop": (i, [C.POINTER(vp), d, d, d, i, d, d, d]),
        "htr_mix_create_airmix": (i, [C.POINTER(vp), d, d, d, i, ip, dp]),
        "htr_mix_create_species": (i, [C.POINTER(vp), i, dp, fp, ip, dp, dp, dp, ip, dp]),
        "htr_mix_add_reaction": (i, [vp, i, d, f, f, i, i, ip, fp, i, ip, fp, i, ip, fp]),
        "htr_mix_finalize": (i, [vp, d, d, d, dp]),
        "htr_mix_nspec": (i, [vp]), "htr_mix_refs": (i, [vp, dp]), "htr_mix_destroy": (None, [vp]),
        "htr_update_properties": (i, [G, F, vp, vp]),
        "htr_update_conserved": (i, [G, F, vp, vp]),
        "htr_update_primitive": (i, [G, F, vp, vp]),
        "htr_ducros_sensor": (i, [G, F, d, vp]),
        "htr_shock_sensor": (i, [i, G, F, vp, vp]),
        "htr_euler_flux": (i, [i, i, G, F, d, d, d, vp, D, vp]),
        "htr_diffusion_flux": (i, [i, G, F, vp, vp]),
        "htr_chem_source": (i, [G, F, vp, vp]),
        "htr_chem_update": (i, [G, F, d, i, vp, D, vp]),
        "htr_max_spectral_radius": (i, [G, F, vp, dp, vp]),
        "htr_max_spectral_radius_noalloc": (i, [G, F, vp, vp, vp]),
        "htr_rk_update": (i, [G, F, i, i, d, vp]),
        "htr_rk_init": (i, [G, F, i, vp]),
        "htr_full_rhs": (i, [G, F, i, d, d, d, i, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def check(rc):
    if rc != 0:
        raise HtrError(rc, lib().htr_last_error().decode())
    return rc


def init(device=0):
    check(lib().htr_init(device))


from .api import *  # noqa: E402,F401,F403
