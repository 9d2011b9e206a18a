"""ctypes binding of libcudaem_amd.so (the C ABI of include/cudaem_amd.h).

Loading fails loudly: there is no CPU fallback.  Device buffers are passed as
torch tensors (their data_ptr()); every shape, dtype and device is checked on
the host before a kernel is launched.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudaem_amd.so")

EM_OK = 0
EM_ERR_INVALID = 1
EM_ERR_UNSUPPORTED = 2
EM_ERR_ALLOC = 3
EM_ERR_HIP = 4
_ERRNAMES = {1: "EM_ERR_INVALID", 2: "EM_ERR_UNSUPPORTED", 3: "EM_ERR_ALLOC", 4: "EM_ERR_HIP"}

BORDER_MIN, BORDER_MAX, BORDER_MIDDLE = 0, 1, 2

LENS_FIELDS = [
    "c_10", "c_12", "phi_12", "c_21", "phi_21", "c_23", "phi_23", "c_30", "c_32", "phi_32",
    "c_34", "phi_34", "c_41", "phi_41", "c_43", "phi_43", "c_45", "phi_45", "c_50", "c_52",
    "phi_52", "c_54", "phi_54", "c_56", "phi_56", "m", "inner_aper_mrad", "outer_aper_mrad",
]

# Every entry point declared in include/cudaem_amd.h.
EXPORTS = [
    "em_fft_size_supported", "em_plan_create", "em_plan_destroy", "em_plan_set_stream",
    "em_plan_constants", "em_plan_set_probe_group", "em_plan_group_size", "em_plan_set_pass_timing",
    "em_plan_pass_times", "em_plan_last_error",
    "em_set_element_coefs", "em_set_spline_params", "em_set_element_params", "em_plan_set_fine_slicing",
    "em_plan_set_specimen", "em_slice_info",
    "em_potential_slices", "em_spline_tables", "em_propagator", "em_build_probes",
    "em_plane_wave", "em_fft2d", "em_multislice_forward", "em_multislice_forward_multi", "em_multislice_adjoint",
    "em_detector_sums", "em_detector_sums_noalloc", "em_ctf_image", "em_pctf_image", "em_tcc_image", "em_diffraction", "em_stem", "em_thickness",
    "em_thickness_multi",
    "em_multislice_ragged", "em_time_passes", "em_host_slicing", "em_host_slicing_fine", "em_host_quadrature", "em_host_spline",
    "em_host_constants", "em_plan_get_tuning", "em_plan_set_tuning", "em_plan_autotune",
    "em_plan_set_band_skip",
]


class em_grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("lx", ctypes.c_double),
                ("ly", ctypes.c_double), ("dz", ctypes.c_double), ("bwl", ctypes.c_int)]


class em_beam(ctypes.Structure):
    _fields_ = [("e0_kv", ctypes.c_double), ("theta", ctypes.c_double), ("phi", ctypes.c_double)]


class em_lens(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in LENS_FIELDS]


# partial coherence (em_coherence of include/cudaem_amd.h)
COH_TEMPORAL_SPATIAL, COH_TEMPORAL, COH_SPATIAL = 1, 2, 3


class em_coherence(ctypes.Structure):
    _fields_ = [("dsf_sigma", ctypes.c_double), ("dsf_npoints", ctypes.c_int), ("ssf_beta", ctypes.c_double),
                ("ssf_npoints", ctypes.c_int), ("mode", ctypes.c_int)]


class em_tuning(ctypes.Structure):
    _fields_ = [("row_blocks", ctypes.c_int), ("col_blocks", ctypes.c_int), ("col_tile", ctypes.c_int),
                ("probe_group", ctypes.c_int), ("col_variant", ctypes.c_int), ("row_variant", ctypes.c_int),
                ("persistent", ctypes.c_int), ("row_radix", ctypes.c_int), ("col_radix", ctypes.c_int)]


class em_atoms(ctypes.Structure):
    _fields_ = [("n", ctypes.c_longlong), ("x", ctypes.POINTER(ctypes.c_double)),
                ("y", ctypes.POINTER(ctypes.c_double)), ("z", ctypes.POINTER(ctypes.c_double)),
                ("Z", ctypes.POINTER(ctypes.c_int)), ("occ", ctypes.POINTER(ctypes.c_double))]


_P = ctypes.c_void_p
_I = ctypes.c_int

# 3D potential parameterisations (EM_PT_* of include/cudaem_amd.h)
PT_DOYLE_0_4, PT_PENG_0_4, PT_PENG_0_12, PT_KIRKLAND_0_12, PT_WEICKENMEIER_0_12, PT_LOBATO_0_12 = 1, 2, 3, 4, 5, 6
_D = ctypes.c_double
_F = ctypes.c_float
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PLL = ctypes.POINTER(ctypes.c_longlong)

_SIGS = {
    "em_fft_size_supported": ([_I], _I),
    "em_plan_create": ([ctypes.POINTER(em_grid), ctypes.POINTER(em_beam), _I, ctypes.POINTER(_P)], _I),
    "em_plan_destroy": ([_P], _I),
    "em_plan_set_stream": ([_P, _P], _I),
    "em_plan_constants": ([_P, _PD, _PD, _PD], _I),
    "em_plan_set_probe_group": ([_P, _I], _I),
    "em_plan_group_size": ([_P, _I, ctypes.POINTER(_I)], _I),
    "em_plan_set_pass_timing": ([_P, _I], _I),
    "em_plan_pass_times": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "em_plan_last_error": ([_P], ctypes.c_char_p),
    "em_set_element_coefs": ([_P, _I, _I, _PD, _PD, _I, _PD, _PD], _I),
    "em_set_spline_params": ([_P, _I, _D, _D, _D], _I),
    "em_set_element_params": ([_P, _I, _I, _PD, _PD], _I),
    "em_plan_set_fine_slicing": ([_P, _I, _I], _I),
    "em_plan_set_specimen": ([_P, ctypes.POINTER(em_atoms), _I, _PI], _I),
    "em_slice_info": ([_P, _I, _PD, _PD, _PLL, _PLL], _I),
    "em_potential_slices": ([_P, _I, _I, _P], _I),
    "em_spline_tables": ([_P, _I, _PI, _PI, _PI, _PD, _PD, _PD, _PD, _PD], _I),
    "em_propagator": ([_P, _P], _I),
    "em_build_probes": ([_P, ctypes.POINTER(em_lens), _D, _PD, _PD, _I, _P], _I),
    "em_plane_wave": ([_P, _I, _P], _I),
    "em_fft2d": ([_P, _P, _I, _I], _I),
    "em_multislice_forward": ([_P, _P, _P, _I, _P, _I, _P, _I], _I),
    "em_multislice_forward_multi": ([_P, _P, _P, _P, _I], _I),
    "em_multislice_adjoint": ([_P, _P, _P, _I, _P, _I, _P, _I, _P, _P, _P], _I),
    "em_detector_sums": ([_P, _P, _I, _I, _PD, _PD, _F, _I, _P], _I),
    "em_detector_sums_noalloc": ([_P, _P, _I, _I, _PD, _PD, _F, _I, _P], _I),
    "em_ctf_image": ([_P, ctypes.POINTER(em_lens), _P, _I, _P, _P, _P, _F, _I], _I),
    "em_pctf_image": ([_P, ctypes.POINTER(em_lens), ctypes.POINTER(em_coherence), _P, _I, _P, _P, _P, _F, _I], _I),
    "em_tcc_image": ([_P, ctypes.POINTER(em_lens), ctypes.POINTER(em_coherence), _P, _I, _P, _P, _P, _F, _I], _I),
    "em_diffraction": ([_P, _P, _I, _P, _F, _I], _I),
    "em_stem": ([_P, ctypes.POINTER(em_lens), _I, _PD, _PD, _I, _PD, _PD, _F, _I, _P], _I),
    "em_thickness": ([_P, _P, _I, _P, _I, _D, _D, _I, _F, _I, _P, _PI], _I),
    "em_thickness_multi": ([_P, _P, _I, _P, _I, _D, _D, _I, _F, _I, _P, _PI], _I),
    "em_multislice_ragged": ([_I, ctypes.POINTER(_P), ctypes.POINTER(_P), _PI, ctypes.POINTER(_P), _PI,
                              _I, _PD, _PD, _I, ctypes.POINTER(_P), _P], _I),
    "em_time_passes": ([_P, _P, _P, _I, _I, ctypes.POINTER(_F), ctypes.POINTER(_F)], _I),
    "em_host_slicing": ([ctypes.c_longlong, _PD, _D, _I, _I, _PI, _PD, _PLL, _PLL], _I),
    "em_host_slicing_fine": ([ctypes.c_longlong, _PD, _D, _I, _D, _I, _PI, _PD, _PLL, _PLL], _I),
    "em_host_quadrature": ([_I, _I, _PD, _PD], _I),
    "em_host_spline": ([_I, _I, _D, _D, _D, _PD, _PD, _PD, _PD, _PD], _I),
    "em_host_constants": ([_D, _PD, _PD, _PD], _I),
    "em_plan_get_tuning": ([_P, ctypes.POINTER(em_tuning)], _I),
    "em_plan_set_tuning": ([_P, ctypes.POINTER(em_tuning)], _I),
    "em_plan_autotune": ([_P, _I, _I, _PD], _I),
    "em_plan_set_band_skip": ([_P, _D, _PI, _PI], _I),
}

_lib = None


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cudaem_amd.build` "
                "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class EmError(RuntimeError):
    pass


def check(status, plan=None, what=""):
    if status != EM_OK:
        msg = _ERRNAMES.get(status, str(status))
        detail = ""
        if plan:
            detail = lib().em_plan_last_error(plan).decode(errors="replace")
        raise EmError(f"{what}: {msg} {detail}".strip())
