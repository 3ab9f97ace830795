"""ctypes binding of libppm_hip.so (the C ABI declared in include/ppm_api.h).

The library is built in-tree by ``python -m spammm_amd.build``.  There is no
fallback: if the library or a GPU is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libppm_hip.so")

PPM_OK = 0
PPM_ERR_INVALID_ARG = 1
PPM_ERR_HIP = 2
PPM_ERR_UNSUPPORTED = 3
PPM_MODEL_LJ = 0
PPM_MODEL_MORSE = 1
PPM_DIRECT_MAX_ATOMS = 128
PPM_BATCH_NOALLOC_MAX_MOL = 65536
PPM_FFT_CORRELATE = 0
PPM_FFT_CONVOLVE = 1
PPM_FFT_POISSON = 2
PPM_FFT_POISSON_CONVOLVE = 3
PPM_EWALD_FULL = 0
PPM_EWALD_VACUUM = 1

_F4 = ctypes.c_float * 4


class PpmGridSpec(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("reserved", ctypes.c_int),
                ("p0", _F4), ("dA", _F4), ("dB", _F4), ("dC", _F4),
                ("dinvA", _F4), ("dinvB", _F4), ("dinvC", _F4)]


class PpmTip(ctypes.Structure):
    _fields_ = [("tipA", _F4), ("tipB", _F4), ("tipC", _F4), ("stiffness", _F4), ("dpos0", _F4),
                ("relax", _F4), ("surfFF", _F4)]


class PpmRelaxCtl(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("f2conv", ctypes.c_float), ("check_before", ctypes.c_int),
                ("scan_ny", ctypes.c_int), ("f32_anchors", ctypes.c_int), ("damped", ctypes.c_int),
                ("lanes", ctypes.c_int)]


EXPORTS = (
    "ppm_version", "ppm_last_error", "ppm_grid_project", "ppm_relax_strokes_tilted", "ppm_relax_strokes",
    "ppm_fe_in_strokes_tilted", "ppm_z_stencil", "ppm_axis_operator", "ppm_relax_adjoint",
    "ppm_grid_adjoint_workspace_size", "ppm_grid_adjoint", "ppm_relax_polish", "ppm_batch_workspace_size",
    "ppm_batch_forward", "ppm_relax_direct_morse", "ppm_direct_morse_vjp_workspace_size", "ppm_direct_morse_vjp",
    "ppm_grid_project_add", "ppm_grid_project_acc64", "ppm_grid_adjoint_add", "ppm_grid_project_f64",
    "ppm_relax_strokes_tilted_f64", "ppm_fdbm_workspace_size", "ppm_fdbm_fft_combine", "ppm_fdbm_pauli_scale",
    "ppm_fdbm_pad_roll", "ppm_fdbm_flip3", "ppm_fdbm_dispersion", "ppm_fdbm_compose_gradient",
    "ppm_ewald_num_g", "ppm_ewald_system_bytes", "ppm_ewald_prepare", "ppm_ewald_eval", "ppm_ewald_brute",
    "ppm_ewald_cluster", "ppm_ewald_grid_workspace_size", "ppm_ewald_tip_grid", "ppm_ewald_kept_count_offset",
    "ppm_relax_state_diag", "ppm_relax_strokes_2d", "ppm_grid_project_window", "ppm_relax_strokes_tilted_window",
    "ppm_batch_forward_noalloc", "ppm_fdbm_release_plans", "ppm_relax_strokes_tilted_df", "ppm_stream_create",
    "ppm_stream_destroy",
)

_lib = None


def f4(v) -> _F4:
    a = [float(x) for x in list(v)[:4]]
    a += [0.0] * (4 - len(a))
    return _F4(*a)


def load():
    """Load libppm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m spammm_amd.build` "
                           "(the AFM path has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    P = ctypes.POINTER
    L.ppm_version.restype = i
    L.ppm_version.argtypes = []
    L.ppm_last_error.restype = ctypes.c_char_p
    L.ppm_last_error.argtypes = []
    L.ppm_grid_project.restype = i
    L.ppm_grid_project.argtypes = [i, vp, vp, i, P(PpmGridSpec), _F4, _F4, i, i, P(ctypes.c_float * 3),
                                   P(ctypes.c_float * 3), i, vp, vp]
    L.ppm_grid_project_acc64.restype = i
    L.ppm_grid_project_acc64.argtypes = L.ppm_grid_project.argtypes
    L.ppm_grid_project_add.restype = i
    L.ppm_grid_project_add.argtypes = [i, vp, vp, i, P(PpmGridSpec), _F4, _F4, i, i, P(ctypes.c_float * 3),
                                       P(ctypes.c_float * 3), i, vp, vp, vp]
    D4, D3 = ctypes.c_double * 4, ctypes.c_double * 3
    L.ppm_grid_project_f64.restype = i
    L.ppm_grid_project_f64.argtypes = [i, vp, vp, i, P(PpmGridSpec), D4, D4, i, i, P(D3), P(D3), i, vp, vp]
    L.ppm_relax_strokes_tilted_f64.restype = i
    L.ppm_relax_strokes_tilted_f64.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), P(PpmRelaxCtl), vp, vp, vp]
    I3 = ctypes.c_int * 3
    F3 = ctypes.c_float * 3
    L.ppm_fdbm_workspace_size.restype = ctypes.c_size_t
    L.ppm_fdbm_workspace_size.argtypes = [i, i, i]
    L.ppm_fdbm_fft_combine.restype = i
    L.ppm_fdbm_fft_combine.argtypes = [i, i, i, i, f, vp, vp, f, vp, vp, ctypes.c_size_t, vp]
    L.ppm_fdbm_pauli_scale.restype = i
    L.ppm_fdbm_pauli_scale.argtypes = [vp, ctypes.c_longlong, f, f, f, vp]
    L.ppm_fdbm_pad_roll.restype = i
    L.ppm_fdbm_pad_roll.argtypes = [vp, P(I3), vp, P(I3), P(I3), P(I3), vp]
    L.ppm_fdbm_flip3.restype = i
    L.ppm_fdbm_flip3.argtypes = [vp, vp, P(I3), vp]
    L.ppm_fdbm_dispersion.restype = i
    L.ppm_fdbm_dispersion.argtypes = [vp, i, P(F3), f, P(I3), f, i, vp, vp]
    L.ppm_fdbm_compose_gradient.restype = i
    L.ppm_fdbm_compose_gradient.argtypes = [vp, vp, vp, P(I3), f, vp, vp]
    D2 = ctypes.c_double * 2
    sz = ctypes.c_size_t
    L.ppm_ewald_num_g.restype = i
    L.ppm_ewald_num_g.argtypes = [i]
    L.ppm_ewald_system_bytes.restype = sz
    L.ppm_ewald_system_bytes.argtypes = [i, i]
    L.ppm_ewald_kept_count_offset.restype = sz
    L.ppm_ewald_kept_count_offset.argtypes = [i, i]
    L.ppm_ewald_prepare.restype = i
    L.ppm_ewald_prepare.argtypes = [vp, vp, i, P(D2), P(D2), i, ctypes.c_double, vp, sz, vp]
    L.ppm_ewald_eval.restype = i
    L.ppm_ewald_eval.argtypes = [i, vp, i, vp, sz, i, i, P(D2), P(D2), vp, vp]
    L.ppm_ewald_brute.restype = i
    L.ppm_ewald_brute.argtypes = [vp, i, vp, i, P(D2), P(D2), i, i, vp, vp]
    L.ppm_ewald_cluster.restype = i
    L.ppm_ewald_cluster.argtypes = [vp, i, vp, i, vp, vp]
    L.ppm_ewald_grid_workspace_size.restype = sz
    L.ppm_ewald_grid_workspace_size.argtypes = [i, i]
    L.ppm_ewald_tip_grid.restype = i
    L.ppm_ewald_tip_grid.argtypes = [vp, sz, i, i, P(D2), P(D2), P(PpmGridSpec), _F4, _F4, i, vp, sz, vp, vp]
    L.ppm_grid_adjoint_add.restype = i
    L.ppm_grid_adjoint_add.argtypes = [i, vp, vp, i, P(PpmGridSpec), _F4, _F4, i, i, P(ctypes.c_float * 3),
                                       P(ctypes.c_float * 3), i, vp, vp, vp, vp, ctypes.c_size_t, i, vp, vp]
    L.ppm_relax_strokes_tilted.restype = i
    L.ppm_relax_strokes_tilted.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), P(PpmRelaxCtl), vp, vp, vp]
    L.ppm_relax_strokes.restype = i
    L.ppm_relax_strokes.argtypes = [vp, P(PpmGridSpec), vp, i, i, _F4, _F4, _F4, _F4, P(PpmRelaxCtl), vp, vp, vp,
                                    vp]
    L.ppm_fe_in_strokes_tilted.restype = i
    L.ppm_fe_in_strokes_tilted.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), _F4, _F4, vp, vp]
    L.ppm_z_stencil.restype = i
    L.ppm_z_stencil.argtypes = [vp, i, i, i, i, vp, vp, i, i, vp, vp]
    ll = ctypes.c_longlong
    L.ppm_axis_operator.restype = i
    L.ppm_axis_operator.argtypes = [vp, ll, ll, ll, _F4, i, i, i, i, vp, vp, i, i, f, i, vp, vp]
    L.ppm_relax_adjoint.restype = i
    L.ppm_relax_adjoint.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), vp, vp, vp, vp, vp]
    L.ppm_grid_project_window.restype = i
    L.ppm_grid_project_window.argtypes = [i, vp, vp, i, P(PpmGridSpec), i, i, _F4, _F4, i, i, P(ctypes.c_float * 3),
                                          P(ctypes.c_float * 3), i, vp, vp]
    L.ppm_relax_strokes_tilted_window.restype = i
    L.ppm_relax_strokes_tilted_window.argtypes = [vp, P(PpmGridSpec), i, i, vp, i, i, P(PpmTip), P(PpmRelaxCtl), vp,
                                                  vp, vp, vp]
    L.ppm_relax_strokes_2d.restype = i
    L.ppm_relax_strokes_2d.argtypes = [vp, P(PpmGridSpec), vp, i, i, f, f, f, f, P(PpmRelaxCtl), vp, vp, vp, vp]
    L.ppm_relax_state_diag.restype = i
    L.ppm_relax_state_diag.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), vp, vp, vp]
    L.ppm_batch_workspace_size.restype = ctypes.c_size_t
    L.ppm_batch_workspace_size.argtypes = [i]
    L.ppm_batch_forward.restype = i
    L.ppm_batch_forward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_size_t, i, i, vp]
    L.ppm_relax_strokes_tilted_df.restype = i
    L.ppm_relax_strokes_tilted_df.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), P(PpmRelaxCtl), vp, vp, vp,
                                              vp, i, i, vp, vp]
    L.ppm_fdbm_release_plans.restype = i
    L.ppm_fdbm_release_plans.argtypes = []
    L.ppm_stream_create.restype = i
    L.ppm_stream_create.argtypes = [P(vp)]
    L.ppm_stream_destroy.restype = i
    L.ppm_stream_destroy.argtypes = [vp]
    L.ppm_batch_forward_noalloc.restype = i
    L.ppm_batch_forward_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i, i, vp]
    L.ppm_relax_direct_morse.restype = i
    L.ppm_relax_direct_morse.argtypes = [vp, vp, i, vp, i, i, P(PpmTip), _F4, _F4, P(PpmRelaxCtl), vp, vp, vp]
    L.ppm_direct_morse_vjp_workspace_size.restype = ctypes.c_size_t
    L.ppm_direct_morse_vjp_workspace_size.argtypes = [i, i, i]
    L.ppm_direct_morse_vjp.restype = i
    L.ppm_direct_morse_vjp.argtypes = [vp, vp, i, vp, i, i, P(PpmTip), _F4, _F4, vp, vp, vp, vp, vp, ctypes.c_size_t,
                                       vp, vp]
    L.ppm_relax_polish.restype = i
    L.ppm_relax_polish.argtypes = [vp, P(PpmGridSpec), vp, i, i, P(PpmTip), i, vp, vp, vp]
    L.ppm_grid_adjoint_workspace_size.restype = ctypes.c_size_t
    L.ppm_grid_adjoint_workspace_size.argtypes = [i, P(PpmGridSpec)]
    L.ppm_grid_adjoint.restype = i
    L.ppm_grid_adjoint.argtypes = [i, vp, vp, i, P(PpmGridSpec), _F4, _F4, i, i, P(ctypes.c_float * 3),
                                   P(ctypes.c_float * 3), i, vp, vp, vp, ctypes.c_size_t, i, vp, vp]
    _lib = L
    return L


def check(rc: int) -> None:
    """Map a PPM status code to the exception type the reference raises for that failure."""
    if rc == PPM_OK:
        return
    msg = load().ppm_last_error().decode(errors="replace")
    if rc == PPM_ERR_INVALID_ARG:
        raise ValueError(msg)
    raise RuntimeError(msg)
