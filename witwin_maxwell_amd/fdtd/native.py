"""ctypes binding of the C ABI in include/maxwell_fdtd.h.

This is the drop-in boundary: the reference reaches its kernels through
``solver.fdtd_module.<CamelName>(**kw).launchRaw()`` (fdtd/cuda/backend.py:81-115)
over torch ops; here the same calls go through ``extern "C"`` entry points with
raw device pointers and the current torch stream.  Loading fails loudly when the
library is missing -- there is no CPU or PyTorch fallback.
"""

from __future__ import annotations

import ctypes
import os

import torch

from .. import build as _build

MW_MAX_POLES = 4
MW_MAX_TERMS = 16
MW_MAX_PLANES = 16

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
vp = ctypes.c_void_p


class MwGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("bc_lo", ctypes.c_int32 * 3), ("bc_hi", ctypes.c_int32 * 3),
                ("inv_e", vp * 3), ("inv_h", vp * 3)]


class MwCoeffs(ctypes.Structure):
    _fields_ = [("decay", vp * 6), ("curl", vp * 6), ("decay_value", ctypes.c_float * 6),
                ("curl_value", ctypes.c_float * 6)]


class MwCpml(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("inv_kappa_e", vp * 3), ("b_e", vp * 3), ("c_e", vp * 3),
                ("inv_kappa_h", vp * 3), ("b_h", vp * 3), ("c_h", vp * 3),
                ("psi", vp * 12), ("psi_low", ctypes.c_int32 * 12), ("psi_high_start", ctypes.c_int32 * 12),
                ("psi_high_len", ctypes.c_int32 * 12)]


class MwFields(ctypes.Structure):
    _fields_ = [("e", vp * 3), ("h", vp * 3)]


class MwPole(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("decay", ctypes.c_float), ("restoring", ctypes.c_float),
                ("drive_scale", ctypes.c_float), ("drive", vp), ("polarization", vp), ("current", vp)]


class MwAde(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int32 * 3), ("dt", ctypes.c_float), ("inv_eps", vp * 3),
                ("poles", (MwPole * MW_MAX_POLES) * 3), ("active", vp * 3), ("weight", vp * 3)]


class MwTerm(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("component", ctypes.c_int32), ("offset", ctypes.c_int32 * 3),
                ("shape", ctypes.c_int32 * 3), ("patch", vp), ("delay", vp), ("signal_table", vp),
                ("wave_kind", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("t_offset", ctypes.c_double), ("frequency", ctypes.c_double), ("fwidth", ctypes.c_double),
                ("amplitude", ctypes.c_double), ("phase", ctypes.c_double), ("delay0", ctypes.c_double),
                ("values", vp), ("line", vp), ("line_axis", ctypes.c_int32), ("line_base", ctypes.c_int32),
                ("line_step", ctypes.c_int32), ("line_len", ctypes.c_int32)]


class MwAuxLine(ctypes.Structure):
    _fields_ = [("n_e", ctypes.c_int32), ("pad", ctypes.c_int32), ("e", vp), ("h", vp), ("decay_e", vp),
                ("curl_e", vp), ("decay_h", vp), ("curl_h", vp), ("source", vp)]


class MwPlaneObs(ctypes.Structure):
    _fields_ = [("component", ctypes.c_int32), ("axis", ctypes.c_int32), ("index", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("re", vp), ("im", vp)]


class MwProgram(ctypes.Structure):
    _fields_ = [("grid", MwGrid), ("coeffs", MwCoeffs), ("cpml", MwCpml), ("ade", MwAde), ("fields", MwFields),
                ("dt", ctypes.c_double), ("total_steps", ctypes.c_int32), ("has_ade", ctypes.c_int32),
                ("n_h_terms", ctypes.c_int32), ("n_e_terms", ctypes.c_int32),
                ("h_terms", ctypes.POINTER(MwTerm)), ("e_terms", ctypes.POINTER(MwTerm)),
                ("nf_obs", ctypes.c_int32), ("n_planes", ctypes.c_int32), ("n_points", ctypes.c_int32),
                ("z_tiling", ctypes.c_int32), ("obs_e_cos", vp), ("obs_e_sin", vp), ("obs_h_cos", vp),
                ("obs_h_sin", vp), ("planes", ctypes.POINTER(MwPlaneObs)), ("point_component", vp),
                ("point_linear", vp), ("point_re", vp), ("point_im", vp), ("nf_dft", ctypes.c_int32),
                ("shell_chunk", ctypes.c_int32), ("dft_cos", vp), ("dft_sin", vp), ("dft_re", vp * 3), ("dft_im", vp * 3),
                ("fused", ctypes.c_int32), ("xchunk", ctypes.c_int32), ("fields_alt", MwFields), ("psi_h_alt", vp * 6),
                ("h_terms_dev", vp), ("e_terms_dev", vp), ("aux", MwAuxLine),
                ("plane_stage", vp), ("plane_stage_floats", ctypes.c_int64), ("plane_stage_depth", ctypes.c_int32),
                ("pad_stage", ctypes.c_int32)]


class MwWindow(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int32 * 3), ("hi", ctypes.c_int32 * 3), ("capacity", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("e_store", vp), ("h_store", vp), ("g_decay", vp * 3), ("g_curl", vp * 3)]


class MwAdjoint(ctypes.Structure):
    _fields_ = [("fields", MwFields), ("psi", vp * 12), ("polarization", (vp * MW_MAX_POLES) * 3),
                ("current", (vp * MW_MAX_POLES) * 3), ("g_point_re", vp), ("g_point_im", vp),
                ("g_plane_re", vp * MW_MAX_PLANES), ("g_plane_im", vp * MW_MAX_PLANES),
                ("g_dft_re", vp * 3), ("g_dft_im", vp * 3)]


STRUCTS = (MwGrid, MwCoeffs, MwCpml, MwFields, MwPole, MwAde, MwTerm, MwPlaneObs, MwProgram, MwWindow, MwAdjoint)

EXPORTS = ("mw_update_magnetic", "mw_update_electric", "mw_add_source_patch", "mw_add_time_shifted_patch",
           "mw_accumulate_dft", "mw_accumulate_plane", "mw_accumulate_points", "mw_run_steps", "mw_capture_steps",
           "mw_reverse_steps", "mw_batch_create", "mw_batch_run", "mw_batch_destroy",
           "mw_adjoint_batch_create", "mw_adjoint_batch_capture", "mw_adjoint_batch_reverse",
           "mw_adjoint_batch_destroy", "mw_batch_run_noalloc", "mw_update_pole_current",
           "mw_apply_polarization_current", "mw_update_aux_magnetic", "mw_update_aux_electric", "mw_last_error",
           "mw_abi_version", "mw_abi_struct_sizes")

_LIB = None


class NativeError(RuntimeError):
    pass


def load_library(path=None):
    """Load libmaxwell_fdtd.so (after torch, so both share torch's HIP runtime)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or _build.library_path()
    if not os.path.exists(path):
        raise NativeError(f"native FDTD library not built: {path} (run __graft_entry__.build() or "
                          f"python -m witwin_maxwell_amd.build)")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    lib.mw_update_magnetic.argtypes = [P(MwGrid), P(MwCoeffs), P(MwCpml), P(MwFields), ctypes.c_int32, vp]
    lib.mw_update_electric.argtypes = [P(MwGrid), P(MwCoeffs), P(MwCpml), P(MwAde), P(MwFields), ctypes.c_int32, vp]
    i32 = ctypes.c_int32
    lib.mw_add_source_patch.argtypes = [vp, i32, i32, i32, vp, i32, i32, i32, i32, i32, i32, ctypes.c_float, vp]
    dbl = ctypes.c_double
    lib.mw_add_time_shifted_patch.argtypes = [vp, i32, i32, i32, vp, vp, i32, i32, i32, i32, i32, i32, dbl, i32,
                                              dbl, dbl, dbl, dbl, dbl, vp]
    lib.mw_accumulate_dft.argtypes = [vp, ctypes.c_int64, i32, vp, vp, vp, vp, vp]
    lib.mw_accumulate_plane.argtypes = [vp, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp]
    lib.mw_accumulate_points.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, vp]
    lib.mw_run_steps.argtypes = [P(MwProgram), i32, i32, vp]
    lib.mw_capture_steps.argtypes = [P(MwProgram), i32, i32, P(MwWindow), vp]
    lib.mw_reverse_steps.argtypes = [P(MwProgram), P(MwAdjoint), P(MwWindow), i32, i32, vp]
    lib.mw_batch_create.argtypes = [P(P(MwProgram)), i32, i32, P(vp)]
    lib.mw_batch_run.argtypes = [vp, i32, i32, i32, vp]
    lib.mw_batch_destroy.argtypes = [vp]
    lib.mw_batch_run_noalloc.argtypes = [P(P(MwProgram)), i32, i32, i32, i32, vp]
    f32 = ctypes.c_float
    i64 = ctypes.c_int64
    lib.mw_update_pole_current.argtypes = [i32, vp, vp, vp, vp, i64, f32, f32, f32, vp]
    lib.mw_apply_polarization_current.argtypes = [vp, vp, vp, i64, f32, vp]
    lib.mw_update_aux_magnetic.argtypes = [vp, vp, vp, vp, i32, vp]
    lib.mw_update_aux_electric.argtypes = [vp, vp, vp, vp, i32, i32, f32, vp]
    lib.mw_adjoint_batch_create.argtypes = [vp, P(P(MwProgram)), P(P(MwAdjoint)), P(P(MwWindow)), i32, P(vp)]
    lib.mw_adjoint_batch_capture.argtypes = [vp, i32, i32, vp]
    lib.mw_adjoint_batch_reverse.argtypes = [vp, i32, i32, vp]
    lib.mw_adjoint_batch_destroy.argtypes = [vp]
    lib.mw_last_error.restype = ctypes.c_char_p
    lib.mw_abi_struct_sizes.argtypes = [P(ctypes.c_int64), i32]
    lib.mw_abi_struct_sizes.restype = i32
    sizes = (ctypes.c_int64 * 16)()
    n = lib.mw_abi_struct_sizes(sizes, 16)
    for i, st in enumerate(STRUCTS[:n]):
        if ctypes.sizeof(st) != sizes[i]:
            raise NativeError(f"ABI layout mismatch for {st.__name__}: python {ctypes.sizeof(st)} vs C {sizes[i]}")
    if path == _build.library_path():
        _LIB = lib
    return lib


def check(rc, lib=None):
    if rc != 0:
        lib = lib or load_library()
        raise RuntimeError(lib.mw_last_error().decode())


def ptr(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise RuntimeError("native FDTD kernels need CUDA (HIP) tensors")
    if not t.is_contiguous():
        raise RuntimeError("native FDTD kernels need contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def current_stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
