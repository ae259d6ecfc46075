"""ctypes binding of libtaichirtm_hip.so (C ABI declared in include/taichirtm_hip.h).

The product path has no CPU fallback: loading fails loudly when the library
is missing, and device calls need a visible HIP device.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtaichirtm_hip.so")

ABI_VERSION = 7
OK = 0
ERR_ARG, ERR_HIP, ERR_UNSUPPORTED = -1, -2, -3
PSV, SH = 1, 2
KERNEL_AUTO, KERNEL_SIMPLE, KERNEL_FUSED, KERNEL_BLOCKED, KERNEL_WAVE = 0, 1, 2, 3, 4
NOALLOC_SCRATCH_BYTES = 64 * 1024 * 1024  # TRTM_NOALLOC_SCRATCH_BYTES

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_f32 = ctypes.c_float


class StepDesc(ctypes.Structure):
    """Mirror of trtm_step_desc."""
    _fields_ = [
        ("struct_size", _i32), ("nx", _i32), ("nz", _i32), ("nt", _i32),
        ("dt", _f32), ("inv_dx", _f32), ("inv_dz", _f32), ("src_factor", _f32),
        ("fw", _i32), ("systems", _i32), ("kernel", _i32), ("tile_x", _i32),
        ("block_steps", _i32), ("snap_base", _i32),
        ("sponge_table", _p),
        ("lam", _p), ("mu", _p), ("mxz", _p), ("myx", _p), ("myz", _p),
        ("inv_rho", _p), ("inv_rho_u", _p), ("inv_rho_w", _p),
        ("surface_matrix", _p),
        ("u", _p), ("v", _p), ("w", _p), ("sxx", _p), ("szz", _p), ("sxz", _p),
        ("syx", _p), ("syz", _p),
        ("n_inj", _i32), ("n_inj_groups", _i32),
        ("inj_loc", _p), ("inj_order", _p), ("inj_group_start", _p),
        ("inj_u", _p), ("inj_v", _p), ("inj_w", _p),
        ("n_rec", _i32), ("tile_z", _i32),
        ("rec_loc", _p), ("rec_u", _p), ("rec_v", _p), ("rec_w", _p),
        ("check_interval", _i32), ("step_offset", _i32),
        ("check_flags", _p),
        ("isnap", _i32), ("num_snaps", _i32),
        ("snap_u", _p), ("snap_v", _p), ("snap_w", _p),
        ("isnaps_host", _p),
        ("image_u", _p), ("image_v", _p), ("image_w", _p),
        ("workspace", _p), ("workspace_bytes", ctypes.c_int64),
        ("snap_stride", ctypes.c_int64), ("snap_sponge", _i32), ("reserved_snap", _i32),
    ]


_f64 = ctypes.c_double


class FwiDesc(ctypes.Structure):
    """Mirror of trtm_fwi_desc."""
    _fields_ = [
        ("struct_size", _i32), ("real_bytes", _i32), ("nx", _i32), ("nz", _i32), ("nt", _i32),
        ("fw", _i32), ("systems", _i32), ("n_src", _i32), ("n_src_groups", _i32),
        ("n_rec", _i32),
        ("dt", _f64), ("inv_dx", _f64), ("inv_dz", _f64), ("src_factor", _f64),
        ("sponge_table", _p),
        ("lam", _p), ("mu", _p), ("mxz", _p), ("myx", _p), ("myz", _p), ("inv_rho", _p),
        ("inv_rho_u", _p), ("inv_rho_w", _p),
        ("u", _p), ("v", _p), ("w", _p), ("sxx", _p), ("szz", _p), ("sxz", _p), ("syx", _p),
        ("syz", _p),
        ("pu", _p), ("pv", _p), ("pw", _p), ("psxx", _p), ("pszz", _p), ("psxz", _p),
        ("psyx", _p), ("psyz", _p),
        ("src_loc", _p), ("src_order", _p), ("src_group_start", _p),
        ("wav_u", _p), ("wav_v", _p), ("wav_w", _p),
        ("rec_loc", _p), ("rec_order", _p), ("rec_group_start", _p),
        ("n_rec_groups", _i32), ("hist_base", _i32),
        ("rec_u", _p), ("rec_v", _p), ("rec_w", _p),
        ("res_u", _p), ("res_v", _p), ("res_w", _p),
        ("hist_len", _i32), ("hist_every", _i32),
        ("hist", _p),
        ("g_lam", _p), ("g_mu", _p), ("g_mxz", _p), ("g_myx", _p), ("g_myz", _p), ("g_ir", _p),
        ("g_iru", _p), ("g_irw", _p),
        ("cell_map", _p),
        ("kernel", _i32), ("tile_x", _i32), ("tile_z", _i32), ("block_steps", _i32),
        ("workspace", _p),
        ("dif_lam", _p), ("dif_mu", _p), ("dif_mxz", _p), ("dif_myx", _p), ("dif_myz", _p),
        ("dif_inv_rho", _p), ("dif_inv_rho_u", _p), ("dif_inv_rho_w", _p),
        ("du", _p), ("dv", _p), ("dw", _p), ("dsxx", _p), ("dszz", _p), ("dsxz", _p),
        ("dsyx", _p), ("dsyz", _p),
        ("scat_workspace", _p),
    ]


# name -> (restype, argtypes) for every entry point of include/taichirtm_hip.h
SIGNATURES = {
    "trtm_abi_version": (ctypes.c_int, []),
    "trtm_last_error": (ctypes.c_char_p, []),
    "trtm_prepare_medium": (ctypes.c_int, [_i32, _i32] + [_p] * 10 + [_p]),
    "trtm_sponge_table": (ctypes.c_int, [_i32, _f32, _p]),
    "trtm_sponge_fill": (ctypes.c_int, [_i32, _i32, _i32, _p, _p, _p]),
    "trtm_plan_injection": (ctypes.c_int, [_i32, _i32, _i32, _p, _p, _p, _p]),
    "trtm_workspace_bytes": (ctypes.c_int64, [ctypes.POINTER(StepDesc)]),
    "trtm_batch_workspace_bytes": (ctypes.c_int64, [_i32]),
    "trtm_block_config": (ctypes.c_int, [ctypes.POINTER(StepDesc), _p]),
    "trtm_block_configs": (ctypes.c_int, [_i32, _i32, _p, _i32]),
    "trtm_forward_steps_batch": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _p, ctypes.c_int64,
                                                _i32, _i32, _i32, _p]),
    "trtm_backward_steps_batch": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _p,
                                                 ctypes.c_int64, _i32, _i32, _p]),
    "trtm_forward_steps": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _i32, _i32, _p]),
    "trtm_backward_steps": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _i32, _p]),
    "trtm_forward_steps_noalloc": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _i32, _i32,
                                                  _p]),
    "trtm_check_finite": (ctypes.c_int, [_i32, _i32, _p, _p, _p, _i32, _p, _p]),
    "trtm_prepare_medium_f64": (ctypes.c_int, [_i32, _i32] + [_p] * 10 + [_p]),
    "trtm_fwi_prepare": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _p]),
    "trtm_fwi_workspace_bytes": (ctypes.c_int64, [ctypes.POINTER(FwiDesc)]),
    "trtm_fwi_block_configs": (ctypes.c_int, [_i32, _i32, _p, _i32]),
    "trtm_fwi_forward": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _i32, _i32, _p]),
    "trtm_fwi_forward_batch": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _i32, _i32, _i32, _p]),
    "trtm_fwi_adjoint_batch": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _i32, _i32, _i32, _p]),
    "trtm_fwi_forward_scattered_batch": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _i32, _i32, _i32,
                                                        _p]),
    "trtm_fwi_adjoint": (ctypes.c_int, [ctypes.POINTER(FwiDesc), _i32, _i32, _p]),
    "trtm_fwi_chain_gradient": (ctypes.c_int, [_i32, _i32] + [_p] * 15 + [_p]),
    "trtm_transpose": (ctypes.c_int, [_i32, _i32, _p, _p, _p]),
    "trtm_plan_create": (ctypes.c_int, [ctypes.POINTER(StepDesc), _i32, _p, ctypes.c_int64, _i32,
                                        _i32, _i32, _i32, _p, ctypes.POINTER(_p)]),
    "trtm_plan_launch": (ctypes.c_int, [_p, _p]),
    "trtm_plan_destroy": (ctypes.c_int, [_p]),
}

_LIB = None


class HIPError(RuntimeError):
    pass


def load():
    """Load the library (once) and check its ABI version."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HIPError(
                "libtaichirtm_hip.so is not built (%s); run `python -m taichirtm_amd.build`"
                % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.trtm_abi_version() != ABI_VERSION:
            raise HIPError("libtaichirtm_hip ABI %d, expected %d"
                           % (lib.trtm_abi_version(), ABI_VERSION))
        _LIB = lib
    return _LIB


def check(rc, what):
    if rc < 0:
        msg = load().trtm_last_error().decode(errors="replace")
        if rc == ERR_ARG:
            raise ValueError("%s: %s" % (what, msg))
        raise HIPError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args), name)
