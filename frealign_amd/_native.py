"""ctypes binding of libfrealign_hip.so (the C-ABI in include/frealign_hip.h).

The library is built in-tree by ``python -m frealign_amd.build``.  There is no
fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfrealign_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "frealign_hip.h")

_lib = None

vp, ci, cf, cd = C.c_void_p, C.c_int, C.c_float, C.c_double
pi_, pf, pd = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double)

SIGNATURES = {
    "fr_version": ([], C.c_char_p),
    "fr_last_error": ([], C.c_char_p),
    "fr_create": ([ci, C.POINTER(vp)], ci),
    "fr_destroy": ([vp], ci),
    "fr_set_stream": ([vp, vp], ci),
    "fr_synchronize": ([vp], ci),
    "fr_search_grid": ([cf, cf, cf, C.c_char, ci, vp, ci, vp, ci, pi_, pi_], ci),
    "fr_point_group": ([C.c_char, ci, vp, vp, pi_, pi_], ci),
    "fr_rotation": ([ci, vp, cf, vp, vp], ci),
    "fr_lcg_random": ([pi_], cf),
    "fr_fft2d_r2c": ([vp, ci, ci, vp], ci),
    "fr_fft2d_c2r": ([vp, ci, ci, vp], ci),
    "fr_prep_batch": ([vp, ci, ci, vp, vp, vp, vp, ci, vp, vp, vp], ci),
    "fr_pextract_batch": ([vp, ci, ci, vp, vp, vp, vp, vp], ci),
    "fr_cc3m_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, cf, vp, vp, vp], ci),
    "fr_cc3_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, cf, vp, vp, vp], ci),
    "fr_presb_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp], ci),
    "fr_sigma_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, ci, cf, vp, cf, vp, vp, vp], ci),
    "fr_plan_create": ([vp, ci, ci, vp, vp, vp, ci, ci, ci, cf, vp, vp, ci, vp, vp, cf, cf, cf,
                        vp, vp, vp, vp, vp, vp, C.POINTER(vp)], ci),
    "fr_plan_images": ([vp], vp),
    "fr_plan_load": ([vp, vp], ci),
    "fr_plan_run": ([vp, vp, vp, vp, vp], ci),
    "fr_plan_destroy": ([vp], ci),
    "fr_set_cc3m_variant": ([vp, ci, ci, ci], ci),
    "fr_autotune_cc3m": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, cf, vp, vp, vp,
                          C.c_char_p, ci], ci),
    "fr_powell_minimize": ([ci, vp, vp, cf, ci, ci, vp, vp, vp, vp], ci),
    "fr_refine_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, ci, cf, vp, vp, vp, ci, vp],
                        ci),
    "fr_search_list_batch": ([vp, ci, ci, ci, cf, ci, vp, ci, ci, vp, vp, ci, vp, ci, ci, vp, vp,
                              vp], ci),
    "fr_search_ragged_batch": ([vp, ci, ci, ci, cf, ci, vp, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci,
                                vp, vp, vp], ci),
    "fr_build_lists": ([vp, ci, ci, vp, vp, cf, ci, vp, vp, ci, ci, ci, ci, vp, vp, vp, vp], ci),
    "fr_rescore_batch": ([vp, ci, ci, ci, vp, vp, vp, vp, vp, ci, vp, ci, ci, ci, cf, vp, vp, vp,
                          vp], ci),
    "fr_cnn_param_count": ([ci, ci, ci, ci, C.POINTER(C.c_longlong)], ci),
    "fr_cnn_init": ([vp, ci, ci, ci, ci, C.c_ulonglong, vp, vp], ci),
    "fr_cnn_forward": ([vp, ci, ci, ci, ci, vp, ci, vp, vp], ci),
    "fr_cnn_forward_bf16": ([vp, ci, ci, ci, ci, vp, ci, vp, vp], ci),
    "fr_cnn_gradients": ([vp, ci, ci, ci, ci, vp, ci, vp, vp, vp, vp], ci),
    "fr_cnn_gradients_bf16": ([vp, ci, ci, ci, ci, vp, ci, vp, vp, vp, vp], ci),
    "fr_cnn_adam": ([vp, ci, ci, ci, ci, vp, vp, vp, vp, vp, ci, cf], ci),
    "fr_cnn_propose": ([vp, ci, ci, vp, ci, vp, vp], ci),
    "fr_cnn_images": ([vp, ci, ci, ci, vp, vp, vp], ci),
    "fr_cnn_synth": ([vp, ci, ci, ci, vp, vp, vp, vp, cf, C.c_ulonglong, vp], ci),
    "fr_orient_nearest": ([vp, ci, vp, ci, vp, ci, vp, vp], ci),
    "fr_orient_within": ([vp, ci, vp, ci, vp, ci, vp, cf, vp], ci),
    "fr_insert_batch": ([vp, ci, ci, vp, vp, vp, vp, ci, vp, vp, cf, cf, cf, vp, vp, vp, vp, vp,
                         vp], ci),
    "fr_insert_batch_se": ([vp, ci, ci, vp, vp, vp, vp, ci, vp, vp, cf, cf, cf, vp, vp, vp, vp,
                            vp, vp, vp, vp, vp, vp, vp], ci),
    "fr_finalize": ([vp, ci, vp, vp, vp, vp, vp, vp, ci, cf, cf, cf, cf, cf, ci, cf, vp, vp,
                     ci, vp, vp, vp, vp, vp, vp, vp], ci),
    "fr_find_fpart": ([vp, ci, C.c_double, C.c_double, vp, vp, vp, vp, ci, cf, cf, cf, cf, cf,
                       vp, vp, ci, vp], ci),
    "fr_sheltest": ([vp, ci, ci, vp, vp, vp, vp, cf, vp, vp, vp, vp, pi_], ci),
    "fr_local_search_batch": ([vp, ci, ci, vp, vp, vp, vp, vp, ci, cf, ci, cf, ci, ci, cf, vp,
                               vp, vp, vp, vp], ci),
    "fr_ctfrefine_batch": ([vp, ci, ci, vp, vp, vp, ci, vp, ci, cf, vp, vp, vp, vp, cf, cf, cf,
                            vp, vp, vp], ci),
    "fr_magrefine_batch": ([vp, ci, ci, vp, vp, vp, ci, vp, vp, vp, vp, cf, cf, cf, vp, vp, vp],
                           ci),
    "fr_mrc_info": ([C.c_char_p, vp, vp, vp, vp, vp], ci),
    "fr_mrc_read": ([C.c_char_p, C.c_longlong, C.c_longlong, vp], ci),
    "fr_mrc_write": ([C.c_char_p, ci, ci, ci, cf, vp, C.c_char_p], ci),
    "fr_par_format": ([ci] + [cf] * 6 + [ci] + [cf] * 8 + [C.c_char_p, ci], ci),
    "fr_par_parse": ([C.c_char_p, vp, vp], ci),
    "fr_calc_occ": ([ci, ci, cf, vp, vp, vp, vp], ci),
    "fr_sheltest_noalloc": ([vp, ci, ci, vp, vp, vp, vp, cf, vp, vp, vp, vp, pi_], ci),
    "fr_storeshift": ([vp, ci, cd, cd, ci, vp, vp, vp, vp, vp, vp], ci),
    "fr_combine_arrays": ([vp, ci, C.c_longlong, ci, ci, ci, vp, vp, vp, vp, ci, vp, vp, vp,
                           C.POINTER(C.c_longlong), pd, pd], ci),
    "fr_shiftvol": ([vp, ci, pd, cd, cf, ci, vp, vp, vp, vp, ci, vp, cf, vp, cf, ci], ci),
    "fr_correct3d": ([vp, ci, vp, ci, ci, vp, ci], ci),
    "fr_fft3d_r2c": ([vp, ci, vp], ci),
    "fr_fft3d_c2r": ([vp, ci, vp], ci),
    "fr_mask3d": ([vp, ci, vp, cf, cf, cf, C.POINTER(C.c_float), C.POINTER(C.c_float),
                   C.POINTER(C.c_float)], ci),
    "fr_fpart_from_mw": ([ci, cf, cf, cf, vp, C.POINTER(C.c_float)], ci),
    "fr_profile_enable": ([vp, ci], ci),
    "fr_profile_read": ([vp, ci, pd, C.POINTER(C.c_longlong)], ci),
    "fr_search_layout": ([vp, ci, ci, ci, pi_, pi_], ci),
    "fr_search_bank": ([vp, ci, ci, ci, vp, ci, vp, ci, vp], ci),
    "fr_search_batch": ([vp, ci, ci, ci, cf, ci, vp, ci, ci, ci, vp, vp, ci, vp, vp, vp], ci),
    "fr_ccp_view_batch": ([vp, ci, ci, ci, vp, vp, vp, vp, ci, ci, cf, cf, ci, vp, vp, vp], ci),
    "fr_psearch": ([vp, ci, ci, vp, vp, vp, vp, cf, ci, ci, vp, vp, ci, ci, vp, cf, vp, ci, pi_,
                    vp, C.c_char], ci),
}


class NativeError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is missing: build it with `python -m frealign_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().fr_last_error().decode(errors="replace")
        raise NativeError(f"{what}: {msg}" if what else msg)


def declared_symbols() -> list[str]:
    """Every function name declared in include/frealign_hip.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(fr_[a-z0-9_]+)\s*\(", txt))
    # the per-call entry points carry the reference's Fortran symbols (name_)
    names |= set(re.findall(r"\b([a-z][a-z0-9]*_)\s*\(", txt))
    return sorted(names)
