"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the
flang-compiled reference routines (oracle/_ref/libfrealign_ref.so).

Test infrastructure only: the product path (frealign_amd) never imports this.
Arrays follow the reference's storage (see oracle/fr_oracle.h): complex data
is interleaved float32/float64, half-planes are JC x NSAM with x fastest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfrealign_ref.so")

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
c64p = np.ctypeslib.ndpointer(dtype=np.complex64, flags="C_CONTIGUOUS")
c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
cf, ci, cd = C.c_float, C.c_int, C.c_double
pf, pi_, pd = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_double)


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_or = None
_ref = None


def oracle():
    global _or
    if _or is None:
        lib = _load(ORACLE_SO)
        if lib is None:
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        _declare_oracle(lib)
        _or = lib
    return _or


def ref():
    """The flang-compiled reference routines, or None when not built."""
    global _ref
    if _ref is None:
        lib = _load(REF_SO)
        if lib is None:
            return None
        _declare_ref(lib)
        _ref = lib
    return _ref


def _declare_oracle(L):
    L.or_random.argtypes = [pi_]
    L.or_random.restype = cf
    L.or_dm.argtypes = [cf, cf, cf, cf, C.c_void_p, f32p]
    L.or_getsymmat.argtypes = [C.c_char, ci, f32p, i32p, pi_]
    L.or_getsymmat.restype = ci
    L.or_searchang.argtypes = [cf, cf, cf, C.c_char, ci, i32p, ci, C.c_void_p, pi_, pf]
    L.or_searchang.restype = ci
    L.or_va04_testfn.argtypes = [ci, f32p, C.c_void_p]
    L.or_va04_testfn.restype = cf
    L.or_va04a.argtypes = [f32p, f32p, ci, cf, ci, ci, ci, C.c_void_p, C.c_void_p, pi_]
    L.or_va04a.restype = cf
    L.or_calcsinc.argtypes = [f32p, ci]
    L.or32_prefine.argtypes = [ci, cf, c64p, c64p, ci, ci, f32p, i32p, cf, f32p, c64p, cf, cf, cf,
                               cf, f32p, C.c_void_p, pi_]
    L.or32_prefine.restype = cf
    L.or32_ctfrefine.argtypes = [ci, ci, c64p, f32p, f32p, ci, cf, cf, c64p, f32p, f32p, cf, cf,
                                 cf, f32p, f32p, pi_]
    L.or32_ctfrefine.restype = cf
    L.or32_magrefine.argtypes = [ci, ci, c64p, f32p, f32p, c64p, f32p, f32p, cf, cf, cf, f32p,
                                 f32p, pi_]
    L.or32_magrefine.restype = cf
    L.or_combine_counters.argtypes = [ci, ci, ci, C.POINTER(C.c_int64), f64p, f64p,
                                      C.POINTER(C.c_int64), pd, pd]
    L.or_rfft2_f.argtypes = [ci, f32p]
    L.or_irfft2_f.argtypes = [ci, f32p]
    L.or_rfft2_d.argtypes = [ci, f64p]
    L.or_irfft2_d.argtypes = [ci, f64p]
    for pre, rp, cp in (("or32_", f32p, c64p), ("or64_", f64p, c128p)):
        g = lambda n: getattr(L, pre + n)  # noqa: E731
        g("ctf").argtypes = [cf] * 8 + [ci, ci, cf, cf, rp]
        g("maskcos").argtypes = [ci, rp, cf, cf, cf, cf, cf]
        g("maskcos2").argtypes = [ci, rp, cf, cf, cf, cf, cf]
        g("pweight").argtypes = [ci, cp, cp, f32p, cp, ci]
        g("ctfapply").argtypes = [ci, c64p] + [cf] * 15 + [f32p, ci, cp, cp, cp]
        g("ainterpo3ds").argtypes = [ci, c64p, cf, cf, cf, rp]
        g("pextract").argtypes = [ci, cp, c64p, cf, cf, cf, cp, f32p]
        g("cc3m").argtypes = [ci, cf, cp, c64p, ci, ci, cf, cf, cf, cf, cf, cf, f32p, cp,
                              cf, cf, cf, cf, f32p]
        g("cc3m").restype = cd
        g("cc3").argtypes = [ci, cf, cp, c64p, ci, ci, cf, cf, cf, cf, cf, cf, cp, cf, cf, cf, cf,
                             f32p]
        g("cc3").restype = cd
        g("sigma").argtypes = [ci, cf, c64p, cp, cf, cf, cf, cf, cf, cp, ci, cf, cf, f32p, cf,
                               f32p, rp]
        g("presb").argtypes = [ci, cf, cp, c64p, cf, cf, cf, cf, cf, rp, cp, cf, cf, cf, cf, f32p]
        g("pinsert").argtypes = [ci, cf, cf, cp, cp, cp, rp, pd, cf, cf, cf, cf, cf, cf, cf, cf,
                                 ci, i32p, f32p, C.POINTER(C.c_int64), cf, cf, f32p, ci]
        g("pinsert_trilin").argtypes = [ci, ci, cp, cp, cf, cp, rp, pd, cf, cf,
                                        C.POINTER(C.c_int64), f32p, cf, cf, cf, cf, ci]
        g("pinsert_se").argtypes = [ci, cf, cf, cp, cp, cp, rp, pd, cf, cf, cf, cf, cf, cf, cf,
                                    cf, ci, i32p, f32p, C.POINTER(C.c_int64), cf, cf, f32p, ci,
                                    cp, i32p, rp, rp, rp]
        g("pinsert_trilin_se").argtypes = [ci, ci, cp, cp, cf, cp, rp, pd, cf, cf,
                                           C.POINTER(C.c_int64), f32p, cf, cf, cf, cf, ci, cp,
                                           i32p, rp, rp, rp]
        g("storeshift").argtypes = [ci, cd, cd, ci, cp, rp, cp, rp, cp, cp]
        g("combine_arrays").argtypes = [ci, C.c_longlong, ci, ci, ci, cp, rp, cp, rp]
        g("shiftvol").argtypes = [ci, pd, cd, cf, ci, cp, cp, rp, rp, ci, f32p, cf, f32p, cf]
        g("correct3d").argtypes = [ci, rp, ci, ci, f32p]
        g("mask3d").argtypes = [ci, rp, cf, cf, cf, pf, pf, pf]
        g("sheltest").argtypes = [ci, ci, cp, cp, rp, rp, cf, i32p, i32p, f64p, rp]
        g("sheltest").restype = ci
        g("ccp").argtypes = [ci, cf, cp, c64p, cf, ci, ci, cf, cf, pf, pf, pf, i32p, cp, ci, f32p]
        g("ccp").restype = cd
        g("psearch").argtypes = [ci, cf, cp, c64p, cf, ci, ci, pf, pf, pf, pf, pf, f32p, ci, ci,
                                 i32p, cf, cp, f32p, ci, pi_, f32p]
        g("psearch").restype = cd


def _declare_ref(L):
    L.ref_searchang.argtypes = [pf, pf, pf, pi_, pi_, C.c_char_p, pi_, f32p, pi_, i32p, pi_]
    L.ref_ctf.argtypes = [pf] * 8 + [pi_, pi_, pf, pf, f32p]
    L.ref_va04a.argtypes = [f32p, f32p, ci, cf, ci, ci, ci, pf, pi_]
    L.ref_va04a.restype = None
    L.ref_ainterpo3ds.argtypes = [pi_, c64p, pf, pf, pf, f32p]
    L.ref_random.argtypes = [pi_, pf]
    L.ref_maskcos.argtypes = [pi_, f32p, pf, pf, pf, pf, pf, pi_]
    L.ref_pweight.argtypes = [pi_, c64p, f32p, c64p, pi_]
    L.ref_pextract.argtypes = [pi_, c64p, c64p, pf, pf, pf, c64p, f32p]
    L.ref_pinsert.argtypes = [pi_, pf, pf, c64p, c64p, c64p, f32p, pd, pf, pf, pf, pf, pf, pf,
                              pf, pf, pi_, i32p, f32p, C.POINTER(C.c_int64), pf, pf, f32p, pi_]
    L.ref_pinsert_se.argtypes = L.ref_pinsert.argtypes + [c64p, i32p, f32p, f32p, f32p]
    L.ref_getsymmat.argtypes = [C.c_char_p, pi_, f32p, pi_, i32p, pi_]
    L.ref_storeshift.argtypes = [pi_, pd, pd, pi_, c64p, f32p, c64p, f32p, c64p, c64p]
    L.ref_sheltest.argtypes = [pi_, pi_, c64p, c64p, f32p, f32p, pf, i32p, i32p, f64p, f32p,
                               f32p, pi_]


def byref_f(x):
    return C.byref(C.c_float(x))


def byref_i(x):
    return C.byref(C.c_int(x))


IDENTITY = np.eye(3, dtype=np.float32).reshape(-1, order="F").copy()


# ---------------------------------------------------------------- helpers
def searchang(dang, asym="C", nasym=1, mask=(1, 1, 1, 1, 1), rmm=0.1, ri=50.0, lib="oracle"):
    m = np.asarray(mask, dtype=np.int32)
    iq = C.c_int(0)
    if lib == "oracle":
        L = oracle()
        dout = C.c_float(0)
        n = L.or_searchang(rmm, ri, dang, asym.encode(), nasym, m, 1, None, C.byref(iq),
                           C.byref(dout))
        sang = np.zeros(3 * max(n, 1), dtype=np.float32)
        L.or_searchang(rmm, ri, dang, asym.encode(), nasym, m, 1,
                       sang.ctypes.data_as(C.c_void_p), C.byref(iq), C.byref(dout))
        return sang[: 3 * n].reshape(n, 3), iq.value
    L = ref()
    ns = C.c_int(0)
    dummy = np.zeros(3, dtype=np.float32)
    L.ref_searchang(byref_f(rmm), byref_f(ri), byref_f(dang), C.byref(ns), C.byref(iq),
                    asym.encode(), byref_i(nasym), dummy, byref_i(1), m, byref_i(0))
    n = ns.value
    sang = np.zeros(3 * max(n, 1), dtype=np.float32)
    ns = C.c_int(0)
    L.ref_searchang(byref_f(rmm), byref_f(ri), byref_f(dang), C.byref(ns), C.byref(iq),
                    asym.encode(), byref_i(nasym), sang, byref_i(1), m, byref_i(1))
    return sang[: 3 * n].reshape(n, 3), iq.value


def getsymmat(asym, nasym, lib="oracle"):
    symop = np.zeros(9 * 8, dtype=np.float32)
    jsym = np.zeros(8, dtype=np.int32)
    ired = C.c_int(0)
    if lib == "oracle":
        nsym = oracle().or_getsymmat(asym.encode(), nasym, symop, jsym, C.byref(ired))
    else:
        ns = C.c_int(0)
        ref().ref_getsymmat(asym.encode(), byref_i(nasym), symop, C.byref(ns), jsym,
                            C.byref(ired))
        nsym = ns.value
    return nsym, symop, jsym, ired.value
