"""The library's VA04A (fr_powell_minimize, the minimiser inside
fr_refine_batch) against the oracle's restatement, which is itself pinned bit
for bit to the flang-compiled va04.f (tests/test_oracle_ref.py): same final
parameters, objective value and number of evaluations, bit for bit, on the
closed-form test objectives.  Runs on the CPU (no device needed)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from frealign_amd import api


def _oracle(L, x, e, fn, escale, icon, maxit):
    x = np.array(x, np.float32)
    which = C.c_int(fn)
    nev = C.c_int(0)
    f = L.or_va04a(x, np.array(e, np.float32), len(x), escale, 0, icon, maxit,
                   C.cast(L.or_va04_testfn, C.c_void_p), C.cast(C.pointer(which), C.c_void_p),
                   C.byref(nev))
    return x, np.float32(f), nev.value


@pytest.mark.parametrize("fn", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 4, 5, 6])
@pytest.mark.parametrize("icon,maxit,escale", [(1, 50, 100.0), (2, 50, 100.0), (1, 3, 10.0)])
def test_powell_matches_oracle(fn, n, icon, maxit, escale):
    L = O.oracle()
    r = np.random.default_rng(fn * 31 + n)
    x0 = r.uniform(-0.6, 0.6, n).astype(np.float32)
    e = (np.array([0.1, 0.1, 0.1, 0.01, 0.01, 0.001], np.float32)[:n] * r.uniform(0.5, 2.0, n)
         ).astype(np.float32)
    xo, fo, no = _oracle(L, x0, e, fn, escale, icon, maxit)
    which = C.c_int(fn)
    fnp = api.POWELL_FN(C.cast(L.or_va04_testfn, C.c_void_p).value)
    xg, fg, ng = api.powell_minimize(fnp, x0, e, escale, icon, maxit,
                                     user=C.cast(C.pointer(which), C.c_void_p))
    assert ng == no
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    assert fg.view(np.uint32) == fo.view(np.uint32)


def test_powell_python_objective():
    # a Python objective through the same entry point: the minimum of a shifted bowl
    x, f, n = api.powell_minimize(lambda v: float(np.sum((v - np.array([0.3, -0.2])) ** 2)),
                                  [0.0, 0.0], [0.1, 0.1], icon=2)
    assert np.allclose(x, [0.3, -0.2], atol=1e-3) and f < 1e-6 and n > 3


def test_powell_rejects_bad_sizes():
    with pytest.raises(Exception):
        api.powell_minimize(lambda v: 0.0, np.zeros(7), np.ones(7))
