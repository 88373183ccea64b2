"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host-side geometry is bit-identical to the oracle
(and therefore to the reference).  No device compute is called."""
import numpy as np
import pytest

import oracle_lib as O
from frealign_amd import _native as N
from frealign_amd import api


def test_library_exports_every_declared_symbol():
    L = N.lib()
    declared = N.declared_symbols()
    assert len(declared) >= 18
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert b"gfx950" in L.fr_version()


def test_errors_are_reported_not_swallowed():
    with pytest.raises(N.NativeError):
        api.point_group("X", 3)


@pytest.mark.parametrize("asym,nasym,dang", [("C", 1, 5.0), ("D", 7, 5.0), ("C", 1, 15.0),
                                             ("O", 0, 7.5), ("I", 0, 5.0), ("T", 0, 15.0)])
def test_search_grid_matches_oracle(asym, nasym, dang):
    a, iq = api.search_grid(dang, asym, nasym)
    b, iq2 = O.searchang(dang, asym, nasym, lib="oracle")
    assert iq == iq2
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("asym,nasym", [("C", 1), ("D", 7), ("T", 0), ("O", 0), ("I", 0)])
def test_point_group_matches_oracle(asym, nasym):
    x = api.point_group(asym, nasym)
    y = O.getsymmat(asym, nasym)
    assert x[0] == y[0] and x[3] == y[3]
    assert np.array_equal(x[2][: x[0]], y[2][: y[0]])
    assert np.array_equal(x[1].view(np.int32), y[1].view(np.int32))


def test_rotation_matches_oracle():
    rng = np.random.default_rng(3)
    eul = rng.uniform(-np.pi, np.pi, (200, 3)).astype(np.float32)
    for amag in (0.0, 1.0, 1.03):
        dm = api.rotation(eul, amag=amag)
        for k in range(len(eul)):
            want = np.zeros(9, np.float32)
            O.oracle().or_dm(*[float(v) for v in eul[k]], amag, O.IDENTITY.ctypes.data, want)
            assert np.array_equal(dm[k].view(np.int32), want.view(np.int32))


def test_lcg_matches_oracle():
    import ctypes as C
    s = C.c_int(-17)
    v = -17
    for _ in range(20):
        a = O.oracle().or_random(C.byref(s))
        b, v = api.lcg_random(v)
        assert v == s.value and np.float32(a) == np.float32(b)


def test_fpart_from_mw_closed_form():
    """FIND_FPART's MW != 0 branch (find_fpart.f:141-149), a host function.
    find_fpart.f needs FFTW and is not built here, so the checks are what the
    quantity is: the particle's share of the box volume.  A particle of
    MW kDa at DALT Da/A^3 whose volume equals the box gives 1; FPART is linear
    in MW and inversely proportional to the box volume; SSNR(1..JC) is divided
    by it; MW = 0 (the iterative search) is refused."""
    import ctypes as C

    L = N.lib()

    def run(nsam, mw, dalt, psize):
        ssnr = np.arange(1, nsam // 2 + 3, dtype=np.float32)
        f = C.c_float(0.0)
        N.check(L.fr_fpart_from_mw(nsam, mw, dalt, psize, ssnr.ctypes.data, C.byref(f)), "fpart")
        return f.value, ssnr

    f, ssnr = run(10, 0.81, 0.81, 1.0)  # 810 Da at 0.81 Da/A^3 fill 1000 A^3 = 10^3 A^3
    assert abs(f - 1.0) < 1e-6
    f1, _ = run(64, 50.0, 0.81, 1.5)
    f2, s2 = run(64, 100.0, 0.81, 1.5)
    f3, _ = run(128, 100.0, 0.81, 1.5)
    assert abs(f2 / f1 - 2.0) < 1e-5 and abs(f2 / f3 - 8.0) < 1e-5
    assert abs(f2 - 100e3 / 0.81 / 96.0 ** 3) < 1e-6 * f2
    jc = 64 // 2 + 1
    assert np.allclose(s2[:jc], np.arange(1, jc + 1) / np.float32(f2), rtol=1e-6)
    assert np.array_equal(s2[jc:], np.arange(jc + 1, 64 // 2 + 3, dtype=np.float32))
    fneg, _ = run(64, -50.0, 0.81, 1.5)
    assert fneg == f1  # FPART = ABS(FPART)
    with pytest.raises(N.NativeError):
        run(64, 0.0, 0.81, 1.5)


@pytest.mark.parametrize("call,message", [
    ("a3d3", "A3D3 (frealign_hip): the NONSYM insertion is not supported"),
    ("ctfapply", "CTFAPPLY (frealign_hip): the MASKENV envelope (XSTD < 0) is not supported"),
    ("pinsert_trilin", "PINSERT_TRILIN (frealign_hip): Ewald-sphere insertion (IEWALD != 0) is not supported"),
    ("prefine", "PREFINE (frealign_hip): only IVFLAG = 0 is supported"),
])
def test_percall_unsupported_cases_stop(tmp_path, call, message):
    """The per-call symbols stop an unsupported case as the reference's STOPs
    do: the message on stdout and exit status 1, before any device work (so
    this runs without a GPU)."""
    import os
    import subprocess
    import sys

    script = f"""
import ctypes as C
import numpy as np
from frealign_amd import _native as N
lib = N.lib()
i = lambda v: C.byref(C.c_int(v))
f = lambda v: C.byref(C.c_float(v))
buf = np.zeros(1 << 16, np.float32)
d = np.zeros(64, np.float64)
k = np.zeros(8, np.int64)
which = {call!r}
if which == "a3d3":
    args = [i(1), i(0), i(32), i(1), f(12.0), buf.ctypes, i(1), i(1), f(1.0), buf.ctypes,
            buf.ctypes] + [buf.ctypes] * 5 + [f(0.0)] * 6 + [f(5.0), f(40.0), buf.ctypes,
            buf.ctypes, d.ctypes, buf.ctypes, buf.ctypes, d.ctypes, i(40), i(1), buf.ctypes, i(1),
            i(1), f(0.0), i(0), i(1), i(8), buf.ctypes, buf.ctypes, i(0), k.ctypes, i(0), i(2),
            i(2), i(1), i(1), f(1.0), f(1.0), f(1.5), C.c_char_p(b"C1 "), f(0.0), f(0.0), i(0),
            i(0), k.ctypes, d.ctypes, d.ctypes, buf.ctypes, buf.ctypes, i(1), buf.ctypes,
            buf.ctypes, buf.ctypes, C.byref(C.c_int(0)), buf.ctypes, i(40), i(40), buf.ctypes,
            i(1), i(1), C.c_size_t(3)]
    assert len(args) == 73
    lib.a3d3_(*args)
elif which == "ctfapply":
    args = [i(32), buf.ctypes] + [f(0.0)] * 10 + [buf.ctypes, buf.ctypes, buf.ctypes, f(1.0),
            f(10.0), f(6.0), f(49.0), f(169.0), f(0.0), buf.ctypes, buf.ctypes, buf.ctypes,
            f(0.0), f(0.0), f(0.0), buf.ctypes, f(-1.0), i(0), f(0.0), f(0.0),
            C.c_char_p(b"C1 "), buf.ctypes, buf.ctypes, buf.ctypes, i(1), buf.ctypes,
            C.c_size_t(3)]
    assert len(args) == 39
    lib.ctfapply_(*args)
elif which == "pinsert_trilin":
    args = [i(32), f(1.0), i(100), buf.ctypes, buf.ctypes, f(1.0), buf.ctypes, buf.ctypes,
            d.ctypes, f(0.0), f(0.0), buf.ctypes, buf.ctypes, buf.ctypes, buf.ctypes, buf.ctypes,
            buf.ctypes, i(1), i(1), f(0.0), i(1), i(0), k.ctypes, i(17), buf.ctypes, f(0.0),
            i(16), i(32 * 17), f(0.0), f(0.0), f(33.0), i(1), buf.ctypes]
    assert len(args) == 33
    lib.pinsert_trilin_(*args)
else:
    args = [i(32), i(1), f(1.0), i(1), i(8)] + [f(0.0)] * 6 + [buf.ctypes, f(0.0), i(0),
            buf.ctypes, buf.ctypes, f(0.0), f(0.0), i(1)] + [f(0.0)] * 5 + [buf.ctypes, f(1.0)] + \\
           [f(1.0)] * 8 + [buf.ctypes, buf.ctypes, buf.ctypes, i(0), buf.ctypes, buf.ctypes,
            buf.ctypes, i(1), buf.ctypes, buf.ctypes, buf.ctypes, buf.ctypes, buf.ctypes, i(0),
            f(0.0), f(0.0), f(0.3), f(0.3), f(0.01), C.c_char_p(b"C1 ")] + [f(0.0)] * 5 + \\
           [buf.ctypes, buf.ctypes, buf.ctypes, C.c_size_t(3)]
    assert len(args) == 63
    lib.prefine_(*args)
print("returned")
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(N.__file__)))
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path), env={**os.environ, "PYTHONPATH": root})
    assert r.returncode == 1, r.stdout + r.stderr
    assert message in r.stdout
    assert "returned" not in r.stdout
