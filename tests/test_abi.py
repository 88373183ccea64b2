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
