"""Pin the CPU oracle to the reference's own routines, compiled by flang from
the reference sources (oracle/_ref).  These are the FFT-free hot-path
routines; FFT-using ones are pinned in test_oracle_math.py.

FP32 oracle vs reference: integer outputs (grid, symmetry orders, counts)
must be bit-exact; float outputs are compared bit-exactly where the routine
is a straight FP32 restatement and to 1e-6 relative where flang's complex
arithmetic may order operations differently.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

rng = np.random.default_rng(1234)


def half_volume(nsam, seed):
    r = np.random.default_rng(seed)
    v = r.standard_normal((nsam, nsam, nsam)).astype(np.float32)
    f = np.fft.rfftn(v).astype(np.complex64)  # (z, y, x//2+1), x fastest
    return np.ascontiguousarray(f).reshape(-1)


def half_plane(nsam, seed):
    r = np.random.default_rng(seed)
    jc = nsam // 2 + 1
    return (r.standard_normal(nsam * jc) + 1j * r.standard_normal(nsam * jc)).astype(np.complex64)


@pytest.mark.parametrize("asym,nasym", [("C", 1), ("D", 7), ("C", 7), ("T", 0), ("O", 0), ("I", 0)])
@pytest.mark.parametrize("dang", [5.0, 7.5, 15.0])
def test_searchang_bitexact(ref_lib, asym, nasym, dang):
    a, iqa = O.searchang(dang, asym, nasym, lib="oracle")
    b, iqb = O.searchang(dang, asym, nasym, lib="ref")
    assert iqa == iqb
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


def test_searchang_counts(ref_lib):
    # the view counts quoted for the workloads, from the reference itself
    assert len(O.searchang(5.0, "D", 7, lib="ref")[0]) == 130
    assert len(O.searchang(5.0, "C", 1, lib="ref")[0]) == 1626
    assert len(O.searchang(15.0, "C", 1, lib="ref")[0]) == 178


@pytest.mark.parametrize("asym,nasym", [("C", 1), ("C", 5), ("D", 7), ("T", 0), ("O", 0), ("I", 0)])
def test_getsymmat(ref_lib, asym, nasym):
    r = O.getsymmat(asym, nasym, lib="ref")
    o = O.getsymmat(asym, nasym)
    assert r[0] == o[0] and r[3] == o[3]
    assert np.array_equal(r[2][: r[0]], o[2][: o[0]])
    assert np.array_equal(r[1].view(np.int32), o[1].view(np.int32))


def test_random_lcg(ref_lib):
    L = O.oracle()
    for seed in (-1, -37, -59, 12345):
        s1, s2 = C.c_int(seed), C.c_int(seed)
        for _ in range(50):
            a = L.or_random(C.byref(s1))
            out = C.c_float(0)
            ref_lib.ref_random(C.byref(s2), C.byref(out))
            assert s1.value == s2.value
            assert np.float32(a).view(np.int32) == np.float32(out.value).view(np.int32)


@pytest.mark.parametrize("wgh2", [0.07, -0.07, -1.0])
def test_ctf_bitexact(ref_lib, wgh2):
    L = O.oracle()
    cs, wl, wgh1 = 2.7e7, 0.01969, float(np.sqrt(1 - 0.07 ** 2))
    out_o = np.zeros(2, np.float32)
    out_r = np.zeros(2, np.float32)
    for trial in range(200):
        ix, iy = int(rng.integers(-64, 65)), int(rng.integers(-64, 65))
        df1, df2 = float(rng.uniform(1e4, 3e4)), float(rng.uniform(1e4, 3e4))
        ang = float(rng.uniform(-3, 3))
        th = float(rng.uniform(1e-4, 2e-4))
        tx, ty = (0.0, 0.0) if trial % 2 else (float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)))
        L.or32_ctf(cs, wl, wgh1, wgh2, df1, df2, ang, th, ix, iy, tx, ty, out_o)
        args = [O.byref_f(v) for v in (cs, wl, wgh1, wgh2, df1, df2, ang, th)]
        ref_lib.ref_ctf(*args, O.byref_i(ix), O.byref_i(iy), O.byref_f(tx), O.byref_f(ty), out_r)
        assert np.array_equal(out_o.view(np.int32), out_r.view(np.int32)), (ix, iy)


@pytest.mark.parametrize("nsam", [32, 64])
@pytest.mark.parametrize("origin", [0, 1])
def test_maskcos_bitexact(ref_lib, nsam, origin):
    L = O.oracle()
    img = rng.standard_normal(nsam * (nsam + 2)).astype(np.float32)
    a, b = img.copy(), img.copy()
    ri = nsam * 0.35
    rih, rii = ri - 3.0, ri + 3.0
    args = (float(rih * rih), float(rii * rii), float(rih), 6.0, 1.0)
    (L.or32_maskcos2 if origin else L.or32_maskcos)(nsam, a, *args)
    ref_lib.ref_maskcos(O.byref_i(nsam), b, *[O.byref_f(v) for v in args], O.byref_i(origin))
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("nsam", [32, 64])
@pytest.mark.parametrize("ictf", [0, 1])
def test_pweight(ref_lib, nsam, ictf):
    """Includes the aliased Nyquist row and the JC+1 bin quirk."""
    L = O.oracle()
    jc = nsam // 2 + 1
    spec = half_plane(nsam, 7)
    ctff = np.concatenate([half_plane(nsam, 8), half_plane(nsam, 9)]) * np.complex64(0.3)
    ssnr = rng.uniform(0.1, 2.0, jc + 1).astype(np.float32)
    a, b = spec.copy(), spec.copy()
    L.or32_pweight(nsam, a, a, ssnr, ctff, ictf)
    ref_lib.ref_pweight(O.byref_i(nsam), b, ssnr, ctff, O.byref_i(ictf))
    np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-7)
    exact = np.mean(a.view(np.int32) == b.view(np.int32))
    assert exact > 0.99


@pytest.mark.parametrize("nsam", [16, 32])
def test_ainterpo3ds(ref_lib, nsam):
    L = O.oracle()
    vol = half_volume(nsam, 3)
    out_o = np.zeros(2, np.float32)
    out_r = np.zeros(2, np.float32)
    h = nsam / 2
    pts = list(rng.uniform(-h, h, (300, 3)))
    # integer and boundary coordinates: negative integers, x at the JC-1 clamp
    pts += [np.array(p, dtype=np.float64) for p in
            [(-2.0, -3.0, 1.0), (0.0, -h, h - 0.5), (h - 0.25, 2.0, -1.0), (-h + 0.1, 0.0, 0.0),
             (0.0, 0.0, 0.0), (3.0, -4.0, -5.0)]]
    for p in pts:
        x, y, z = (float(np.float32(v)) for v in p)
        L.or32_ainterpo3ds(nsam, vol, x, y, z, out_o)
        ref_lib.ref_ainterpo3ds(O.byref_i(nsam), vol, O.byref_f(x), O.byref_f(y),
                                O.byref_f(z), out_r)
        assert np.array_equal(out_o.view(np.int32), out_r.view(np.int32)), (x, y, z)


@pytest.mark.parametrize("nsam", [16, 32])
def test_pextract(ref_lib, nsam):
    L = O.oracle()
    vol = half_volume(nsam, 5)
    jc = nsam // 2 + 1
    ctff = np.concatenate([half_plane(nsam, 10), half_plane(nsam, 11)])
    for trial in range(5):
        phi, theta, psi = (float(v) for v in rng.uniform(-np.pi, np.pi, 3))
        a = np.zeros(nsam * jc, np.complex64)
        b = np.zeros(nsam * jc, np.complex64)
        L.or32_pextract(nsam, a, vol, phi, theta, psi, ctff, O.IDENTITY)
        ref_lib.ref_pextract(O.byref_i(nsam), b, vol, O.byref_f(phi), O.byref_f(theta),
                             O.byref_f(psi), ctff, O.IDENTITY)
        assert np.array_equal(a.view(np.int32), b.view(np.int32))


def _insert(lib, which, nsam, spec, ctff, par, sym, occ=1.0, rrec=0.5):
    jc = nsam // 2 + 1
    b3 = np.zeros(nsam * nsam * jc, np.complex64)
    t3 = np.zeros(nsam * nsam * jc, np.float32)
    ttd = C.c_double(0.0)
    nk = C.c_int64(0)
    nsym, symop, jsym, _ = sym
    phi, theta, psi, shx, shy = par
    presa, pbc, boff, psize = 0.6, 5.0, 40.0, 1.5
    if which == "oracle":
        lib.or32_pinsert(nsam, rrec, 1.0, spec, ctff, b3, t3, C.byref(ttd), phi, theta, psi,
                         shx, shy, presa, pbc, boff, nsym, jsym, symop, C.byref(nk), psize, occ,
                         O.IDENTITY, 0)
    else:
        f = O.byref_f
        lib.ref_pinsert(O.byref_i(nsam), f(rrec), f(1.0), spec, ctff, b3, t3, C.byref(ttd),
                        f(phi), f(theta), f(psi), f(shx), f(shy), f(presa), f(pbc), f(boff),
                        O.byref_i(nsym), jsym, symop, C.byref(nk), f(psize), f(occ), O.IDENTITY,
                        O.byref_i(0))
    return b3, t3, ttd.value, nk.value


@pytest.mark.parametrize("nsam,asym,nasym", [(16, "C", 1), (32, "C", 1), (32, "D", 7), (32, "O", 0)])
def test_pinsert(ref_lib, nsam, asym, nasym):
    """Voxel set and count bit-exact (NKSUM), values to FP32 rounding."""
    L = O.oracle()
    sym = O.getsymmat(asym, nasym)
    spec = half_plane(nsam, 21)
    ctff = np.concatenate([half_plane(nsam, 22), half_plane(nsam, 23)]) * np.complex64(0.5)
    for trial in range(3):
        par = tuple(float(v) for v in rng.uniform(-np.pi, np.pi, 3)) + \
            tuple(float(v) for v in rng.uniform(-0.3, 0.3, 2))
        bo, to, tto, nko = _insert(L, "oracle", nsam, spec, ctff, par, sym)
        br, tr, ttr, nkr = _insert(ref_lib, "ref", nsam, spec, ctff, par, sym)
        assert nko == nkr
        # the FP32 restatement reproduces the reference bit for bit
        assert np.array_equal(bo.view(np.int32), br.view(np.int32))
        assert np.array_equal(to.view(np.int32), tr.view(np.int32))
        assert tto == ttr


def test_pinsert_integer_angles_exact(ref_lib):
    """Integer-aligned orientations land on grid points: hits the INT()/negative rule."""
    L = O.oracle()
    nsam = 16
    sym = O.getsymmat("C", 1)
    spec = half_plane(nsam, 31)
    ctff = np.concatenate([half_plane(nsam, 32), half_plane(nsam, 33)])
    for par in [(0.0, 0.0, 0.0, 0.0, 0.0), (np.pi / 2, np.pi / 2, 0.0, 0.1, -0.2),
                (np.pi, 0.0, np.pi / 2, 0.0, 0.0), (0.0, np.pi, 0.0, 0.0, 0.0)]:
        par = tuple(float(np.float32(v)) for v in par)
        bo, to, tto, nko = _insert(L, "oracle", nsam, spec, ctff, par, sym)
        br, tr, ttr, nkr = _insert(ref_lib, "ref", nsam, spec, ctff, par, sym)
        assert nko == nkr
        assert np.array_equal(to.view(np.int32), tr.view(np.int32))
        assert np.array_equal(bo.view(np.int32), br.view(np.int32))


def ref_pinsert_trilin(ref_lib, nsam, irrec2, spec, ctff, occ, b3, t3, ttd, shx, shy, nk, dm,
                       tmp, dx, dy, st, ip):
    """The reference's own PINSERT_TRILIN (pinsert_trilin.f:8-244), every
    argument by reference; ASUM..KSUM, SINCLUT, THET and NS1 are unused with
    IEWALD = 0, NNSTAT = 0."""
    jc = nsam // 2 + 1
    f, i = O.byref_f, O.byref_i
    dummy = np.zeros(4, np.float32)
    ns1 = np.zeros(4 * nsam, np.int32)
    ref_lib.pinsert_trilin_(i(nsam), f(1.0), i(irrec2), spec.ctypes, ctff.ctypes, f(occ),
                            b3.ctypes, t3.ctypes, C.byref(ttd), f(shx), f(shy), dummy.ctypes,
                            dummy.ctypes, dummy.ctypes, dummy.ctypes, ns1.ctypes, dummy.ctypes,
                            i(1), i(1), f(0.0), i(0), i(0), C.byref(nk), i(jc), dm.ctypes, f(tmp),
                            i(nsam // 2), i(nsam * jc), f(dx), f(dy), f(st), i(ip), ns1.ctypes)


def _stats(nv):
    return (np.zeros(nv, np.int32), np.zeros(nv, np.float32), np.zeros(nv, np.float32),
            np.zeros(nv, np.float32))


def _reference_volume(nsam, seed):
    """A random half-volume transform, every fifth coefficient zero (PDIFF's 0 case)."""
    r = np.random.default_rng(seed)
    nv = nsam * nsam * (nsam // 2 + 1)
    c = (r.normal(size=nv) + 1j * r.normal(size=nv)).astype(np.complex64)
    c[::5] = 0
    return c


@pytest.mark.parametrize("nsam,imp", [(16, 1), (32, 3)])
def test_pinsert_trilin_se_stats(ref_lib, nsam, imp):
    """PINSERT_TRILIN_SE with NNSTAT = 1 (pinsert_trilin_se.f:186-194): per
    stripe, the oracle is bit exact to the reference routine in B3DF, T3DF,
    TTD, NKSUM and the per-voxel KSUM, VSUM, PSUM, ASUM; KSUM sums to NKSUM,
    and B3DF/T3DF equal PINSERT_TRILIN's."""
    L = O.oracle()
    jc = nsam // 2 + 1
    nv = nsam * nsam * jc
    spec = half_plane(nsam, 51)
    ctff = np.concatenate([half_plane(nsam, 52), half_plane(nsam, 53)]) * np.complex64(0.5)
    c3 = _reference_volume(nsam, 54)
    dm = np.zeros(9, np.float32)
    L.or_dm(-1.1, 0.8, 2.6, 1.0, None, dm)
    irrec2 = (nsam * 2 // 5) ** 2
    tmp, occ, shx, shy, dx, dy = -0.003, 0.9, -0.05, 0.13, 0.01, -0.02
    st = np.float32(2 * jc - 1) / np.float32(imp)
    f, i = O.byref_f, O.byref_i
    for ip in range(1, imp + 1):
        bo, br, bp = (np.zeros(nv, np.complex64) for _ in range(3))
        to, tr, tp = (np.zeros(nv, np.float32) for _ in range(3))
        tdo, tdr, tdp = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
        nko, nkr, nkp = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        so, sr = _stats(nv), _stats(nv)
        L.or32_pinsert_trilin_se(nsam, irrec2, spec, ctff, occ, bo, to, C.byref(tdo), shx, shy,
                                 C.byref(nko), dm, tmp, dx, dy, float(st), ip, c3, *so)
        ns1 = np.zeros(4 * nsam, np.int32)
        ref_lib.pinsert_trilin_se_(i(nsam), f(1.0), i(irrec2), spec.ctypes, ctff.ctypes, f(occ),
                                   br.ctypes, tr.ctypes, C.byref(tdr), f(shx), f(shy),
                                   sr[3].ctypes, sr[1].ctypes, sr[2].ctypes, c3.ctypes,
                                   sr[0].ctypes, np.zeros(4, np.float32).ctypes, i(1), i(1),
                                   f(0.0), i(0), i(1), C.byref(nkr), i(jc), dm.ctypes, f(tmp),
                                   i(nsam // 2), i(nsam * jc), f(dx), f(dy), f(float(st)), i(ip),
                                   ns1.ctypes)
        L.or32_pinsert_trilin(nsam, irrec2, spec, ctff, occ, bp, tp, C.byref(tdp), shx, shy,
                              C.byref(nkp), dm, tmp, dx, dy, float(st), ip)
        assert nko.value == nkr.value == nkp.value > 0
        assert int(so[0].sum()) == nko.value
        assert np.array_equal(bo.view(np.int32), br.view(np.int32))
        assert np.array_equal(to.view(np.int32), tr.view(np.int32))
        assert np.array_equal(bo.view(np.int32), bp.view(np.int32)) and tdo.value == tdp.value
        assert tdo.value == tdr.value
        assert np.array_equal(so[0], sr[0])
        for a, b in zip(so[1:], sr[1:]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
        hit = so[0] > 0
        assert np.all(so[3][hit] > 0) and np.all(so[2] <= np.float32(np.pi) * so[0] + 1e-5)
        assert np.any(so[2][hit] > 0) and np.any(hit[::5])
        assert np.all(so[2][::5] == 0)


@pytest.mark.parametrize("nsam,asym,nasym", [(16, "C", 1), (32, "D", 7)])
def test_pinsert_se(ref_lib, nsam, asym, nasym):
    """PINSERT with NNSTAT = 1 over every symmetry copy (pinsert.f:137-176 ->
    PINSERT_TRILIN_SE): the oracle's volumes, counters and statistics are bit
    exact to the reference's."""
    L = O.oracle()
    nsym, symop, jsym, _ = O.getsymmat(asym, nasym)
    jc = nsam // 2 + 1
    nv = nsam * nsam * jc
    spec = half_plane(nsam, 61)
    ctff = np.concatenate([half_plane(nsam, 62), half_plane(nsam, 63)]) * np.complex64(0.5)
    c3 = _reference_volume(nsam, 64)
    presa, pbc, boff, psize, occ, rrec = 0.6, 5.0, 40.0, 1.5, 0.8, 0.45
    f = O.byref_f
    for trial in range(2):
        phi, theta, psi = (float(v) for v in rng.uniform(-np.pi, np.pi, 3))
        shx, shy = (float(v) for v in rng.uniform(-0.3, 0.3, 2))
        out = []
        for which in ("oracle", "ref"):
            b3 = np.zeros(nv, np.complex64)
            t3 = np.zeros(nv, np.float32)
            ttd, nk = C.c_double(0.0), C.c_int64(0)
            s = _stats(nv)
            if which == "oracle":
                L.or32_pinsert_se(nsam, rrec, 1.0, spec, ctff, b3, t3, C.byref(ttd), phi, theta,
                                  psi, shx, shy, presa, pbc, boff, nsym, jsym, symop,
                                  C.byref(nk), psize, occ, O.IDENTITY, 0, c3, *s)
            else:
                ref_lib.ref_pinsert_se(O.byref_i(nsam), f(rrec), f(1.0), spec, ctff, b3, t3,
                                       C.byref(ttd), f(phi), f(theta), f(psi), f(shx), f(shy),
                                       f(presa), f(pbc), f(boff), O.byref_i(nsym), jsym, symop,
                                       C.byref(nk), f(psize), f(occ), O.IDENTITY, O.byref_i(0),
                                       c3, *s)
            out.append((b3, t3, ttd.value, nk.value, s))
        (bo, to, tdo, nko, so), (br, tr, tdr, nkr, sr) = out
        assert nko == nkr == int(so[0].sum())
        assert np.array_equal(bo.view(np.int32), br.view(np.int32))
        assert np.array_equal(to.view(np.int32), tr.view(np.int32)) and tdo == tdr
        assert np.array_equal(so[0], sr[0])
        for a, b in zip(so[1:], sr[1:]):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("nsam,imp", [(16, 1), (32, 3), (64, 4)])
def test_pinsert_trilin_stripes(ref_lib, nsam, imp):
    """One PINSERT_TRILIN call per stripe IP = 1..IMP (ST = (2JC-1)/IMP, as
    pinsert.f:95 sets it) with a DX, DY phase offset: the stripe oracle is bit
    exact to the reference routine, the stripes partition the plane (their sum
    equals the whole-plane call), and the last stripe's I range is checked
    against the reference formula by an empty-stripe case."""
    L = O.oracle()
    jc = nsam // 2 + 1
    nv = nsam * nsam * jc
    spec = half_plane(nsam, 41)
    ctff = np.concatenate([half_plane(nsam, 42), half_plane(nsam, 43)]) * np.complex64(0.5)
    dm = np.zeros(9, np.float32)
    L.or_dm(0.7, 1.9, -2.3, 1.0, None, dm)
    irrec2 = (nsam * 2 // 5) ** 2
    tmp, occ, shx, shy, dx, dy = -0.004, 0.8, 0.11, -0.07, 0.02, 0.05
    st = np.float32(2 * jc - 1) / np.float32(imp)
    tot_b = np.zeros(nv, np.complex64)
    tot_k = 0
    for ip in range(1, imp + 1):
        bo, br = np.zeros(nv, np.complex64), np.zeros(nv, np.complex64)
        to, tr = np.zeros(nv, np.float32), np.zeros(nv, np.float32)
        tdo, tdr = C.c_double(0.0), C.c_double(0.0)
        nko, nkr = C.c_int64(0), C.c_int64(0)
        L.or32_pinsert_trilin(nsam, irrec2, spec, ctff, occ, bo, to, C.byref(tdo), shx, shy,
                              C.byref(nko), dm, tmp, dx, dy, float(st), ip)
        ref_pinsert_trilin(ref_lib, nsam, irrec2, spec, ctff, occ, br, tr, tdr, shx, shy, nkr,
                           dm, tmp, dx, dy, float(st), ip)
        assert nko.value == nkr.value
        assert np.array_equal(bo.view(np.int32), br.view(np.int32))
        assert np.array_equal(to.view(np.int32), tr.view(np.int32))
        assert tdo.value == tdr.value
        tot_b += bo
        tot_k += nko.value
    whole = np.zeros(nv, np.complex64)
    tw = np.zeros(nv, np.float32)
    td, nk = C.c_double(0.0), C.c_int64(0)
    L.or32_pinsert_trilin(nsam, irrec2, spec, ctff, occ, whole, tw, C.byref(td), shx, shy,
                          C.byref(nk), dm, tmp, dx, dy, float(2 * jc - 1), 1)
    assert nk.value == tot_k > 0
    assert np.allclose(tot_b, whole, rtol=1e-5, atol=1e-5)
    # a stripe past the plane (IP = IMP + 1) inserts nothing in both
    bo, br = np.zeros(nv, np.complex64), np.zeros(nv, np.complex64)
    to, tr = np.zeros(nv, np.float32), np.zeros(nv, np.float32)
    nko, nkr = C.c_int64(0), C.c_int64(0)
    L.or32_pinsert_trilin(nsam, irrec2, spec, ctff, occ, bo, to, C.byref(C.c_double()), shx, shy,
                          C.byref(nko), dm, tmp, dx, dy, float(st), imp + 1)
    ref_pinsert_trilin(ref_lib, nsam, irrec2, spec, ctff, occ, br, tr, C.c_double(), shx, shy,
                       nkr, dm, tmp, dx, dy, float(st), imp + 1)
    assert nko.value == nkr.value == 0


@pytest.mark.parametrize("nsam", [16, 32])
def test_storeshift(ref_lib, nsam):
    L = O.oracle()
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    bo, co = np.zeros(n, np.complex64), np.zeros(n, np.complex64)
    br, cr = np.zeros(n, np.complex64), np.zeros(n, np.complex64)
    L.or32_storeshift(nsam, 0.37, 0.41, 14, a, s, d, v, bo, co)
    ref_lib.ref_storeshift(O.byref_i(nsam), C.byref(C.c_double(0.37)), C.byref(C.c_double(0.41)),
                           O.byref_i(14), a, s, d, v, br, cr)
    np.testing.assert_allclose(bo, br, rtol=1e-6)
    np.testing.assert_allclose(co, cr, rtol=1e-6)


@pytest.mark.parametrize("nsam,istp", [(16, 1), (32, 1), (32, 3)])
def test_sheltest_fsc(ref_lib, nsam, istp):
    L = O.oracle()
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (a + 0.7 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    npo, npto = np.zeros(jc + 2, np.int32), np.zeros(jc + 2, np.int32)
    dfo, pso = np.zeros(jc + 2), np.zeros(jc + 2, np.float32)
    nst = L.or32_sheltest(nsam, istp, a, d, s, v, 1.0, npo, npto, dfo, pso)
    npr, nptr = np.zeros(jc + 2, np.int32), np.zeros(jc + 2, np.int32)
    dfr, fsc, psr = np.zeros(jc + 2), np.zeros(jc + 2, np.float32), np.zeros(jc + 2, np.float32)
    nstr = C.c_int(0)
    ref_lib.ref_sheltest(O.byref_i(nsam), O.byref_i(istp), a, d, s, v, O.byref_f(1.0), npr, nptr,
                         dfr, fsc, psr, C.byref(nstr))
    assert nst == nstr.value
    assert np.array_equal(npo[:nst], npr[:nst]) and np.array_equal(npto[:nst], nptr[:nst])
    np.testing.assert_allclose(dfo[:nst], dfr[:nst], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(pso[:nst], psr[:nst], rtol=1e-5)


# ----------------------------------------------------- VA04A (va04.f:8-369)
def _or_va04a(L, x, e, fn, escale, icon, maxit):
    x = np.array(x, np.float32)
    which = C.c_int(fn)
    nev = C.c_int(0)
    f = L.or_va04a(x, np.array(e, np.float32), len(x), escale, 0, icon, maxit,
                   C.cast(L.or_va04_testfn, C.c_void_p), C.cast(C.pointer(which), C.c_void_p),
                   C.byref(nev))
    return x, np.float32(f), nev.value


@pytest.mark.parametrize("fn", [1, 2, 3])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 6])
@pytest.mark.parametrize("icon,maxit,escale", [(1, 50, 100.0), (2, 50, 100.0), (1, 3, 10.0)])
def test_va04a_bit_exact(ref_lib, fn, n, icon, maxit, escale):
    """The oracle's Powell minimiser reproduces the reference's iterates bit for
    bit: final X, F and the number of objective evaluations."""
    L = O.oracle()
    r = np.random.default_rng(fn * 100 + n)
    x0 = r.uniform(-0.6, 0.6, n).astype(np.float32)
    e = (np.array([0.1, 0.1, 0.1, 0.01, 0.01, 0.001], np.float32)[:n] * r.uniform(0.5, 2.0, n)
         ).astype(np.float32)
    xo, fo, no = _or_va04a(L, x0, e, fn, escale, icon, maxit)
    xr = x0.copy()
    fr = C.c_float(0)
    nr = C.c_int(0)
    ref_lib.ref_va04a(xr, e.copy(), n, escale, icon, maxit, fn, C.byref(fr), C.byref(nr))
    print(f"fn {fn} n {n}: F {fo:.6g} after {no} evaluations (reference {nr.value})")
    assert no == nr.value
    assert np.array_equal(xo.view(np.uint32), xr.view(np.uint32))
    assert np.float32(fr.value).view(np.uint32) == fo.view(np.uint32)


def test_va04a_minimises():
    """Closed form: the coupled quadratic's minimiser solves a 2x2 linear system."""
    L = O.oracle()
    x, f, _ = _or_va04a(L, [0.5, -0.5], [0.1, 0.1], 1, 100.0, 2, 50)
    # f = 0.7 a^2 + 1.4 b^2 + 0.4 a b with a = x1-0.1, b = x2-0.2: minimum at a = b = 0
    assert abs(x[0] - 0.1) < 1e-3 and abs(x[1] - 0.2) < 1e-3 and f < 1e-6


# ------------------------------------------------- the volume finalisation
def test_calcsinc_bitexact(ref_lib):
    L = O.oracle()
    a = np.zeros(180, np.float32)
    b = np.zeros(180, np.float32)
    L.or_calcsinc(a, 180)
    ref_lib.calcsinc_(b.ctypes, O.byref_i(180))
    assert np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("nsam,interp,ipad", [(16, 1, 1), (32, 1, 1), (32, 0, 1), (32, 1, 2)])
def test_correct3d_bitexact(ref_lib, nsam, interp, ipad):
    L = O.oracle()
    lut = np.zeros(180, np.float32)
    L.or_calcsinc(lut, 180)
    v = rng.standard_normal(nsam * nsam * (nsam + 2)).astype(np.float32)
    a, b = v.copy(), v.copy()
    L.or32_correct3d(nsam, a, interp, ipad, lut)
    ref_lib.correct3d_(O.byref_i(nsam), lut.ctypes, b.ctypes, O.byref_i(interp), O.byref_i(ipad))
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    # the padding columns are untouched, the centre voxel (all |a| < 0.02) too
    assert np.array_equal(a.reshape(nsam, nsam, nsam + 2)[:, :, nsam:],
                          v.reshape(nsam, nsam, nsam + 2)[:, :, nsam:])
    c = nsam // 2
    assert a.reshape(nsam, nsam, nsam + 2)[c, c, c] == v.reshape(nsam, nsam, nsam + 2)[c, c, c]


@pytest.mark.parametrize("nsam,fpart,ifsc", [(16, 0.0, 0), (32, 0.0, 0), (32, 0.3, 0), (32, 0.3, 1)])
def test_shiftvol_bitexact(ref_lib, nsam, fpart, ifsc):
    """Both branches of the final-map filter (plain and Wiener), with the
    half-map sum (IFSC = 0) or without; FSC entries at +-1 take A1 = 1000 and
    zero PSSNR entries zero their shells."""
    L = O.oracle()
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    s[::97] = 0.0
    v[::97] = 0.0
    nsh = 2 * nsam
    pssnr = rng.uniform(0.1, 2.0, nsh).astype(np.float32)
    pssnr[5] = 0.0
    fsc = rng.uniform(-0.2, 0.99, nsh).astype(np.float32)
    fsc[3], fsc[4] = 1.0, -1.0
    rlim, iredun, fmask = 0.4, 14, 0.3
    ao, do, so = a.copy(), d.copy(), s.copy()
    ar, dr, sr = a.copy(), d.copy(), s.copy()
    std_o, std_r = C.c_double(0.37), C.c_double(0.37)
    L.or32_shiftvol(nsam, C.byref(std_o), 0.41, rlim, iredun, ao, do, so, v, ifsc, pssnr, fpart,
                    fsc, fmask)
    asum = np.zeros(4, np.float32)
    ref_lib.shiftvol_(O.byref_i(nsam), C.byref(std_r), C.byref(C.c_double(0.41)), O.byref_i(jc),
                      O.byref_f(rlim), O.byref_i(nsam // 2), O.byref_i(iredun), ar.ctypes,
                      dr.ctypes, sr.ctypes, v.ctypes, asum.ctypes, asum.ctypes, O.byref_i(0),
                      O.byref_i(ifsc), pssnr.ctypes, O.byref_f(fpart), fsc.ctypes,
                      O.byref_f(fmask))
    assert std_o.value == std_r.value
    assert np.array_equal(ao.view(np.int32), ar.view(np.int32))
    assert np.array_equal(so.view(np.int32), sr.view(np.int32))
    assert np.array_equal(do.view(np.int32), dr.view(np.int32))
    assert np.count_nonzero(ao) > n // 10


@pytest.mark.parametrize("ifsc,imp,nvol1", [(0, 1, 1), (0, 4, 2), (1, 3, 3)])
def test_combine_arrays_bitexact(ref_lib, ifsc, imp, nvol1):
    L = O.oracle()
    nsam, nn1 = 16, 20
    jc = nsam // 2 + 1
    nn3 = nn1 * nn1 * (nn1 // 2 + 1)
    nvol2 = nvol1
    a = (rng.standard_normal(nn3 * nvol1) + 1j * rng.standard_normal(nn3 * nvol1)).astype(np.complex64)
    d = (rng.standard_normal(nn3 * nvol2) + 1j * rng.standard_normal(nn3 * nvol2)).astype(np.complex64)
    s = rng.uniform(0, 3, nn3 * nvol1).astype(np.float32)
    v = rng.uniform(0, 3, nn3 * nvol2).astype(np.float32)
    nksum1 = rng.integers(0, 10 ** 9, max(imp, nvol1)).astype(np.int64)
    ttd1 = rng.uniform(0, 10, max(imp, nvol1))
    ttd2 = rng.uniform(0, 10, max(imp, nvol1))
    out = []
    for which in ("oracle", "ref"):
        aa, dd, ss, vv = a.copy(), d.copy(), s.copy(), v.copy()
        nk, sd, vd = C.c_int64(7), C.c_double(0.5), C.c_double(0.25)
        if which == "oracle":
            L.or32_combine_arrays(nsam, nn3, nvol1, nvol2, ifsc, aa, ss, dd, vv)
            L.or_combine_counters(ifsc, imp, nvol1, nksum1.ctypes.data_as(C.POINTER(C.c_int64)),
                                  ttd1, ttd2, C.byref(nk), C.byref(sd), C.byref(vd))
        else:
            z = np.zeros(4, np.float32)
            iz = np.zeros(4, np.int32)
            i = O.byref_i
            ref_lib.combine_arrays_(i(nsam), z.ctypes, z.ctypes, z.ctypes, iz.ctypes, aa.ctypes,
                                    ss.ctypes, C.byref(sd), dd.ctypes, vv.ctypes, C.byref(vd),
                                    i(nn1), i(0), C.byref(nk), i(ifsc), i(imp), i(max(imp, nvol1)),
                                    i(nvol1), i(nvol2), nksum1.ctypes, ttd1.ctypes, ttd2.ctypes,
                                    i(nn1), i(nn3))
        out.append((aa, ss, dd, vv, nk.value, sd.value, vd.value))
    o, r = out
    for x, y in zip(o[:4], r[:4]):
        assert np.array_equal(x.view(np.int32), y.view(np.int32))
    assert o[4:] == r[4:]
    del jc


@pytest.mark.parametrize("nsam,rio,ric,halfw", [(16, 6.0, 0.0, 3.0), (32, 12.0, 0.0, 6.0),
                                                (32, 20.0, 4.0, 6.0), (32, 9.0, 2.0, 1.0)])
def test_mask3d_bitexact(ref_lib, nsam, rio, ric, halfw):
    """MASK3D: outer cosine edge (and the inner mask for RIC > 0), the edge
    shell's mean outside, FMASK/AVE/STD; RIO beyond NSAM/2 is clamped."""
    L = O.oracle()
    v = (rng.standard_normal(nsam * nsam * (nsam + 2)) + 0.3).astype(np.float32)
    a, b = v.copy(), v.copy()
    fo, ao, so = C.c_float(), C.c_float(), C.c_float()
    fr, ar, sr = C.c_float(), C.c_float(), C.c_float()
    L.or32_mask3d(nsam, a, rio, ric, halfw, C.byref(fo), C.byref(ao), C.byref(so))
    ref_lib.mask3d_(O.byref_i(nsam), b.ctypes, O.byref_f(rio), O.byref_f(ric), O.byref_f(halfw),
                    C.byref(fr), C.byref(ar), C.byref(sr))
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert (fo.value, ao.value, so.value) == (fr.value, ar.value, sr.value)
    assert 0.0 < fo.value < 1.0 and not np.array_equal(a, v)


@pytest.mark.parametrize("nsam", [16, 32])
def test_volmeasure_vs_reference(ref_lib, nsam):
    """VOLMEASURE (volmeasure.f:8-69) against the restatement in
    oracle/fpart_oracle.py on a smooth two-blob map with noise: the same
    fraction (the restatement sums the classes in FP64, the reference in FP32
    voxel order; on these maps no voxel sits within that rounding of the
    threshold)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import fpart_oracle as FO
    r = np.random.default_rng(40 + nsam)
    z, y, x = np.meshgrid(*(np.arange(nsam) - nsam / 2,) * 3, indexing="ij")
    m = (np.exp(-((x - 2) ** 2 + y ** 2 + z ** 2) / (0.05 * nsam * nsam))
         + 0.6 * np.exp(-((x + 3) ** 2 + (y - 2) ** 2 + z ** 2) / (0.02 * nsam * nsam))
         + 0.05 * r.standard_normal(x.shape))
    v = np.zeros((nsam, nsam, nsam + 2), np.float32)
    v[:, :, :nsam] = m
    fr = C.c_float()
    ref_lib.volmeasure_(O.byref_i(nsam), v.ctypes, C.byref(fr))
    got = FO.volmeasure(v, nsam)
    assert got == fr.value
    assert 0.02 < got < 0.9
