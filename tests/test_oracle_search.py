"""Pin the oracle's global search -- CCP with PAD and ROTATE2D (ccp.f:8-275,
pad.f:8-68, rotate2d.f:8-77) and PSEARCH (psearch.f:8-121) -- without the
FFT library the reference needs (absent here, so its copies cannot be built).

* CCP: an independent FP64 restatement.  The padded spectrum is numpy's
  rfft2 of the padded image (pinned against numpy elsewhere), the quadrants
  are 90-degree rotations of the rotated spectrum extended by Hermitian
  symmetry (not ccp.f's index-permutation table), and the correlation map is
  a direct DFT evaluated at every shift (not an inverse FFT).  It reproduces
  the score (3x3 box / 9, normalised by sqrt(sum|T|^2) sqrt(sum|X|^2)), the
  argmax order (x outer, first maximum, wrap on the NSAMR grid) and the
  returned in-plane angle and shift.
* PSEARCH: TESTPAR equals independent per-view CCP calls on the B-factor
  weighted particle, jittered as psearch.f:61-64, ranked by strict '>' with
  the CC > 0 admission of psearch.f:42-45,89-90.

Both run on the CPU; the oracle's FP32 geometry (DM, NINT of the rotated
coordinates) is reproduced with the same libm cosf/sinf/expf.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O

L = O.oracle()
_libm = C.CDLL("libm.so.6")
for _n in ("cosf", "sinf", "expf"):
    getattr(_libm, _n).argtypes = [C.c_float]
    getattr(_libm, _n).restype = C.c_float
PI_F = np.float32(3.1415926535897)
f32 = np.float32


def nint(v):
    """Fortran NINT of an FP32 value: round half away from zero."""
    v = float(v)
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def blob_volume(nsam, seed, nblob=12):
    """C3DF of a positive blob phantom (origin-shifted half-volume), complex64."""
    r = np.random.default_rng(seed)
    c = nsam // 2
    g = np.arange(nsam) - c
    vol = np.zeros((nsam,) * 3)
    for _ in range(nblob):
        cen = r.uniform(-nsam * 0.25, nsam * 0.25, 3)
        s = r.uniform(1.5, 3.0)
        a = r.uniform(0.5, 1.0)
        ez, ey, ex = (np.exp(-(g - cen[k]) ** 2 / (2 * s * s)) for k in (2, 1, 0))
        vol += a * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    f = np.fft.rfftn(vol)
    n = np.arange(nsam)
    l = np.arange(nsam // 2 + 1)
    sg = 1 - 2 * ((n[:, None, None] + n[None, :, None] + l[None, None, :]) % 2)
    return np.ascontiguousarray((f * sg).astype(np.complex64)).reshape(-1)


def rotation_dm(phi, theta, psi, amag):
    dm = np.zeros(9, np.float32)
    L.or_dm(phi, theta, psi, amag, None, dm)
    # the oracle's DM against R = Rz(psi) Ry(theta) Rz(phi) (ccp.f:81-102)
    cf, sf, ct, st, cp, sp = (math.cos(phi), math.sin(phi), math.cos(theta), math.sin(theta),
                              math.cos(psi), math.sin(psi))
    r = np.array([[cp, sp, 0], [-sp, cp, 0], [0, 0, 1]]) @ \
        np.array([[ct, 0, -st], [0, 1, 0], [st, 0, ct]]) @ \
        np.array([[cf, sf, 0], [-sf, cf, 0], [0, 0, 1]])
    assert np.allclose(dm.reshape(3, 3, order="F").T, r / abs(amag), atol=1e-6)
    return dm


def ccp_fp64(nsam, amag, spec, vol, dangin, maxr1, maxr2, phi, theta, psi, mask, ctff, iquadmax):
    """CCP restated in FP64 with the FP32 geometry of the reference."""
    jc = nsam // 2 + 1
    nsamh = nsam // 2
    nsamrh = nsamh // 2 if (nsamh // 2) * 2 == nsamh else (nsamh + 1) // 2
    nsamr = 2 * nsamrh
    spec = spec.reshape(nsam, jc)          # [M][L], M wraps J
    cr = ctff[:nsam * jc].reshape(nsam, jc)
    cl = ctff[nsam * jc:].reshape(nsam, jc)
    dm = rotation_dm(phi, theta, psi, amag)
    # the CTF-multiplied section on the annulus (ccp.f:105-131)
    sec = np.zeros((nsam, jc), np.complex128)
    out = np.zeros(2)
    for i in range(0, jc - 1):
        for j in range(-jc + 1, jc):
            it = i * i + j * j
            if maxr1 * maxr1 <= it < maxr2 * maxr2 and abs(j) != jc - 1:
                x3 = f32(dm[0] * f32(i)) + f32(dm[3] * f32(j))
                y3 = f32(dm[1] * f32(i)) + f32(dm[4] * f32(j))
                z3 = f32(dm[2] * f32(i)) + f32(dm[5] * f32(j))
                L.or64_ainterpo3ds(nsam, vol, float(x3), float(y3), float(z3), out)
                c = complex(cr[j % nsam, i]) + complex(cl[j % nsam, i]).conjugate()
                sec[j % nsam, i] = complex(out[0], out[1]) * c
    # PAD (pad.f): the image, padded to 4 NSAM with the mean of its central
    # cross.  A pixel whose source slot lies in a padding column of the
    # 4 NSAM array (the two slots after each row) never received the +CIRC
    # of pad.f:124-128, so it lands with -CIRC.
    n4 = 4 * nsam
    x = np.fft.irfft2(sec, s=(nsam, nsam))            # [y][x], = c2r(sec) / NSAM^2
    on_cross = np.zeros((nsam, nsam), bool)
    for k in (nsam // 2 - 1, nsam // 2):
        on_cross[k, :] = True
        on_cross[:, k] = True
    circ = x[on_cross].sum() / (4 * nsam - 4)
    pad = np.full((n4, n4), circ)
    for yy in range(nsam):
        for xx in range(nsam):
            src = xx + (nsam + 2) * yy
            bug = (src % (n4 + 2)) >= n4
            py = yy if yy < nsamh else yy + n4 - nsam
            px = xx if xx < nsamh else xx + n4 - nsam
            pad[py, px] = x[yy, xx] - (circ if bug else 0.0)
    fpad = np.fft.rfft2(pad)                            # [M][L], L = 0..2 NSAM
    npsi = nint(f32(90.0) / f32(dangin if dangin != 0 else 0.25))
    np_, iq = (npsi, iquadmax) if mask[2] else (1, 1)
    # correlation region on the NSAMR grid (ccp.f:160-174)
    reg = [(i, j) for i in range(0, nsamrh + 1) for j in range(-nsamrh, nsamrh)
           if 1 <= i * i + j * j < maxr2 * maxr2 and abs(j) != jc - 1]
    ri = np.array([p[0] for p in reg])
    rj = np.array([p[1] for p in reg])
    samp = spec[rj % nsam, ri].astype(np.complex128)
    sum2 = float(np.sum(np.abs(samp) ** 2))
    w = np.where((ri == 0) | (2 * ri == nsamr), 1.0, 2.0)
    n = np.arange(nsamr)
    ph1 = np.exp(2j * np.pi * np.outer(ri, n) / nsamr)  # [point][n1]
    ph2 = np.exp(2j * np.pi * np.outer(rj, n) / nsamr)  # [point][n2]
    best = (-1e30, None)
    for ipsi in range(1, np_ + 1):
        psis = f32(f32(f32(f32(ipsi - 1) * PI_F) / f32(2.0)) / f32(npsi))
        c, s = f32(_libm.cosf(psis)), f32(_libm.sinf(psis))
        u = {}
        for i in range(0, nsamh + 1):
            for j in range(-nsamh, nsamh + 1):
                if i * i + j * j < (nsamh - 1) ** 2:
                    xr = f32(c * f32(i)) + f32(f32(-s) * f32(j))
                    yr = f32(s * f32(i)) + f32(c * f32(j))
                    lx, my = nint(f32(4.0) * xr), nint(f32(4.0) * yr)
                    v = fpad[my % n4, lx] if lx >= 0 else np.conj(fpad[(-my) % n4, -lx])
                    u[(i, j % nsam)] = v

        def ufull(a, b):
            if a > 0 or (a == 0 and (0, b % nsam) in u):
                return u.get((a, b % nsam), 0.0)
            return np.conj(u.get((-a, (-b) % nsam), 0.0))
        rots = [lambda i, j: (i, j), lambda i, j: (-j, i), lambda i, j: (-i, -j),
                lambda i, j: (j, -i)]
        for iquad in range(1, iq + 1):
            t = np.array([ufull(*rots[iquad - 1](i, j)) for i, j in reg], np.complex128)
            ssum = float(np.sum(np.abs(t) ** 2))
            cc = samp * np.conj(t) * w
            cmap = np.real(np.einsum("p,pa,pb->ab", cc, ph1, ph2))  # [n1][n2]
            ccm, ixm, iym = -1e30, 0, 0
            for ix in range(-6 * mask[3], 6 * mask[3] + 1):
                for iy in range(-6 * mask[4], 6 * mask[4] + 1):
                    v = cmap[ix % nsamr, iy % nsamr]
                    if v > ccm:
                        ccm, ixm, iym = v, ix, iy
            box = sum(cmap[(ixm + a) % nsamr, (iym + b) % nsamr] for a in (-1, 0, 1)
                      for b in (-1, 0, 1))
            score = box / math.sqrt(ssum) / math.sqrt(sum2) / 9
            if score > best[0]:
                psim = f32(f32(f32(psi) + f32(f32(f32(iquad - 1) * PI_F) / f32(2.0))) + psis)
                sx = f32(f32(f32(ixm) / f32(nsamr)) * PI_F) * f32(2.0)
                sy = f32(f32(f32(iym) / f32(nsamr)) * PI_F) * f32(2.0)
                best = (score, (psim, f32(sx), f32(sy), ipsi, iquad, ixm, iym))
    return best


def _particle(nsam, vol, eul, ctff, ixs, iys, noise, seed):
    jc = nsam // 2 + 1
    sec = np.zeros(nsam * jc, np.complex64)
    L.or32_pextract(nsam, sec, vol, *eul, ctff, O.IDENTITY)
    l = np.tile(np.arange(jc), nsam)
    m = np.repeat(np.arange(nsam), jc)
    mm = np.where(m >= jc, m - nsam, m)
    nsamr = nsam // 2
    spec = sec * np.exp(-1j * 2 * np.pi * (l * ixs + mm * iys) / nsamr)
    r = np.random.default_rng(seed)
    scale = np.abs(spec).mean()
    spec = spec + noise * scale * (r.standard_normal(spec.shape) + 1j * r.standard_normal(spec.shape))
    return spec.astype(np.complex64)


def _ctff(nsam, seed):
    jc = nsam // 2 + 1
    r = np.random.default_rng(seed)
    return (0.6 + 0.3 * r.standard_normal(2 * nsam * jc)
            + 0.1j * r.standard_normal(2 * nsam * jc)).astype(np.complex64)


@pytest.mark.parametrize("nsam,maxr2,dang,eul,shift,noise",
                         [(16, 4, 30.0, (0.4, 0.9, 0.3), (2, -1), 0.3),
                          (32, 8, 30.0, (-1.1, 1.4, 2.0), (1, 3), 0.5),
                          (32, 8, 18.0, (2.6, 0.5, -0.9), (-6, 5), 0.2)])
def test_ccp_matches_fp64_restatement(nsam, maxr2, dang, eul, shift, noise):
    vol = blob_volume(nsam, 3 + nsam)
    ctff = _ctff(nsam, 5)
    spec = _particle(nsam, vol, eul, ctff, *shift, noise, 9)
    mask = np.array([1, 1, 1, 1, 1], np.int32)
    # score the particle against a few views near and away from the truth
    views = [(eul[0], eul[1], 0.0), (eul[0] + 0.5, eul[1] - 0.3, 0.0), (0.2, 2.0, 0.0)]
    for phi, theta, psi in views:
        for pre, ct, tol in (("or64_", np.complex128, 1e-9), ("or32_", np.complex64, 2e-4)):
            p, sx, sy = C.c_float(psi), C.c_float(0.0), C.c_float(0.0)
            got = getattr(L, pre + "ccp")(nsam, 1.0, spec.astype(ct), vol, dang, 1, maxr2, phi,
                                          theta, C.byref(p), C.byref(sx), C.byref(sy), mask,
                                          ctff.astype(ct), 4, O.IDENTITY)
            want, arg = ccp_fp64(nsam, 1.0, spec, vol, dang, 1, maxr2, phi, theta, psi, mask,
                                 ctff, 4)
            assert abs(got - want) <= tol * abs(want), (pre, got, want)
            # same in-plane angle and shift (FP32 bit patterns)
            assert np.float32(p.value) == arg[0] and np.float32(sx.value) == arg[1] \
                and np.float32(sy.value) == arg[2], (pre, (p.value, sx.value, sy.value), arg)


def test_ccp_planted_shift_wraps_first_max():
    """At NSAM = 16 the NSAMR grid is 8 wide, so IX = -6..6 visits each map
    column twice; a planted shift of +2 is reported as -6 (visited first)."""
    nsam, maxr2 = 16, 4
    vol = blob_volume(nsam, 4)
    ctff = np.full(2 * nsam * (nsam // 2 + 1), 0.5, np.complex64)
    eul = (0.7, 1.1, 0.0)
    spec = _particle(nsam, vol, eul, ctff, 2, 0, 0.0, 1)
    mask = np.array([1, 1, 1, 1, 1], np.int32)
    p, sx, sy = C.c_float(0.0), C.c_float(0.0), C.c_float(0.0)
    L.or64_ccp(nsam, 1.0, spec.astype(np.complex128), vol, 90.0, 1, maxr2, eul[0], eul[1],
               C.byref(p), C.byref(sx), C.byref(sy), mask, ctff.astype(np.complex128), 4,
               O.IDENTITY)
    assert abs(sx.value - (-6 / 8) * 2 * math.pi) < 1e-6
    assert sy.value == 0.0  # IY = 0 has no alias in -6..6


def _psearch_case(nsam, sign, seed):
    vol = blob_volume(nsam, seed)
    ctff = _ctff(nsam, seed + 1)
    eul = (0.9, 1.2, 0.4)
    spec = sign * _particle(nsam, vol, eul, ctff, 1, -1, 0.4, seed + 2)
    return vol, ctff, spec.astype(np.complex64)


@pytest.mark.parametrize("sign,mask345", [(1.0, 1), (-1.0, 0)])
def test_psearch_is_per_view_ccp_ranked(sign, mask345):
    """TESTPAR = the IPMAX best views by strict '>' among CC > 0; a negated
    particle without the in-plane and shift searches scores most views <= 0,
    which PSEARCH never admits (their TESTPAR columns keep their previous
    contents)."""
    nsam, maxr1, maxr2, rbfact = 32, 1, 8, 0.002
    jc = nsam // 2 + 1
    vol, ctff, spec = _psearch_case(nsam, sign, 21)
    sang, iquadmax = O.searchang(30.0, "C", 1)
    nsang = len(sang)
    mask = np.array([1, 1, mask345, mask345, mask345], np.int32)
    ipmax = 12
    testpar = np.full(6 * ipmax, 7.0, np.float32)
    iran = C.c_int(-13)
    phi, theta, psi = C.c_float(0.1), C.c_float(0.2), C.c_float(0.3)
    dshx, dshy = C.c_float(0.0), C.c_float(0.0)
    ccmax = L.or32_psearch(nsam, 1.0, spec, vol, 30.0, maxr1, maxr2, C.byref(phi),
                           C.byref(theta), C.byref(psi), C.byref(dshx), C.byref(dshy),
                           np.ascontiguousarray(sang.reshape(-1)), nsang, iquadmax, mask,
                           rbfact, ctff, testpar, ipmax, C.byref(iran), O.IDENTITY)
    # the independent composition
    cb = spec.reshape(nsam, jc).copy()
    for i in range(0, nsam // 2 + 1):
        for j in range(-(nsam // 2), nsam // 2 + 1):
            w = f32(_libm.expf(f32(-f32(rbfact * f32(i * i + j * j)))))
            v = cb[j % nsam, i]
            cb[j % nsam, i] = complex(f32(f32(v.real) * w), f32(f32(v.imag) * w))
    cb = cb.reshape(-1)
    s = C.c_int(-13)
    r = [f32(L.or_random(C.byref(s))) for _ in range(3)]
    d = f32(f32(f32(30.0) / f32(180.0)) * PI_F)
    jit = [f32(f32(f32(r[k] - f32(0.5)) * d) * f32(mask[k])) for k in range(3)]
    dccp = max(f32(f32(30.0) / f32(5.0)), f32(1.0))
    rows = []
    for v in range(nsang):
        x = [f32(f32(f32(mask[k]) * f32(sang[v, k] + jit[k])) + f32((0.1, 0.2, 0.3)[k]))
             for k in range(3)]
        x = [f32(a - f32(2.0) * PI_F) if a > PI_F else a for a in x]
        p, sx, sy = C.c_float(x[2]), C.c_float(0.0), C.c_float(0.0)
        cc = f32(L.or32_ccp(nsam, 1.0, cb, vol, float(dccp), maxr1, maxr2, float(x[0]),
                            float(x[1]), C.byref(p), C.byref(sx), C.byref(sy), mask, ctff,
                            iquadmax, O.IDENTITY))
        pp = f32(p.value)
        pp = f32(pp - f32(2.0) * PI_F) if pp > PI_F else pp
        rows.append((cc, [x[0], x[1], pp, f32(sx.value), f32(sy.value)]))
    admitted = [k for k in range(nsang) if rows[k][0] > 0]
    order = sorted(admitted, key=lambda k: -float(rows[k][0]))  # stable: earlier view on ties
    tp = testpar.reshape(ipmax, 6)
    for q in range(ipmax):
        if q < len(order):
            cc, par = rows[order[q]]
            assert tp[q, 5] == cc and np.array_equal(tp[q, :5], np.array(par, np.float32)), q
        else:
            assert tp[q, 5] == 0.0 and np.all(tp[q, :5] == 7.0), q
    # the global best (any sign) is PSEARCH's returned orientation
    kb = int(np.argmax([float(rw[0]) for rw in rows]))
    assert f32(ccmax) == rows[kb][0]
    assert (f32(phi.value), f32(theta.value), f32(psi.value)) == tuple(rows[kb][1][:3])
    if sign < 0:
        assert len(admitted) < ipmax, "the negated particle should leave TESTPAR slots empty"
    else:
        assert len(order) >= ipmax
