"""Pin the oracle's FFT-using routines (the reference's copies need FFTW,
which is not in this image, so they cannot be compiled here) by what the
mathematics fixes: the FFT against numpy.fft, closed forms of CC3M and PRESB
on self-correlation, invariances, and CCP/PSEARCH recovering planted
parameters.  Each check is chosen so that a plausible mistake (a dropped
conjugate, a wrong shift sign or bin, a transposed operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O

L = O.oracle()
rng = np.random.default_rng(77)


def real_img(nsam, seed):
    r = np.random.default_rng(seed)
    a = np.zeros((nsam, nsam + 2))
    a[:, :nsam] = r.standard_normal((nsam, nsam))
    return a


@pytest.mark.parametrize("nsam", [16, 32, 48, 64, 96, 160])
def test_rfft2_matches_numpy(nsam):
    a = real_img(nsam, nsam)
    want = np.fft.rfft2(a[:, :nsam])
    d = a.copy().reshape(-1)
    L.or_rfft2_d(nsam, d)
    got = d.view(np.complex128).reshape(nsam, nsam // 2 + 1)
    assert np.max(np.abs(got - want)) < 1e-10 * np.abs(want).max()
    L.or_irfft2_d(nsam, d)
    back = d.reshape(nsam, nsam + 2)[:, :nsam] / nsam ** 2
    assert np.max(np.abs(back - a[:, :nsam])) < 1e-12


def test_irfft2_ignores_imaginary_dc_and_nyquist_columns():
    """c2r row-column semantic: Im of the x=0 and x=NSAM/2 columns (after the
    column pass) is discarded -- exactly numpy.fft.irfft2's behaviour."""
    nsam = 32
    jc = nsam // 2 + 1
    h = (rng.standard_normal((nsam, jc)) + 1j * rng.standard_normal((nsam, jc)))
    want = np.fft.irfft2(h, s=(nsam, nsam)) * nsam * nsam
    d = np.zeros(nsam * (nsam + 2))
    d.view(np.complex128)[:] = h.reshape(-1)
    L.or_irfft2_d(nsam, d)
    got = d.reshape(nsam, nsam + 2)[:, :nsam]
    assert np.max(np.abs(got - want)) < 1e-9 * np.abs(want).max()


def _volume(nsam, seed):
    r = np.random.default_rng(seed)
    v = r.standard_normal((nsam,) * 3)
    return np.ascontiguousarray(np.fft.rfftn(v).astype(np.complex64)).reshape(-1)


def _masked_projection(nsam, vol, eul, ctff, maxr1, maxr2, ring):
    """The AINPO of cc3m.f:74-115 built from pinned pieces: PEXTRACT (pinned
    to the reference) restricted to the annulus, scaled, c2r, MASKCOS2
    (pinned to the reference), r2c -- all in FP64."""
    jc = nsam // 2 + 1
    sec = np.zeros(nsam * jc, np.complex128)
    L.or64_pextract(nsam, sec, vol, *eul, ctff, O.IDENTITY)
    i = np.arange(jc)[None, :]
    m = np.arange(nsam)[:, None]
    j = np.where(m >= jc, m - nsam, m)
    it = (i * i + j * j).reshape(-1)
    keep = (it >= maxr1 * maxr1) & (it < maxr2 * maxr2)
    buf = np.where(keep, sec, 0) / nsam / nsam
    d = np.zeros(nsam * (nsam + 2))
    d.view(np.complex128)[:] = buf
    L.or_irfft2_d(nsam, d)
    L.or64_maskcos2(nsam, d, *ring, 1.0)
    L.or_rfft2_d(nsam, d)
    return d.view(np.complex128).copy(), it, keep


@pytest.mark.parametrize("nsam", [32, 64])
def test_cc3m_self_correlation_closed_form(nsam):
    """SPEC = the masked CTF projection, FSCT >= 0.95, RBFACT = 0: every shell
    gives R_k = sqrt(sum |P|^2)_k = a_k, so CC3M = sum a_k / sqrt(N sum a_k^2)."""
    jc = nsam // 2 + 1
    vol = _volume(nsam, 3)
    ctff = (0.7 + 0.3 * (rng.standard_normal(2 * nsam * jc)
                         + 1j * rng.standard_normal(2 * nsam * jc))).astype(np.complex128)
    eul = (0.3, 1.1, -0.7)
    maxr1, maxr2 = 2, nsam // 2 - 4
    ri = 0.4 * nsam
    ring = ((ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0)
    ainpo, it, keep = _masked_projection(nsam, vol, eul, ctff, maxr1, maxr2, ring)
    spec = ainpo.copy()
    fsct = np.ones(nsam // 2 + 1, np.float32)
    got = L.or64_cc3m(nsam, 1.0, spec, vol, maxr1, maxr2, *eul, 0.0, 0.0, 0.0, fsct, ctff,
                      *ring, O.IDENTITY)
    kbin = (np.sqrt(it.astype(np.float32)) + 0.5).astype(int)
    a = []
    for k in range(1, nsam // 2):
        sel = keep & (kbin == k)
        if sel.any():
            a.append(np.sqrt(np.sum(np.abs(ainpo[sel]) ** 2)))
    a = np.array(a)
    want = a.sum() / math.sqrt(len(a) * (a ** 2).sum())
    assert abs(got - want) < 1e-9, (got, want)
    # a dropped conjugate or wrong CTF side would break the self-match: score < 1 strictly
    assert 0 < got <= 1.0
    # FP32 restatement agrees to FP32 rounding
    s32 = L.or32_cc3m(nsam, 1.0, spec.astype(np.complex64), vol, maxr1, maxr2, *eul, 0.0, 0.0,
                      0.0, fsct, ctff.astype(np.complex64), *ring, O.IDENTITY)
    assert abs(s32 - got) < 1e-5


def test_cc3m_scale_invariance_and_shift_convention():
    nsam = 32
    jc = nsam // 2 + 1
    vol = _volume(nsam, 8)
    ctff = (0.6 + 0.1 * rng.standard_normal(2 * nsam * jc)).astype(np.complex128)
    eul = (-0.4, 0.9, 2.0)
    maxr1, maxr2 = 1, 12
    ring = ((9.8) ** 2, (15.8) ** 2, 9.8, 6.0)
    ainpo, it, keep = _masked_projection(nsam, vol, eul, ctff, maxr1, maxr2, ring)
    fsct = np.ones(nsam // 2 + 1, np.float32)
    base = L.or64_cc3m(nsam, 1.0, ainpo, vol, maxr1, maxr2, *eul, 0.0, 0.0, 0.0, fsct, ctff,
                       *ring, O.IDENTITY)
    scaled = L.or64_cc3m(nsam, 1.0, 3.7 * ainpo, vol, maxr1, maxr2, *eul, 0.0, 0.0, 0.0, fsct,
                         ctff, *ring, O.IDENTITY)
    assert abs(base - scaled) < 1e-12
    # a particle shifted by -s is re-aligned by CC3M at +s: SAMP = SPEC e^{i(SHX I + SHY J)}
    i = np.tile(np.arange(jc), nsam)
    m = np.repeat(np.arange(nsam), jc)
    j = np.where(m >= jc, m - nsam, m)
    sx, sy = 0.37, -0.81
    moved = ainpo * np.exp(-1j * (sx * i + sy * j))
    back = L.or64_cc3m(nsam, 1.0, moved, vol, maxr1, maxr2, *eul, sx, sy, 0.0, fsct, ctff, *ring,
                       O.IDENTITY)
    wrong = L.or64_cc3m(nsam, 1.0, moved, vol, maxr1, maxr2, *eul, -sx, -sy, 0.0, fsct, ctff,
                        *ring, O.IDENTITY)
    assert abs(back - base) < 1e-9 and wrong < base - 1e-3


def test_presb_unit_correlation():
    """PRESB of a particle equal to the masked projection: every shell = 1."""
    nsam = 32
    jc = nsam // 2 + 1
    vol = _volume(nsam, 9)
    ctff = (0.8 + 0.1 * rng.standard_normal(2 * nsam * jc)).astype(np.complex128)
    eul = (1.0, 0.5, -1.2)
    ri = 12.0
    ring = ((ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0)
    ainpo, it, keep = _masked_projection(nsam, vol, eul, ctff, 1, nsam // 2, ring)
    rbin = np.zeros(2 * nsam)
    L.or64_presb(nsam, 1.0, ainpo, vol, *eul, 0.0, 0.0, rbin, ctff, *ring, O.IDENTITY)
    vals = rbin[0:nsam:2][1:nsam // 2]
    assert np.allclose(vals[vals != 0], 1.0, atol=1e-10)


def test_ctfapply_composition():
    """CTFAPPLY = phase ramp x table, c2r, scale, MASKCOS, r2c, (-1)^(l+m),
    PWEIGHT -- built from the individually pinned routines in FP64."""
    nsam = 32
    jc = nsam // 2 + 1
    img = real_img(nsam, 5).reshape(-1)
    spec = img.copy()
    L.or_rfft2_d(nsam, spec)
    spec_c = spec.view(np.complex128).astype(np.complex64)
    pars = dict(shx=0.21, shy=-0.13, cs=2.7e7, wl=0.0197, wgh1=0.99755, wgh2=0.07, df1=15000.0,
                df2=14500.0, ang=0.3, th=0.0197 / (2.0 * nsam), amag=1.0, rih=9.0, halfw=6.0)
    pars = {k: float(np.float32(v)) for k, v in pars.items()}  # the routine takes REALs
    ri2, ri3 = 81.0, 225.0
    pssnr = np.linspace(1.0, 0.1, jc + 1).astype(np.float32)
    ctff = np.zeros(2 * nsam * jc, np.complex128)
    outc = np.zeros(nsam * jc, np.complex128)
    datc = np.zeros(nsam * jc, np.complex128)
    L.or64_ctfapply(nsam, spec_c, pars["shx"], pars["shy"], pars["cs"], pars["wl"], pars["wgh1"],
                    pars["wgh2"], pars["df1"], pars["df2"], pars["ang"], pars["th"], pars["amag"],
                    pars["rih"], pars["halfw"], ri2, ri3, pssnr, 1, ctff, outc, datc)
    # rebuild
    l = np.tile(np.arange(jc), nsam)
    m = np.repeat(np.arange(nsam), jc)
    mm = np.where(m >= jc, m - nsam, m)
    ramp = np.exp(1j * (pars["shx"] * l + pars["shy"] * mm))
    tab = np.zeros(2 * nsam * jc, np.complex128)
    out2 = np.zeros(2)
    for k in range(nsam * jc):
        L.or64_ctf(pars["cs"], pars["wl"], pars["wgh1"], pars["wgh2"], pars["df1"], pars["df2"],
                   pars["ang"], pars["th"], int(l[k]), int(mm[k]), 0.0, 0.0, out2)
        tab[k] = out2[0] + 1j * out2[1]
        L.or64_ctf(pars["cs"], pars["wl"], pars["wgh1"], pars["wgh2"], pars["df1"], pars["df2"],
                   pars["ang"], pars["th"], -int(l[k]), -int(mm[k]), 0.0, 0.0, out2)
        tab[k + nsam * jc] = out2[0] + 1j * out2[1]
    assert np.allclose(ctff, tab)
    sh = spec_c.astype(np.complex128) * ramp
    assert np.allclose(datc * ((-1.0) ** (l + mm)), sh, atol=1e-9)
    d = np.zeros(nsam * (nsam + 2))
    d.view(np.complex128)[:] = sh
    L.or_irfft2_d(nsam, d)
    d.reshape(nsam, nsam + 2)[:, :nsam] /= nsam * nsam
    L.or64_maskcos(nsam, d, ri2, ri3, pars["rih"], pars["halfw"], 1.0)
    L.or_rfft2_d(nsam, d)
    w = d.view(np.complex128) * ((-1.0) ** (l + mm))
    L.or64_pweight(nsam, w, w, pssnr, tab, 1)
    assert np.allclose(outc, w, rtol=1e-9, atol=1e-9 * np.abs(w).max())


def test_ccp_recovers_planted_shift_and_view():
    """A particle equal to the CTF projection of the volume at a grid view,
    shifted by whole NSAMR pixels, is found by CCP at that view's in-plane
    angle (within one in-plane step) with exactly that shift (ccp.f:229-256)."""
    nsam = 64
    jc = nsam // 2 + 1
    vol = _volume(nsam, 12)
    ctff = np.full(2 * nsam * jc, 0.5, np.complex64)
    phi, theta, psi = 0.6, 1.0, math.radians(20.0)
    sec = np.zeros(nsam * jc, np.complex64)
    L.or32_pextract(nsam, sec, vol, phi, theta, psi, ctff, O.IDENTITY)
    nsamr = nsam // 2
    ixs, iys = 2, -3  # NSAMR pixels
    l = np.tile(np.arange(jc), nsam)
    m = np.repeat(np.arange(nsam), jc)
    mm = np.where(m >= jc, m - nsam, m)
    spec = (sec * np.exp(-1j * 2 * np.pi * (l * ixs + mm * iys) / nsamr)).astype(np.complex64)
    mask = np.array([1, 1, 1, 1, 1], np.int32)
    p, sx, sy = C.c_float(0.0), C.c_float(0), C.c_float(0)
    cc = L.or32_ccp(nsam, 1.0, spec, vol, 5.0, 1, nsam // 4, phi, theta, C.byref(p), C.byref(sx),
                    C.byref(sy), mask, ctff, 4, O.IDENTITY)
    assert abs(sx.value - ixs / nsamr * 2 * math.pi) < 1e-6
    assert abs(sy.value - iys / nsamr * 2 * math.pi) < 1e-6
    assert abs(math.remainder(p.value - psi, 2 * math.pi)) < math.radians(5.0 + 0.1)
    assert cc > 0.1


# ------------------------------------------------------------------ SIGMA
def _sigma_case(nsam, seed, beta=1.0, noise=0.0):
    """A particle made from the section itself: DATC = beta * section * ramp
    (+ white noise), with the ramp exp(+i(SHX I + SHY J)) that SIGMA undoes."""
    jc = nsam // 2 + 1
    r = np.random.default_rng(seed)
    vol = _volume(nsam, seed)
    ctff = (0.6 + 0.2 * r.standard_normal((2, nsam, jc))).astype(np.complex64)
    eul = r.uniform(-1, 1, 3).astype(np.float32)
    sec = np.zeros(nsam * jc, np.complex64)
    L.or32_pextract(nsam, sec, vol, *[float(x) for x in eul], ctff.reshape(-1), O.IDENTITY)
    sec = sec.reshape(nsam, jc)
    shx, shy = np.float32(0.11), np.float32(-0.07)
    m = np.arange(nsam)
    mm = np.where(m >= jc, m - nsam, m)
    l = np.arange(jc)
    ph = (shx * l[None, :] + shy * mm[:, None]).astype(np.float64)
    datc = (beta * sec * np.exp(1j * ph)).astype(np.complex64)
    if noise:
        datc = (datc + noise * (r.standard_normal(datc.shape) + 1j * r.standard_normal(datc.shape))
                ).astype(np.complex64)
    return vol, datc, ctff, eul, shx, shy


def _sigma(pre, nsam, vol, datc, ctff, eul, shx, shy, signset=0.0, sigp=1.0, dmask=None,
           rclas=0.4):
    ct, rt = (np.complex64, np.float32) if pre == "or32_" else (np.complex128, np.float64)
    out = np.zeros(3, rt)
    dm = np.array(dmask if dmask is not None else [nsam / 2 + 1, nsam / 2 + 1, nsam / 2 + 1,
                                                  nsam * 0.35], np.float32)
    getattr(L, pre + "sigma")(nsam, 1.0, vol, datc.reshape(-1).astype(ct), *[float(x) for x in eul],
                              float(shx), float(shy), ctff.reshape(-1).astype(ct), 1, rclas,
                              float(signset), dm, float(sigp), O.IDENTITY, out)
    return out


def test_sigma_noise_free_particle_has_no_noise():
    """A particle equal to the CTF-weighted section: the fitted scale A is 1, so
    the noise variance SIGMA2 vanishes; scaling the particle by beta leaves
    SIGMA2N unchanged (S2OLD ~ beta^2, A ~ beta, sqrt(SUM1 SUM2) ~ beta)."""
    nsam = 32
    vol, datc, ctff, eul, shx, shy = _sigma_case(nsam, 5)
    s1 = _sigma("or64_", nsam, vol, datc, ctff, eul, shx, shy)
    vol, datc2, ctff, eul, shx, shy = _sigma_case(nsam, 5, beta=2.0)
    s2 = _sigma("or64_", nsam, vol, datc2, ctff, eul, shx, shy)
    _, datcn, *_ = _sigma_case(nsam, 5, noise=0.5)
    sn = _sigma("or64_", nsam, vol, datcn, ctff, eul, shx, shy)
    assert sn[1] > 0
    assert s1[1] < 1e-8 * sn[1] and s2[1] < 1e-8 * sn[1]
    assert abs(s2[0] - s1[0]) < 1e-9 * abs(s1[0])
    # the FP32 restatement agrees with FP64 up to the cancellation in
    # SUM2 - A^2 SUM1 (FP32 images, FP64 sums: 4e-4 here)
    s32 = _sigma("or32_", nsam, vol, datcn, ctff, eul, shx, shy)
    assert np.allclose(s32[:2], sn[:2], rtol=2e-3)


def test_sigma_logpc_closed_form():
    """With SIGNSET inside the accepted band, SIGMA2OLD = SIGNSET^2; with SIGP =
    SIGNSET the likelihood count is the number of pixels in the mask disc, so
    LOGPC = -N (SIGMA2/SIGNSET^2 + log(2 pi SIGNSET^2)) / 2 with N counted here."""
    nsam = 32
    vol, datc, ctff, eul, shx, shy = _sigma_case(nsam, 9, noise=0.3)
    base = _sigma("or64_", nsam, vol, datc, ctff, eul, shx, shy)
    # SIGMA2OLD of this particle without SIGNSET: recover a SIGNSET in range
    rad = nsam * 0.35
    dmask = [nsam / 2 + 1, nsam / 2 + 1, nsam / 2 + 1, rad]
    # the mask centre is the box centre: XM = YM = 1 (rotation of the zero vector)
    n = 0
    for Lx in range(1, nsam + 1):
        dx = Lx - 1.0 - (nsam if Lx >= nsam // 2 else 0)
        for My in range(1, nsam + 1):
            dy = My - 1.0 - (nsam if My >= nsam // 2 else 0)
            if dx * dx + dy * dy <= rad * rad:
                n += 1
    out = _sigma("or64_", nsam, vol, datc, ctff, eul, shx, shy, signset=1.0, sigp=1.0, dmask=dmask)
    # SIGNSET = 1 is accepted only when sqrt(SIGMA2OLD) is within 1/3..3 of it; pick one that is
    for signset in np.geomspace(1e-6, 1e6, 121):
        o = _sigma("or64_", nsam, vol, datc, ctff, eul, shx, shy, signset=signset, sigp=signset,
                   dmask=dmask)
        want = -n * o[1] / signset ** 2 / 2 - n * math.log(2 * math.pi * signset ** 2) / 2
        if abs(o[2] - want) <= 1e-9 * abs(want):
            break
    else:
        pytest.fail("no SIGNSET inside the accepted band reproduced the closed form")
    assert base[1] == o[1] and out is not None
