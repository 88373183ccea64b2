"""FIND_FPART restated in numpy (test infrastructure only: only tests/ may
import this module).

find_fpart.f:8-150 with FPART_CC (fpart_cc.f:8-94) and VOLMEASURE
(volmeasure.f:8-69).  Volumes: half-volume transforms indexed [n][m][l]
(complex64 A, D; float32 S, V), real volumes in the padded layout
[n][m][NSAM + 2] that MASK3D works on.  The 3D c2r is numpy's inverse real
FFT times N^3 (FFTW_BWD is unnormalised); arithmetic is FP64 with FP32
rounding where the reference stores REAL values; MASK3D is passed in (the
C oracle's or32_mask3d).  The class sums of VOLMEASURE are taken in FP64
(the reference adds them in FP32 in voxel order).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _miller(nsam: int):
    jc = nsam // 2 + 1
    idx = np.arange(nsam)
    mm = np.where(idx >= jc, idx - nsam, idx)
    n3, m3, l3 = np.meshgrid(mm, mm, np.arange(jc), indexing="ij")
    return l3, m3, n3


def _c2r(h: np.ndarray, nsam: int) -> np.ndarray:
    """FFTW_BWD of a half-volume into the padded real layout (float32)."""
    r = np.fft.irfftn(h.astype(np.complex128), s=(nsam,) * 3, axes=(0, 1, 2)) * float(nsam) ** 3
    out = np.zeros((nsam, nsam, nsam + 2), np.float32)
    out[:, :, :nsam] = r.astype(np.float32)
    return out


def volmeasure(cv: np.ndarray, nsam: int) -> float:
    """volmeasure.f:17-60: ten two-class mean iterations from 0, then the
    fraction of voxels at or above the last threshold."""
    x = cv[:, :, :nsam].astype(np.float64)
    thresh = F32(0.0)
    d1 = d2 = F32(0.0)
    for _ in range(10):
        hi = x >= thresh
        n2, n1 = int(hi.sum()), int((~hi).sum())
        d2 = F32(x[hi].sum() / n2) if n2 else F32(x[hi].sum())
        d1 = F32(x[~hi].sum() / n1) if n1 else F32(x[~hi].sum())
        thresh = F32(F32(d2 - d1) / F32(2.0) + d1)
    thresh = F32(F32(d2 - d1) / F32(2.0) + d1)
    return float(F32(F32(int((x >= thresh).sum())) / F32(nsam ** 3)))


def fpart_cc(nsam, d, v, fpart, ssnr, fsc, bv, mave):
    """fpart_cc.f:31-79: the correlation over the voxels where B != 0 of B
    with the real-space D / (V + FPART / SSNR(I)) (shells with SSNR 0 or FSC
    <= 0.5 left out)."""
    l3, m3, n3 = _miller(nsam)
    sg = np.where((l3 + m3 + n3) % 2 != 0, F32(-1.0), F32(1.0))
    r2 = (l3 * l3 + m3 * m3 + n3 * n3).astype(np.float32)
    sh = (np.sqrt(r2) + F32(0.5)).astype(np.int64)
    ok = np.zeros(sh.shape, bool)
    inr = sh < len(ssnr)
    shc = np.minimum(sh, len(ssnr) - 1)
    ok[inr] = (np.asarray(ssnr, np.float32)[shc][inr] != 0) & (
        np.asarray(fsc, np.float32)[shc][inr] > 0.5)
    den = (v + F32(fpart) / np.where(ok, np.asarray(ssnr, np.float32)[shc], F32(1.0))).astype(
        np.float32)
    n3f = F32(float(nsam) ** 3)
    c = np.where(ok, (d * sg / den / n3f).astype(np.complex64), np.complex64(0))
    cv = _c2r(c, nsam)[:, :, :nsam].astype(np.float64)
    b = bv[:, :, :nsam]
    sel = b != 0
    n = int(sel.sum())
    s1, s2 = cv[sel].sum(), (cv[sel].astype(np.float32) ** 2).astype(np.float64).sum()
    if n:
        s1, s2 = s1 / n, s2 / n
    mc = F32(s1)
    db = (b[sel] - F32(mave)).astype(np.float32)
    dc = (cv[sel].astype(np.float32) - mc).astype(np.float32)
    t1 = (db * dc).astype(np.float64).sum()
    t2 = (db * db).astype(np.float64).sum()
    t3 = (dc * dc).astype(np.float64).sum()
    return float(F32(t1 / np.sqrt(t2 * t3)))


def find_fpart(nsam, std, vtd, a, d, s, v, iredun, ri, ric, psize, mw, dalt, fsc, pssnr,
               mask3d):
    """find_fpart.f:37-149; returns (FPART, PSSNR / FPART).  mask3d(vol, ri,
    ric, halfw) masks the padded real volume in place and returns (FMASK,
    MAVE, MSTD)."""
    if std == 0.0 or vtd == 0.0:
        raise ValueError("STD or VTD zero")
    jc = nsam // 2 + 1
    if mw == 0.0:
        l3, m3, n3 = _miller(nsam)
        sg = np.where((l3 + m3 + n3) % 2 != 0, F32(-1.0), F32(1.0))
        n3d = float(nsam) ** 3
        den1 = s.astype(np.float64) + std * iredun
        b = ((a * sg).astype(np.complex128) / den1 / n3d).astype(np.complex64)
        den2 = (s + v).astype(np.float32).astype(np.float64) + (std + vtd) * iredun
        c = ((a + d) * sg).astype(np.complex128) / den2 / n3d
        frad2 = F32(F32(F32(2.0) * F32(psize)) / F32(30.0) * F32(nsam))
        g = np.exp(-(l3 * l3 + m3 * m3 + n3 * n3).astype(np.float32) / frad2).astype(np.float32)
        c = (c.astype(np.complex64) * g).astype(np.complex64)
        bv = _c2r(b, nsam)
        cv = _c2r(c, nsam)
        fmask, mave, mstd = mask3d(cv, ri, ric, 1.0)
        lim = F32(F32(0.2) * F32(nsam ** 3)) * F32(fmask)
        x = cv[:, :, :nsam]
        tt = None
        for k in range(16):
            tt = F32(F32(mave) + F32(F32(1.5) - F32(k) * F32(0.1)) * F32(mstd))
            if F32(int((x >= tt).sum())) > lim:
                break
        b3 = bv[:, :, :nsam]
        b3[x < tt] = 0.0
        sel = x >= tt
        n = int(sel.sum())
        s1 = b3[sel].astype(np.float64).sum()
        if n:
            s1 = s1 / n
        mb = F32(s1)
        f = F32(volmeasure(cv, nsam))
        df1 = F32(-0.2)
        ccmax = fpart_cc(nsam, d, v, f, pssnr, fsc, bv, mb)
        while True:
            f = F32(f + df1)
            cc = fpart_cc(nsam, d, v, f, pssnr, fsc, bv, mb)
            if cc > ccmax:
                ccmax = cc
                continue
            f = F32(f - df1)
            if abs(df1) > F32(0.005):
                df1 = F32(-df1 / F32(3.0))
                continue
            break
    else:
        box = F32(F32(psize) * F32(nsam))
        f = F32(F32(F32(1000.0) * F32(mw)) / F32(dalt) / F32(box * box * box))
    f = F32(abs(f))
    out = np.asarray(pssnr, np.float32).copy()
    if f != 0:
        out[:jc] = (out[:jc] / f).astype(np.float32)
    return float(f), out
