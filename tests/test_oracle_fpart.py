"""Pins of the FIND_FPART restatement (oracle/fpart_oracle.py), on the CPU.

  * VOLMEASURE on a volume of two levels returns the fraction of the upper
    level exactly (the two-class iteration settles between them);
  * FPART_CC is 1 when D and V are built so that D / (V + FPART / SSNR(I))
    equals B voxel for voxel -- which holds only with the reference's shell
    index I = INT(SQRT(r^2) + 0.5) + 1, sign (-1)^(l+m+n) and 1/N^3 -- and it
    is unchanged by a positive scale of D;
  * the MW branch is the closed form 1000 MW / DALT / (PSIZE NSAM)^3, and
    PSSNR is divided by FPART in its first NSAM/2 + 1 shells.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import fpart_oracle as FO  # noqa: E402


def test_volmeasure_two_levels():
    nsam = 16
    r = np.random.default_rng(1)
    v = np.zeros((nsam, nsam, nsam + 2), np.float32)
    hi = r.random((nsam, nsam, nsam)) < 0.3
    v[:, :, :nsam] = np.where(hi, 2.0, -1.0)
    assert FO.volmeasure(v, nsam) == np.float32(hi.sum()) / np.float32(nsam ** 3)


def test_fpart_cc_closed_form():
    nsam, fpart = 16, 0.35
    jc = nsam // 2 + 1
    r = np.random.default_rng(2)
    a = (r.standard_normal((nsam, nsam, jc)) + 1j * r.standard_normal((nsam, nsam, jc))
         ).astype(np.complex64)
    s = r.uniform(1.0, 2.0, (nsam, nsam, jc)).astype(np.float32)
    std, iredun = 0.3, 2
    ssnr = r.uniform(0.5, 4.0, jc + 8).astype(np.float32)
    fsc = np.full(jc + 8, 0.9, np.float32)
    # B as FIND_FPART forms it (find_fpart.f:57), in real space
    l3, m3, n3 = FO._miller(nsam)
    sg = np.where((l3 + m3 + n3) % 2 != 0, -1.0, 1.0)
    b = (a * sg / (s.astype(np.float64) + std * iredun) / float(nsam) ** 3).astype(np.complex64)
    bv = FO._c2r(b, nsam)
    mb = np.float32(bv[:, :, :nsam].mean())
    # V so that V + FPART / SSNR(I) = S + STD IREDUN at every voxel, D = A
    sh = (np.sqrt((l3 * l3 + m3 * m3 + n3 * n3).astype(np.float32)) + 0.5).astype(int)
    v = (s + std * iredun - np.float32(fpart) / ssnr[sh]).astype(np.float32)
    cc = FO.fpart_cc(nsam, a, v, fpart, ssnr, fsc, bv, mb)
    assert abs(cc - 1.0) < 1e-5, cc
    assert abs(FO.fpart_cc(nsam, a * np.complex64(3.0), v, fpart, ssnr, fsc, bv, mb) - cc) < 1e-5
    # a wrong shell table breaks it
    cc2 = FO.fpart_cc(nsam, a, v, fpart, np.roll(ssnr, 1), fsc, bv, mb)
    assert cc2 < 1.0 - 1e-3


def test_mw_closed_form():
    nsam, psize, mw, dalt = 64, 1.5, 400.0, 0.81
    pssnr = np.linspace(5, 1, nsam // 2 + 4).astype(np.float32)
    f, out = FO.find_fpart(nsam, 1.0, 1.0, None, None, None, None, 1, 0.0, 0.0, psize, mw, dalt,
                           None, pssnr, None)
    want = 1000.0 * mw / dalt / (psize * nsam) ** 3
    assert abs(f - want) < 1e-6 * want
    jc = nsam // 2 + 1
    assert np.allclose(out[:jc], pssnr[:jc] / np.float32(f), rtol=1e-6)
    assert np.array_equal(out[jc:], pssnr[jc:])
