"""A Fortran host calling PEXTRACT, CC3M, PRESB and SIGMA with the reference's
own symbols and argument lists, linked against libfrealign_hip.so in place of
the reference routines, against the FP32 oracle on the same inputs.

Bars as the batched parity tests: PEXTRACT 1e-5 relative L2, CC3M 1e-4
relative, PRESB 2e-5 absolute per shell (unwritten shells untouched), SIGMA
1e-3 relative (SIGMA2 carries a cancellation) and LOGPC 1e-4."""
import math
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
LIB = os.path.join(ROOT, "frealign_amd", "libfrealign_hip.so")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_percall_vs_oracle(tmp_path):
    L = O.oracle()
    nsam, maxr1, maxr2 = 64, 2, 20
    jc = nsam // 2 + 1
    ic = nsam * jc
    r = np.random.default_rng(11)
    vol = np.ascontiguousarray(np.fft.rfftn(r.standard_normal((nsam,) * 3)).astype(np.complex64))
    spec = (r.standard_normal((nsam, jc)) + 1j * r.standard_normal((nsam, jc))).astype(np.complex64)
    datc = (r.standard_normal((nsam, jc)) + 1j * r.standard_normal((nsam, jc))).astype(np.complex64)
    ctff = (0.5 + 0.2 * r.standard_normal((2, nsam, jc))).astype(np.complex64)
    fsct = np.zeros(jc, np.float32)
    fsct[: jc // 2] = 1.0
    rbin0 = np.full(nsam, 7.0, np.float32)
    phi, theta, psi, shx, shy, amag = 0.3, 1.1, -0.7, 0.05, -0.08, 1.01
    ri = 0.4 * nsam
    ri2, ri3, rih, halfw = (ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0
    rbfact, rclas, sigp, signset = 0.001, 0.3, 0.01, 0.0
    dmask = np.array([nsam / 2 + 2, nsam / 2 - 1, nsam / 2 + 1, 0.35 * nsam], np.float32)
    sm = O.IDENTITY.astype(np.float32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([nsam, maxr1, maxr2], np.int32).tofile(f)
        np.array([phi, theta, psi, shx, shy, amag, ri2, ri3, rih, halfw, rbfact, rclas, sigp,
                  signset], np.float32).tofile(f)
        dmask.tofile(f)
        sm.tofile(f)
        for a in (vol, spec, datc, ctff, fsct, rbin0):
            a.tofile(f)
    exe = tmp_path / "fr_percall_check"
    rr = subprocess.run([FLANG, "-O1", os.path.join(ROOT, "frealign_amd", "fortran",
                                                   "fr_percall_check.f90"), LIB,
                         f"-Wl,-rpath,{os.path.dirname(LIB)}", "-o", str(exe)],
                        capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    print("running the Fortran host", flush=True)
    rr = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True,
                        timeout=120)
    assert rr.returncode == 0, rr.stdout + rr.stderr
    assert "FR_PERCALL_CHECK done" in rr.stdout
    raw = np.fromfile(fout, dtype=np.uint8)
    off = 0
    outc = raw[off:off + 8 * ic].view(np.complex64)
    off += 8 * ic
    ccv = raw[off:off + 4].view(np.float32)[0]
    off += 4
    rbin = raw[off:off + 4 * nsam].view(np.float32)
    off += 4 * nsam
    s2n, s2, lpc = raw[off:off + 12].view(np.float32)
    # oracle
    w = np.zeros(ic, np.complex64)
    L.or32_pextract(nsam, w, vol.reshape(-1), phi, theta, psi, ctff.reshape(-1), sm)
    e = np.linalg.norm(outc - w) / np.linalg.norm(w)
    print(f"PEXTRACT rel L2 {e:.1e}")
    assert e < 1e-5
    c32 = L.or32_cc3m(nsam, amag, spec.reshape(-1), vol.reshape(-1), maxr1, maxr2, phi, theta,
                      psi, shx, shy, rbfact, fsct, ctff.reshape(-1), ri2, ri3, rih, halfw, sm)
    print(f"CC3M {ccv:.7f} oracle {c32:.7f}")
    assert abs(ccv - c32) <= 1e-4 * abs(c32)
    rb = rbin0.copy()
    L.or32_presb(nsam, amag, spec.reshape(-1), vol.reshape(-1), phi, theta, psi, shx, shy, rb,
                 ctff.reshape(-1), ri2, ri3, rih, halfw, sm)
    assert np.array_equal(rbin[1::2], rb[1::2])
    assert np.max(np.abs(rbin[0::2] - rb[0::2])) < 2e-5
    out = np.zeros(3, np.float32)
    L.or32_sigma(nsam, amag, vol.reshape(-1), datc.reshape(-1), phi, theta, psi, shx, shy,
                 ctff.reshape(-1), maxr1, rclas, signset, dmask, sigp, sm, out)
    got = np.array([s2n, s2, lpc])
    rel = np.abs(got - out) / np.abs(out)
    print(f"SIGMA {got} oracle {out} rel {rel}")
    assert rel[0] < 1e-3 and rel[1] < 1e-3 and rel[2] < 1e-4
    assert math.isfinite(ccv)
