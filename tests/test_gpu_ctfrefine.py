"""CTFREFINE and MAGREFINE on the GPU (fr_ctfrefine_batch, fr_magrefine_batch)
against the oracle's (ctfrefine.f / magrefine.f over CALCFX's CTF and
magnification sections with CC3; VA04A bit-exact to va04.f).

The two objectives agree to FP32 rounding (CTFAPPLY and CC3 each ~1e-6).
Where the two Powell runs take the same path (the same number of objective
evaluations) the bar is the final objective within 1e-3 relative, the
defocus changes within 20 A, the astigmatism angle within 0.05 rad and the
magnification within 2e-3.  With free astigmatism (FASTIG) a particle whose
defocus is nearly isotropic leaves the angle direction flat, and VA04A's
1-radian steps at ESCALE 1000 (ctfrefine.f:32-33) let rounding-level
differences send the two runs down different paths: there each end point,
re-evaluated by the oracle, must lie below the start and within 1 % of the
other's objective.  Groups: a film of four particles (SIG2 0) and two single
particles with the FD prior."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, synth  # noqa: E402

DEV = "cuda"


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def _case(ctx, nsam=64, psize=2.0, n=6, seed=61):
    ri = 0.4 * nsam
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 40, 0.375 * nsam, seed=seed, device=DEV)
    a3df = synth.half_volume(vol)
    parts = synth.make_particles(ctx, a3df, n, seed=seed + 1, optics=optics, snr=5.0, ri=ri,
                                 shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    imgs = parts.images.clone()
    ctx.ctfapply(nsam, imgs, parts.ppar(), pssnr, gpar)  # leaves the r2c spectra in imgs
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    return a3df, imgs, parts, gpar, pssnr, ring


@pytest.mark.parametrize("fastig", [False, True])
def test_ctfrefine_vs_oracle(ctx, fastig):
    L = O.oracle()
    nsam = 64
    a3df, spectra, parts, gpar, pssnr, ring = _case(ctx, nsam)
    n = spectra.shape[0]
    ppar = parts.ppar().cpu().numpy()
    ppar[:, 0:2] = parts.shift          # CTFAPPLY re-centres the particle
    off = np.array([250.0, 250.0, 250.0, 250.0, -180.0, 120.0], np.float32)
    ppar[:, 2] += off
    ppar[:, 3] += off * (0.6 if fastig else 1.0)
    groups = [4, 1, 1]
    sig2 = np.array([0.0, 0.5, 0.5], np.float32)
    rmax1, rmax2, rbfact, dfstd = 1.0 / nsam, 14.0 / nsam + 1e-4, 0.0, 150.0
    print(f"[ctfrefine] fastig={fastig}: groups {groups}", flush=True)
    out, nev = ctx.ctfrefine(nsam, spectra, ppar, parts.euler, groups, a3df, pssnr, gpar,
                             rmax1=rmax1, rmax2=rmax2, rbfact=rbfact, ring=ring, fastig=fastig,
                             dfstd=dfstd, sig2=sig2)
    jc = nsam // 2 + 1
    spec_h = np.ascontiguousarray(spectra.cpu().numpy().reshape(n, -1).view(np.complex64))
    vol = a3df.cpu().numpy().reshape(-1)
    ps = pssnr.cpu().numpy()
    p0 = 0
    for g, sz in enumerate(groups):
        x = np.zeros(3, np.float32)
        ne = C.c_int(0)
        f = L.or32_ctfrefine(nsam, sz, np.ascontiguousarray(spec_h[p0:p0 + sz]),
                             np.ascontiguousarray(ppar[p0:p0 + sz]),
                             np.ascontiguousarray(parts.euler[p0:p0 + sz]), int(fastig), dfstd,
                             float(sig2[g]), vol, ps, gpar, rmax1, rmax2, rbfact, ring, x,
                             C.byref(ne))
        print(f"[ctfrefine] group {g}: gpu {out[g]} ({nev[g]} evals), oracle {x} F {f:.5f} "
              f"({ne.value} evals)")
        if not fastig:
            assert out[g, 1] == out[g, 0] and out[g, 2] == 0.0
        if nev[g] == ne.value:
            assert abs(out[g, 3] - f) <= 1e-3 * abs(f)
            assert np.max(np.abs(out[g, :2] - x[:2])) < 20.0
            assert abs(out[g, 2] - x[2]) < 0.05
        else:
            def objective(xx):
                pp = np.ascontiguousarray(ppar[p0:p0 + sz]).copy()
                pp[:, 2] += xx[0]
                pp[:, 3] += xx[1]
                pp[:, 4] += xx[2]
                # CALCFX's CTF section at xx through the oracle's CTFAPPLY and CC3
                tot = np.float32(0.0)
                for k in range(sz):
                    ctf = np.zeros(2 * nsam * jc, np.complex64)
                    oc = np.zeros(nsam * jc, np.complex64)
                    dc = np.zeros(nsam * jc, np.complex64)
                    q = pp[k]
                    L.or32_ctfapply(nsam, spec_h[p0 + k], float(q[0]), float(q[1]),
                                    *[float(v) for v in gpar[:4]], float(q[2]), float(q[3]),
                                    float(q[4]), float(gpar[4]), float(q[5]),
                                    *[float(v) for v in gpar[5:9]], ps, 0, ctf, oc, dc)
                    mr1 = int(np.float32(nsam) * np.float32(rmax1) * np.float32(abs(q[5])))
                    mr2 = int(np.float32(nsam) * np.float32(rmax2) * np.float32(abs(q[5])))
                    e = parts.euler[p0 + k]
                    cc = np.float32(L.or32_cc3(nsam, float(q[5]), oc, vol, mr1, mr2,
                                               float(e[0]), float(e[1]), float(e[2]), 0.0, 0.0,
                                               rbfact, ctf, *[float(v) for v in ring],
                                               O.IDENTITY))
                    fdv = np.float32(sig2[g]) * (-(np.float32(xx[0]) ** 2) / np.float32(2)
                                                 / np.float32(dfstd) ** 2
                                                 - (np.float32(xx[1]) ** 2) / np.float32(2)
                                                 / np.float32(dfstd) ** 2)
                    tot = np.float32(tot - cc - fdv)
                return float(tot / np.float32(sz))

            f0, fg, fo = objective((0, 0, 0)), objective(out[g, :3]), objective(x)
            print(f"[ctfrefine] group {g}: paths parted; objective start {f0:.5f}, "
                  f"gpu end {fg:.5f}, oracle end {fo:.5f}")
            assert fg <= f0 + 1e-6 and fo <= f0 + 1e-6
            assert abs(fg - out[g, 3]) <= 1e-4 * abs(fg)
            assert abs(fg - fo) <= 1e-2 * abs(fo)
        p0 += sz


def test_magrefine_vs_oracle(ctx):
    L = O.oracle()
    nsam = 64
    a3df, spectra, parts, gpar, pssnr, ring = _case(ctx, nsam, seed=71)
    n = spectra.shape[0]
    ppar = parts.ppar().cpu().numpy()
    ppar[:, 0:2] = parts.shift
    ppar[:3, 5] = 1.015
    ppar[3:, 5] = 0.99
    groups = [3, 3]
    rmax1, rmax2 = 1.0 / nsam, 14.0 / nsam + 1e-4
    out, nev = ctx.ctfrefine(nsam, spectra, ppar, parts.euler, groups, a3df, pssnr, gpar,
                             rmax1=rmax1, rmax2=rmax2, rbfact=0.0, ring=ring, magnification=True)
    jc = nsam // 2 + 1
    spec_h = np.ascontiguousarray(spectra.cpu().numpy().reshape(n, -1).view(np.complex64))
    vol = a3df.cpu().numpy().reshape(-1)
    ps = pssnr.cpu().numpy()
    for g, p0 in enumerate((0, 3)):
        x = np.zeros(3, np.float32)
        ne = C.c_int(0)
        f = L.or32_magrefine(nsam, 3, np.ascontiguousarray(spec_h[p0:p0 + 3]),
                             np.ascontiguousarray(ppar[p0:p0 + 3]),
                             np.ascontiguousarray(parts.euler[p0:p0 + 3]), vol, ps, gpar, rmax1,
                             rmax2, 0.0, ring, x, C.byref(ne))
        print(f"[magrefine] group {g}: gpu AMAG {out[g, 0]:.5f} F {out[g, 3]:.5f} ({nev[g]}), "
              f"oracle {x[0]:.5f} F {f:.5f} ({ne.value})")
        assert abs(out[g, 3] - f) <= 1e-3 * abs(f)
        assert abs(out[g, 0] - x[0]) < 2e-3
    assert jc > 0
