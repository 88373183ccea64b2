"""fr_refine_batch (PREFINE on the GPU) against the oracle's PREFINE: the
oracle's VA04A (bit-exact to va04.f) minimising CALCFX built on the oracle's
FP32 CC3M.  The two objectives agree to FP32 rounding (CC3M parity ~1e-6), so
the Powell paths may part after many steps: the bar is the final PRESA within
1e-3 relative and the parameters within 0.01 rad / 0.005 phase units, plus
self-consistency (PRESA equals CC3M + FSH at the returned parameters) and
improvement over the start."""
import ctypes as C
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, synth  # noqa: E402

DEV = "cuda"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI = np.float32(3.1415926535897)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def fsh(sig2, x, y, xm, ym, sx, sy):
    f = np.float32
    return f(sig2) * (-(f(x) - f(xm)) * (f(x) - f(xm)) / f(2) / (f(sx) * f(sx))
                      - (f(y) - f(ym)) * (f(y) - f(ym)) / f(2) / (f(sy) * f(sy)))


def test_refine_vs_oracle(ctx):
    L = O.oracle()
    nsam, psize, ri = 64, 2.0, 26.0
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 30, 24.0, seed=8, device=DEV)
    a3df = synth.half_volume(vol)
    n = 4
    parts = synth.make_particles(ctx, a3df, n, seed=9, optics=optics, snr=2.0, ri=ri, shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    fsct = synth.default_fsct(nsam, psize).astype(np.float32)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    maxr1, maxr2, rbfact = 2, 14, 0.0
    rng = np.random.default_rng(5)
    par = np.zeros((n, 6), np.float32)
    par[:, 0:3] = parts.euler + np.radians(rng.normal(0, 1.5, (n, 3))).astype(np.float32)
    par[:, 3:5] = 0.0
    prior = np.zeros((n, 5), np.float32)
    prior[:, 2:4] = 0.3
    prior[:, 4] = 0.01
    mask = np.ones(5, np.int32)
    amag = np.ones(n, np.float32)
    got, nev = ctx.refine(nsam, a3df, outc, ctff, np.arange(n), amag, par, mask, maxr1, maxr2,
                          rbfact, torch.from_numpy(fsct).to(DEV), ring, prior)
    vol_h = a3df.cpu().numpy().reshape(-1)
    spec_h = outc.cpu().numpy()
    ctff_h = ctff.cpu().numpy()

    def cc3m(k, x):
        return np.float32(L.or32_cc3m(nsam, 1.0, spec_h[k].reshape(-1), vol_h, maxr1, maxr2,
                                      float(x[0]), float(x[1]), float(x[2]), float(x[3]),
                                      float(x[4]), rbfact, fsct, ctff_h[k].reshape(-1),
                                      *[float(v) for v in ring], O.IDENTITY))

    for k in range(n):
        def objective(nn, px, _u, k=k):
            x = np.ctypeslib.as_array(px, shape=(nn,)).copy()
            rf = -cc3m(k, x) - fsh(prior[k, 4], x[3], x[4], prior[k, 0], prior[k, 1],
                                  prior[k, 2], prior[k, 3])
            for j, lim in ((0, PI), (1, PI), (2, PI), (3, np.float32(1)), (4, np.float32(1))):
                if abs(x[j]) > lim:
                    rf = rf + (lim - abs(x[j])) * (lim - abs(x[j]))
            return float(rf)
        cb = C.CFUNCTYPE(C.c_float, C.c_int, C.POINTER(C.c_float), C.c_void_p)(objective)
        x = par[k, :5].copy()
        ne = C.c_int(0)
        f = L.or_va04a(x, np.array([0.1, 0.1, 0.1, 0.01, 0.01], np.float32), 5, 100.0, 0, 1, 50,
                       C.cast(cb, C.c_void_p), None, C.byref(ne))
        presa_o = -f
        start = cc3m(k, par[k]) + fsh(prior[k, 4], 0, 0, 0, 0, prior[k, 2], prior[k, 3])
        # self-consistency of the GPU result
        again = cc3m(k, got[k]) + fsh(prior[k, 4], got[k, 3], got[k, 4], 0, 0, prior[k, 2],
                                      prior[k, 3])
        print(f"item {k}: start {start:.5f}  gpu {got[k, 5]:.5f} ({nev[k]} evals)  "
              f"oracle {presa_o:.5f} ({ne.value} evals)  "
              f"|dangles| {np.max(np.abs(got[k, :3] - x[:3])):.1e} |dshift| "
              f"{np.max(np.abs(got[k, 3:5] - x[3:5])):.1e}")
        assert abs(again - got[k, 5]) < 1e-4 * max(1.0, abs(got[k, 5]))
        assert got[k, 5] >= start - 1e-6
        assert abs(got[k, 5] - presa_o) < 1e-3 * max(abs(presa_o), 1e-2)
        assert np.max(np.abs(got[k, :3] - x[:3])) < 0.01
        assert np.max(np.abs(got[k, 3:5] - x[3:5])) < 0.005


def test_prefine_percall_matches_batch(ctx):
    """prefine_ (the reference's PREFINE symbol and argument list, called by
    reference as a Fortran LMAIN would) gives the batched refinement's answer
    bit for bit, for all five parameters and for an angles-only MASK; the
    batched refinement is held to the oracle's VA04A + CC3M above."""
    lib = C.CDLL(os.path.join(ROOT, "frealign_amd", "libfrealign_hip.so"))
    nsam, psize, ri = 64, 2.0, 26.0
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 30, 24.0, seed=8, device=DEV)
    a3df = synth.half_volume(vol)
    n = 2
    parts = synth.make_particles(ctx, a3df, n, seed=19, optics=optics, snr=2.0, ri=ri, shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    fsct = synth.default_fsct(nsam, psize).astype(np.float32)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    maxr1, maxr2, rbfact = 2, 14, 0.0
    rng = np.random.default_rng(15)
    start = np.zeros((n, 6), np.float32)
    start[:, 0:3] = parts.euler + np.radians(rng.normal(0, 1.5, (n, 3))).astype(np.float32)
    xm, ym, shx, shy, sx, sy, sig2 = 0.03, -0.02, 0.01, 0.015, 0.3, 0.3, 0.01
    f32 = np.float32
    prior = np.tile(np.array([f32(xm) - f32(shx), f32(ym) - f32(shy), sx, sy, sig2], f32), (n, 1))
    vol_h = np.ascontiguousarray(a3df.cpu().numpy().reshape(-1))
    spec_h = outc.cpu().numpy()
    ctff_h = ctff.cpu().numpy()
    f, i = O.byref_f, O.byref_i
    dummy = np.zeros(16, np.float32)
    idummy = np.zeros(16, np.int32)
    for mask in (np.ones(5, np.int32), np.array([1, 1, 1, 0, 0], np.int32)):
        got, _ = ctx.refine(nsam, a3df, outc, ctff, np.arange(n), np.ones(n, np.float32),
                            start.copy(), mask, maxr1, maxr2, rbfact,
                            torch.from_numpy(fsct).to(DEV), ring, prior)
        for k in range(n):
            x = [C.c_float(float(v)) for v in start[k, :5]]
            presa, npar, amag = C.c_float(0.0), C.c_int(-1), C.c_float(1.0)
            spec_k = np.ascontiguousarray(spec_h[k].reshape(-1))
            ctff_k = np.ascontiguousarray(ctff_h[k].reshape(-1))
            args = [i(nsam), i(1), C.byref(amag), i(maxr1), i(maxr2)] + [C.byref(v) for v in x] + [
                C.byref(presa), mask.ctypes, f(rbfact), C.byref(npar), vol_h.ctypes,
                dummy.ctypes, f(shx), f(shy), i(1)] + [f(0.0)] * 5 + [
                ctff_k.ctypes, f(1.0), f(float(ring[2])), f(float(ring[3])), f(float(ring[0])),
                f(float(ring[1])), f(0.0), f(0.0), f(0.0), f(0.0), dummy.ctypes, idummy.ctypes,
                dummy.ctypes, i(0), dummy.ctypes, dummy.ctypes, dummy.ctypes, i(0),
                dummy.ctypes, spec_k.ctypes, dummy.ctypes, dummy.ctypes, fsct.ctypes, i(0),
                f(xm), f(ym), f(sx), f(sy), f(sig2), C.c_char_p(b"C1 ")] + [f(0.0)] * 5 + [
                dummy.ctypes, dummy.ctypes, O.IDENTITY.ctypes, C.c_size_t(3)]
            assert len(args) == 63
            lib.prefine_(*args)
            mine = np.array([v.value for v in x] + [presa.value], np.float32)
            print(f"mask {mask.tolist()} item {k}: prefine_ {mine} batch {got[k]}")
            assert npar.value == int(mask.sum())
            assert amag.value == 1.0
            assert np.array_equal(mine.view(np.int32), got[k].view(np.int32))
