"""FIND_FPART on the GPU (fr_find_fpart) against the numpy restatement
(oracle/fpart_oracle.py, MASK3D from the C oracle) on half-maps of a blob
phantom with noise: the MW = 0 search at 32^3 and 64^3 ends at the same
FPART (both runs take the same steps, so within 2e-3: the thresholds and the
correlations differ only by summation order) and PSSNR is divided by it; the
MW closed form is exact.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import fpart_oracle as FO  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, synth  # noqa: E402


def _maps(nsam, seed):
    r = np.random.default_rng(seed)
    vol = synth.blob_phantom(nsam, 12, 0.3 * nsam, seed=seed, device="cuda")
    sig = synth.half_volume(vol).cpu().numpy()
    jc = nsam // 2 + 1
    sh = (nsam, nsam, jc)
    noise = lambda: (r.standard_normal(sh) + 1j * r.standard_normal(sh)).astype(np.complex64)  # noqa
    scale = np.abs(sig).mean()
    s = r.uniform(0.8, 1.2, sh).astype(np.float32) * np.float32(40.0)
    v = r.uniform(0.8, 1.2, sh).astype(np.float32) * np.float32(40.0)
    a = (sig * s + 3.0 * scale * noise()).astype(np.complex64)
    d = (sig * v + 3.0 * scale * noise()).astype(np.complex64)
    fsc = np.clip(np.linspace(1.0, -0.1, jc + 3), -1, 1).astype(np.float32)
    pssnr = np.linspace(8.0, 0.1, jc + 3).astype(np.float32)
    return a, d, s, v, fsc, pssnr


def _mask3d(L, nsam):
    def run(vol, ri, ric, halfw):
        flat = vol.reshape(-1)
        fm, av, sd = C.c_float(), C.c_float(), C.c_float()
        L.or32_mask3d(nsam, flat, ri, ric, halfw, C.byref(fm), C.byref(av), C.byref(sd))
        return fm.value, av.value, sd.value
    return run


@pytest.mark.parametrize("nsam", [32, 64])
def test_find_fpart_search_vs_oracle(nsam):
    ctx = api.Context(0)
    L = O.oracle()
    a, d, s, v, fsc, pssnr = _maps(nsam, 7 + nsam)
    std, vtd, iredun, ri, ric, psize = 0.8, 0.9, 1, 0.4 * nsam, 0.0, 2.0
    t = lambda x: torch.from_numpy(x).cuda()  # noqa: E731
    print(f"[fpart] nsam={nsam}", flush=True)
    fg, pg = ctx.find_fpart(nsam, std, vtd, t(a), t(d), t(s), t(v), iredun, ri, ric, psize, 0.0,
                            0.81, fsc, pssnr)
    fo, po = FO.find_fpart(nsam, std, vtd, a, d, s, v, iredun, ri, ric, psize, 0.0, 0.81, fsc,
                           pssnr, _mask3d(L, nsam))
    print(f"[fpart] nsam={nsam}: FPART gpu {fg:.5f} oracle {fo:.5f}")
    assert 0.0 < fg < 1.0
    assert abs(fg - fo) <= 2e-3
    jc = nsam // 2 + 1
    assert np.allclose(pg[:jc] * np.float32(fg), pssnr[:jc], rtol=1e-5)
    assert np.array_equal(pg[jc:], pssnr[jc:])


def test_find_fpart_mw_closed_form():
    ctx = api.Context(0)
    nsam, psize, mw, dalt = 64, 1.5, 300.0, 0.81
    a, d, s, v, fsc, pssnr = _maps(nsam, 3)
    t = lambda x: torch.from_numpy(x).cuda()  # noqa: E731
    fg, pg = ctx.find_fpart(nsam, 0.5, 0.5, t(a), t(d), t(s), t(v), 1, 20.0, 0.0, psize, mw, dalt,
                            fsc, pssnr)
    fo, po = FO.find_fpart(nsam, 0.5, 0.5, None, None, None, None, 1, 20.0, 0.0, psize, mw, dalt,
                           fsc, pssnr, None)
    assert fg == pytest.approx(fo, rel=1e-6)
    assert np.allclose(pg, po, rtol=1e-6)
