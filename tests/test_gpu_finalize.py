"""fr_finalize (the end of a refinement cycle, frealign_v9.f:790-950, in one
call) against the same chain composed from the separate entry points
(STORESHIFT, 3D c2r, CORRECT3D, MASK3D, 3D r2c, SHELTEST, FIND_FPART,
SHIFTVOL, the final map's c2r, CORRECT3D and MASK3D), each of which has its
own parity test against the oracle: the same kernels in the same order, so
the final map, the FSC table's numbers, FMASK and FPART agree bit for bit.
With and without the Wiener filter, and with a negative RI (the final map
left unmasked)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api  # noqa: E402

DEV = "cuda"


def _maps(nsam, seed):
    rng = np.random.default_rng(seed)
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    zz, yy, xx = np.meshgrid(np.fft.fftfreq(nsam), np.fft.fftfreq(nsam), np.fft.rfftfreq(nsam),
                             indexing="ij")
    env = np.exp(-((xx ** 2 + yy ** 2 + zz ** 2) * nsam ** 2) / 300.0).reshape(-1)
    sig = env * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    noise = lambda: 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))  # noqa: E731
    a = (sig * 40 + noise()).astype(np.complex64)
    d = (sig * 40 + noise()).astype(np.complex64)
    s = rng.uniform(20, 60, n).astype(np.float32)
    v = rng.uniform(20, 60, n).astype(np.float32)
    vols = api.HalfMaps.zeros(nsam)
    shp = (nsam, nsam, jc)
    for dst, src in ((vols.num[0], a), (vols.num[1], d), (vols.ctf2[0], s), (vols.ctf2[1], v)):
        dst.copy_(torch.from_numpy(src.reshape(shp)).to(DEV))
    vols.nksum.copy_(torch.tensor([4000, 4200], dtype=torch.int64))
    vols.ttd.copy_(torch.tensor([7400.0, 8300.0], dtype=torch.float64))
    return vols


def _composed(ctx, nsam, vols, iredun, psize, ri, ric, mw, dalt, ffilt, rlim):
    nk = float(vols.nksum.sum().item())
    ttd = vols.ttd.cpu().numpy()
    std, vtd = ttd[0] / nk * 0.1, ttd[1] / nk * 0.1
    lut = np.zeros(180, np.float32)
    O.oracle().or_calcsinc(lut, 180)
    b, c = ctx.storeshift(nsam, std, vtd, iredun, vols.num[0], vols.ctf2[0], vols.num[1],
                          vols.ctf2[1])
    fmask = 1.0
    for k, vol in enumerate((b, c)):
        ctx.fft3d(nsam, vol, False)
        ctx.correct3d(nsam, vol, 1, 1, lut)
        fm, _, _ = ctx.mask3d(nsam, vol, abs(ri), ric, 6.0)
        if k == 0:
            fmask = fm
        ctx.fft3d(nsam, vol, True)
    st = ctx.sheltest(nsam, b, c, vols.ctf2[0], vols.ctf2[1], fmask=fmask)
    fsc = st["fsc"].astype(np.float32)
    fpart, pssnr = ctx.find_fpart(nsam, std, vtd, vols.num[0], vols.num[1], vols.ctf2[0],
                                  vols.ctf2[1], iredun, abs(ri), ric, psize, mw, dalt, fsc,
                                  st["pssnr"])
    a = vols.num[0].clone()
    s = vols.ctf2[0].clone()
    ctx.shiftvol(nsam, std, vtd, rlim, iredun, a, vols.num[1], s, vols.ctf2[1], 0, pssnr,
                 fpart if ffilt else 0.0, fsc, fmask)
    ctx.fft3d(nsam, a, False)
    ctx.correct3d(nsam, a, 1, 1, lut)
    if ri > 0:
        ctx.mask3d(nsam, a, ri, ric, 6.0)
    final = a.view(torch.float32).reshape(nsam, nsam, nsam + 2)[..., :nsam]
    return dict(final=final, b=b, c=c, fsc=fsc, pssnr=pssnr, npix=st["npix"],
                npixt=st["npixt"], fmask=fmask, fpart=fpart)


@pytest.mark.parametrize("nsam,ri,mw,ffilt", [(64, 24.0, 0.0, True), (64, -24.0, 0.0, False),
                                              (48, 18.0, 150.0, True)])
def test_finalize_equals_composed_chain(nsam, ri, mw, ffilt):
    ctx = api.Context(0)
    kw = dict(iredun=1, psize=1.5, ri=ri, ric=0.0, mw=mw, dalt=0.81, ffilt=ffilt, rlim=0.45)
    ref = _composed(ctx, nsam, _maps(nsam, 5), **kw)
    got = ctx.finalize(nsam, _maps(nsam, 5), **kw)
    torch.cuda.synchronize()
    print(f"finalize nsam={nsam} ri={ri} mw={mw} ffilt={ffilt}: FMASK {got['fmask']:.4f} "
          f"FPART {got['fpart']:.4f}, {len(got['fsc'])} shells")
    assert got["fmask"] == ref["fmask"] and got["fpart"] == ref["fpart"]
    # these maps have no compact particle, so the FPART_CC walk may go past 1
    assert got["fpart"] > 0.0
    for k in ("fsc", "pssnr", "npix", "npixt"):
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])[:len(got[k])]), k
    for k in ("final", "b", "c"):
        x, y = got[k].contiguous(), ref[k].contiguous()
        assert torch.equal(x.view(torch.int32) if x.dtype == torch.float32 else x, y.view(
            torch.int32) if y.dtype == torch.float32 else y), k
