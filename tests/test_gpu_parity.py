"""GPU parity: the HIP path (through libfrealign_hip.so) against the CPU oracle
on identical seeded inputs.

Bars (written in each test):
  - Miller-index mapping / voxel set / NKSUM / grid indices: bit-exact
  - CC3M / PRESB scores: 1e-4 relative to the FP32 oracle
  - inserted numerator and CTF^2 volumes: 1e-5 relative L2 (summation order)
  - every result's error is also reported against the FP64 oracle (or64_*)
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api  # noqa: E402

rng = np.random.default_rng(20261202)
DEV = "cuda"


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, dtype=np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def half_volume(nsam, seed):
    r = np.random.default_rng(seed)
    v = r.standard_normal((nsam, nsam, nsam)).astype(np.float32)
    return np.ascontiguousarray(np.fft.rfftn(v).astype(np.complex64))


def gpar_for(nsam, psize=1.5):
    wl = 12.26 / np.sqrt(300000.0 + 0.9785 * 300000.0 ** 2 / 1e6)
    ri = 0.4 * nsam
    halfw = 6.0
    rih, rii = ri - halfw / 2, ri + halfw / 2
    w = 0.07
    return np.array([2.7e7, wl, np.sqrt(1 - w * w), w, wl / (psize * nsam), rih, halfw, rih * rih,
                     rii * rii, 0.0, 0.0], np.float32)


def pssnr_for(nsam, ri=None):
    jc = nsam // 2 + 1
    ri = 0.4 * nsam if ri is None else ri
    p = np.zeros(jc + 1, np.float32)
    for i in range(2, nsam // 2 + 1):  # card10.f:81-90
        p[i - 1] = np.exp(-((i - 1) / nsam * ri / nsam) ** 2)
    p[jc] = p[jc - 1]
    return p


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("nsam", [64, 96, 128, 160, 256])
def test_fft_r2c_c2r(ctx, nsam):
    nimg = 3
    img = rng.standard_normal((nimg, nsam, nsam)).astype(np.float32)
    buf = np.zeros((nimg, nsam, nsam + 2), np.float32)
    buf[:, :, :nsam] = img
    d = torch.from_numpy(buf).to(DEV)
    ctx.fft_r2c(nsam, d)
    got = d.cpu().numpy().view(np.complex64).reshape(nimg, nsam, nsam // 2 + 1)
    want = np.fft.rfft2(img.astype(np.float64))
    err = rel_l2(got, want)
    print(f"nsam={nsam} r2c rel L2 vs FP64 = {err:.2e}")
    assert err < 2e-6
    ctx.fft_c2r(nsam, d)
    back = d.cpu().numpy()[:, :, :nsam] / (nsam * nsam)
    assert rel_l2(back, img) < 2e-6


# ------------------------------------------------------------- CTFAPPLY
@pytest.mark.parametrize("nsam", [64, 96, 128, 256])
def test_ctfapply_vs_oracle(ctx, nsam):
    nb = 3
    jc = nsam // 2 + 1
    L = O.oracle()
    g = gpar_for(nsam)
    ps = pssnr_for(nsam)
    img = rng.standard_normal((nb, nsam, nsam)).astype(np.float32)
    buf = np.zeros((nb, nsam, nsam + 2), np.float32)
    buf[:, :, :nsam] = img
    ppar = np.zeros((nb, 8), np.float32)
    ppar[:, 0:2] = rng.uniform(-0.3, 0.3, (nb, 2))
    ppar[:, 2] = rng.uniform(10000, 30000, nb)
    ppar[:, 3] = ppar[:, 2] + rng.uniform(0, 1000, nb)
    ppar[:, 4] = rng.uniform(-1, 1, nb)
    ppar[:, 5] = 1.0
    imgs = torch.from_numpy(buf.copy()).to(DEV)
    ctff, outc, datc = ctx.ctfapply(nsam, imgs, torch.from_numpy(ppar).to(DEV),
                                    torch.from_numpy(ps).to(DEV), g)
    ctff, outc, datc = ctff.cpu().numpy(), outc.cpu().numpy(), datc.cpu().numpy()
    # the images hold their r2c spectra afterwards (lmain.f:459, in place),
    # the same bits as the FFT entry point gives
    ref = torch.from_numpy(buf.copy()).to(DEV)
    ctx.fft_r2c(nsam, ref)
    assert torch.equal(imgs.view(torch.int32), ref.view(torch.int32))
    # PWEIGHT's shell sums run in a fixed order: a rerun gives the same bits
    again = ctx.ctfapply(nsam, torch.from_numpy(buf.copy()).to(DEV),
                         torch.from_numpy(ppar).to(DEV), torch.from_numpy(ps).to(DEV), g)
    for x, y in zip((ctff, outc, datc), again):
        assert np.array_equal(x.view(np.uint32), y.cpu().numpy().view(np.uint32))
    for b in range(nb):
        spec = buf[b].copy().reshape(-1)
        L.or_rfft2_f(nsam, spec)
        specc = spec.view(np.complex64).copy()
        res = {}
        for pre, ct in (("or32_", np.complex64), ("or64_", np.complex128)):
            cf_ = np.zeros(2 * nsam * jc, ct)
            oc = np.zeros(nsam * jc, ct)
            dc = np.zeros(nsam * jc, ct)
            getattr(L, pre + "ctfapply")(nsam, specc, *[float(x) for x in ppar[b, :2]], float(g[0]),
                                         float(g[1]), float(g[2]), float(g[3]),
                                         *[float(x) for x in ppar[b, 2:5]], float(g[4]),
                                         float(ppar[b, 5]), float(g[5]), float(g[6]), float(g[7]),
                                         float(g[8]), ps, 1, cf_, oc, dc)
            res[pre] = (cf_, oc, dc)
        for k, name in enumerate(("ctff", "outc", "datc")):
            gpu = (ctff, outc, datc)[k][b].reshape(-1)
            e32 = rel_l2(gpu, res["or32_"][k])
            e64 = rel_l2(gpu, res["or64_"][k])
            print(f"nsam={nsam} particle {b} {name}: rel L2 vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
            assert e32 < 1e-5, name


@pytest.mark.parametrize("nsam", [64, 96, 256])
def test_ctfapply_from_spectrum_equals_prep(ctx, nsam):
    """CTFAPPLY from a particle's spectrum (the per-call ctfapply_, which the
    defocus refinement's path shares) against fr_prep_batch from the image, on
    the spectrum fr_prep_batch leaves: CTFF, OUTC and DATC agree bit for bit
    (the same transforms and point operations, whether the column pass
    starts from the image's row transforms or from the whole spectrum)."""
    from frealign_amd import _native as N
    nb = 2
    jc = nsam // 2 + 1
    g = gpar_for(nsam)
    ps = pssnr_for(nsam)
    buf = np.zeros((nb, nsam, nsam + 2), np.float32)
    buf[:, :, :nsam] = rng.standard_normal((nb, nsam, nsam))
    ppar = np.zeros((nb, 8), np.float32)
    ppar[:, 0:2] = rng.uniform(-0.3, 0.3, (nb, 2))
    ppar[:, 2] = rng.uniform(10000, 30000, nb)
    ppar[:, 3] = ppar[:, 2] + rng.uniform(0, 1000, nb)
    ppar[:, 4] = rng.uniform(-1, 1, nb)
    ppar[:, 5] = 1.0
    imgs = torch.from_numpy(buf).to(DEV)
    ctff, outc, datc = [t.cpu().numpy() for t in ctx.ctfapply(
        nsam, imgs, torch.from_numpy(ppar).to(DEV), torch.from_numpy(ps).to(DEV), g)]
    spec = imgs.cpu().numpy()
    lib = N.lib()
    fp = lambda v: C.byref(C.c_float(float(v)))  # noqa: E731
    for b in range(nb):
        sp = np.ascontiguousarray(spec[b])
        cf_ = np.zeros(4 * nsam * jc, np.float32)
        od = np.zeros(2 * nsam * jc, np.float32)
        dc = np.zeros(2 * nsam * jc, np.float32)
        scratch = np.zeros(4 * nsam * jc, np.float32)
        args = [C.byref(C.c_int(nsam)), sp.ctypes, fp(ppar[b, 0]), fp(ppar[b, 1]), fp(g[0]),
                fp(g[1]), fp(g[2]), fp(g[3]), fp(ppar[b, 2]), fp(ppar[b, 3]), fp(ppar[b, 4]),
                fp(g[4]), cf_.ctypes, od.ctypes, scratch.ctypes, fp(ppar[b, 5]), fp(g[5]),
                fp(g[6]), fp(g[7]), fp(g[8]), fp(0.0), scratch.ctypes, dc.ctypes,
                scratch.ctypes, fp(0.0), fp(0.0), fp(0.0), scratch.ctypes, fp(0.0),
                C.byref(C.c_int(0)), fp(g[9]), fp(g[10]), C.c_char_p(b"C1 "), ps.ctypes,
                scratch.ctypes, scratch.ctypes, C.byref(C.c_int(1)), scratch.ctypes,
                C.c_size_t(3)]
        lib.ctfapply_(*args)
        for name, got, want in (("ctff", cf_, ctff[b]), ("outc", od, outc[b]),
                                ("datc", dc, datc[b])):
            assert np.array_equal(got.view(np.uint32), np.ascontiguousarray(want).reshape(
                -1).view(np.uint32)), f"nsam={nsam} particle {b} {name}"


# ------------------------------------------------------------- PEXTRACT
@pytest.mark.parametrize("nsam", [32, 64])
def test_pextract_vs_oracle(ctx, nsam):
    jc = nsam // 2 + 1
    L = O.oracle()
    vol = half_volume(nsam, 41)
    n = 6
    eul = rng.uniform(-np.pi, np.pi, (n, 3)).astype(np.float32)
    eul[0] = [0, 0, 0]
    eul[1] = np.float32([np.pi / 2, np.pi / 2, 0])
    ctff = (rng.standard_normal((n, 2, nsam, jc)) + 1j * rng.standard_normal((n, 2, nsam, jc))
            ).astype(np.complex64)
    dm = api.rotation(eul, amag=0.0)
    out = ctx.pextract(nsam, torch.from_numpy(vol).to(DEV), torch.from_numpy(dm).to(DEV),
                       torch.from_numpy(ctff).to(DEV)).cpu().numpy()
    for k in range(n):
        w32 = np.zeros(nsam * jc, np.complex64)
        w64 = np.zeros(nsam * jc, np.complex128)
        L.or32_pextract(nsam, w32, vol.reshape(-1), *[float(x) for x in eul[k]],
                        ctff[k].reshape(-1), O.IDENTITY)
        L.or64_pextract(nsam, w64, vol.reshape(-1), *[float(x) for x in eul[k]],
                        ctff[k].reshape(-1).astype(np.complex128), O.IDENTITY)
        g = out[k].reshape(-1)
        # same support and same voxels: a wrong corner would be an O(1) error
        assert np.array_equal(g != 0, w32 != 0)
        e32, e64 = rel_l2(g, w32), rel_l2(g, w64)
        worst = np.max(np.abs(g - w32) / (np.abs(w32) + 1e-3 * np.abs(w32).max()))
        print(f"nsam={nsam} item {k}: rel L2 vs FP32 {e32:.2e} vs FP64 {e64:.2e} worst {worst:.1e}")
        assert e32 < 1e-5 and worst < 1e-4


# ----------------------------------------------------------------- CC3M
# (64, 24), (128, 32), (256, 32), (160, 24): band-pruned DFT path (MAXR2 <= 32,
# NSAM % 32 == 0), with AMAG = 1 (ring mean from the precomputed table) or
# jittered (ring pass on the matrix cores); (96, 40): the full-FFT path
# (MAXR2 > 32, NSAM = 2^5 x 3)
@pytest.mark.parametrize("nsam,maxr2,amag_jitter", [(64, 24, 0.0), (128, 32, 0.02), (256, 32, 0.02),
                                                    (160, 24, 0.0), (96, 40, 0.02)])
def test_cc3m_vs_oracle(ctx, nsam, maxr2, amag_jitter):
    jc = nsam // 2 + 1
    L = O.oracle()
    vol = half_volume(nsam, 43)
    npart, nitem = 2, 8
    spec = (rng.standard_normal((npart, nsam, jc)) + 1j * rng.standard_normal((npart, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.5 + 0.1 * rng.standard_normal((npart, 2, nsam, jc))).astype(np.complex64)
    pidx = rng.integers(0, npart, nitem).astype(np.int32)
    eul = rng.uniform(-np.pi, np.pi, (nitem, 3)).astype(np.float32)
    shift = rng.uniform(-0.2, 0.2, (nitem, 2)).astype(np.float32)
    amag = (1.0 + amag_jitter * rng.uniform(-1, 1, nitem)).astype(np.float32)
    dm = np.concatenate([api.rotation(eul[k:k + 1], amag=float(amag[k])) for k in range(nitem)])
    maxr1, rbfact = 2, 0.001
    fsct = np.zeros(nsam // 2 + 1, np.float32)
    fsct[: nsam // 4] = 1.0
    ri = 0.4 * nsam
    ring = np.array([(ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0], np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    got = ctx.cc3m(nsam, t(vol), t(spec), t(ctff), t(pidx), t(dm), t(shift), t(amag), maxr1,
                   maxr2, rbfact, t(fsct), ring).cpu().numpy()
    for k in range(nitem):
        p = pidx[k]
        args = [float(x) for x in eul[k]] + [float(shift[k, 0]), float(shift[k, 1]), rbfact]
        s32 = L.or32_cc3m(nsam, float(amag[k]), spec[p].reshape(-1), vol.reshape(-1), maxr1, maxr2, *args,
                          fsct, ctff[p].reshape(-1), *[float(x) for x in ring], O.IDENTITY)
        s64 = L.or64_cc3m(nsam, float(amag[k]), spec[p].reshape(-1).astype(np.complex128), vol.reshape(-1),
                          maxr1, maxr2, *args, fsct, ctff[p].reshape(-1).astype(np.complex128),
                          *[float(x) for x in ring], O.IDENTITY)
        r32 = abs(got[k] - s32) / abs(s32)
        print(f"item {k}: gpu {got[k]:.7f} or32 {s32:.7f} or64 {s64:.7f} rel32 {r32:.1e} "
              f"rel64 {abs(got[k] - s64) / abs(s64):.1e}")
        # 1e-4 relative; a score near zero is a difference of shell terms of
        # order 1e-2, so its error is bounded by 1e-4 of 1e-3 instead
        assert abs(got[k] - s32) <= 1e-4 * max(abs(s32), 1e-3)


# ----------------------------------------------------------------- PRESB
@pytest.mark.parametrize("nsam", [64, 96])
def test_presb_vs_oracle(ctx, nsam):
    jc = nsam // 2 + 1
    L = O.oracle()
    vol = half_volume(nsam, 47)
    npart, nitem = 2, 6
    spec = (rng.standard_normal((npart, nsam, jc)) + 1j * rng.standard_normal((npart, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.5 + 0.1 * rng.standard_normal((npart, 2, nsam, jc))).astype(np.complex64)
    pidx = rng.integers(0, npart, nitem).astype(np.int32)
    eul = rng.uniform(-np.pi, np.pi, (nitem, 3)).astype(np.float32)
    shift = rng.uniform(-0.2, 0.2, (nitem, 2)).astype(np.float32)
    amag = (1.0 + 0.02 * rng.uniform(-1, 1, nitem)).astype(np.float32)
    dm = np.concatenate([api.rotation(eul[k:k + 1], amag=float(amag[k])) for k in range(nitem)])
    ri = 0.4 * nsam
    ring = np.array([(ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0], np.float32)
    # sentinel: entries PRESB does not write must come back unchanged
    rbin0 = np.full((nitem, nsam), 7.0, np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    got = ctx.presb(nsam, t(vol), t(spec), t(ctff), t(pidx), t(dm), t(shift), t(amag), ring,
                    rbin=t(rbin0)).cpu().numpy()
    for k in range(nitem):
        p = pidx[k]
        args = [float(x) for x in eul[k]] + [float(shift[k, 0]), float(shift[k, 1])]
        w32 = rbin0[k].copy()
        L.or32_presb(nsam, float(amag[k]), spec[p].reshape(-1), vol.reshape(-1), *args, w32,
                     ctff[p].reshape(-1), *[float(x) for x in ring], O.IDENTITY)
        w64 = rbin0[k].astype(np.float64)
        L.or64_presb(nsam, float(amag[k]), spec[p].reshape(-1).astype(np.complex128),
                     vol.reshape(-1), *args, w64, ctff[p].reshape(-1).astype(np.complex128),
                     *[float(x) for x in ring], O.IDENTITY)
        # the written set (count slots = 1) and the untouched sentinels are exact
        assert np.array_equal(got[k][1::2], w32[1::2])
        assert np.array_equal(got[k][0::2] == 7.0, w32[0::2] == 7.0)
        e32 = np.max(np.abs(got[k][0::2] - w32[0::2]))
        e64 = np.max(np.abs(got[k][0::2] - w64[0::2]))
        print(f"nsam={nsam} item {k}: max |dRBIN| vs or32 {e32:.1e} vs or64 {e64:.1e}")
        # shell correlations lie in [-1, 1]; FP32 shell sums over <= 2*pi*r points
        assert e32 < 2e-5


# ------------------------------------------------------------ insertion
def _oracle_insert(L, pre, nsam, datc, ctff, recb, half, sym, pbc, boff, psize):
    jc = nsam // 2 + 1
    ct, rt = (np.complex64, np.float32) if pre == "or32_" else (np.complex128, np.float64)
    num = [np.zeros(nsam * nsam * jc, ct) for _ in range(2)]
    c2 = [np.zeros(nsam * nsam * jc, rt) for _ in range(2)]
    ttd = [C.c_double(0.0), C.c_double(0.0)]
    nk = [C.c_int64(0), C.c_int64(0)]
    nsym, symop, jsym, _ = sym
    for b in range(recb.shape[0]):
        r = recb[b]
        h = int(half[b])
        on = np.float32(r[6]) / np.float32(r[7])
        getattr(L, pre + "pinsert")(nsam, float(r[9]), float(r[5]), datc[b].reshape(-1).astype(ct),
                                    ctff[b].reshape(-1).astype(ct), num[h], c2[h], C.byref(ttd[h]),
                                    float(r[0]), float(r[1]), float(r[2]), float(r[3]), float(r[4]),
                                    float(r[8]), pbc, boff, nsym, jsym, symop, C.byref(nk[h]),
                                    psize, float(on), O.IDENTITY, 0)
    return num, c2, [t.value for t in ttd], [k.value for k in nk]


@pytest.mark.parametrize("nsam,asym,nasym,aligned", [(32, "C", 1, False), (64, "C", 1, False),
                                                     (64, "D", 7, False), (64, "C", 1, True),
                                                     (40, "D", 4, True)])
def test_insert_vs_oracle(ctx, nsam, asym, nasym, aligned):
    """PINSERT for a batch against the oracle: NKSUM and the voxel set exact,
    the sums to FP32 rounding.  `aligned`: planes along the axes and on the
    diagonals (in-plane axes with zero components) at magnifications other
    than 1, so that points fall on the voxel grid and the plane geometry is
    not orthonormal."""
    jc = nsam // 2 + 1
    L = O.oracle()
    nb = 6
    datc = (rng.standard_normal((nb, nsam, jc)) + 1j * rng.standard_normal((nb, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.6 + 0.2 * (rng.standard_normal((nb, 2, nsam, jc))
                         + 1j * rng.standard_normal((nb, 2, nsam, jc)))).astype(np.complex64)
    recb = np.zeros((nb, 10), np.float32)
    recb[:, 0:3] = rng.uniform(-np.pi, np.pi, (nb, 3))
    recb[0, 0:3] = [0.0, 0.0, 0.0]  # integer-aligned plane: exercises the INT() rules
    recb[:, 3:5] = rng.uniform(-0.3, 0.3, (nb, 2))
    recb[:, 5] = 1.0
    recb[:, 6] = 1.0
    recb[:, 7] = rng.uniform(0.5, 2.0, nb)
    recb[:, 8] = rng.uniform(0.2, 0.8, nb)
    recb[:, 9] = 0.5 if nsam < 64 else 0.45
    if aligned:
        q = np.float32(np.pi / 2)
        recb[:, 0:3] = [[0, 0, 0], [q, 0, 0], [0, q, 0], [q, q, 0], [0, q, q], [q / 2, q, 0]]
        recb[:, 5] = [1.0, 0.97, 1.03, 0.95, 1.05, 1.0]
    half = (np.arange(nb) % 2).astype(np.int32)
    sym = api.point_group(asym, nasym)
    pbc, boff, psize = 5.0, 40.0, 1.5
    vols = api.HalfMaps.zeros(nsam)
    ctx.insert(nsam, torch.from_numpy(datc).to(DEV), torch.from_numpy(ctff).to(DEV), recb, half,
               sym, pbc, boff, psize, vols)
    torch.cuda.synchronize()
    n32, c32, t32, k32 = _oracle_insert(L, "or32_", nsam, datc, ctff, recb, half, sym, pbc, boff,
                                        psize)
    n64, c64, t64, _ = _oracle_insert(L, "or64_", nsam, datc, ctff, recb, half, sym, pbc, boff,
                                      psize)
    nk = vols.nksum.cpu().numpy()
    td = vols.ttd.cpu().numpy()
    for h in range(2):
        gnum = vols.num[h].cpu().numpy().reshape(-1)
        gc2 = vols.ctf2[h].cpu().numpy().reshape(-1)
        assert nk[h] == k32[h], "NKSUM must be bit-exact"
        assert np.array_equal(gc2 != 0, c32[h] != 0), "voxel set must be identical"
        e_num, e_c2 = rel_l2(gnum, n32[h]), rel_l2(gc2, c32[h])
        print(f"{asym}{nasym} nsam={nsam} half {h}: NKSUM {nk[h]}; numerator rel L2 {e_num:.2e} "
              f"(FP64 {rel_l2(gnum, n64[h]):.2e}); CTF^2 {e_c2:.2e} (FP64 {rel_l2(gc2, c64[h]):.2e});"
              f" TTD rel {abs(td[h] - t32[h]) / t32[h]:.1e}")
        assert e_num < 1e-5 and e_c2 < 1e-5
        # TTD sums the same FP32 CTF^2 terms in FP64: its bar follows CTF^2's
        assert abs(td[h] - t32[h]) <= 1e-6 * abs(t32[h])


@pytest.mark.parametrize("nsam,asym,nasym", [(32, "C", 1), (64, "D", 7), (128, "C", 1)])
def test_insert_se_vs_oracle(ctx, nsam, asym, nasym):
    """The _SE insertion (NNSTAT = 1, pinsert_trilin_se.f:186-194) against
    the oracle's PINSERT_SE, itself bit exact to the reference routine: the
    per-voxel KSUM exact, VSUM, ASUM, PSUM within 1e-5 relative L2 (the
    voxel's terms summed in another order), and the half-maps, NKSUM and
    TTD bit-identical to the insertion without statistics."""
    jc = nsam // 2 + 1
    L = O.oracle()
    nb = 5
    nv = nsam * nsam * jc
    datc = (rng.standard_normal((nb, nsam, jc)) + 1j * rng.standard_normal((nb, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.6 + 0.2 * (rng.standard_normal((nb, 2, nsam, jc))
                         + 1j * rng.standard_normal((nb, 2, nsam, jc)))).astype(np.complex64)
    ref = (rng.standard_normal(nv) + 1j * rng.standard_normal(nv)).astype(np.complex64)
    ref[::7] = 0
    recb = np.zeros((nb, 10), np.float32)
    recb[:, 0:3] = rng.uniform(-np.pi, np.pi, (nb, 3))
    recb[0, 0:3] = [0.0, 0.0, 0.0]
    recb[:, 3:5] = rng.uniform(-0.3, 0.3, (nb, 2))
    recb[:, 5] = 1.0
    recb[:, 6] = 1.0
    recb[:, 7] = rng.uniform(0.5, 2.0, nb)
    recb[:, 8] = rng.uniform(0.2, 0.8, nb)
    recb[:, 9] = 0.45
    half = (np.arange(nb) % 2).astype(np.int32)
    sym = api.point_group(asym, nasym)
    pbc, boff, psize = 5.0, 40.0, 1.5
    print(f"[insert_se] nsam={nsam} {asym}{nasym}", flush=True)
    dd, dc = torch.from_numpy(datc).to(DEV), torch.from_numpy(ctff).to(DEV)
    vols = api.HalfMaps.zeros(nsam)
    stats = api.VoxelStats.zeros(nsam, torch.from_numpy(ref).to(DEV).reshape(nsam, nsam, jc))
    ctx.insert(nsam, dd, dc, recb, half, sym, pbc, boff, psize, vols, stats=stats)
    plain = api.HalfMaps.zeros(nsam)
    ctx.insert(nsam, dd, dc, recb, half, sym, pbc, boff, psize, plain)
    torch.cuda.synchronize()
    for h in range(2):
        assert np.array_equal(vols.num[h].cpu().numpy().view(np.uint32),
                              plain.num[h].cpu().numpy().view(np.uint32))
        assert np.array_equal(vols.ctf2[h].cpu().numpy().view(np.uint32),
                              plain.ctf2[h].cpu().numpy().view(np.uint32))
    assert np.array_equal(vols.nksum.cpu().numpy(), plain.nksum.cpu().numpy())
    assert np.array_equal(vols.ttd.cpu().numpy().view(np.uint64),
                          plain.ttd.cpu().numpy().view(np.uint64))
    # the oracle, particle by particle in order, both halves into one set of statistics
    nsym, symop, jsym, _ = sym
    ks = np.zeros(nv, np.int32)
    vs, ps, as_ = (np.zeros(nv, np.float32) for _ in range(3))
    for b in range(nb):
        r = recb[b]
        on = np.float32(r[6]) / np.float32(r[7])
        L.or32_pinsert_se(nsam, float(r[9]), float(r[5]), datc[b].reshape(-1),
                          ctff[b].reshape(-1), np.zeros(nv, np.complex64),
                          np.zeros(nv, np.float32), C.byref(C.c_double(0.0)), float(r[0]),
                          float(r[1]), float(r[2]), float(r[3]), float(r[4]), float(r[8]), pbc,
                          boff, nsym, jsym, symop, C.byref(C.c_int64(0)), psize, float(on),
                          O.IDENTITY, 0, ref, ks, vs, ps, as_)
    gk = stats.ksum.cpu().numpy().reshape(-1)
    assert np.array_equal(gk, ks), "KSUM must be exact"
    assert int(gk.sum()) == int(vols.nksum.sum().item())
    errs = [rel_l2(g.cpu().numpy().reshape(-1), o) for g, o in
            ((stats.vsum, vs), (stats.psum, ps), (stats.asum, as_))]
    print(f"[insert_se] KSUM total {gk.sum()}; rel L2 VSUM {errs[0]:.1e} PSUM {errs[1]:.1e} "
          f"ASUM {errs[2]:.1e}")
    assert max(errs) < 1e-5


def test_insert_deterministic_support(ctx):
    """The insertion sums in a fixed order: two runs give the same bits
    (volumes, NKSUM and TTD)."""
    nsam, nb = 64, 4
    jc = nsam // 2 + 1
    datc = (rng.standard_normal((nb, nsam, jc)) + 1j * rng.standard_normal((nb, nsam, jc))
            ).astype(np.complex64)
    ctff = np.ones((nb, 2, nsam, jc), np.complex64) * 0.5
    recb = np.zeros((nb, 10), np.float32)
    recb[:, 0:3] = rng.uniform(-np.pi, np.pi, (nb, 3))
    recb[:, 5:8] = 1.0
    recb[:, 8] = 0.5
    recb[:, 9] = 0.5
    half = np.zeros(nb, np.int32)
    sym = api.point_group("C", 1)
    outs = []
    for _ in range(2):
        v = api.HalfMaps.zeros(nsam)
        ctx.insert(nsam, torch.from_numpy(datc).to(DEV), torch.from_numpy(ctff).to(DEV), recb,
                   half, sym, 5.0, 40.0, 1.5, v)
        outs.append((v.num[0].cpu().numpy(), v.nksum.cpu().numpy(), v.ctf2[0].cpu().numpy(),
                     v.ttd.cpu().numpy()))
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][2].view(np.uint32), outs[1][2].view(np.uint32))
    assert np.array_equal(outs[0][3].view(np.uint64), outs[1][3].view(np.uint64))


# ------------------------------------------------------------ statistics
@pytest.mark.parametrize("nsam,istp", [(32, 1), (64, 1), (64, 3)])
def test_sheltest_vs_oracle(ctx, nsam, istp):
    L = O.oracle()
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (a + 0.8 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    t = lambda x: torch.from_numpy(x).to(DEV)  # noqa: E731
    got = ctx.sheltest(nsam, t(a), t(d), t(s), t(v), istp=istp)
    res = {}
    for pre, ct, rt in (("or32_", np.complex64, np.float32), ("or64_", np.complex128, np.float64)):
        npo, npto = np.zeros(jc + 2, np.int32), np.zeros(jc + 2, np.int32)
        dfo, pso = np.zeros(jc + 2), np.zeros(jc + 2, rt)
        nst = getattr(L, pre + "sheltest")(nsam, istp, a.astype(ct), d.astype(ct), s.astype(rt),
                                          v.astype(rt), 1.0, npo, npto, dfo, pso)
        res[pre] = (nst, npo, npto, dfo, pso)
    nst, npo, npto, dfo, pso = res["or32_"]
    assert len(got["fsc"]) == nst
    assert np.array_equal(got["npix"], npo[:nst]) and np.array_equal(got["npixt"], npto[:nst])
    # FP32 products accumulated in FP64, as sheltest.f does: only the summation
    # order differs from the FP32 restatement
    err32 = np.max(np.abs(got["fsc"] - dfo[:nst]))
    err64 = np.max(np.abs(got["fsc"] - res["or64_"][3][:nst]))
    print(f"nsam={nsam} istp={istp}: max |FSC_gpu - FSC_fp32| = {err32:.1e}, vs FP64 {err64:.1e}")
    assert err32 < 1e-6 and err64 < 1e-6
    # PSSNR = 2|F|/(1-|F|) * npix / sum(S+V): its relative error is the FSC's
    # absolute error times 1/(|F|(1-|F|)) (large where |F| nears 1), plus the
    # FP32 summation order of S+V
    f = np.abs(dfo[:nst])
    cond = 1.0 / np.maximum(f * (1.0 - f), 1e-12)
    bound = np.abs(pso[:nst]) * (1e-5 + 2e-6 * cond)
    assert np.all(np.abs(got["pssnr"] - pso[:nst]) <= bound), (got["pssnr"], pso[:nst])


def test_storeshift_vs_oracle(ctx):
    L = O.oracle()
    nsam = 32
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    t = lambda x: torch.from_numpy(x).to(DEV)  # noqa: E731
    b, c = ctx.storeshift(nsam, 0.37, 0.41, 14, t(a), t(s), t(d), t(v))
    bo, co = np.zeros(n, np.complex64), np.zeros(n, np.complex64)
    L.or32_storeshift(nsam, 0.37, 0.41, 14, a, s, d, v, bo, co)
    assert rel_l2(b.cpu().numpy(), bo) < 1e-6 and rel_l2(c.cpu().numpy(), co) < 1e-6


# ------------------------------------------------------------------ SIGMA
@pytest.mark.parametrize("nsam,signset", [(32, 0.0), (64, 0.0), (64, 0.01)])
def test_sigma_vs_oracle(ctx, nsam, signset):
    """fr_sigma_batch against the FP32 oracle (sigma.f restated): SIGMA2N and
    SIGMA2 within 1e-3 relative (SIGMA2 carries the cancellation SUM2 -
    A^2 SUM1 of FP32 images), LOGPC within 1e-4 relative; errors also printed
    against the FP64 oracle."""
    L = O.oracle()
    jc = nsam // 2 + 1
    vol = half_volume(nsam, 53)
    npart, nitem = 2, 5
    datc = (rng.standard_normal((npart, nsam, jc)) + 1j * rng.standard_normal((npart, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.5 + 0.2 * rng.standard_normal((npart, 2, nsam, jc)) +
            0.05j * rng.standard_normal((npart, 2, nsam, jc))).astype(np.complex64)
    pidx = rng.integers(0, npart, nitem).astype(np.int32)
    par = np.zeros((nitem, 5), np.float32)
    par[:, 0:3] = rng.uniform(-np.pi, np.pi, (nitem, 3))
    par[:, 3:5] = rng.uniform(-0.2, 0.2, (nitem, 2))
    amag = (1.0 + 0.02 * rng.uniform(-1, 1, nitem)).astype(np.float32)
    dmask = np.array([nsam / 2 + 3, nsam / 2 - 2, nsam / 2 + 1, 0.35 * nsam], np.float32)
    # SIGP near the noise sigma so that the likelihood count N1*SIGMA2OLD/SIGP^2
    # is of the order of the disc (SIGNSET 0.01 lands in the accepted band)
    rclas, maxr1, sigp = 0.3, 1, 0.01
    ss = np.full(nitem, signset, np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    got = ctx.sigma(nsam, t(vol), t(datc), t(ctff), pidx, par, amag, maxr1, rclas, dmask, sigp,
                    ss if signset else None)
    for k in range(nitem):
        p = pidx[k]
        res = {}
        for pre, ct, rt in (("or32_", np.complex64, np.float32), ("or64_", np.complex128, np.float64)):
            out = np.zeros(3, rt)
            getattr(L, pre + "sigma")(nsam, float(amag[k]), vol.reshape(-1),
                                      datc[p].reshape(-1).astype(ct), *[float(x) for x in par[k]],
                                      ctff[p].reshape(-1).astype(ct), maxr1, rclas, float(ss[k]),
                                      dmask, sigp, O.IDENTITY, out)
            res[pre] = out
        r32 = np.abs(got[k] - res["or32_"]) / np.abs(res["or32_"])
        r64 = np.abs(got[k] - res["or64_"]) / np.abs(res["or64_"])
        print(f"nsam {nsam} item {k}: gpu {got[k]} or32 {res['or32_']} rel32 {r32} rel64 {r64}")
        assert r32[0] < 1e-3 and r32[1] < 1e-3 and r32[2] < 1e-4


# ------------------------------------------------- volume finalisation
@pytest.mark.parametrize("ifsc,imp,nvol1", [(0, 4, 2), (1, 3, 3), (0, 1, 1)])
def test_combine_arrays_vs_oracle(ctx, ifsc, imp, nvol1):
    """COMBINE_ARRAYS: the copies add in copy order, so the sums are the
    oracle's bit for bit (the oracle is bit-exact to the reference's
    combine_arrays.f); the counters as combine_arrays.f:53-67."""
    L = O.oracle()
    nsam, nn1 = 64, 72
    nn3 = nn1 * nn1 * (nn1 // 2 + 1)
    nvol2 = nvol1
    a = (rng.standard_normal(nn3 * nvol1) + 1j * rng.standard_normal(nn3 * nvol1)).astype(np.complex64)
    d = (rng.standard_normal(nn3 * nvol2) + 1j * rng.standard_normal(nn3 * nvol2)).astype(np.complex64)
    s = rng.uniform(0, 3, nn3 * nvol1).astype(np.float32)
    v = rng.uniform(0, 3, nn3 * nvol2).astype(np.float32)
    m = max(imp, nvol1)
    nksum1 = rng.integers(0, 10 ** 9, m).astype(np.int64)
    ttd1, ttd2 = rng.uniform(0, 10, m), rng.uniform(0, 10, m)
    ga, gs, gd, gv = (torch.from_numpy(x.copy()).to(DEV) for x in (a, s, d, v))
    nk, sd, vd = ctx.combine_arrays(nsam, nn3, nvol1, nvol2, ifsc, ga, gs, gd, gv, imp, nksum1,
                                    ttd1, ttd2, 7, 0.5, 0.25)
    torch.cuda.synchronize()
    L.or32_combine_arrays(nsam, nn3, nvol1, nvol2, ifsc, a, s, d, v)
    onk, osd, ovd = C.c_int64(7), C.c_double(0.5), C.c_double(0.25)
    L.or_combine_counters(ifsc, imp, nvol1, nksum1.ctypes.data_as(C.POINTER(C.c_int64)), ttd1,
                          ttd2, C.byref(onk), C.byref(osd), C.byref(ovd))
    for g, o in ((ga, a), (gs, s), (gd, d), (gv, v)):
        assert np.array_equal(g.cpu().numpy().view(np.int32), o.view(np.int32))
    assert (nk, sd, vd) == (onk.value, osd.value, ovd.value)


@pytest.mark.parametrize("nsam,fpart,ifsc", [(64, 0.0, 0), (64, 0.3, 0), (128, 0.3, 1)])
def test_shiftvol_vs_oracle(ctx, nsam, fpart, ifsc):
    """SHIFTVOL (plain and Wiener-filtered final map): 1e-6 relative L2 against
    the FP32 oracle (bit-exact to the reference's shiftvol.f); the support
    (zeros outside RLIM and on empty shells) identical."""
    L = O.oracle()
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    s = rng.uniform(0, 3, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    s[::97] = 0.0
    v[::97] = 0.0
    nsh = nsam
    pssnr = rng.uniform(0.1, 2.0, nsh).astype(np.float32)
    pssnr[5] = 0.0
    fsc = rng.uniform(-0.2, 0.99, nsh).astype(np.float32)
    fsc[3], fsc[4] = 1.0, -1.0
    rlim, iredun, fmask = 0.4, 14, 0.3
    ga, gd, gs, gv = (torch.from_numpy(x.copy()).to(DEV) for x in (a, d, s, v))
    std_g = ctx.shiftvol(nsam, 0.37, 0.41, rlim, iredun, ga, gd, gs, gv, ifsc, pssnr, fpart, fsc,
                         fmask)
    torch.cuda.synchronize()
    std_o = C.c_double(0.37)
    L.or32_shiftvol(nsam, C.byref(std_o), 0.41, rlim, iredun, a, d, s, v, ifsc, pssnr, fpart, fsc,
                    fmask)
    assert std_g == std_o.value
    got = ga.cpu().numpy()
    assert np.array_equal(got != 0, a != 0)
    e = rel_l2(got, a)
    exact = np.mean(got.view(np.int32) == a.view(np.int32))
    print(f"SHIFTVOL nsam={nsam} FPART={fpart} IFSC={ifsc}: rel L2 {e:.1e}, bit-equal {exact:.4f}")
    assert e < 1e-6
    assert np.array_equal(gs.cpu().numpy().view(np.int32), s.view(np.int32))


@pytest.mark.parametrize("nsam,interp,ipad", [(64, 1, 1), (64, 0, 1), (128, 1, 2)])
def test_correct3d_vs_oracle(ctx, nsam, interp, ipad):
    """CORRECT3D: 1e-6 relative L2 against the FP32 oracle (bit-exact to the
    reference's correct3d.f and boxft_lut.f); padding columns untouched."""
    L = O.oracle()
    lut = np.zeros(180, np.float32)
    L.or_calcsinc(lut, 180)
    vol = rng.standard_normal(nsam * nsam * (nsam + 2)).astype(np.float32)
    g = torch.from_numpy(vol.copy()).to(DEV)
    ctx.correct3d(nsam, g, interp, ipad, lut)
    torch.cuda.synchronize()
    want = vol.copy()
    L.or32_correct3d(nsam, want, interp, ipad, lut)
    got = g.cpu().numpy()
    e = rel_l2(got, want)
    exact = np.mean(got.view(np.int32) == want.view(np.int32))
    print(f"CORRECT3D nsam={nsam} INTERP={interp} IPAD={ipad}: rel L2 {e:.1e}, bit-equal {exact:.4f}")
    assert e < 1e-6
    pad = got.reshape(nsam, nsam, nsam + 2)[:, :, nsam:]
    assert np.array_equal(pad, vol.reshape(nsam, nsam, nsam + 2)[:, :, nsam:])


@pytest.mark.parametrize("nsam,rio,ric,halfw", [(64, 24.0, 0.0, 6.0), (64, 40.0, 6.0, 6.0),
                                                (128, 50.0, 0.0, 6.0)])
def test_mask3d_vs_oracle(ctx, nsam, rio, ric, halfw):
    """MASK3D: the masked volume 1e-6 relative L2 against the FP32 oracle
    (bit-exact to the reference's mask3d.f); FMASK, AVE, STD to 1e-6 relative
    (FP64 sums in a different order); padding columns untouched."""
    L = O.oracle()
    vol = (rng.standard_normal(nsam * nsam * (nsam + 2)) + 0.3).astype(np.float32)
    g = torch.from_numpy(vol.copy()).to(DEV)
    fm, av, sd = ctx.mask3d(nsam, g, rio, ric, halfw)
    torch.cuda.synchronize()
    want = vol.copy()
    fo, ao, so = C.c_float(), C.c_float(), C.c_float()
    L.or32_mask3d(nsam, want, rio, ric, halfw, C.byref(fo), C.byref(ao), C.byref(so))
    got = g.cpu().numpy()
    e = rel_l2(got, want)
    exact = np.mean(got.view(np.int32) == want.view(np.int32))
    print(f"MASK3D nsam={nsam} RIO={rio} RIC={ric}: rel L2 {e:.1e}, bit-equal {exact:.4f}; "
          f"FMASK {fm} / {fo.value}, AVE {av} / {ao.value}, STD {sd} / {so.value}")
    assert e < 1e-6
    for x, y in ((fm, fo.value), (av, ao.value), (sd, so.value)):
        assert abs(x - y) <= 1e-6 * abs(y)
    pad = got.reshape(nsam, nsam, nsam + 2)[:, :, nsam:]
    assert np.array_equal(pad, vol.reshape(nsam, nsam, nsam + 2)[:, :, nsam:])


@pytest.mark.parametrize("nsam", [64, 96, 128])
def test_fft3d_r2c_c2r(ctx, nsam):
    """The 3D r2c against numpy's FP64 rfftn (2e-6 relative L2, as the 2D
    transforms), and c2r(r2c(x)) / NSAM^3 back to x."""
    vol = rng.standard_normal((nsam, nsam, nsam)).astype(np.float32)
    buf = np.zeros((nsam, nsam, nsam + 2), np.float32)
    buf[:, :, :nsam] = vol
    d = torch.from_numpy(buf).to(DEV)
    ctx.fft3d(nsam, d, True)
    torch.cuda.synchronize()
    got = d.cpu().numpy().view(np.complex64).reshape(nsam, nsam, nsam // 2 + 1)
    want = np.fft.rfftn(vol.astype(np.float64))
    err = rel_l2(got, want)
    print(f"nsam={nsam} 3D r2c rel L2 vs FP64 = {err:.2e}")
    assert err < 2e-6
    ctx.fft3d(nsam, d, False)
    torch.cuda.synchronize()
    back = d.cpu().numpy()[:, :, :nsam] / float(nsam) ** 3
    assert rel_l2(back, vol) < 2e-6


def test_finalisation_chain_fsc(ctx):
    """The half-map finalisation of frealign_v9.f:809-869 end to end on the
    device -- STORESHIFT, 3D c2r, CORRECT3D, MASK3D, 3D r2c, SHELTEST -- against
    the same chain through the FP32 oracle (numpy's FP64 transforms standing
    in for FFTW).  The per-shell FSC agrees to 1e-5 and so does FSC = 0.143's
    crossing shell."""
    L = O.oracle()
    rng = np.random.default_rng(808)  # this test's own draws, whatever ran before
    nsam = 64
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    # two half-maps sharing a signal that falls off with resolution
    zz, yy, xx = np.meshgrid(np.fft.fftfreq(nsam), np.fft.fftfreq(nsam), np.fft.rfftfreq(nsam),
                             indexing="ij")
    env = np.exp(-((xx ** 2 + yy ** 2 + zz ** 2) * nsam ** 2) / 300.0).reshape(-1)
    sig = env * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    a = (sig + 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    d = (sig + 0.3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    s = rng.uniform(0.5, 3, n).astype(np.float32)
    v = rng.uniform(0.5, 3, n).astype(np.float32)
    std, vtd, iredun = 0.37, 0.41, 1
    lut = np.zeros(180, np.float32)
    L.or_calcsinc(lut, 180)
    rio, ric, halfw = 24.0, 0.0, 6.0
    # device chain
    t = lambda x: torch.from_numpy(x.copy()).to(DEV)  # noqa: E731
    b, c = ctx.storeshift(nsam, std, vtd, iredun, t(a), t(s), t(d), t(v))
    for vol in (b, c):
        ctx.fft3d(nsam, vol, False)
        ctx.correct3d(nsam, vol, 1, 1, lut)
        ctx.mask3d(nsam, vol, rio, ric, halfw)
        ctx.fft3d(nsam, vol, True)
    got = ctx.sheltest(nsam, b, c, t(s), t(v))
    torch.cuda.synchronize()
    # oracle chain
    bo, co = np.zeros(n, np.complex64), np.zeros(n, np.complex64)
    L.or32_storeshift(nsam, std, vtd, iredun, a, s, d, v, bo, co)
    outs = []
    for h in (bo, co):
        real = np.fft.irfftn(h.reshape(nsam, nsam, jc).astype(np.complex128), s=(nsam,) * 3,
                             axes=(0, 1, 2))
        pad = np.zeros((nsam, nsam, nsam + 2), np.float32)
        pad[:, :, :nsam] = (real * float(nsam) ** 3).astype(np.float32)  # FFTW_BWD: unnormalised
        pad = pad.reshape(-1)
        L.or32_correct3d(nsam, pad, 1, 1, lut)
        fm, av, sd = C.c_float(), C.c_float(), C.c_float()
        L.or32_mask3d(nsam, pad, rio, ric, halfw, C.byref(fm), C.byref(av), C.byref(sd))
        spec = np.fft.rfftn(pad.reshape(nsam, nsam, nsam + 2)[:, :, :nsam].astype(np.float64))
        outs.append(spec.astype(np.complex64).reshape(-1))
    npo, npto = np.zeros(jc + 2, np.int32), np.zeros(jc + 2, np.int32)
    dfo, pso = np.zeros(jc + 2), np.zeros(jc + 2, np.float32)
    nst = L.or32_sheltest(nsam, 1, outs[0], outs[1], s, v, 1.0, npo, npto, dfo, pso)
    fsc_g, fsc_o = got["fsc"], dfo[:nst]
    err = np.max(np.abs(fsc_g - fsc_o))
    # the first shell past DC (entry 0) whose FSC falls below 0.143
    cross = lambda f: 1 + (int(np.argmax(f[1:] < 0.143)) if np.any(f[1:] < 0.143)  # noqa: E731
                           else len(f) - 1)
    print(f"finalisation chain: max |FSC_gpu - FSC_oracle| = {err:.1e}; FSC(0.143) shell "
          f"{cross(fsc_g)} / {cross(fsc_o)}; FSC[1:6] {np.round(fsc_g[1:6], 4)}")
    assert err < 1e-5
    assert cross(fsc_g) == cross(fsc_o)
    assert np.mean(fsc_g[1:6]) > 0.5 and 0 < cross(fsc_g) < nst
