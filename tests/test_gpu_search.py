"""GPU global search: bank, GEMM scores and candidates against independent
checkers, and recovery of known orientations.

Bars: bank entries match the oracle's AINTERPO3DS to 1e-6 relative; FP32
GEMM scores match an FP64 numpy evaluation of the same correlation to 1e-4
relative and select the same view wherever the score gap exceeds that; BF16
scores match to 2e-2 relative; on clean synthetic particles the search
recovers the true orientation within one grid step (modulo symmetry).
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, search, synth  # noqa: E402

DEV = "cuda"


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def annulus(nsam, maxr1, maxr2):
    jc = nsam // 2 + 1
    lo = max(maxr1 * maxr1, 1)
    out = []
    for i in range(0, jc - 1):
        for j in range(-jc + 1, jc):
            it = i * i + j * j
            if lo <= it < maxr2 * maxr2 and abs(j) != jc - 1:
                out.append((i, j))
    return np.array(out)


def numpy_scores(nsam, kidx, bank_c, outc, ctff, rbfact, nsamr):
    """FP64 evaluation of the CCP score (ccp.f:194-252) for every orientation:
    explicit 15x15 shift maps, argmax over -6..6 (x outer), 3x3 box / 9 /
    sqrt(sum|cT|^2 sum|X|^2)."""
    jc = nsam // 2 + 1
    I, J = kidx[:, 0], kidx[:, 1]
    rows = np.where(J < 0, J + nsam, J)
    w = np.exp(-rbfact * (I * I + J * J))
    X = outc[rows, I].astype(np.complex128) * w
    c = (ctff[0][rows, I] + np.conj(ctff[1][rows, I])).astype(np.complex128)
    wk = np.where(I > 0, 2.0, 1.0)
    A = wk * X * np.conj(c)
    s = np.arange(-7, 8)
    ph = np.exp(1j * 2 * np.pi * (I[:, None, None] * s[None, :, None] + J[:, None, None] * s[None, None, :]) / nsamr)
    As = A[:, None, None] * ph  # (k, sx, sy)
    cc = np.einsum("kxy,ok->oxy", As, np.conj(bank_c)).real  # (o, 15, 15)
    # SUM2 over 1 <= r^2 < MAXR2^2 on the half plane (includes the low ring)
    return cc, A, c


def epilogue(cc, nrm, s2):
    o = cc.shape[0]
    inner = cc[:, 1:14, 1:14].reshape(o, -1)
    idx = np.argmax(inner, axis=1)  # first maximum in (x outer, y inner) order
    bx, by = idx // 13 + 1, idx % 13 + 1
    box = np.zeros(o)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            box += cc[np.arange(o), bx + dx, by + dy]
    return box / np.sqrt(nrm) / np.sqrt(s2) / 9.0, bx * 15 + by


def test_search_bank_matches_oracle(ctx):
    nsam, maxr1, maxr2 = 64, 2, 14
    L = O.oracle()
    rng = np.random.default_rng(5)
    v = rng.standard_normal((nsam,) * 3).astype(np.float32)
    vol = np.ascontiguousarray(np.fft.rfftn(v).astype(np.complex64))
    eul = rng.uniform(-np.pi, np.pi, (37, 3)).astype(np.float32)
    dm = api.rotation(eul, amag=1.0)
    bank = ctx.search_bank(nsam, maxr1, maxr2, torch.from_numpy(vol).to(DEV),
                           torch.from_numpy(dm).to(DEV), dtype=torch.float32).cpu().numpy()
    kidx = annulus(nsam, maxr1, maxr2)
    kc, kpad = ctx.search_layout(nsam, maxr1, maxr2)
    assert kc == len(kidx) and kpad >= 2 * kc and kpad % 32 == 0
    out = np.zeros(2, np.float32)
    for o in range(0, 37, 6):
        d = dm[o]
        for k in range(0, kc, 7):
            i, j = (float(x) for x in kidx[k])
            x3 = np.float32(np.float32(d[0] * np.float32(i)) + np.float32(d[3] * np.float32(j)))
            y3 = np.float32(np.float32(d[1] * np.float32(i)) + np.float32(d[4] * np.float32(j)))
            z3 = np.float32(np.float32(d[2] * np.float32(i)) + np.float32(d[5] * np.float32(j)))
            L.or32_ainterpo3ds(nsam, vol.reshape(-1), float(x3), float(y3), float(z3), out)
            g = bank[o, 2 * k:2 * k + 2]
            assert np.allclose(g, out, rtol=1e-5, atol=1e-5 * np.abs(vol).max() * 1e-3), (o, k)
    assert np.all(bank[:, 2 * kc:] == 0)


def sep_element(i, j):
    """Complex element of grid point (I, J) in a row of the separable bank."""
    jg = j + 32
    q = (i // 16) * 4 + jg // 16
    lane = ((jg % 16) // 4) * 16 + i % 16
    return (q * 64 + lane) * 4 + jg % 4


@pytest.mark.parametrize("maxr1,maxr2", [(2, 14), (1, 32)])
def test_separable_bank_equals_annulus_bank(ctx, maxr1, maxr2):
    """The separable layout holds the annulus bank's BF16 values, bit for bit,
    at the documented positions, and zeros elsewhere."""
    nsam = 128
    rng = np.random.default_rng(6)
    v = rng.standard_normal((nsam,) * 3).astype(np.float32)
    vol = torch.from_numpy(np.ascontiguousarray(np.fft.rfftn(v).astype(np.complex64))).to(DEV)
    eul = rng.uniform(-np.pi, np.pi, (21, 3)).astype(np.float32)
    dm = torch.from_numpy(api.rotation(eul, amag=1.0)).to(DEV)
    ann = ctx.search_bank(nsam, maxr1, maxr2, vol, dm, dtype=torch.bfloat16)
    sep = ctx.search_bank(nsam, maxr1, maxr2, vol, dm, dtype=torch.bfloat16, separable=True)
    assert sep.shape == (21, 4096)
    ann = ann.view(torch.int16).cpu().numpy().astype(np.int32)
    sep = sep.view(torch.int16).cpu().numpy().astype(np.int32)
    kidx = annulus(nsam, maxr1, maxr2)
    e = np.array([sep_element(i, j) for i, j in kidx])
    assert len(set(e.tolist())) == len(e)
    kc = len(kidx)
    assert np.array_equal(sep[:, 2 * e], ann[:, 0:2 * kc:2])
    assert np.array_equal(sep[:, 2 * e + 1], ann[:, 1:2 * kc:2])
    rest = np.ones(2048, bool)
    rest[e] = False
    assert np.all(sep[:, 0::2][:, rest] == 0) and np.all(sep[:, 1::2][:, rest] == 0)


# BF16 runs the separable two-stage search (the whole I < 32, |J| < 32 grid
# at 128, MAXR2 = 32)
@pytest.mark.parametrize("dtype,tol,nsam,maxr2", [(torch.float32, 1e-4, 64, 14),
                                                  (torch.bfloat16, 2e-2, 64, 14),
                                                  (torch.bfloat16, 2e-2, 128, 32)])
def test_search_scores_vs_numpy(ctx, dtype, tol, nsam, maxr2):
    maxr1, rbfact = 2, 0.0005
    jc = nsam // 2 + 1
    rng = np.random.default_rng(11)
    v = rng.standard_normal((nsam,) * 3).astype(np.float32)
    vol = np.ascontiguousarray(np.fft.rfftn(v).astype(np.complex64))
    a3df = torch.from_numpy(vol).to(DEV)
    gs = search.GlobalSearch(ctx, a3df, dang=15.0, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                             rbfact=rbfact, dtype=dtype, iran=-7)
    npart = 3
    outc = (rng.standard_normal((npart, nsam, jc)) + 1j * rng.standard_normal((npart, nsam, jc))
            ).astype(np.complex64)
    ctff = (0.4 + 0.2 * rng.standard_normal((npart, 2, nsam, jc))).astype(np.complex64)
    top_o, top_r, top_s = gs.candidates(torch.from_numpy(outc).to(DEV), torch.from_numpy(ctff).to(DEV),
                                        topk=8)
    top_o, top_r, top_s = top_o.cpu().numpy(), top_r.cpu().numpy(), top_s.cpu().numpy()
    bank = gs.bank.float().cpu().numpy().astype(np.float64)
    kidx = annulus(nsam, maxr1, maxr2)
    kc = len(kidx)
    bank_c = bank[:, 0:2 * kc:2] + 1j * bank[:, 1:2 * kc:2]
    for p in range(npart):
        cc, A, c = numpy_scores(nsam, kidx, bank_c, outc[p], ctff[p], rbfact, gs.nsamr)
        nrm = (np.abs(c) ** 2 * np.abs(bank_c) ** 2).sum(1)
        # SUM2: 1 <= r^2 < MAXR2^2, half plane incl. both signs of J at I=0
        s2 = 0.0
        for i in range(0, maxr2 + 1):
            for j in range(-maxr2, maxr2 + 1):
                it = i * i + j * j
                if 1 <= it < maxr2 * maxr2:
                    s2 += abs(outc[p][j % nsam, i] * math.exp(-rbfact * it)) ** 2
        score, row = epilogue(cc, nrm, s2)
        # per-view first maximum, then the best views
        g = gs.group
        sv = score.reshape(-1, g)
        vbest = np.argmax(sv, axis=1)
        vscore = sv[np.arange(sv.shape[0]), vbest]
        order = np.argsort(-vscore, kind="stable")[:8]
        want_o = order * g + vbest[order]
        rel = np.abs(top_s[p] - vscore[order]) / np.abs(vscore[order]).max()
        print(f"{dtype} particle {p}: max rel score error {rel.max():.2e}")
        assert rel.max() < tol
        if dtype == torch.float32:
            gaps = np.abs(np.diff(vscore[order]))
            for q in range(8):
                clear = (q == 0 or gaps[q - 1] > 1e-4 * abs(vscore[order[0]])) and \
                        (q == 7 or gaps[q] > 1e-4 * abs(vscore[order[0]]))
                if clear:
                    assert top_o[p, q] == want_o[q]
                    assert top_r[p, q] == row[want_o[q]]


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_search_recovers_orientation(ctx, dtype):
    """Clean-ish synthetic particles: the best view lies within one grid step."""
    nsam, psize = 64, 2.0
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 30, 24.0, seed=3, device=DEV)
    a3df = synth.half_volume(vol)
    ri = 26.0
    parts = synth.make_particles(ctx, a3df, 12, seed=4, optics=optics, snr=2.0, ri=ri, shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, datc = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    maxr2 = 16
    gs = search.GlobalSearch(ctx, a3df, dang=7.5, asym="C", nasym=1, maxr1=1, maxr2=maxr2,
                             rbfact=0.0, dtype=dtype, iran=-11)
    fsct = synth.default_fsct(nsam, psize)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    res = gs.run(outc, ctff, fsct, ring, topk=16)
    err = search.angular_error_deg(res.euler, parts.euler)
    print(f"{dtype}: angular errors (deg) {np.round(err, 1)}; CC3M {np.round(res.score, 3)}")
    assert np.mean(err < 7.5 * 1.5) >= 0.9
