"""The GPU global search against the oracle's PSEARCH (psearch.f:8-121 ->
ccp.f:8-275 with PAD and ROTATE2D) on the same prepared particles, the same
SEARCHANG grid and the same RANDOM jitter.

The GEMM searches score each in-plane angle from a section extracted at that
angle; CCP rotates one padded section by nearest-neighbour lookup.  The two
scores differ by a few percent, so neighbouring views trade places at the top:
the bar for the GEMM search is that PSEARCH's best view lies within one grid
step (modulo the point group) of one of the GEMM's top-8 candidates, which go
on to FP32 rescoring; the top-1 agreement and the score gaps are printed.
fr_psearch, the reference-exact path, is held to TESTPAR slot by slot.
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

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


def oracle_psearch(nsam, outc, ctff, a3df, dang, maxr1, maxr2, rbfact, asym, nasym, iran,
                   ipmax):
    """or32_psearch per particle (threads: the oracle releases the GIL)."""
    L = O.oracle()
    sang, iquadmax = api.search_grid(dang, asym, nasym)
    flat = np.ascontiguousarray(sang.reshape(-1))
    mask = np.ones(5, np.int32)

    def one(p):
        tp = np.zeros(6 * ipmax, np.float32)
        ir = C.c_int(iran)
        ph, th, ps, dx, dy = (C.c_float(0.0) for _ in range(5))
        L.or32_psearch(nsam, 1.0, outc[p].reshape(-1), a3df, dang, maxr1, maxr2, C.byref(ph),
                       C.byref(th), C.byref(ps), C.byref(dx), C.byref(dy), flat, len(sang),
                       iquadmax, mask, rbfact, ctff[p].reshape(-1), tp, ipmax, C.byref(ir),
                       O.IDENTITY)
        return tp.reshape(ipmax, 6)

    with ThreadPoolExecutor(8) as ex:
        return np.stack(list(ex.map(one, range(outc.shape[0]))))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_search_best_view_matches_oracle_psearch(ctx, dtype):
    nsam, psize, ri = 64, 2.0, 26.0
    dang, maxr1, maxr2, rbfact, iran = 15.0, 1, 14, 0.0005, -5
    optics = synth.Optics(psize=psize)
    sym = api.point_group("D", 7)
    # a phantom with enough structure at the 9 A band and particles at SNR 10,
    # so that the best view is determined by the data, not by the noise
    vol = synth.blob_phantom(nsam, 40, 24.0, seed=7, sym=sym, device=DEV)
    a3df = synth.half_volume(vol)
    npart = 16
    parts = synth.make_particles(ctx, a3df, npart, seed=8, optics=optics, snr=10.0, ri=ri,
                                 shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    print(f"[psearch] gpu search {dtype}, {npart} particles at {nsam}^2, D7 {dang} deg",
          flush=True)
    gs = search.GlobalSearch(ctx, a3df, dang=dang, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                             rbfact=rbfact, dtype=dtype, iran=iran)
    top_o, top_r, top_s = gs.candidates(outc, ctff, topk=8)
    top_o, top_s = top_o.cpu().numpy(), top_s.cpu().numpy()
    print("[psearch] oracle PSEARCH", flush=True)
    tp = oracle_psearch(nsam, outc.cpu().numpy(), ctff.cpu().numpy(),
                        a3df.cpu().numpy().reshape(-1), dang, maxr1, maxr2, rbfact, "D", 7, iran,
                        8)
    assert (top_o[:, 0] >= 0).all() and (tp[:, 0, 5] > 0).all()
    gpu_best = gs.euler[top_o[:, 0]]
    ora_best = tp[:, 0, :3]
    ops = synth.symmetry_matrices(sym)
    err = search.angular_error_deg(gpu_best, ora_best, ops)
    gap = np.abs(top_s[:, 0] - tp[:, 0, 5]) / np.abs(tp[:, 0, 5])
    print(f"[psearch] {dtype}: best-view distance to PSEARCH (deg) {np.round(err, 1)}")
    print(f"[psearch] {dtype}: relative score gap GEMM vs CCP {np.round(gap, 3)}")
    et_g = search.angular_error_deg(gpu_best, parts.euler, ops)
    et_o = search.angular_error_deg(ora_best, parts.euler, ops)
    print(f"[psearch] {dtype}: GEMM best to truth (deg) {np.round(et_g, 1)}")
    print(f"[psearch] {dtype}: PSEARCH best to truth (deg) {np.round(et_o, 1)}")
    in8 = [min(search.angular_error_deg(gs.euler[top_o[p_, :]], np.repeat(ora_best[p_:p_ + 1], 8, 0),
                                        ops)) for p_ in range(npart)]
    print(f"[psearch] {dtype}: PSEARCH best to nearest GEMM top-8 (deg) {np.round(in8, 1)}")
    # one grid step: the view spacing (15 deg) and CCP's in-plane step (3 deg).
    # The candidates go on to FP32 rescoring, as PSEARCH's TESTPAR goes on to
    # PREFINE (lmain.f:615-642): PSEARCH's best view must be among them.
    step = dang + 3.0
    print(f"[psearch] {dtype}: GEMM top-1 within one step of PSEARCH's best: "
          f"{np.mean(err <= step):.3f}")
    assert np.mean(np.asarray(in8) <= step) >= 0.99


# ------------------------------------------------- the reference-exact search
_libm = C.CDLL("libm.so.6")
_libm.expf.argtypes = [C.c_float]
_libm.expf.restype = C.c_float


def _iran_states(iran, n):
    """IRAN before each of n PSEARCH calls (3 RANDOM draws per call)."""
    out, s = [], int(iran)
    for _ in range(n):
        out.append(s)
        for _ in range(3):
            _, s = api.lcg_random(s)
    return out


def _case(ctx, nsam, psize, ri, npart, seed, asym="D", nasym=7, snr=1.0, nblob=12):
    optics = synth.Optics(psize=psize)
    sym = api.point_group(asym, nasym)
    vol = synth.blob_phantom(nsam, nblob, 0.375 * nsam, seed=seed, sym=sym, device=DEV)
    a3df = synth.half_volume(vol)
    parts = synth.make_particles(ctx, a3df, npart, seed=seed + 1, optics=optics, snr=snr, ri=ri,
                                 shift_a=2.0 * psize)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    return a3df, outc, ctff, sym, parts


def test_ccp_views_match_oracle_ccp(ctx):
    """fr_ccp_view_batch against or32_ccp view by view: the CCMAX of each view
    within 1e-4 relative and the same in-plane step, quadrant and shift
    wherever the view's best is not tied within that tolerance."""
    nsam, maxr1, maxr2, rbfact = 64, 1, 14, 0.0005
    a3df, outc, ctff, sym, _ = _case(ctx, nsam, 2.0, 26.0, 3, seed=31)
    sang, iq = api.search_grid(15.0, "D", 7)
    sang = sang[:10]
    mask = np.ones(5, np.int32)
    dang = 3.0
    np_, nv = 3, len(sang)
    dm = np.zeros((np_, nv, 9), np.float32)
    for p in range(np_):
        dm[p] = api.rotation(sang, amag=1.0)
    score, code = ctx.ccp_views(nsam, a3df, outc, ctff, torch.from_numpy(dm).to(DEV),
                                maxr1=maxr1, maxr2=maxr2, rbfact=rbfact, dang=dang, iquadmax=iq,
                                mask=mask)
    score, code = score.cpu().numpy(), code.cpu().numpy()
    L = O.oracle()
    jc = nsam // 2 + 1
    oc, cf = outc.cpu().numpy(), ctff.cpu().numpy()
    vol = a3df.cpu().numpy().reshape(-1)
    nsamr = nsam // 2
    same = 0
    for p in range(np_):
        cb = oc[p].reshape(nsam, jc).copy()
        for i in range(0, nsam // 2 + 1):
            for j in range(-(nsam // 2), nsam // 2 + 1):
                w = np.float32(_libm.expf(np.float32(-np.float32(rbfact * np.float32(i * i + j * j)))))
                v = cb[j % nsam, i]
                cb[j % nsam, i] = complex(np.float32(np.float32(v.real) * w),
                                          np.float32(np.float32(v.imag) * w))
        cb = cb.reshape(-1)
        for v in range(nv):
            ps, sx, sy = C.c_float(float(sang[v, 2])), C.c_float(0.0), C.c_float(0.0)
            cc = L.or32_ccp(nsam, 1.0, cb, vol, dang, maxr1, maxr2, float(sang[v, 0]),
                            float(sang[v, 1]), C.byref(ps), C.byref(sx), C.byref(sy), mask,
                            cf[p].reshape(-1), iq, O.IDENTITY)
            rel = abs(score[p, v] - cc) / abs(cc)
            assert rel < 1e-4, (p, v, score[p, v], cc)
            c = int(code[p, v])
            ipsi, iquad, row = c // 900, (c // 225) % 4, c % 225
            psim = np.float32(np.float32(sang[v, 2]) + np.float32(np.float32(iquad) * np.float32(
                3.1415926535897) / np.float32(2.0)) + np.float32(np.float32(ipsi) * np.float32(
                    3.1415926535897) / np.float32(2.0) / np.float32(90.0 / dang)))
            gx = (row // 15 - 7) / nsamr * 2 * np.pi
            gy = (row % 15 - 7) / nsamr * 2 * np.pi
            ok = abs(psim - ps.value) < 1e-5 and abs(gx - sx.value) < 1e-5 and \
                abs(gy - sy.value) < 1e-5
            same += ok
    print(f"[ccp] views with the oracle's in-plane step, quadrant and shift: {same}/{np_ * nv}")
    assert same >= 0.95 * np_ * nv


@pytest.mark.parametrize("nsam,dang,maxr2,npart", [(64, 15.0, 14, 12), (256, 5.0, 32, 3)])
def test_exact_psearch_matches_oracle(ctx, nsam, dang, maxr2, npart):
    """fr_psearch (PSEARCH -> CCP -> ROTATE2D/PAD on the GPU) against
    or32_psearch on the same particles, grid and IRAN sequence: TESTPAR's
    orientations and shifts bit for bit and its scores within 1e-4 relative,
    slot by slot, except where the oracle's neighbouring scores are tied within
    that tolerance; the best view identical for every particle."""
    psize = 128.0 / nsam
    a3df, outc, ctff, sym, _ = _case(ctx, nsam, psize, 0.4 * nsam, npart, seed=41)
    maxr1, rbfact, iran, ipmax = 1, 0.0005, -17, 8
    sang, iq = api.search_grid(dang, "D", 7)
    print(f"[psearch-exact] {npart} particles at {nsam}^2, {len(sang)} views", flush=True)
    par, tp, ir_out = ctx.psearch(nsam, a3df, outc, ctff, dangin=dang, maxr1=maxr1, maxr2=maxr2,
                                  rbfact=rbfact, sang=sang, iquadmax=iq, ipmax=ipmax, iran=iran,
                                  asym="D")
    states = _iran_states(iran, npart + 1)
    assert ir_out == states[-1]
    L = O.oracle()
    oc, cf = outc.cpu().numpy(), ctff.cpu().numpy()
    vol = a3df.cpu().numpy().reshape(-1)
    flat = np.ascontiguousarray(sang.reshape(-1))
    mask = np.ones(5, np.int32)

    def one(p):
        t = np.zeros(6 * ipmax, np.float32)
        ir = C.c_int(states[p])
        ph, th, ps, dx, dy = (C.c_float(0.0) for _ in range(5))
        L.or32_psearch(nsam, 1.0, oc[p].reshape(-1), vol, dang, maxr1, maxr2, C.byref(ph),
                       C.byref(th), C.byref(ps), C.byref(dx), C.byref(dy), flat, len(sang), iq,
                       mask, rbfact, cf[p].reshape(-1), t, ipmax, C.byref(ir), O.IDENTITY)
        return t.reshape(ipmax, 6), (ph.value, th.value, ps.value, dx.value, dy.value)

    print("[psearch-exact] oracle PSEARCH", flush=True)
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(one, range(npart)))
    slots = agree = 0
    for p in range(npart):
        ot, obest = res[p]
        assert np.array_equal(par[p, :3], np.array(obest[:3], np.float32)), (p, par[p], obest)
        for q in range(ipmax):
            if ot[q, 5] == 0.0:
                assert tp[p, q, 5] == 0.0
                continue
            slots += 1
            near = [abs(ot[q, 5] - ot[r, 5]) <= 1e-4 * abs(ot[q, 5]) for r in (q - 1, q + 1)
                    if 0 <= r < ipmax]
            rel = abs(tp[p, q, 5] - ot[q, 5]) / abs(ot[q, 5])
            if any(near):
                continue
            assert rel < 1e-4, (p, q, tp[p, q], ot[q])
            agree += np.array_equal(tp[p, q, :5], ot[q, :5])
    print(f"[psearch-exact] TESTPAR slots with identical angles and shifts: {agree}/{slots}")
    assert agree >= 0.95 * slots
