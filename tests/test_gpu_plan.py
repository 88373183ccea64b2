"""Graph-captured batches (fr_plan_*) and the CC3M autotuner.

A plan's replay must equal the eager sequence fr_prep_batch ->
fr_cc3m_batch -> fr_insert_batch on the same batch bit for bit (NKSUM, CC3M
scores, TTD and the half-maps): every kernel on the path sums in a fixed
order.  Box sizes 64 and 128 take the fused preparation (its ring sums in
the plan's own buffer), 96 the one-kernel-per-step form.  The graph is
replayed several times on different batches, with eager work between the
replays, and every replay is checked.  The
autotuner's choice must be one of its measured candidates, and each CC3M
variant it can choose must give the same scores (1e-6 between persistent
grid sizes of the fused kernel, whose shell sums use LDS float atomics; 1e-4
against the FFT path).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

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


def _records(parts, rng, nb):
    recb = np.zeros((nb, 10), np.float32)
    recb[:, 0:3] = parts.euler + np.radians(rng.normal(0, 2, (nb, 3))).astype(np.float32)
    recb[:, 3:5] = parts.shift
    recb[:, 5] = 1.0
    recb[:, 6] = 1.0
    recb[:, 6][::5] = 0.0  # some particles carry no occupancy: nothing inserted
    recb[:, 7] = 1.0
    recb[:, 8] = 0.5
    recb[:, 9] = 0.45
    half = (np.arange(nb) % 2).astype(np.int32)
    return recb, half


@pytest.mark.parametrize("nsam,asym,nasym", [(64, "C", 1), (96, "D", 7), (128, "C", 2)])
def test_plan_replays_equal_eager(ctx, nsam, asym, nasym):
    psize, ri, nb = 2.0, 0.4 * nsam, 8
    optics = synth.Optics(psize=psize)
    sym = api.point_group(asym, nasym)
    vol = synth.blob_phantom(nsam, 20, 0.38 * nsam, seed=3, sym=sym, device=DEV)
    a3df = synth.half_volume(vol)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    fsct = torch.from_numpy(synth.default_fsct(nsam, psize).astype(np.float32)).to(DEV)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    maxr1, maxr2, rbfact = 2, min(20, nsam // 2 - 2), 0.0
    vols_p = api.HalfMaps.zeros(nsam)
    plan = api.Plan(ctx, nsam, nb, a3df, pssnr, gpar, maxr1, maxr2, rbfact, fsct, ring, sym,
                    5.0, 40.0, psize, vols_p)
    vols_e = api.HalfMaps.zeros(nsam)
    rng = np.random.default_rng(nsam)
    for rep in range(3):
        parts = synth.make_particles(ctx, a3df, nb, seed=100 + rep, optics=optics, snr=1.0, ri=ri)
        recb, half = _records(parts, rng, nb)
        ppar = parts.ppar().cpu().numpy()
        score_p = plan.run(parts.images, ppar, recb, half)
        # eager: the same three entry points
        ctff, outc, datc = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
        dm = np.concatenate([api.rotation(recb[b:b + 1, 0:3], amag=float(recb[b, 5]))
                             for b in range(nb)])
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
        score_e = ctx.cc3m(nsam, a3df, outc, ctff, t(np.arange(nb, dtype=np.int32)), t(dm),
                           t(recb[:, 3:5]), t(recb[:, 5]), maxr1, maxr2, rbfact, fsct,
                           ring).cpu().numpy()
        ctx.insert(nsam, datc, ctff, recb, half, sym, 5.0, 40.0, psize, vols_e)
        torch.cuda.synchronize()
        # every kernel on the path sums in a fixed order: replay and eager agree bitwise
        assert np.array_equal(score_p.view(np.uint32), score_e.view(np.uint32)), rep
        assert np.array_equal(vols_p.nksum.cpu().numpy(), vols_e.nksum.cpu().numpy()), rep
        tp, te = vols_p.ttd.cpu().numpy(), vols_e.ttd.cpu().numpy()
        assert np.array_equal(tp.view(np.uint64), te.view(np.uint64)), rep
        for h in range(2):
            for a, b in ((vols_p.num[h], vols_e.num[h]), (vols_p.ctf2[h], vols_e.ctf2[h])):
                a, b = a.cpu().numpy().ravel(), b.cpu().numpy().ravel()
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (rep, h)
        print(f"nsam {nsam} {asym}{nasym} replay {rep}: max |dscore| "
              f"{np.max(np.abs(score_p - score_e)):.1e}, NKSUM "
              f"{vols_p.nksum.cpu().numpy().tolist()}")
    plan.close()


def test_autotune_cc3m(ctx):
    nsam, psize, ri = 64, 2.0, 26.0
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 20, 24.0, seed=4, device=DEV)
    a3df = synth.half_volume(vol)
    n = 300
    parts = synth.make_particles(ctx, a3df, 16, seed=5, optics=optics, snr=1.0, ri=ri)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    rng = np.random.default_rng(1)
    pidx = rng.integers(0, 16, n).astype(np.int32)
    eul = rng.uniform(-np.pi, np.pi, (n, 3)).astype(np.float32)
    dm = api.rotation(eul, amag=1.0)
    shift = rng.uniform(-0.1, 0.1, (n, 2)).astype(np.float32)
    amag = np.ones(n, np.float32)
    fsct = torch.from_numpy(synth.default_fsct(nsam, psize).astype(np.float32)).to(DEV)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    args = (nsam, a3df, outc, ctff, t(pidx), t(dm), t(shift), t(amag), 2, 14, 0.0, fsct, ring)
    log = ctx.autotune_cc3m(*args)
    print(log)
    assert log["nsam"] == nsam and 0 <= log["choice"] < len(log["candidates"])
    assert all(c["ms"] > 0 for c in log["candidates"])
    chosen = ctx.cc3m(*args).cpu().numpy()
    # every variant gives the same scores: bitwise across the fused kernel's grid
    # sizes (one workgroup per item, fixed-order sums), to FP32 rounding against
    # the FFT path
    fused = []
    for c in log["candidates"]:
        ctx.set_cc3m_variant(nsam, c["path"] == "fused",
                             c["workgroups"] if c["workgroups"] != n else 0)
        got = ctx.cc3m(*args).cpu().numpy()
        if c["path"] == "fused":
            fused.append(got)
        assert np.allclose(got, chosen, rtol=1e-4, atol=1e-6), c
    for got in fused[1:]:
        assert np.array_equal(got.view(np.uint32), fused[0].view(np.uint32))
