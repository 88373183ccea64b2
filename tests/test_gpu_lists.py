"""Ragged candidate lists of the learned search, on the GPU.

Bars (bit for bit where the paths run the same kernels on the same inputs):
- fr_search_ragged_batch gives exactly fr_search_list_batch's answer on each
  particle's list; for a list that is the whole bank in order with the
  bank's view group it gives fr_search_batch's answer up to the summation
  order of the normalisation (scores to 1e-5, the same best view);
- fr_build_lists builds the lists a numpy construction builds from the same
  classes, probabilities and cells;
- fr_rescore_batch gives the scores of CC3M at the host-built candidate
  items, and the winner numpy's argmax picks;
- LearnedProposal.run equals the composition of those entry points with
  the lists built on the host.
"""
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, proposal, search, synth  # noqa: E402

DEV = "cuda"
rng = np.random.default_rng(0x11575)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(ctx):
    nsam = 64
    vol = synth.blob_phantom(nsam, 20, 24.0, 0x52, sym=api.point_group("D", 7), device=DEV)
    a3df = synth.half_volume(vol)
    gs = search.GlobalSearch(ctx, a3df, dang=10.0, asym="D", nasym=7, maxr1=1, maxr2=14,
                             rbfact=0.0, dtype=torch.bfloat16, iran=5)
    optics = synth.Optics(psize=2.0)
    parts = synth.make_particles(ctx, a3df, 12, 21, optics, snr=0.5, ri=26.0)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, 26.0)).to(DEV)
    ctff, outc, datc = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr,
                                    optics.gpar(nsam, 26.0))
    return nsam, a3df, gs, optics, parts, outc, ctff, datc


def ragged(lists, groups):
    loff = np.zeros(len(lists) + 1, np.int32)
    loff[1:] = np.cumsum([len(x) for x in lists])
    cand = np.concatenate(lists + [np.zeros(0, np.int32)]).astype(np.int32)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.int32)).to(DEV)  # noqa: E731
    return t(loff), t(np.asarray(groups, np.int32)), t(cand if len(cand) else np.full(1, -1))


@pytest.mark.parametrize("kind", ["annulus", "separable"])
def test_ragged_equals_exhaustive_and_uniform_lists(ctx, setup, kind):
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    bank = gs.bank if kind == "annulus" else gs.bank_sep
    npart = outc.shape[0]
    topk = 8
    o_d, r_d, s_d = [x.cpu().numpy() for x in ctx.search(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff, topk=topk, group=gs.group)]
    bank_rows = torch.arange(gs.norient, dtype=torch.int32, device=DEV).repeat(npart, 1)
    o_f, r_f, s_f = [x.cpu().numpy() for x in ctx.search_list(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff, bank_rows.contiguous(),
        gs.group, topk=topk)]
    if kind == "separable":
        # one kernel, the norm inside it either way: the same bits
        assert np.array_equal(o_f, o_d) and np.array_equal(r_f, r_d)
        assert np.array_equal(s_f.view(np.uint32), s_d.view(np.uint32))
    assert np.array_equal(o_f[:, 0], o_d[:, 0])
    assert np.allclose(s_f, s_d, rtol=1e-5, atol=0)
    # every particle: 6 random views (with their whole in-plane groups)
    nview = gs.norient // gs.group
    rows = []
    for p in range(npart):
        views = rng.choice(nview, 6, replace=False)
        rows.append(np.concatenate([np.arange(v * gs.group, (v + 1) * gs.group) for v in views]))
    rows = np.array(rows, np.int32)
    o_u, r_u, s_u = [x.cpu().numpy() for x in ctx.search_list(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff,
        torch.from_numpy(rows).to(DEV), gs.group, topk=topk)]
    # ragged: particles 0, 5, 9 search the whole bank, 7 has an empty list,
    # the others their uniform-list row (some cut to fewer views)
    full = {0, 5, 9}
    lists, groups = [], []
    for p in range(npart):
        if p in full:
            lists.append(np.arange(gs.norient, dtype=np.int32))
        elif p == 7:
            lists.append(np.zeros(0, np.int32))
        else:
            lists.append(rows[p])
        groups.append(gs.group)
    loff, grp, cand = ragged(lists, groups)
    o_r, r_r, s_r = [x.cpu().numpy() for x in ctx.search_ragged(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff, loff, grp, cand, topk)]
    for p in range(npart):
        if p in full:
            ref = (o_f[p], r_f[p], s_f[p])
        elif p == 7:
            assert np.all(o_r[p] == -1) and np.all(r_r[p] == -1)
            continue
        else:
            ref = (o_u[p], r_u[p], s_u[p])
        assert np.array_equal(o_r[p], ref[0]), p
        assert np.array_equal(r_r[p], ref[1]), p
        assert np.array_equal(s_r[p].view(np.uint32), ref[2].view(np.uint32)), p
    if kind == "separable":
        # the fallback particles flagged (info[p, 0] = 0) are searched four to
        # a workgroup against shared templates: the same bits
        info = np.ones((npart, 2), np.int32)
        for p in full:
            info[p, 0] = 0
        o_i, r_i, s_i = [x.cpu().numpy() for x in ctx.search_ragged(
            nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff, loff, grp, cand, topk,
            info=torch.from_numpy(info).to(DEV))]
        assert np.array_equal(o_i, o_r) and np.array_equal(r_i, r_r)
        assert np.array_equal(s_i.view(np.uint32), s_r.view(np.uint32))
        # every particle falling back (three workgroup quads)
        lf, gf, cf_ = ragged([np.arange(gs.norient, dtype=np.int32)] * npart, [gs.group] * npart)
        o_a, r_a, s_a = [x.cpu().numpy() for x in ctx.search_ragged(
            nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, outc, ctff, lf, gf, cf_, topk,
            info=torch.zeros((npart, 2), dtype=torch.int32, device=DEV))]
        assert np.array_equal(o_a, o_f) and np.array_equal(r_a, r_f)
        assert np.array_equal(s_a.view(np.uint32), s_f.view(np.uint32))


@pytest.mark.parametrize("kind", ["annulus", "separable"])
def test_ragged_capacity_is_respected(ctx, setup, kind):
    """Entries at or beyond the capacity are ignored: a list cut by the
    capacity gives the answer of the shorter list."""
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    bank = gs.bank if kind == "annulus" else gs.bank_sep
    o1, c1 = outc[:2].contiguous(), ctff[:2].contiguous()
    lists = [np.arange(3 * gs.group, dtype=np.int32),
             np.arange(5 * gs.group, dtype=np.int32)]
    loff, grp, cand = ragged(lists, [gs.group, gs.group])
    short = cand[:3 * gs.group + 2 * gs.group].clone()  # particle 1 keeps 2 views
    o_c, r_c, s_c = [x.cpu().numpy() for x in ctx.search_ragged(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, o1, c1, loff, grp, short, 4)]
    loff2, grp2, cand2 = ragged([lists[0], lists[1][:2 * gs.group]], [gs.group, gs.group])
    o_s, r_s, s_s = [x.cpu().numpy() for x in ctx.search_ragged(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, o1, c1, loff2, grp2, cand2, 4)]
    assert np.array_equal(o_c, o_s) and np.array_equal(r_c, r_s)
    assert np.array_equal(s_c.view(np.uint32), s_s.view(np.uint32))


def host_lists(lp, top, prob, threshold, norient, bank_group):
    lists, groups, conf, valid = [], [], [], []
    for row, pr in zip(top, prob):
        segs = [lp.cells[c] for c in row if 0 <= c < lp.classes.n]
        x = np.concatenate(segs) if segs else np.zeros(0, np.int32)
        ok = pr[0] >= threshold and len(x) > 0
        if not ok:
            x = np.arange(norient, dtype=np.int32)
        lists.append(x.astype(np.int32))
        groups.append(lp.cfg.group if ok else bank_group)
        conf.append(int(ok))
        valid.append(int((x >= 0).sum()))
    return lists, groups, conf, valid


def test_build_lists_vs_numpy(ctx, setup):
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    sym = synth.symmetry_matrices(api.point_group("D", 7))
    cfg = proposal.ProposalConfig(S=32, channels=8, nconv=3, class_deg=30.0, topk=4,
                                  threshold=0.5, group=16)
    lp = proposal.LearnedProposal(ctx, gs, sym, "D", 7, cfg, seed=5)
    npart, ktop = 40, cfg.topk
    top = np.stack([rng.choice(lp.classes.n, ktop, replace=False) for _ in range(npart)])
    top[3, 1] = -1  # an unused slot
    prob = np.sort(rng.uniform(0, 1, (npart, ktop)), axis=1)[:, ::-1].astype(np.float32)
    prob[4, 0] = cfg.threshold  # exactly at the threshold: confident
    t = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dt)).to(DEV)  # noqa: E731
    loff, grp, cand, info = ctx.build_lists(t(top, np.int32), t(prob, np.float32), cfg.threshold,
                                            lp.cell_off_dev, lp.cells_dev, cfg.group, gs.norient,
                                            gs.group, lp.maxlen)
    lists, groups, conf, valid = host_lists(lp, top, prob, cfg.threshold, gs.norient, gs.group)
    loff, grp, cand, info = (x.cpu().numpy() for x in (loff, grp, cand, info))
    assert conf[4] == 1 and 0 < sum(conf) < npart
    assert np.array_equal(grp, groups)
    assert np.array_equal(info[:, 0], conf) and np.array_equal(info[:, 1], valid)
    for p in range(npart):
        assert np.array_equal(cand[loff[p]:loff[p + 1]], lists[p]), p
    # a small capacity clips each list to whole groups
    small = 3 * gs.group
    loff2, grp2, cand2, info2 = ctx.build_lists(t(top, np.int32), t(prob, np.float32),
                                                cfg.threshold, lp.cell_off_dev, lp.cells_dev,
                                                cfg.group, gs.norient, gs.group, small)
    loff2, cand2 = loff2.cpu().numpy(), cand2.cpu().numpy()
    for p in range(npart):
        g = groups[p]
        n = min(len(lists[p]), small // g * g)
        assert loff2[p + 1] - loff2[p] == n
        assert np.array_equal(cand2[loff2[p]:loff2[p + 1]], lists[p][:n])


def test_rescore_matches_host_items(ctx, setup):
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    fsct = synth.default_fsct(nsam, optics.psize)
    ring = optics.gpar(nsam, 26.0)[[7, 8, 5, 6]]
    o, r, _ = gs.candidates(outc, ctff, 6)
    o_h = o.cpu().numpy()
    o_h[2, 3] = -1  # empty candidates
    o_h[4, :] = -1
    o = torch.from_numpy(o_h).to(DEV)
    sc_h, _ = gs.rescore(outc, ctff, o, r, fsct, ring)
    sc_d, best = ctx.rescore(nsam, a3df, outc, ctff, o, r, gs.dm, gs.nsamr, gs.maxr1, gs.maxr2,
                             gs.rbfact, gs.fsct_device(fsct), ring)
    sc_d, best = sc_d.cpu().numpy(), best.cpu().numpy()
    assert np.array_equal(sc_d.view(np.uint32), sc_h.astype(np.float32).view(np.uint32))
    assert np.array_equal(best, np.argmax(sc_h, axis=1))


def test_learned_run_equals_composed_host_path(ctx, setup):
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    sym = synth.symmetry_matrices(api.point_group("D", 7))
    cfg = proposal.ProposalConfig(S=32, channels=8, nconv=3, class_deg=30.0, topk=4,
                                  threshold=0.0, group=16)
    lp = proposal.LearnedProposal(ctx, gs, sym, "D", 7, cfg, seed=9)
    fsct = synth.default_fsct(nsam, optics.psize)
    ring = optics.gpar(nsam, 26.0)[[7, 8, 5, 6]]
    top, prob = lp.propose(datc, ctff)
    top_h, prob_h = top.cpu().numpy(), prob.cpu().numpy()
    # an untrained network: put the threshold at the median confidence so that
    # both kinds of list occur
    lp.cfg = dataclasses.replace(cfg, threshold=float(np.median(prob_h[:, 0])))
    res, conf = lp.run(outc, ctff, datc, fsct, ring, topk=8)
    lists, groups, conf_h, valid = host_lists(lp, top_h, prob_h, lp.cfg.threshold, gs.norient,
                                              gs.group)
    assert np.array_equal(conf, np.array(conf_h, bool)) and 0 < conf.sum() < len(conf)
    assert lp.last_valid == sum(v for v, c in zip(valid, conf_h) if c)
    npart = outc.shape[0]
    o_all = np.zeros((npart, 8), np.int32)
    r_all = np.zeros((npart, 8), np.int32)
    s_all = np.zeros((npart, 8), np.float32)
    for p in range(npart):
        one = slice(p, p + 1)
        if conf_h[p]:
            cand = torch.from_numpy(lists[p][None, :].copy()).to(DEV)
            o, r, s = ctx.search_list(nsam, gs.maxr1, gs.maxr2, gs.rbfact, gs.search_bank,
                                      outc[one].contiguous(), ctff[one].contiguous(), cand,
                                      cfg.group, topk=8)
        else:
            cand = torch.arange(gs.norient, dtype=torch.int32, device=DEV)[None, :].contiguous()
            o, r, s = ctx.search_list(nsam, gs.maxr1, gs.maxr2, gs.rbfact, gs.search_bank,
                                      outc[one].contiguous(), ctff[one].contiguous(), cand,
                                      gs.group, topk=8)
        o_all[p], r_all[p], s_all[p] = o.cpu().numpy()[0], r.cpu().numpy()[0], s.cpu().numpy()[0]
    assert np.array_equal(res.cand_o, o_all)
    assert np.array_equal(res.cand_score.view(np.uint32), s_all.view(np.uint32))
    sc, shifts = gs.rescore(outc, ctff, torch.from_numpy(o_all).to(DEV),
                            torch.from_numpy(r_all).to(DEV), fsct, ring)
    best = np.argmax(sc, axis=1)
    ar = np.arange(npart)
    assert np.array_equal(res.view, o_all[ar, best])
    assert np.array_equal(res.score.view(np.uint32), sc[ar, best].astype(np.float32).view(np.uint32))
    assert np.array_equal(res.shift.view(np.uint32), shifts[ar, best].view(np.uint32))
    assert np.array_equal(res.euler, gs.euler[np.maximum(o_all[ar, best], 0)])


@pytest.mark.parametrize("kind", ["annulus", "separable"])
def test_invalid_entries_score_as_empty_columns(ctx, setup, kind):
    """List entries outside the bank (-1, >= norient) are empty columns: they
    never reach the top-K, and the valid entries keep their scores."""
    nsam, a3df, gs, optics, parts, outc, ctff, datc = setup
    bank = gs.bank if kind == "annulus" else gs.bank_sep
    o1, c1 = outc[:3].contiguous(), ctff[:3].contiguous()
    g = gs.group
    clean = [np.arange(v * g, (v + 1) * g, dtype=np.int32) for v in (2, 5, 9)]
    dirty = []
    for p, rows in enumerate(clean):
        bad = np.full(g, -1 if p != 1 else gs.norient + 7, np.int32)
        dirty.append(np.concatenate([bad, rows, bad]))
    loff, grp, cand = ragged(clean, [g] * 3)
    loff2, grp2, cand2 = ragged(dirty, [g] * 3)
    o_c, r_c, s_c = [x.cpu().numpy() for x in ctx.search_ragged(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, o1, c1, loff, grp, cand, 4)]
    o_d, r_d, s_d = [x.cpu().numpy() for x in ctx.search_ragged(
        nsam, gs.maxr1, gs.maxr2, gs.rbfact, bank, o1, c1, loff2, grp2, cand2, 4)]
    # one view per particle is valid: the best entry is that view's, same bits
    assert np.array_equal(o_d[:, 0], o_c[:, 0]) and np.array_equal(r_d[:, 0], r_c[:, 0])
    assert np.array_equal(s_d[:, 0].view(np.uint32), s_c[:, 0].view(np.uint32))
    assert np.all((o_d[:, 1:] == -1) | (s_d[:, 1:] < -1e37))
