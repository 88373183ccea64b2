"""GPU kernels of the learned orientation proposal against the FP64 oracle
(oracle/cnn_oracle.py, itself pinned by finite differences) and numpy.

Bars: logits, losses and every gradient segment within 1e-5 relative L2 of
the oracle evaluated on the same BF16-rounded weights; Adam's master copy
within 1e-6; class assignment identical except at FP32 near-ties; network
input images within 1e-4 of a numpy restatement; the list search equal to the
exhaustive search on the same orientations.  The training test shows what
the trained network gains over the search without it at the same cost.
"""
import math
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import cnn_oracle as CO  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, proposal, search, synth  # noqa: E402

DEV = "cuda"
rng = np.random.default_rng(20261202)


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


SHAPES = [(32, 8, 3, 10), (16, 4, 1, 3), (64, 16, 4, 37)]


# the last case has 40 x 32^2 first-layer outputs per channel: the bias
# gradient is summed over three splits
@pytest.mark.parametrize("shape,B", [(s, 6) for s in SHAPES] + [(SHAPES[2], 40)])
def test_forward_and_gradients_vs_oracle(ctx, shape, B):
    S, C, L, K = shape
    net = proposal.Network(ctx, shape, seed=7, device=DEV)
    img = torch.from_numpy(rng.standard_normal((B, S, S)).astype(np.float32)).to(DEV)
    lab = rng.integers(0, K, B).astype(np.int32)
    wq = net.wq.cpu().numpy()
    # the working copy holds BF16 values for weights and exact biases
    wm = CO.weight_mask(shape)
    assert np.array_equal(CO.bf16_round(wq[wm]), wq[wm])
    logits = net.forward(img).cpu().numpy()
    ref, _ = CO.forward(wq, shape, img.cpu().numpy())
    e = rel_l2(logits, ref)
    print(f"{shape}: logits rel L2 {e:.1e}")
    assert e < 1e-5
    loss = ctx.cnn_gradients(shape, net.wq, img, torch.from_numpy(lab).to(DEV), net.grad)
    rl, rg = CO.loss_grad(wq, shape, img.cpu().numpy(), lab)
    assert np.max(np.abs(loss.cpu().numpy() - rl)) < 1e-5 * max(1.0, np.abs(rl).max())
    g = net.grad.cpu().numpy()
    offs, fw, fb, F, n = CO.layout(*shape)
    segs = [(f"conv{l}.w", w) for l, (w, b, ci) in enumerate(offs)] + \
           [(f"conv{l}.b", b) for l, (w, b, ci) in enumerate(offs)] + [("fc.w", fw), ("fc.b", fb)]
    for name, (a, b) in segs:
        e = rel_l2(g[a:b], rg[a:b])
        print(f"  {name}: grad rel L2 {e:.1e}")
        assert e < 1e-5, name


@pytest.mark.parametrize("shape,B", [(SHAPES[0], 6), (SHAPES[2], 40), ((64, 64, 4, 424), 6),
                                     ((128, 64, 5, 424), 6)])
def test_mixed_precision_gradients_vs_oracle(ctx, shape, B):
    """fr_cnn_gradients_bf16 (the training step: BF16 GEMM operands, FP32
    accumulation) against the FP64 oracle with the same operand rounding
    (loss_grad with bf16_ops): losses within 1e-3 and every gradient segment
    within 5e-3 relative L2.  The differences are the FP32 accumulation order
    and the operands whose FP32 value sits at a BF16 rounding boundary (one
    BF16 step apart in the two); they compound through the layers into the
    first convolution's gradient, a strongly cancelling sum over every
    position.  The distance to the exact FP64 gradient (the price of the
    mixed precision itself, ~1e-1 at the trained shape) is printed beside it."""
    S, C, L, K = shape
    net = proposal.Network(ctx, shape, seed=11, device=DEV)
    img = torch.from_numpy(rng.standard_normal((B, S, S)).astype(np.float32)).to(DEV)
    lab = rng.integers(0, K, B).astype(np.int32)
    wq = net.wq.cpu().numpy()
    loss = ctx.cnn_gradients(shape, net.wq, img, torch.from_numpy(lab).to(DEV), net.grad,
                             bf16=True)
    rl, rg = CO.loss_grad(wq, shape, img.cpu().numpy(), lab, bf16_ops=True)
    _, xg = CO.loss_grad(wq, shape, img.cpu().numpy(), lab)
    el = rel_l2(loss.cpu().numpy(), rl)
    print(f"{shape} B={B}: loss rel L2 {el:.1e}")
    assert el < 1e-3
    g = net.grad.cpu().numpy()
    offs, fw, fb, F, n = CO.layout(*shape)
    segs = [(f"conv{l}.w", w) for l, (w, b, ci) in enumerate(offs)] + \
           [(f"conv{l}.b", b) for l, (w, b, ci) in enumerate(offs)] + [("fc.w", fw), ("fc.b", fb)]
    for name, (a, b) in segs:
        e = rel_l2(g[a:b], rg[a:b])
        print(f"  {name}: grad rel L2 {e:.1e} (to the FP64 gradient {rel_l2(g[a:b], xg[a:b]):.1e})")
        assert e < 5e-3, name


def test_adam_vs_oracle(ctx):
    shape = (32, 8, 3, 10)
    net = proposal.Network(ctx, shape, seed=9, device=DEV)
    n = net.nparam
    master = net.master.cpu().numpy().astype(np.float64)
    m = np.zeros(n)
    v = np.zeros(n)
    wm = CO.weight_mask(shape)
    for step in range(1, 4):
        gr = (rng.standard_normal(n) * 0.01).astype(np.float32)
        net.grad.copy_(torch.from_numpy(gr))
        net.step = step - 1
        ctx.cnn_adam(shape, net.master, net.wq, net.grad, net.m, net.v, step, 1e-3)
        master, wq, m, v = CO.adam(master, gr.astype(np.float64), m, v, step, 1e-3, wm)
    gm = net.master.cpu().numpy()
    assert np.max(np.abs(gm - master)) < 1e-6 * max(1.0, np.abs(master).max())
    gq = net.wq.cpu().numpy()
    same = np.mean(gq == CO.bf16_round(gm) * wm + gm * (~wm))
    assert same == 1.0
    # and the FP64 master lands on the same BF16 values except at rounding ties
    assert np.mean(CO.bf16_round(master.astype(np.float32))[wm] == gq[wm]) > 0.999


def test_propose_vs_numpy(ctx):
    B, K, k = 9, 50, 8
    z = rng.standard_normal((B, K)).astype(np.float32) * 3
    top, prob = ctx.cnn_propose(torch.from_numpy(z).to(DEV), k)
    e = np.exp(z.astype(np.float64) - z.max(axis=1, keepdims=True))
    p = e / e.sum(axis=1, keepdims=True)
    want = np.argsort(-p, axis=1, kind="stable")[:, :k]
    assert np.array_equal(top.cpu().numpy(), want)
    assert np.allclose(prob.cpu().numpy(), np.take_along_axis(p, want, 1), rtol=1e-5)


def _crop_flip_numpy(nsam, S, datc, ctff):
    """numpy restatement of the network input: crop, sign-flip, centre, c2r,
    normalise."""
    jc = nsam // 2 + 1
    js = S // 2 + 1
    out = []
    for b in range(datc.shape[0]):
        h = np.zeros((S, js), np.complex128)
        c = ctff[b, 0] + np.conj(ctff[b, 1])
        for m in range(S):
            j = m if m < js else m - S
            for l in range(js):
                if l < S // 2 and -S // 2 < j < S // 2:
                    row = j % nsam
                    sg = -1.0 if c[row, l].real < 0 else 1.0
                    if (l + j) % 2:
                        sg = -sg
                    h[m, l] = datc[b, row, l] * sg
        # c2r semantics of the library: Im of the l = 0 column is ignored
        x = np.fft.irfft2(h, s=(S, S)) * S * S
        x = (x - x.mean()) / x.std()
        out.append(x)
    return np.array(out)


def test_cnn_images_vs_numpy(ctx):
    nsam, S, B = 64, 32, 3
    jc = nsam // 2 + 1
    datc = (rng.standard_normal((B, nsam, jc)) + 1j * rng.standard_normal((B, nsam, jc))
            ).astype(np.complex64)
    ctff = (rng.standard_normal((B, 2, nsam, jc)) + 0.3j * rng.standard_normal((B, 2, nsam, jc))
            ).astype(np.complex64)
    # numpy's irfft2 is ifft over rows then irfft over columns, dropping Im at
    # l = 0: the library's row-column c2r convention
    got = ctx.cnn_images(nsam, S, torch.from_numpy(datc).to(DEV),
                         torch.from_numpy(ctff).to(DEV)).cpu().numpy()
    want = _crop_flip_numpy(nsam, S, datc, ctff)
    err = np.max(np.abs(got - want))
    print(f"max |img - numpy| {err:.1e}")
    assert err < 1e-4


def _oracle_section(L, nsam, S, vol, eul, optics, spar, gpar):
    """Noise-free training image from the oracle's PEXTRACT and CTF."""
    jc = nsam // 2 + 1
    js = S // 2 + 1
    ctff = np.zeros((2, nsam, jc), np.complex64)
    for m in range(nsam):
        j = m if m < jc else m - nsam
        for l in range(jc):
            o = np.zeros(2, np.float32)
            L.or32_ctf(float(gpar[0]), float(gpar[1]), float(gpar[2]), float(gpar[3]),
                       float(spar[2]), float(spar[3]), float(spar[4]), float(gpar[4]), l, j, 0.0,
                       0.0, o)
            ctff[0, m, l] = o[0] + 1j * o[1]
            L.or32_ctf(float(gpar[0]), float(gpar[1]), float(gpar[2]), float(gpar[3]),
                       float(spar[2]), float(spar[3]), float(spar[4]), float(gpar[4]), -l, -j,
                       0.0, 0.0, o)
            ctff[1, m, l] = o[0] + 1j * o[1]
    sec = np.zeros(nsam * jc, np.complex64)
    L.or32_pextract(nsam, sec, vol.reshape(-1), *[float(x) for x in eul], ctff.reshape(-1),
                    O.IDENTITY)
    sec = sec.reshape(nsam, jc)
    datc = np.zeros((1, nsam, jc), np.complex64)
    for m in range(nsam):
        j = m if m < jc else m - nsam
        for l in range(jc):
            ph = -(np.float32(spar[0]) * l + np.float32(spar[1]) * j)
            datc[0, m, l] = sec[m, l] * np.exp(1j * ph)
    return _crop_flip_numpy(nsam, S, datc, ctff[None])[0]


def test_cnn_synth_vs_oracle(ctx):
    L = O.oracle()
    nsam, S = 64, 32
    vol = synth.half_volume(synth.blob_phantom(nsam, 12, 20.0, 5, device=DEV)).cpu().numpy()
    optics = synth.Optics(psize=2.0)
    gpar = optics.gpar(nsam, 26.0)
    B = 3
    eul = synth.random_orientations(B, rng)
    dm = api.rotation(eul, amag=0.0)
    spar = np.zeros((B, 8), np.float32)
    spar[:, 0:2] = rng.uniform(-0.2, 0.2, (B, 2))
    spar[:, 2] = 1.5e4
    spar[:, 3] = 1.4e4
    spar[:, 4] = 0.3
    spar[:, 5] = 1e12  # noise-free
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    got = ctx.cnn_synth(nsam, S, t(vol), t(dm), t(spar), gpar, math.pi * 26.0 ** 2, 1).cpu().numpy()
    for b in range(B):
        want = _oracle_section(L, nsam, S, vol, eul[b], optics, spar[b], gpar)
        err = np.max(np.abs(got[b] - want))
        print(f"image {b}: max |synth - oracle| {err:.1e}")
        assert err < 2e-4
    # noise: deterministic per seed, and the clean/noisy correlation falls with the SNR
    spar[:, 5] = 0.05
    n1 = ctx.cnn_synth(nsam, S, t(vol), t(dm), t(spar), gpar, math.pi * 26.0 ** 2, 5).cpu().numpy()
    n1b = ctx.cnn_synth(nsam, S, t(vol), t(dm), t(spar), gpar, math.pi * 26.0 ** 2, 5).cpu().numpy()
    n2 = ctx.cnn_synth(nsam, S, t(vol), t(dm), t(spar), gpar, math.pi * 26.0 ** 2, 6).cpu().numpy()
    assert np.array_equal(n1, n1b) and not np.array_equal(n1, n2)
    spar[:, 5] = 1.0
    h = ctx.cnn_synth(nsam, S, t(vol), t(dm), t(spar), gpar, math.pi * 26.0 ** 2, 5).cpu().numpy()
    c_lo = np.mean([np.corrcoef(n1[b].ravel(), got[b].ravel())[0, 1] for b in range(B)])
    c_hi = np.mean([np.corrcoef(h[b].ravel(), got[b].ravel())[0, 1] for b in range(B)])
    print(f"corr(clean, noisy): SNR 0.05 {c_lo:.3f}, SNR 1 {c_hi:.3f}")
    assert 0.0 < c_lo < c_hi < 1.0


def _sym_trace_np(a, b, S):
    A = a.reshape(3, 3).T.astype(np.float64)
    Bm = b.reshape(3, 3).T.astype(np.float64)
    return np.sum((S @ A) * Bm)


def test_orient_classes_vs_numpy(ctx):
    sym = synth.symmetry_matrices(api.point_group("D", 7))
    cls = proposal.OrientationClasses(30.0, "D", 7, sym)
    eul = synth.random_orientations(300, rng)
    dm = api.rotation(eul, amag=0.0)
    S = np.stack(sym)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)  # noqa
    lab = ctx.orient_nearest(t(dm), t(cls.dm), t(S.reshape(-1, 9))).cpu().numpy()
    tr = np.array([[max(_sym_trace_np(dm[i], cls.dm[c], s) for s in S) for c in range(cls.n)]
                   for i in range(len(dm))])
    want = tr.argmax(axis=1)
    bad = lab != want
    # disagreements only where FP32 cannot separate the two traces
    assert np.all(np.abs(tr[bad, want[bad]] - tr[bad, lab[bad]]) < 1e-5)
    assert bad.mean() < 0.01
    mask = ctx.orient_within(t(dm), t(cls.dm), t(S.reshape(-1, 9)), 20.0).cpu().numpy()
    bits = (mask.view(np.uint32)[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1
    bits = bits.reshape(cls.n, -1)[:, :len(dm)].T.astype(bool)
    want_w = tr >= 1 + 2 * math.cos(math.radians(20.0))
    near = np.abs(tr - (1 + 2 * math.cos(math.radians(20.0)))) < 1e-5
    assert np.array_equal(bits[~near], want_w[~near])
    # symmetry-equivalent orientations get the same class
    e2 = eul.copy()
    dm2 = np.stack([(S[3] @ dm[i].reshape(3, 3).T).T.reshape(9) for i in range(len(dm))])
    lab2 = ctx.orient_nearest(t(dm2), t(cls.dm), t(S.reshape(-1, 9))).cpu().numpy()
    assert np.mean(lab2 == lab) > 0.99
    del e2


@pytest.fixture(scope="module")
def small_search(ctx):
    nsam = 64
    vol = synth.blob_phantom(nsam, 20, 24.0, 0x51, sym=api.point_group("D", 7), device=DEV)
    a3df = synth.half_volume(vol)
    gs = search.GlobalSearch(ctx, a3df, dang=10.0, asym="D", nasym=7, maxr1=1, maxr2=14,
                             rbfact=0.0, dtype=torch.bfloat16, iran=7)
    return nsam, a3df, gs


def test_list_search_equals_exhaustive(ctx, small_search):
    nsam, a3df, gs = small_search
    optics = synth.Optics(psize=2.0)
    parts = synth.make_particles(ctx, a3df, 8, 3, optics, snr=0.5, ri=26.0)
    outc, ctff, _ = gs_prep(ctx, nsam, parts, optics)
    o_d, r_d, s_d = gs.candidates(outc, ctff, topk=8)
    # candidate list = every orientation, in bank order, grouped by view
    cand = torch.arange(gs.norient, dtype=torch.int32, device=DEV).repeat(8, 1).contiguous()
    o_l, r_l, s_l = ctx.search_list(nsam, gs.maxr1, gs.maxr2, gs.rbfact, gs.search_bank, outc, ctff,
                                    cand, gs.group, topk=8)
    assert np.array_equal(o_d.cpu().numpy()[:, 0], o_l.cpu().numpy()[:, 0])
    assert np.array_equal(r_d.cpu().numpy()[:, 0], r_l.cpu().numpy()[:, 0])
    assert np.allclose(s_d.cpu().numpy(), s_l.cpu().numpy(), rtol=1e-5, atol=1e-7)
    # a shuffled subset holding the best view: same winner, same score
    best_view = o_d.cpu().numpy()[:, 0] // gs.group
    rows = []
    for p in range(8):
        views = rng.choice(gs.norient // gs.group, 5, replace=False)
        views[0] = best_view[p]
        rng.shuffle(views)
        rows.append(np.concatenate([np.arange(v * gs.group, (v + 1) * gs.group) for v in views]))
    cand2 = torch.from_numpy(np.array(rows, np.int32)).to(DEV)
    o2, r2, s2 = ctx.search_list(nsam, gs.maxr1, gs.maxr2, gs.rbfact, gs.search_bank, outc, ctff, cand2,
                                 gs.group, topk=4)
    assert np.array_equal(o2.cpu().numpy()[:, 0], o_d.cpu().numpy()[:, 0])
    assert np.allclose(s2.cpu().numpy()[:, 0], s_d.cpu().numpy()[:, 0], rtol=1e-5)


def gs_prep(ctx, nsam, parts, optics, ri=26.0):
    """fr_prep_batch of synthetic particles: (OUTC, CTFF, DATC)."""
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, datc = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr,
                                    optics.gpar(nsam, ri))
    return outc, ctff, datc


def test_training_gains_over_no_network(ctx, small_search):
    """Train a small network on the GPU; its top-k proposals must contain the
    true class far more often than chance, and the learned search must find
    the exhaustive search's orientation far more often than a search over the
    same number of randomly chosen classes."""
    nsam, a3df, gs = small_search
    optics = synth.Optics(psize=2.0)
    sym = synth.symmetry_matrices(api.point_group("D", 7))
    cfg = proposal.ProposalConfig(S=32, channels=16, nconv=3, class_deg=30.0, topk=4,
                                  threshold=0.0, group=16)
    lp = proposal.LearnedProposal(ctx, gs, sym, "D", 7, cfg, seed=3)
    rep = lp.train(optics, 26.0, nsamples=256 * 400, batch=256, lr=3e-3, seed=11,
                   snr_range=(0.2, 1.0), log_every=100, progress=print)
    chance = cfg.topk / lp.classes.n
    print(f"classes {lp.classes.n}, held-out top-1 {rep.heldout_top1:.3f} "
          f"top-{cfg.topk} {rep.heldout_topk:.3f} (chance {chance:.3f}), "
          f"{rep.seconds:.1f}s, loss {rep.log[0][1]:.3f} -> {rep.final_loss:.3f}")
    assert rep.final_loss < rep.log[0][1]
    assert rep.heldout_topk > 0.8 and rep.heldout_topk > 4 * chance
    # search ablation on held-out particles (synthetic particles through the
    # full preparation path, not the training generator)
    parts = synth.make_particles(ctx, a3df, 64, 99, optics, snr=0.5, ri=26.0)
    outc, ctff, datc = gs_prep(ctx, nsam, parts, optics)
    fsct = synth.default_fsct(nsam, optics.psize)
    ring = optics.gpar(nsam, 26.0)[[7, 8, 5, 6]]
    truth_cls = ctx.orient_nearest(torch.from_numpy(api.rotation(parts.euler, amag=0.0)).to(DEV),
                                   lp.dmc, lp.sym).cpu().numpy()
    top, prob = lp.propose(datc, ctff)
    hit = np.mean((top.cpu().numpy() == truth_cls[:, None]).any(axis=1))
    ex = gs.run(outc, ctff, fsct, ring, topk=8)
    res, conf = lp.run(outc, ctff, datc, fsct, ring, topk=8)
    # no network: the same number of classes, chosen at random
    rtop = np.stack([rng.choice(lp.classes.n, cfg.topk, replace=False) for _ in range(64)])
    cand = torch.from_numpy(lp.candidates(rtop)).to(DEV)
    o, r, s = ctx.search_list(nsam, gs.maxr1, gs.maxr2, gs.rbfact, gs.search_bank, outc, ctff, cand,
                              cfg.group, topk=8)
    sc, sh = gs.rescore(outc, ctff, o, r, fsct, ring)
    bo = o.cpu().numpy()[np.arange(64), np.argmax(sc, axis=1)]
    step = 1.5 * 10.0
    acc = lambda e: float(np.mean(search.angular_error_deg(e, parts.euler, sym) <= step))  # noqa
    acc_ex, acc_l = acc(ex.euler), acc(res.euler)
    acc_r = acc(gs.euler[np.maximum(bo, 0)])
    print(f"held-out particles: true class in top-{cfg.topk} {hit:.3f}; within {step:g} deg of "
          f"truth: exhaustive {acc_ex:.3f}, learned {acc_l:.3f}, random classes {acc_r:.3f}")
    same_l = acc_l
    same_r = acc_r
    assert hit >= 0.8 and acc_l >= acc_ex - 0.1 and same_l > same_r + 0.3


@pytest.mark.parametrize("shape,B", [((32, 64, 3, 10), 6), ((64, 64, 4, 424), 12),
                                     ((128, 64, 5, 424), 12)])
def test_forward_bf16_vs_oracle(ctx, shape, B):
    """BF16-activation inference against the oracle that rounds the same
    activations (FP64 arithmetic), and its top class against the FP32 forward."""
    S = shape[0]
    net = proposal.Network(ctx, shape, seed=11, device=DEV)
    img = torch.from_numpy(rng.standard_normal((B, S, S)).astype(np.float32)).to(DEV)
    got = ctx.cnn_forward_bf16(shape, net.wq, img).cpu().numpy()
    ref = CO.forward_bf16_act(net.wq.cpu().numpy(), shape, img.cpu().numpy())
    e = rel_l2(got, ref)
    f32 = net.forward(img).cpu().numpy()
    agree = np.mean(np.argmax(got, 1) == np.argmax(f32, 1))
    print(f"{shape}: BF16-activation logits rel L2 {e:.1e} vs oracle, "
          f"{rel_l2(got, f32):.1e} vs FP32; top-1 agreement {agree:.2f}")
    assert e < 1e-3
    assert rel_l2(got, f32) < 3e-2
