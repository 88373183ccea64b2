"""Benchmark: global-search refinement at 256^2 on one MI355X.

Workload (the mode-3 global search the metric is quoted on): a 256^3
D7-symmetric Gaussian-blob phantom at 1.0 A/px; synthetic 256x256 particles
at uniform random orientations, +-3 A shifts, CTF at 300 kV / Cs 2.7 mm /
7% amplitude contrast, defocus U(1,3) um with up to 0.1 um astigmatism, SNR
0.05.  The search grid is SEARCHANG at 5 deg for D7: 130 views x 360
in-plane angles = 46,800 orientations, 15x15 NSAMR shift hypotheses, 8 A
band; the 32 best views are rescored with the FP32 CC3M.

One step = one batch of B particles (1,024 by default) already resident in HBM through
particle preparation (r2c + CTFAPPLY), the search and the FP32 rescoring;
the timed steps cycle over a pool of 5,120 distinct particles.  Searches:
  * learned + BF16 (`value`): the proposal network's 8 most probable
    15-degree classes select the bank orientations to score (the BF16
    search with the shifts applied in two separable matrix-core stages);
    particles whose top class probability is below 0.3 take the exhaustive
    search instead;
  * exhaustive BF16: every orientation, the same separable BF16 search;
  * plain FP32: every orientation, the 225-shift FP32 MFMA GEMM (the
    baseline of the metric);
  * reference-exact PSEARCH on a few particles: CCP with PAD and ROTATE2D as
    the reference computes them.
The network is trained first, on the GPU, on projections of the same
reference; `particles_per_s_including_training` charges its time to the
20,000 particles of the dataset.

Accuracy against the truth and between searches, and the FSC(0.143) of the
half-maps each search's parameters give, are measured on every particle of
the pool.  Also reported: LMAIN's steps after the search on the learned
winners (PREFINE, PRESB, SIGMA) per batch and the rates with them, insertion
GB/s (the refined particles inserted into
the two half-maps with the 14 D7 copies), the wall time of the device
finalisation of the learned half-maps (STORESHIFT, 3D FFTs, CORRECT3D, MASK3D,
SHELTEST at 256^3) with its masked FSC, the mode-1 local refinement at 128^2
(local grid search, PREFINE, insertion of 10,000 particles), the uneven
96/160/256 workload, the roofline of the dominant kernel from HIP events on
its own stream, and a CPU baseline (the oracle's port of PSEARCH/CCP and
PREFINE on a bounded sample, all host cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1024,
                    help="particles per step (the learned, exhaustive and plain FP32 searches "
                         "alike)")
    ap.add_argument("--fp32-steps", type=int, default=4)
    ap.add_argument("--pool", type=int, default=5120,
                    help="distinct particles: the timed steps cycle over them; accuracy and "
                         "FSC are measured on all of them for every search")
    ap.add_argument("--exact-particles", type=int, default=16,
                    help="particles of the reference-exact PSEARCH leg (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: all host cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-views", type=int, default=12,
                    help="views of PSEARCH per CPU-baseline particle (timed, then scaled to all)")
    ap.add_argument("--exh-steps", type=int, default=4, help="timed steps of the exhaustive BF16 search")
    ap.add_argument("--train-samples", type=int, default=102_400,
                    help="synthetic projections the proposal network is trained on (400 Adam "
                         "steps of 256): near the optimum amortised over a 20,000-particle "
                         "dataset, where more samples raise the confident fraction but cost "
                         "more than the fallbacks they save")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="one device copy of this many MiB first (PMC calibration of the "
                         "HBM byte counters against a known size)")
    ap.add_argument("--c2-particles", type=int, default=10000,
                    help="particles of the mode-1 local refinement at 128^2 (0: skip)")
    ap.add_argument("--post-batches", type=int, default=2,
                    help="batches on which LMAIN's steps after the search (PREFINE of the "
                         "winner, PRESB, SIGMA) are timed (0: skip)")
    ap.add_argument("--uneven-batches", type=int, default=1,
                    help="batches of 256 per 'unit' of the 3:2:1 uneven workload (0: skip)")
    return ap.parse_args()


NSAM = 256
PSIZE = 1.0
RI = 115.0           # particle radius (px)
DANG = 5.0           # search step (deg)
RMAX1_A, RMAX2_A = 100.0, 8.0
SNR = 0.05
TOPK = 32
PEAK_BF16 = 2.5e15   # MI355X dense BF16 MFMA, FLOP/s
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12


def cc3m_issued_mfma(nsam, ri3):
    """16x16x4 FP32 MFMA issued per fused CC3M item with MAXR2 = 32 (the
    band-pruned transforms, folded over the cos/sin symmetry of the DFT
    table): the column c2c takes (N/2 / 16) x 4 folded tiles of 16 products,
    the row pass 32 per folded tile (16 forward, 16 back) for each 16-row
    block and each tile that reaches inside the mask radius^2 ri3 (the
    kernel's own test), the column r2c 8 waves x (N/4 / 2 + 1) x 4; 2048
    flops each.  The ring-mean pass (AMAG != 1 only) is not counted."""
    h = nsam // 2

    def lo(v0):
        a = v0 if v0 < h else v0 - nsam
        b = a + 15
        return 0 if a <= 0 <= b else min(a * a, b * b)
    rows = sum(1 for rb in range(nsam // 16) for nt in range(h // 16)
               if not lo(16 * nt) + lo(16 * rb) > ri3)
    return (h // 16) * 4 * 16 + rows * 32 + 8 * (nsam // 8 + 1) * 4


CC3M_FLOPS = cc3m_issued_mfma(NSAM, (RI + 3.0) ** 2) * 2048.0
# the algorithmic unit of one CC3M evaluation: its two NSAM^2 FFTs (the c2r
# and r2c of cc3m.f:113-115), 5 N^2 log2(N^2) flops each
CC3M_FFT_FLOPS = 2 * 5.0 * NSAM * NSAM * math.log2(NSAM * NSAM)
# BF16 MFMA work of the separable search per (particle, template): 16 stage-1
# and 2 stage-2 v_mfma_f32_16x16x32_bf16; the template read is 8 KiB
SEP_FLOPS = 18 * 16 * 16 * 32 * 2.0
SEP_BYTES = 8192.0
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "hbm_traffic_256_d7_pmc.md")


def measured_traffic(key):
    """HBM bytes per launch of one kernel shape, as two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE, each scaled by a plain copy of known size)
    measured them; the summary is a text file under profiles/.  None when that
    shape was not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {key}:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from frealign_amd import api, proposal, search, synth

    dev = "cuda"
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
        del src, dst
    ctx = api.Context(0)
    optics = synth.Optics(psize=PSIZE)
    sym = api.point_group("D", 7)
    mats = synth.symmetry_matrices(sym)
    maxr1 = int(NSAM * PSIZE / RMAX1_A)
    maxr2 = int(NSAM * PSIZE / RMAX2_A)
    t0 = time.time()
    vol = synth.blob_phantom(NSAM, 40, 0.4375 * NSAM, seed=0xC3, sym=sym, device=dev)
    a3df = synth.half_volume(vol)
    del vol
    B = args.batch
    npool = max(B, (args.pool // B) * B)
    nbatch = npool // B
    parts = synth.make_particles(ctx, a3df, npool, seed=0xC3 + 1, optics=optics, snr=SNR, ri=RI,
                                 shift_a=3.0)
    gpar = optics.gpar(NSAM, RI)
    pssnr = torch.from_numpy(synth.default_pssnr(NSAM, RI)).to(dev)
    fsct = synth.default_fsct(NSAM, PSIZE)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    ppar_all = parts.ppar(dev)
    gs16 = search.GlobalSearch(ctx, a3df, dang=DANG, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                               rbfact=0.0, dtype=torch.bfloat16, iran=-0xC3)
    kc, kpad = ctx.search_layout(NSAM, maxr1, maxr2)
    norient = gs16.norient
    setup_s = time.time() - t0
    log(f"setup {setup_s:.1f}s: {norient} orientations, annulus {kc}, batch {B}, "
        f"{npool} distinct particles")

    def batch(i):
        lo = (i * B) % npool
        return parts.images[lo:lo + B], ppar_all[lo:lo + B], lo

    ar = np.arange(B)

    def prep(i, c=None):
        img, pp, lo = batch(i)
        ctff, outc, datc = (c or ctx).ctfapply(NSAM, img.clone(), pp, pssnr, gpar)
        return ctff, outc, datc, lo

    def exhaustive(gs, ctff, outc):
        top_o, top_r, top_s = gs.candidates(outc, ctff, TOPK)
        sc, shifts = gs.rescore(outc, ctff, top_o, top_r, fsct, ring)
        best = np.argmax(sc, axis=1)
        o = top_o.cpu().numpy()
        return o[ar, best], shifts[ar, best]

    def step_exhaustive(gs, i):
        ctff, outc, datc, lo = prep(i)
        return exhaustive(gs, ctff, outc)

    def timed(fn, nsteps, first):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(nsteps):
            fn(first + i)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    # ------------------------------------------------------ proposal network
    cfg = proposal.ProposalConfig()
    lp = proposal.LearnedProposal(ctx, gs16, mats, "D", 7, cfg, seed=0xC5)
    log(f"proposal: {lp.classes.n} classes at {cfg.class_deg:g} deg, "
        f"{int(lp.cell_groups.sum())} cell groups of {cfg.group}; training on "
        f"{args.train_samples} synthetic projections")
    rep = lp.train(optics, RI, nsamples=args.train_samples, batch=256, lr=2e-3, seed=0xC5,
                   progress=log, log_every=100)
    log(f"trained in {rep.seconds:.1f}s: held-out top-1 {rep.heldout_top1:.3f}, "
        f"top-{cfg.topk} {rep.heldout_topk:.3f}")

    # ------------------------------------------------------- learned + BF16
    for i in range(args.warmup):
        ctff, outc, datc, lo = prep(i)
        lp.run(outc, ctff, datc, fsct, ring, TOPK)
    ctx.profile(True)
    for k in ("search_gemm", "search_list", "cc3m", "prep", "cnn"):
        ctx.profile_read(k)
    valid = []

    def learned_pipelined(nsteps, first):
        """The learned steps with batch i+1 enqueued before batch i is
        unpacked on the host (two batches in flight), all inside the timed
        region."""
        pending = None
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(first, first + nsteps):
            ctff, outc, datc, lo = prep(i)
            nxt = lp.launch(outc, ctff, datc, fsct, ring, TOPK)
            if pending is not None:
                lp.finish(pending)
                valid.append(lp.last_valid)
            pending = nxt
        lp.finish(pending)
        valid.append(lp.last_valid)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    dtl = learned_pipelined(args.steps, args.warmup)
    fb_ms, fb_n = ctx.profile_read("search_gemm")
    cnn_ms, cnn_n = ctx.profile_read("cnn")
    gemm_ms, gemm_n = ctx.profile_read("search_list")
    cc3m_ms, cc3m_n = ctx.profile_read("cc3m")
    prep_ms, prep_n = ctx.profile_read("prep")
    ctx.profile(False)
    ppsl = B * args.steps / dtl
    log(f"learned + bf16: {ppsl:.1f} particles/s")

    # the same steps on two library contexts (each its own workspace) and two
    # streams of default priority, consecutive batches alternating: one
    # batch's rescoring overlaps the next one's preparation and search
    ctx2 = api.Context(0)
    ctxs = (ctx, ctx2)
    streams = (torch.cuda.Stream(), torch.cuda.Stream())
    torch.cuda.synchronize()
    for i in range(2):
        with torch.cuda.stream(streams[i]):
            ctff, outc, datc, lo = prep(i, ctxs[i])
            lp.run(outc, ctff, datc, fsct, ring, TOPK, ctx=ctxs[i])
    pending = None
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for i in range(args.warmup, args.warmup + args.steps):
        with torch.cuda.stream(streams[i % 2]):
            ctff, outc, datc, lo = prep(i, ctxs[i % 2])
            nxt = lp.launch(outc, ctff, datc, fsct, ring, TOPK, ctx=ctxs[i % 2])
        if pending is not None:
            lp.finish(pending)
        pending = nxt
    lp.finish(pending)
    torch.cuda.synchronize()
    pps_two = B * args.steps / (time.perf_counter() - t2)
    log(f"learned + bf16, two streams: {pps_two:.1f} particles/s")
    del ctx2, ctxs

    # ------------------------------------------------------- exhaustive BF16
    step_exhaustive(gs16, 0)
    ctx.profile(True)
    ctx.profile_read("search_gemm")
    dt16 = timed(lambda i: step_exhaustive(gs16, i), args.exh_steps, 0)
    gemm16_ms, gemm16_n = ctx.profile_read("search_gemm")
    ctx.profile(False)
    pps16 = B * args.exh_steps / dt16
    log(f"exhaustive bf16: {pps16:.1f} particles/s")

    # ---------------------------------------------------------- plain FP32
    gs32 = search.GlobalSearch(ctx, a3df, dang=DANG, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                               rbfact=0.0, dtype=torch.float32, iran=-0xC3)
    step_exhaustive(gs32, 0)
    dt32 = timed(lambda i: step_exhaustive(gs32, i), args.fp32_steps, 0)
    pps32 = B * args.fp32_steps / dt32
    log(f"plain fp32: {pps32:.1f} particles/s")

    # -------------------------------- reference-exact PSEARCH (PAD, ROTATE2D)
    exact = None
    if args.exact_particles > 0:
        ne = min(args.exact_particles, B)
        sang, iq = api.search_grid(DANG, "D", 7)
        ctff, outc, datc, lo = prep(0)
        oc, cf = outc[:ne].contiguous(), ctff[:ne].contiguous()
        ctx.psearch(NSAM, a3df, oc[:1], cf[:1], dangin=DANG, maxr1=maxr1, maxr2=maxr2, rbfact=0.0,
                    sang=sang, iquadmax=iq, ipmax=TOPK, iran=-0xC3, asym="D")
        torch.cuda.synchronize()
        t = time.perf_counter()
        par_e, tp_e, _ = ctx.psearch(NSAM, a3df, oc, cf, dangin=DANG, maxr1=maxr1, maxr2=maxr2,
                                     rbfact=0.0, sang=sang, iquadmax=iq, ipmax=TOPK, iran=-0xC3,
                                     asym="D")
        dte = time.perf_counter() - t
        o32, _ = exhaustive(gs32, ctff, outc)
        e_e = search.angular_error_deg(par_e[:, :3], parts.euler[lo:lo + ne], mats)
        e_32e = search.angular_error_deg(par_e[:, :3], gs16.euler[np.maximum(o32[:ne], 0)], mats)
        exact = {"workload": "fr_psearch: PSEARCH -> CCP with PAD and ROTATE2D as the reference "
                             "computes them (130 views x 90 in-plane steps x 4 quadrants, NSAMR "
                             "grid shifts, top-32 TESTPAR), FP32",
                 "particles": ne, "particles_per_s": ne / dte,
                 "within_one_step_of_truth": float(np.mean(e_e <= DANG)),
                 "plain_fp32_gemm_within_one_step": float(np.mean(e_32e <= DANG))}
        log(f"reference-exact PSEARCH: {ne / dte:.1f} particles/s")
    del gs32

    # ------------------------------- accuracy and FSC on every pool particle
    gs32 = search.GlobalSearch(ctx, a3df, dang=DANG, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                               rbfact=0.0, dtype=torch.float32, iran=-0xC3)
    npts = 0
    irrec = NSAM // 2
    for x in range(-irrec, irrec + 1):
        for y in range(-irrec, irrec + 1):
            if x * x + y * y <= irrec * irrec:
                npts += 1
    maps = {k: api.HalfMaps.zeros(NSAM) for k in ("learned", "exh16", "fp32")}
    chosen = {k: [] for k in maps}
    confs, valid_eval = [], []
    ins_ms = 0.0

    def insert(vols, o, sh, lo, ctff, datc, profile=False):
        recb = np.zeros((B, 10), np.float32)
        recb[:, 0:3] = gs16.euler[np.maximum(o, 0)]
        recb[:, 3:5] = sh
        recb[:, 5:8] = 1.0
        recb[:, 8] = 0.5
        recb[:, 9] = 0.5
        half = ((np.arange(B) + lo) % 2).astype(np.int32)
        if profile:
            ctx.profile(True)
            ctx.profile_read("insert")
        ctx.insert(NSAM, datc, ctff, recb, half, sym, 5.0, 40.0, PSIZE, vols)
        if profile:
            torch.cuda.synchronize()
            ms, _ = ctx.profile_read("insert")
            ctx.profile(False)
            return ms
        return 0.0

    for i in range(nbatch):
        ctff, outc, datc, lo = prep(i)
        res, conf = lp.run(outc, ctff, datc, fsct, ring, TOPK)
        confs.append(conf)
        valid_eval.append(lp.last_valid)
        chosen["learned"].append(res.view)
        ins_ms += insert(maps["learned"], res.view, res.shift, lo, ctff, datc, profile=True)
        for k, gs in (("exh16", gs16), ("fp32", gs32)):
            o, sh = exhaustive(gs, ctff, outc)
            chosen[k].append(o)
            insert(maps[k], o, sh, lo, ctff, datc)
        if (i + 1) % 5 == 0 or i + 1 == nbatch:
            log(f"evaluation: {(i + 1) * B} of {npool} particles")
    del gs32
    conf = np.concatenate(confs)
    eul = {k: gs16.euler[np.maximum(np.concatenate(v), 0)] for k, v in chosen.items()}
    err = {k: search.angular_error_deg(v, parts.euler[:npool], mats) for k, v in eul.items()}
    err_l32 = search.angular_error_deg(eul["learned"], eul["fp32"], mats)
    err_1632 = search.angular_error_deg(eul["exh16"], eul["fp32"], mats)
    ins_bytes = 216.0 * npts * sym[3] * npool
    ins_gbps = ins_bytes / (ins_ms / 1e3) / 1e9
    log(f"insertion: {ins_gbps:.1f} GB/s")
    res_fsc = {k: fsc_resolution(ctx, v, sym) for k, v in maps.items()}
    fin_ms, res_l_masked = finalisation(ctx, maps["learned"], sym)
    log(f"finalisation (STORESHIFT .. SHELTEST, 256^3): {fin_ms:.1f} ms")
    del maps
    log(f"FSC(0.143): learned {res_fsc['learned']:.2f} A, plain fp32 {res_fsc['fp32']:.2f} A "
        f"({npool} particles)")

    # ------------------------ LMAIN's per-particle steps after the search
    post = None
    if args.post_batches > 0:
        post = after_search(ctx, lp, gs16, prep, a3df, fsct, ring, maxr1, maxr2,
                            min(args.post_batches, nbatch))
        post_ms = sum(post["ms_per_batch"].values())
        post["learned_with_these_particles_per_s"] = B / (dtl / args.steps + post_ms / 1e3)
        post["plain_fp32_with_these_particles_per_s"] = B / (B / pps32 + post_ms / 1e3)
        post["speedup_with_these"] = (post["learned_with_these_particles_per_s"] /
                                      post["plain_fp32_with_these_particles_per_s"])
        log(f"after the search: PREFINE {post['ms_per_batch']['prefine']:.1f}, PRESB "
            f"{post['ms_per_batch']['presb']:.2f}, SIGMA {post['ms_per_batch']['sigma']:.2f} ms "
            f"per batch; learned with them {post['learned_with_these_particles_per_s']:.0f} "
            f"particles/s")

    # ------------------------------------------------------------ roofline
    # dominant kernel of the learned step: the FP32 CC3M rescoring or the search GEMM
    ncand = float(np.mean(valid))
    gemm_avg_s = gemm_ms / 1e3 / max(gemm_n, 1)
    cc3m_avg_s = cc3m_ms / 1e3 / max(cc3m_n, 1)
    if cc3m_avg_s >= gemm_avg_s:
        items = B * TOPK
        flops = items * CC3M_FFT_FLOPS
        achieved = flops / cc3m_avg_s
        roof = {"kernel": "k_cc3m_fused (FP32 MFMA)", "bound": "mfma", "achieved": achieved / 1e12,
                "peak": PEAK_FP32_MFMA / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA,
                "traffic": measured_traffic(f"k_cc3m_fused_{items}"),
                "traffic_source": os.path.relpath(TRAFFIC_FILE, os.path.dirname(TRAFFIC_FILE) + "/.."),
                "launch_ms": cc3m_avg_s * 1e3,
                "algorithmic_unit": "two NSAM^2 FFTs per item, 2 x 5 N^2 log2(N^2) = "
                                    f"{CC3M_FFT_FLOPS / 1e6:.2f} MFLOP at 256^2",
                "algorithmic_flops_per_launch": flops,
                "issued_mfma_TFLOPs": items * CC3M_FLOPS / cc3m_avg_s / 1e12,
                "issued_mfma_frac": items * CC3M_FLOPS / cc3m_avg_s / PEAK_FP32_MFMA}
    else:
        flops = SEP_FLOPS * ncand  # ncand: listed candidates summed over the batch
        achieved = flops / gemm_avg_s
        roof = {"kernel": "k_search_sep (candidate lists)", "bound": "mfma",
                "achieved": achieved / 1e12, "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_BF16, "traffic": None, "launch_ms": gemm_avg_s * 1e3,
                "algorithmic_flops_per_launch": flops}
    assert roof["frac"] <= 1.0, roof
    g16_avg = gemm16_ms / 1e3 / max(gemm16_n, 1)
    exh_pairs = float(B) * norient
    exh_roof = {"kernel": "k_search_sep (exhaustive, separable BF16)", "launch_ms": g16_avg * 1e3,
                "pairs_per_s": exh_pairs / g16_avg,
                "mfma_TFLOPs": exh_pairs * SEP_FLOPS / g16_avg / 1e12,
                "mfma_frac": exh_pairs * SEP_FLOPS / g16_avg / PEAK_BF16,
                "template_GBps": exh_pairs * SEP_BYTES / g16_avg / 1e9}

    # --------------------------------------- mode-1 local refinement, 128^2
    c2 = None
    if args.c2_particles > 0:
        c2 = config2_workload(ctx, args.c2_particles, log)
        log(f"config 2: {c2['particles_per_s']:.1f} particles/s, median error "
            f"{c2['median_angular_error_deg']['refined']:.2f} deg")

    # ------------------------------------------------ uneven box sizes
    uneven = None
    if args.uneven_batches > 0:
        uneven = uneven_workload(ctx, args.uneven_batches, log)

    # --------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.cpu_views, parts, a3df, optics=optics, gpar=gpar, maxr1=maxr1,
                           maxr2=maxr2, nviews_total=norient // gs16.group, ctx=ctx, pssnr=pssnr,
                           fsct=fsct, ring=ring, threads=args.cpu_threads)

    n_ds = 20000  # particles of the mode-3 dataset the network is trained for
    incl = n_ds / (n_ds / ppsl + rep.seconds)
    out = {
        "metric": "particles/s global-search refine @256² (learned+BF16 vs plain FP32 GPU), same FSC",
        "value": ppsl,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dtl / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (D7 blob phantom, CTF + white noise at SNR 0.05)",
        "config": {"workload": "mode-3 global search, 256^2, D7, 5 deg (46,800 orientations), "
                               "8 A band, 15x15 shifts, top-32 FP32 CC3M rescoring",
                   "batch_particles": B, "orientations": norient, "annulus_complex": kc,
                   "proposal": f"CNN {cfg.nconv} conv x {cfg.channels} ch on {cfg.S}^2, "
                               f"{lp.classes.n} classes at {cfg.class_deg:g} deg, top-{cfg.topk}, "
                               f"fallback below {cfg.threshold}"},
        "particles_per_s_including_training": incl,
        "including_training_note": f"network training ({rep.seconds:.2f} s) charged to the "
                                   f"{n_ds} particles of the dataset it serves",
        "plain_fp32_particles_per_s": pps32,
        "speedup_vs_plain_fp32": ppsl / pps32,
        "speedup_vs_plain_fp32_including_training": incl / pps32,
        "exhaustive_bf16_particles_per_s": pps16,
        "reference_exact_psearch": exact,
        "accuracy_particles": npool,
        "learned_within_one_step_of_truth": float(np.mean(err["learned"] <= DANG)),
        "exhaustive_bf16_within_one_step_of_truth": float(np.mean(err["exh16"] <= DANG)),
        "plain_fp32_within_one_step_of_truth": float(np.mean(err["fp32"] <= DANG)),
        "learned_vs_plain_fp32_within_one_step": float(np.mean(err_l32 <= DANG)),
        "exhaustive_bf16_vs_plain_fp32_within_one_step": float(np.mean(err_1632 <= DANG)),
        "fallback_fraction": float(np.mean(~conf)),
        "candidates_per_confident_particle": float(np.sum(valid_eval)) / max(1, int(np.sum(conf))),
        "fsc0143_A": {"learned_bf16": res_fsc["learned"], "exhaustive_bf16": res_fsc["exh16"],
                      "plain_fp32": res_fsc["fp32"], "particles": npool},
        "training": {"samples": rep.samples, "seconds": rep.seconds, "final_loss": rep.final_loss,
                     "heldout_top1": rep.heldout_top1, f"heldout_top{cfg.topk}": rep.heldout_topk},
        "finalisation_256": {"ms": fin_ms, "fsc0143_A_masked_learned": res_l_masked,
                             "steps": "STORESHIFT, 3D c2r, CORRECT3D, MASK3D, 3D r2c per half-map, "
                                      "SHELTEST"},
        "insertion_GBps": ins_gbps,
        "insertion_frac_of_hbm_peak": ins_gbps * 1e9 / PEAK_HBM,
        "learned_two_streams_particles_per_s": pps_two,
        "learned_two_streams_note": "the same learned steps on two contexts and streams, "
                                    "consecutive batches overlapping on the GPU",
        "kernel_ms_per_step": {"search_gemm_lists": gemm_ms / max(args.steps, 1),
                               "search_gemm_fallback": fb_ms / max(args.steps, 1),
                               "cnn_forward": cnn_ms / max(args.steps, 1),
                               "cc3m_rescoring": cc3m_ms / max(args.steps, 1),
                               "prep": prep_ms / max(args.steps, 1)},
        "roofline": roof,
        "exhaustive_gemm": exh_roof,
        "cpu_baseline": cpu,
        "after_search_256": post,
        "local_refinement_128": c2,
        "uneven_batches": uneven,
        "setup_s": setup_s,
    }
    print(json.dumps(out))


def after_search(ctx, lp, gs, prep, a3df, fsct, ring, maxr1, maxr2, nbatch):
    """LMAIN's per-particle steps after a mode-3 search, on the learned
    winners of nbatch batches: PREFINE of the winner (lmain.f:615-642; all
    five parameters, a shift prior of 0.05 phase units), PRESB at the refined
    parameters (lmain.f:731-736) and SIGMA (lmain.f:897-903, the 8 A band and
    the particle disc).  They are the same for the learned and the plain
    search, so they are reported beside `value` with the rates both searches
    would have with them."""
    import torch

    from frealign_amd import api
    dev = "cuda"
    fs_d = torch.from_numpy(np.asarray(fsct, np.float32)).to(dev)
    mask = np.ones(5, np.int32)
    dmask = np.array([NSAM / 2 + 1] * 3 + [RI], np.float32)
    tms = {"prefine": 0.0, "presb": 0.0, "sigma": 0.0}
    nev = []
    npart = 0
    for i in range(nbatch):
        ctff, outc, datc, lo = prep(i)
        res, _ = lp.run(outc, ctff, datc, fsct, ring, TOPK)
        n = outc.shape[0]
        npart += n
        par = np.zeros((n, 6), np.float32)
        par[:, 0:3] = gs.euler[np.maximum(res.view, 0)]
        par[:, 3:5] = res.shift
        prior = np.zeros((n, 5), np.float32)
        prior[:, 2:4] = 0.05
        prior[:, 4] = 0.01
        ones = np.ones(n, np.float32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref, ne = ctx.refine(NSAM, a3df, outc, ctff, np.arange(n), ones, par, mask, maxr1, maxr2,
                             0.0, fs_d, ring, prior)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dm = torch.from_numpy(api.rotation(ref[:, 0:3], amag=1.0)).to(dev)
        sh = torch.from_numpy(np.ascontiguousarray(ref[:, 3:5])).to(dev)
        ctx.presb(NSAM, a3df, outc, ctff, torch.arange(n, dtype=torch.int32, device=dev), dm, sh,
                  torch.ones(n, dtype=torch.float32, device=dev), ring)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ctx.sigma(NSAM, a3df, datc, ctff, np.arange(n), ref[:, 0:5], ones, maxr1,
                  PSIZE / RMAX2_A, dmask, 1.0)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        tms["prefine"] += t1 - t0
        tms["presb"] += t2 - t1
        tms["sigma"] += t3 - t2
        nev.append(ne)
    return {"batches": nbatch, "particles": npart,
            "ms_per_batch": {k: v / nbatch * 1e3 for k, v in tms.items()},
            "prefine_cc3m_evaluations_per_particle": float(np.mean(np.concatenate(nev)))}


def uneven_workload(ctx, units, log):
    """Three seeded stacks of boxes 96 / 160 / 256 (1.3 / 1.1 / 1.0 A/px) in
    ratio 3:2:1, defocus U(0.8, 3.5) um and SNR U(0.03, 0.2) per batch, each
    box size with its own reference, half-maps, autotuned CC3M and
    graph-captured score-and-insert plan; batches interleaved.  Returns
    particles/s per box size and in total, and the autotuner's choices."""
    import torch
    from frealign_amd import api, synth
    dev = "cuda"
    sym = api.point_group("C", 1)
    B = 256
    stacks = [(96, 1.3, 3), (160, 1.1, 2), (256, 1.0, 1)]
    state = {}
    for nsam, psize, ratio in stacks:
        optics = synth.Optics(psize=psize)
        ri = 0.4375 * nsam
        vol = synth.blob_phantom(nsam, 40, 0.4375 * nsam * 0.9, seed=0xC400 + nsam, device=dev)
        a3df = synth.half_volume(vol)
        del vol
        gpar = optics.gpar(nsam, ri)
        pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(dev)
        fsct = torch.from_numpy(synth.default_fsct(nsam, psize).astype(np.float32)).to(dev)
        ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
        maxr1 = max(1, int(nsam * psize / 100.0))
        maxr2 = min(32, int(nsam * psize / 8.0), nsam // 2 - 1)
        rng = np.random.default_rng(0xC400 + nsam)
        batches = []
        for k in range(ratio * units):
            snr = float(rng.uniform(0.03, 0.2))
            parts = synth.make_particles(ctx, a3df, B, seed=0xC4000 + 100 * nsam + k,
                                         optics=optics, snr=snr, ri=ri,
                                         defocus=(0.8e4, 3.5e4))
            recb = np.zeros((B, 10), np.float32)
            recb[:, 0:3] = parts.euler
            recb[:, 3:5] = parts.shift
            recb[:, 5:8] = 1.0
            recb[:, 8] = 0.5
            recb[:, 9] = 0.5
            half = (np.arange(B) % 2).astype(np.int32)
            batches.append((parts.images, parts.ppar().cpu().numpy(), recb, half))
        # autotune the CC3M launch on the first batch of this box size
        img, ppar, recb, half = batches[0]
        ctff, outc, _ = ctx.ctfapply(nsam, img.clone(), torch.from_numpy(ppar).to(dev), pssnr, gpar)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        dm = api.rotation(recb[:, 0:3], amag=1.0)
        tune = ctx.autotune_cc3m(nsam, a3df, outc, ctff, t(np.arange(B, dtype=np.int32)), t(dm),
                                 t(recb[:, 3:5]), t(recb[:, 5]), maxr1, maxr2, 0.0, fsct, ring)
        vols = api.HalfMaps.zeros(nsam)
        plan = api.Plan(ctx, nsam, B, a3df, pssnr, gpar, maxr1, maxr2, 0.0, fsct, ring, sym, 5.0,
                        40.0, psize, vols)
        plan.run(*batches[0])  # warm
        state[nsam] = dict(plan=plan, batches=batches, vols=vols, a3df=a3df, tune=tune,
                           time=0.0, n=0)
        log(f"uneven: box {nsam}: {len(batches)} batches, CC3M choice "
            f"{tune['candidates'][tune['choice']]}")
    order = []
    for k in range(6 * units):
        for nsam, _, ratio in stacks:
            if k < ratio * units:
                order.append((nsam, k))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for nsam, k in order:
        st = state[nsam]
        t1 = time.perf_counter()
        st["plan"].run(*st["batches"][k])
        st["time"] += time.perf_counter() - t1
        st["n"] += B
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    out = {"workload": "score-and-insert at known parameters, boxes 96/160/256 in ratio 3:2:1, "
                       "C1, graph-captured plan per box size",
           "total_particles_per_s": sum(st["n"] for st in state.values()) / total}
    for nsam, st in state.items():
        out[f"box{nsam}"] = {"particles_per_s": st["n"] / st["time"], "particles": st["n"],
                             "cc3m_choice": st["tune"]["candidates"][st["tune"]["choice"]]}
        st["plan"].close()
    log(f"uneven: {out['total_particles_per_s']:.1f} particles/s in total")
    return out


def fsc_resolution(ctx, vols, sym, nsam=NSAM, psize=PSIZE) -> float:
    """FSC(0.143) resolution (A) of the two half-maps: STD/VTD as in
    frealign_v9.f:793-798 (TTD / NKSUM * 0.1), STORESHIFT, SHELTEST; the
    first shell (radius k in Fourier pixels, k >= 2) below 0.143 gives
    NSAM * PSIZE / k."""
    nk = float(vols.nksum.sum().item())
    ttd = vols.ttd.cpu().numpy()
    std, vtd = ttd[0] / nk * 0.1, ttd[1] / nk * 0.1
    b, c = ctx.storeshift(nsam, std, vtd, sym[3], vols.num[0], vols.ctf2[0], vols.num[1],
                          vols.ctf2[1])
    fsc = ctx.sheltest(nsam, b, c, vols.ctf2[0], vols.ctf2[1])["fsc"]
    for k in range(2, len(fsc)):
        if fsc[k] < 0.143:
            return float(nsam * psize / k)
    return float(2.0 * psize)


def config2_workload(ctx, npart, log):
    """Mode-1 local refinement at 128^2 (the FP32 configuration): a 128^3 C1
    blob phantom (40 blobs in a radius-56 px ball, 1.2 A/px), particles at
    SNR 0.1 with +-3 A shifts and the CTF as above, started N(0, 3 deg) in
    each Euler angle and N(0, 2 A) in each shift from the truth.  Per batch of
    256: CTFAPPLY, the local grid search (+-6 steps of 1 deg = 2,197
    orientations, +-4 one-pixel shifts = +-4.8 A, CC3M on a 6 A band), PREFINE
    from its best point, and trilinear insertion into the half-maps.  Returns
    particles/s with every stage inside the timed region, the stage times,
    the angular and shift errors against the truth before and after, and the
    FSC(0.143) of the refined half-maps next to that at the true parameters."""
    import torch
    from frealign_amd import api, search, synth

    dev = "cuda"
    nsam, psize, ri, B = 128, 1.2, 56.0, 256
    nb = max(1, (npart + B - 1) // B)
    sym = api.point_group("C", 1)
    optics = synth.Optics(psize=psize)
    vol = synth.blob_phantom(nsam, 40, ri, seed=0xC2, sym=sym, device=dev)
    a3df = synth.half_volume(vol)
    del vol
    parts = synth.make_particles(ctx, a3df, nb * B, seed=0xC2 + 1, optics=optics, snr=0.1,
                                 ri=ri, shift_a=3.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(dev)
    fsct = torch.from_numpy(synth.default_fsct(nsam, psize).astype(np.float32)).to(dev)
    ring = np.array([gpar[7], gpar[8], gpar[5], gpar[6]], np.float32)
    maxr1, maxr2 = max(1, int(nsam * psize / 100.0)), int(nsam * psize / 6.0)
    rng = np.random.default_rng(0xC2)
    dph = np.float32(2 * math.pi / nsam / psize)  # phase units per A
    start = np.zeros((nb * B, 5), np.float32)
    start[:, :3] = parts.euler + np.radians(rng.normal(0, 3.0, (nb * B, 3))).astype(np.float32)
    start[:, 3:5] = parts.shift + (rng.normal(0, 2.0, (nb * B, 2)) * dph).astype(np.float32)
    sx = np.float32(2.0 * dph)
    lr = search.LocalRefinement(ctx, a3df, maxr1=maxr1, maxr2=maxr2, rbfact=0.0,
                                prior=(0.0, 0.0, sx, sx, 0.01))
    ppar = parts.ppar(dev)

    def recb_of(par):
        r = np.zeros((len(par), 10), np.float32)
        r[:, 0:5] = par[:, 0:5]
        r[:, 5:8] = 1.0
        r[:, 8] = 0.5
        r[:, 9] = 0.5
        return r

    half = (np.arange(B) % 2).astype(np.int32)

    def one(k, vols, stage):
        lo = k * B
        t0 = time.perf_counter()
        ctff, outc, datc = ctx.ctfapply(nsam, parts.images[lo:lo + B].clone(), ppar[lo:lo + B],
                                        pssnr, gpar)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        grid, ref, nev = lr.run(outc, ctff, start[lo:lo + B], fsct, ring)
        t2 = time.perf_counter()
        ctx.insert(nsam, datc, ctff, recb_of(ref), half, sym, 5.0, 40.0, psize, vols)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        stage["prep"] += t1 - t0
        stage["search_and_refine"] += t2 - t1
        stage["insert"] += t3 - t2
        return grid, ref, nev

    stage = {"prep": 0.0, "search_and_refine": 0.0, "insert": 0.0}
    one(0, api.HalfMaps.zeros(nsam), dict(stage))  # warm
    vols = api.HalfMaps.zeros(nsam)
    grids, refs, nevs = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(nb):
        g, r, n = one(k, vols, stage)
        grids.append(g)
        refs.append(r)
        nevs.append(n)
        if (k + 1) % 8 == 0 or k + 1 == nb:
            log(f"config 2: {(k + 1) * B} particles, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    grid, ref = np.concatenate(grids), np.concatenate(refs)
    e_start = search.angular_error_deg(start[:, :3], parts.euler)
    e_grid = search.angular_error_deg(grid[:, :3], parts.euler)
    e_ref = search.angular_error_deg(ref[:, :3], parts.euler)
    sh_err = lambda p: np.hypot(*(p[:, 3:5] - parts.shift).T) / dph  # noqa: E731
    res_ref = fsc_resolution(ctx, vols, sym, nsam, psize)
    del vols
    # the same particles inserted at their true parameters
    vt = api.HalfMaps.zeros(nsam)
    truth = np.concatenate([parts.euler, parts.shift], axis=1)
    for k in range(nb):
        lo = k * B
        ctff, outc, datc = ctx.ctfapply(nsam, parts.images[lo:lo + B].clone(), ppar[lo:lo + B],
                                        pssnr, gpar)
        ctx.insert(nsam, datc, ctff, recb_of(truth[lo:lo + B]), half, sym, 5.0, 40.0, psize, vt)
    res_true = fsc_resolution(ctx, vt, sym, nsam, psize)
    del vt
    n = nb * B
    return {"workload": "mode-1 local refinement, 128^2 C1, 1.2 A/px, SNR 0.1: +-6 deg at 1 deg "
                        "(2,197 orientations) x +-4.8 A at one pixel (81 shifts) scored by CC3M "
                        "on a 6 A band, then PREFINE, then trilinear insertion",
            "particles": n, "particles_per_s": n / dt,
            "stage_s": stage, "cc3m_evaluations_per_particle_prefine": float(np.mean(np.concatenate(nevs))),
            "grid_cc3m_evaluations_per_particle": lr.norient * (2 * lr.nsh + 1) ** 2,
            "median_angular_error_deg": {"start": float(np.median(e_start)),
                                         "grid": float(np.median(e_grid)),
                                         "refined": float(np.median(e_ref))},
            "within_1deg_of_truth_refined": float(np.mean(e_ref <= 1.0)),
            "median_shift_error_A": {"start": float(np.median(sh_err(start))),
                                     "refined": float(np.median(sh_err(ref)))},
            "fsc0143_A": {"refined": res_ref, "true_parameters": res_true}}


def finalisation(ctx, vols, sym):
    """The half-map finalisation of frealign_v9.f:809-869 on the device:
    STORESHIFT, then per half-map the 3D c2r, CORRECT3D (trilinear sinc^2),
    MASK3D at the particle radius, the 3D r2c, and SHELTEST.  Returns its
    wall time (ms, host syncs included) and the FSC(0.143) resolution (A) of
    the masked half-maps."""
    import torch

    nsam = NSAM
    nk = float(vols.nksum.sum().item())
    ttd = vols.ttd.cpu().numpy()
    std, vtd = ttd[0] / nk * 0.1, ttd[1] / nk * 0.1
    j = np.arange(1, 181, dtype=np.float32)
    x = j * np.float32(3.1415926535897) / np.float32(180.0)
    lut = (np.sin(x) / x).astype(np.float32)  # CALCSINC (calcsinc.f:17-19)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b, c = ctx.storeshift(nsam, std, vtd, sym[3], vols.num[0], vols.ctf2[0], vols.num[1],
                          vols.ctf2[1])
    for vol in (b, c):
        ctx.fft3d(nsam, vol, False)
        ctx.correct3d(nsam, vol, 1, 1, lut)
        ctx.mask3d(nsam, vol, RI, 0.0, 6.0)
        ctx.fft3d(nsam, vol, True)
    fsc = ctx.sheltest(nsam, b, c, vols.ctf2[0], vols.ctf2[1])["fsc"]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    res = float(2.0 * PSIZE)
    for k in range(2, len(fsc)):
        if fsc[k] < 0.143:
            res = float(nsam * PSIZE / k)
            break
    return ms, res


def cpu_baseline(nviews, parts, a3df, optics, gpar, maxr1, maxr2, nviews_total, ctx, pssnr, fsct,
                 ring, threads=0):
    """The oracle's port of the reference's CPU path for one mode-3 particle:
    PSEARCH over CCP with PAD and ROTATE2D (psearch.f, ccp.f) on `nviews` of
    the views, then one PREFINE (prefine.f, calcfx.f over VA04A on the port's
    CC3M) from the best view; the particle's time is the PSEARCH time scaled
    to all views plus IPMAX + 1 = 33 times the PREFINE (lmain.f:615-642 refines
    every TESTPAR entry and the input orientation).  One particle per thread,
    all host cores (at most 16) at once, as the reference runs one process per
    core."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import subprocess
    from concurrent.futures import ThreadPoolExecutor

    import oracle_lib as O
    L = O.oracle()
    nsam = NSAM
    ncores = os.cpu_count() or 1
    nt = threads if threads > 0 else min(16, ncores)
    img = parts.images[:nt].clone()
    ctff, outc, _ = ctx.ctfapply(nsam, img, parts.ppar("cuda")[:nt], pssnr, gpar)
    specs = outc.cpu().numpy().reshape(nt, -1).astype(np.complex64)
    cfs = ctff.cpu().numpy().reshape(nt, -1).astype(np.complex64)
    vol = a3df.cpu().numpy().reshape(-1)
    sang, iq = O.searchang(DANG, "D", 7, lib="oracle")
    fs = np.ascontiguousarray(fsct, np.float32)
    prior = np.array([0.0, 0.0, 0.1, 0.1, 0.01], np.float32)

    def one(k):
        sg = np.ascontiguousarray(sang[(k * nviews) % len(sang):][:nviews].reshape(-1))
        nv = sg.size // 3
        mask = np.array([1, 1, 1, 1, 1], np.int32)
        testpar = np.zeros(6 * TOPK, np.float32)
        phi, theta, psi, dx, dy = (C.c_float(0) for _ in range(5))
        iran = C.c_int(-0xC3 - k)
        t = time.perf_counter()
        L.or32_psearch(nsam, 1.0, specs[k], vol, DANG, maxr1, maxr2, C.byref(phi),
                       C.byref(theta), C.byref(psi), C.byref(dx), C.byref(dy), sg, nv, iq, mask,
                       0.0, cfs[k], testpar, TOPK, C.byref(iran), O.IDENTITY)
        t_ps = (time.perf_counter() - t) / nv * nviews_total
        par = np.array([phi.value, theta.value, psi.value, dx.value, dy.value], np.float32)
        ne = C.c_int(0)
        t = time.perf_counter()
        L.or32_prefine(nsam, 1.0, specs[k], vol, maxr1, maxr2, par, mask, 0.0, fs, cfs[k],
                       *[float(v) for v in ring], prior, None, C.byref(ne))
        t_pr = time.perf_counter() - t
        return t_ps, t_pr, ne.value

    t0 = time.perf_counter()
    with ThreadPoolExecutor(nt) as ex:
        res = list(ex.map(one, range(nt)))
    wall = time.perf_counter() - t0
    t_ps = float(np.mean([r[0] for r in res]))
    t_pr = float(np.mean([r[1] for r in res]))
    per_particle = t_ps + (TOPK + 1) * t_pr
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"value": nt / per_particle, "unit": "particles/s", "cores": nt, "kind": "port",
            "nproc": ncores, "cpu_model": model,
            "per_particle_s": {"psearch_all_views": t_ps, "prefine_once": t_pr,
                               "prefine_x33": (TOPK + 1) * t_pr},
            "prefine_cc3m_evaluations": float(np.mean([r[2] for r in res])),
            "sample": f"oracle port of PSEARCH/CCP (PAD, ROTATE2D) on {nviews} of {nviews_total} "
                      f"views and one PREFINE (VA04A over CC3M) per particle, {nt} particles on "
                      f"{nt} threads at once ({wall:.1f} s wall); each particle's time scaled to "
                      f"all views and {TOPK + 1} PREFINEs"}


if __name__ == "__main__":
    main()
