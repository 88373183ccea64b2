"""Benchmark: global-search refinement at 256^2 on one MI355X.

Workload (the mode-3 global search the metric is quoted on): a 256^3
D7-symmetric Gaussian-blob phantom at 1.0 A/px; synthetic 256x256 particles
at uniform random orientations, +-3 A shifts, CTF at 300 kV / Cs 2.7 mm /
7% amplitude contrast, defocus U(1,3) um with up to 0.1 um astigmatism, SNR
0.05.  One step = one batch of B particles already resident in HBM through
particle preparation (r2c + CTFAPPLY), the exhaustive D7 search (SEARCHANG at
5 deg: 130 views x 360 in-plane angles = 46,800 orientations, 15x15 NSAMR
shift hypotheses, 8 A band) as an MFMA GEMM, per-view top-32, and FP32 CC3M
rescoring of the 32 candidates.  `value` is particles/s of the BF16 path;
the plain FP32-GEMM path is timed on the same batch and reported beside it.

Also reported: insertion GB/s (the refined particles inserted into the two
half-maps with the 14 D7 copies), the roofline of the dominant kernel (the
search GEMM) from HIP events on its own stream, and a CPU baseline (the
oracle's port of PSEARCH/CCP on a bounded sample, 1 thread).
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
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-views", type=int, default=16)
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from frealign_amd import api, search, synth

    dev = "cuda"
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
    npool = B * 2
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
    log(f"setup {setup_s:.1f}s: {norient} orientations, annulus {kc}, batch {B}")

    def batch(i):
        lo = (i * B) % npool
        return parts.images[lo:lo + B], ppar_all[lo:lo + B], lo

    def step(gs, i, rescore=True):
        img, pp, lo = batch(i)
        work = img.clone()
        ctff, outc, datc = ctx.ctfapply(NSAM, work, pp, pssnr, gpar)
        top_o, top_r, top_s = gs.candidates(outc, ctff, TOPK)
        res = None
        if rescore:
            sc, shifts = gs.rescore(outc, ctff, top_o, top_r, fsct, ring)
            best = np.argmax(sc, axis=1)
            o = top_o.cpu().numpy()
            res = (o[np.arange(B), best], shifts[np.arange(B), best], lo, ctff, datc)
        return res

    # ---------------------------------------------------------------- BF16
    for i in range(args.warmup):
        step(gs16, i)
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read("search_gemm")
    results = []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        results.append(step(gs16, args.warmup + i))
    torch.cuda.synchronize()
    dt16 = time.perf_counter() - t1
    gemm_ms, gemm_n = ctx.profile_read("search_gemm")
    cc3m_ms, cc3m_n = ctx.profile_read("cc3m")
    prep_ms, prep_n = ctx.profile_read("prep")
    ctx.profile(False)
    pps16 = B * args.steps / dt16
    log(f"bf16 path: {pps16:.1f} particles/s")

    # accuracy of the BF16 path against the truth (modulo D7)
    errs = []
    for o, _, lo, _, _ in results:
        errs.append(search.angular_error_deg(gs16.euler[o], parts.euler[lo:lo + B], mats))
    err = np.concatenate(errs)

    # ---------------------------------------------------------------- FP32
    gs32 = search.GlobalSearch(ctx, a3df, dang=DANG, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                               rbfact=0.0, dtype=torch.float32, iran=-0xC3)
    step(gs32, 0)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    res32 = []
    for i in range(args.fp32_steps):
        res32.append(step(gs32, args.warmup + i))
    torch.cuda.synchronize()
    dt32 = time.perf_counter() - t2
    pps32 = B * args.fp32_steps / dt32
    log(f"fp32 path: {pps32:.1f} particles/s")
    agree = []
    for (o16, _, _, _, _), (o32, _, _, _, _) in zip(results, res32):
        agree.append(search.angular_error_deg(gs16.euler[o16], gs32.euler[o32], mats))
    agree = np.concatenate(agree)
    del gs32

    # ----------------------------------------------------------- insertion
    vols = api.HalfMaps.zeros(NSAM)
    ctx.profile(True)
    ctx.profile_read("insert")
    npts = 0
    irrec = NSAM // 2
    for x in range(-irrec, irrec + 1):
        for y in range(-irrec, irrec + 1):
            if x * x + y * y <= irrec * irrec:
                npts += 1
    ins_parts = 0
    for o, sh, lo, ctff, datc in results[:1]:
        recb = np.zeros((B, 10), np.float32)
        recb[:, 0:3] = gs16.euler[o]
        recb[:, 3:5] = sh
        recb[:, 5] = 1.0
        recb[:, 6] = 1.0
        recb[:, 7] = 1.0
        recb[:, 8] = 0.5
        recb[:, 9] = 0.5
        half = (np.arange(B) + lo) % 2
        ctx.insert(NSAM, datc, ctff, recb, half.astype(np.int32), sym, 5.0, 40.0, PSIZE, vols)
        ins_parts += B
    torch.cuda.synchronize()
    ins_ms, ins_n = ctx.profile_read("insert")
    ctx.profile(False)
    ins_bytes = 216.0 * npts * sym[3] * ins_parts
    ins_gbps = ins_bytes / (ins_ms / 1e3) / 1e9
    log(f"insertion: {ins_gbps:.1f} GB/s")

    # ------------------------------------------------------------ roofline
    gemm_avg_s = gemm_ms / 1e3 / max(gemm_n, 1)
    flops = 2.0 * B * 225 * norient * (2 * kc)
    achieved = flops / gemm_avg_s
    roof = {"kernel": "k_search_gemm<bf16>", "bound": "mfma", "achieved": achieved / 1e12,
            "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_BF16,
            "traffic": None, "launch_ms": gemm_avg_s * 1e3,
            "algorithmic_flops_per_launch": flops}

    # --------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.cpu_views, parts, a3df, ctff_for_cpu=None, optics=optics,
                           gpar=gpar, maxr1=maxr1, maxr2=maxr2, norient_views=norient // gs16.group,
                           ctx=ctx, pssnr=pssnr)

    out = {
        "metric": "particles/s global-search refine @256² (learned+BF16 vs plain FP32 GPU), same FSC",
        "value": pps16,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt16 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (D7 blob phantom, CTF + white noise at SNR 0.05)",
        "config": {"workload": "mode-3 global search, 256^2, D7, 5 deg (46,800 orientations), "
                               "8 A band, 15x15 shifts, top-32 FP32 CC3M rescoring",
                   "batch_particles": B, "orientations": norient, "annulus_complex": kc,
                   "proposal": "none (exhaustive; learned proposal not yet built)"},
        "plain_fp32_particles_per_s": pps32,
        "speedup_vs_plain_fp32": pps16 / pps32,
        "bf16_within_one_step_of_truth": float(np.mean(err <= DANG)),
        "bf16_vs_fp32_same_view_within_one_step": float(np.mean(agree <= DANG)),
        "insertion_GBps": ins_gbps,
        "insertion_frac_of_hbm_peak": ins_gbps * 1e9 / PEAK_HBM,
        "kernel_ms_per_step": {"search_gemm": gemm_ms / max(args.steps, 1),
                               "cc3m_rescoring": cc3m_ms / max(args.steps, 1),
                               "prep": prep_ms / max(args.steps, 1)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    print(json.dumps(out))


def cpu_baseline(nviews, parts, a3df, ctff_for_cpu, optics, gpar, maxr1, maxr2, norient_views,
                 ctx, pssnr):
    """The oracle's PSEARCH/CCP (a port of the reference's CPU path) on one
    particle and `nviews` of the views, 1 thread, extrapolated to all views."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import oracle_lib as O
    import torch
    L = O.oracle()
    nsam = NSAM
    jc = nsam // 2 + 1
    # the particle's CTFAPPLY output from the GPU pipeline (identical inputs)
    img = parts.images[:1].clone()
    ctff, outc, _ = ctx.ctfapply(nsam, img, parts.ppar("cuda")[:1], pssnr, gpar)
    spec = outc[0].cpu().numpy().reshape(-1).astype(np.complex64)
    cf = ctff[0].cpu().numpy().reshape(-1).astype(np.complex64)
    vol = a3df.cpu().numpy().reshape(-1)
    sang, iq = O.searchang(DANG, "D", 7, lib="oracle")
    sang = np.ascontiguousarray(sang[:nviews].reshape(-1))
    mask = np.array([1, 1, 1, 1, 1], np.int32)
    testpar = np.zeros(6 * TOPK, np.float32)
    phi, theta, psi, dx, dy = (C.c_float(0) for _ in range(5))
    iran = C.c_int(-0xC3)
    t = time.perf_counter()
    L.or32_psearch(nsam, 1.0, spec, vol, DANG, maxr1, maxr2, C.byref(phi), C.byref(theta),
                   C.byref(psi), C.byref(dx), C.byref(dy), sang, nviews, iq, mask, 0.0, cf,
                   testpar, TOPK, C.byref(iran), O.IDENTITY)
    dt = time.perf_counter() - t
    per_particle = dt / nviews * norient_views
    return {"value": 1.0 / per_particle, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"oracle PSEARCH/CCP (psearch.f/ccp.f port incl. PAD and ROTATE2D) on 1 "
                      f"particle x {nviews} of {norient_views} views in {dt:.1f} s, extrapolated "
                      f"to the full grid; PREFINE/CC3M rescoring not included"}


if __name__ == "__main__":
    main()
