"""Learned orientation proposal in front of the global search.

A small convolutional network (the kernels of ``csrc/fr_cnn.hip``) maps each
prepared particle to logits over a coarse orientation grid ("classes": the
SEARCHANG views at ``class_deg`` times in-plane angles at the same step).
Every orientation of the search bank belongs to its nearest class under the
point group (a Voronoi cell), so the network's labels and the bank partition
are the same rule.  A particle is then searched only over the cells of its
``topk`` most probable classes (``fr_search_list_batch``); when the largest
class probability is below ``threshold`` it falls back to the exhaustive
search (``fr_search_batch``), which stays the judge of both.

The network is trained on the GPU, by ``fr_cnn_gradients`` and
``fr_cnn_adam``, on CTF-modulated noisy projections of the current reference
made by ``fr_cnn_synth``; nothing here runs on the CPU except bookkeeping.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import api
from . import synth
from .search import GlobalSearch, SearchResult


@dataclass
class ProposalConfig:
    # network input (Fourier crop of the particle): 64^2 and four stride-2
    # convolutions end at 4 x 4 x 64 features; at 15-degree classes it
    # proposes as well as a 128^2 input with five, at a quarter of the cost
    S: int = 64
    channels: int = 64
    nconv: int = 4
    class_deg: float = 15.0
    topk: int = 8
    threshold: float = 0.3  # max softmax below which the exhaustive search runs
    group: int = 16         # in-plane candidates per (class, view) group


class OrientationClasses:
    """SEARCHANG views at class_deg x in-plane angles at class_deg, without
    the orientations that the point group makes equal to an earlier one
    (on the asymmetric unit's boundary the grid holds such pairs)."""

    def __init__(self, class_deg: float, asym: str, nasym: int, sym_mats=None):
        views, _ = api.search_grid(class_deg, asym, nasym)
        npsi = int(round(360.0 / class_deg))
        psi = np.radians(np.arange(npsi) * class_deg).astype(np.float32)
        e = np.zeros((len(views) * npsi, 3), np.float32)
        e[:, 0] = np.repeat(views[:, 0], npsi)
        e[:, 1] = np.repeat(views[:, 1], npsi)
        e[:, 2] = np.tile(psi, len(views))
        e[:, 2] = np.where(e[:, 2] > np.float32(math.pi), e[:, 2] - np.float32(2 * math.pi),
                           e[:, 2]).astype(np.float32)
        dm = api.rotation(e, amag=0.0)
        if sym_mats is not None:
            keep = unique_orientations(dm, sym_mats)
            e, dm = e[keep], dm[keep]
        self.euler = e
        self.dm = dm
        self.n = e.shape[0]


def unique_orientations(dm: np.ndarray, sym_mats, tol_deg: float = 0.01) -> np.ndarray:
    """Indices of the orientations not equal (within tol_deg, under the point
    group) to an earlier one."""
    A = dm.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)  # A[k][j] = dm[3j + k]
    tmin = 1.0 + 2.0 * math.cos(math.radians(tol_deg))
    keep = []
    for i in range(len(A)):
        dup = False
        if keep:
            K = A[keep]
            for s in sym_mats:
                tr = np.sum((s @ A[i])[None] * K, axis=(1, 2))  # trace((S A_i)^T A_k)
                if np.any(tr >= tmin):
                    dup = True
                    break
        if not dup:
            keep.append(i)
    return np.array(keep, dtype=np.int64)


def sym_array(sym_mats) -> np.ndarray:
    return np.ascontiguousarray(np.stack(sym_mats).reshape(-1, 9), dtype=np.float32)


class Network:
    """Parameters (FP32 master, BF16-rounded working copy, Adam moments) on the device."""

    def __init__(self, ctx: api.Context, shape, seed: int, device):
        self.ctx = ctx
        self.shape = tuple(int(x) for x in shape)
        n = api.cnn_param_count(*self.shape)
        self.nparam = n
        z = lambda: torch.zeros(n, dtype=torch.float32, device=device)  # noqa: E731
        self.master, self.wq, self.m, self.v, self.grad = z(), z(), z(), z(), z()
        self.step = 0
        ctx.cnn_init(self.shape, seed, self.master, self.wq)

    def forward(self, img: torch.Tensor) -> torch.Tensor:
        return self.ctx.cnn_forward(self.shape, self.wq, img)

    def infer(self, img: torch.Tensor, ctx: api.Context | None = None) -> torch.Tensor:
        """Logits for proposing: BF16 activations where the shape allows (64
        channels), else the FP32 forward; on `ctx` (its stream and workspace)
        when given."""
        c = ctx or self.ctx
        if self.shape[1] == 64 and self.shape[2] >= 2:
            return c.cnn_forward_bf16(self.shape, self.wq, img)
        return c.cnn_forward(self.shape, self.wq, img)

    def train_step(self, img: torch.Tensor, label: torch.Tensor, lr: float,
                   bf16: bool = True) -> torch.Tensor:
        """One Adam step on a batch; the gradients in mixed precision (BF16
        GEMM operands, FP32 accumulation) unless bf16 is False."""
        loss = self.ctx.cnn_gradients(self.shape, self.wq, img, label, self.grad, bf16=bf16)
        self.step += 1
        self.ctx.cnn_adam(self.shape, self.master, self.wq, self.grad, self.m, self.v, self.step,
                          lr)
        return loss


@dataclass
class TrainReport:
    steps: int
    samples: int
    seconds: float
    final_loss: float
    heldout_top1: float
    heldout_topk: float
    log: list = field(default_factory=list)


class LearnedProposal:
    """Network + class cells of one search bank."""

    def __init__(self, ctx: api.Context, search: GlobalSearch, sym_mats, asym: str, nasym: int,
                 cfg: ProposalConfig = ProposalConfig(), seed: int = 0x5EED):
        self.ctx = ctx
        self.search = search
        self.cfg = cfg
        dev = search.a3df.device
        self.dev = dev
        self.classes = OrientationClasses(cfg.class_deg, asym, nasym, sym_mats)
        self.sym = torch.from_numpy(sym_array(sym_mats)).to(dev)
        self.dmc = torch.from_numpy(self.classes.dm).to(dev)
        self.net = Network(ctx, (cfg.S, cfg.channels, cfg.nconv, self.classes.n), seed, dev)
        self._build_cells()

    # -------------------------------------------------------------- cells
    def _build_cells(self):
        """Bank orientation -> nearest class; per class, its members in groups of
        `group` in-plane candidates of one view (padding -1)."""
        s = self.search
        G = self.cfg.group
        bank_dm = torch.from_numpy(np.ascontiguousarray(s.dm_host)).to(self.dev)
        cls = self.ctx.orient_nearest(bank_dm, self.dmc, self.sym).cpu().numpy()
        self.bank_class = cls
        view = np.arange(s.norient) // s.group
        order = np.lexsort((np.arange(s.norient), view, cls))  # by class, view, index
        cells = []
        for c in range(self.classes.n):
            cells.append([])
        oc, ov = cls[order], view[order]
        start = 0
        n = len(order)
        while start < n:
            end = start
            while end < n and oc[end] == oc[start] and ov[end] == ov[start]:
                end += 1
            members = order[start:end]
            for g0 in range(0, len(members), G):
                grp = np.full(G, -1, np.int32)
                chunk = members[g0:g0 + G]
                grp[:len(chunk)] = chunk
                cells[oc[start]].append(grp)
            start = end
        self.cells = [np.concatenate(c) if c else np.zeros(0, np.int32) for c in cells]
        self.cell_groups = np.array([len(c) // G for c in self.cells])
        # device copies for fr_build_lists: cells back to back, class c at
        # cell_off[c]..cell_off[c+1]
        lens = np.array([len(c) for c in self.cells], np.int64)
        off = np.zeros(len(lens) + 1, np.int64)
        off[1:] = np.cumsum(lens)
        flat = np.concatenate(self.cells) if off[-1] else np.full(1, -1, np.int32)
        self.cell_off_dev = torch.from_numpy(off.astype(np.int32)).to(self.dev)
        self.cells_dev = torch.from_numpy(flat.astype(np.int32)).to(self.dev)
        # any set of distinct classes spans at most all cells; a fallback list is the bank
        self.maxlen = int(max(off[-1], s.norient))

    def candidates(self, top: np.ndarray) -> np.ndarray:
        """Concatenated cells of each particle's proposed classes, padded with -1
        to whole groups (every row has the same length)."""
        lists = [np.concatenate([self.cells[c] for c in row if c >= 0] or
                                [np.zeros(0, np.int32)]) for row in top]
        G = self.cfg.group
        n = max(max((len(x) for x in lists), default=0), G)
        out = np.full((len(lists), n), -1, np.int32)
        for i, x in enumerate(lists):
            out[i, :len(x)] = x
        return out

    # ------------------------------------------------------------ training
    def _sample(self, rng, n: int, optics: synth.Optics, snr_range, defocus, astig, shift_a):
        eul = synth.random_orientations(n, rng)
        dm = api.rotation(eul, amag=0.0)
        nsam = self.search.nsam
        sh_px = rng.uniform(-shift_a, shift_a, (n, 2)) / optics.psize
        spar = np.zeros((n, 8), np.float32)
        spar[:, 0:2] = sh_px / nsam * 2 * math.pi
        df = rng.uniform(*defocus, n)
        ast = rng.uniform(*astig, n)
        spar[:, 2] = df + ast / 2
        spar[:, 3] = df - ast / 2
        spar[:, 4] = rng.uniform(-math.pi, math.pi, n)
        spar[:, 5] = rng.uniform(*snr_range, n)
        return eul, dm, spar

    def make_batch(self, rng, n: int, optics: synth.Optics, ri: float, seed: int,
                   snr_range=(0.03, 0.2), defocus=(1.0e4, 3.0e4), astig=(0.0, 1.0e3),
                   shift_a: float = 3.0):
        """n training images of the current reference and their class labels."""
        nsam = self.search.nsam
        eul, dm, spar = self._sample(rng, n, optics, snr_range, defocus, astig, shift_a)
        dm_t = torch.from_numpy(dm).to(self.dev)
        gpar = optics.gpar(nsam, ri)
        img = self.ctx.cnn_synth(nsam, self.cfg.S, self.search.a3df, dm_t,
                                 torch.from_numpy(spar).to(self.dev), gpar,
                                 math.pi * ri * ri, seed)
        label = self.ctx.orient_nearest(dm_t, self.dmc, self.sym)
        return img, label, eul

    def train(self, optics: synth.Optics, ri: float, nsamples: int = 200_000, batch: int = 256,
              lr: float = 2e-3, seed: int = 0xC5, snr_range=(0.03, 0.2), log_every: int = 100,
              progress=None, heldout: int = 2048) -> TrainReport:
        rng = np.random.default_rng(seed)
        steps = max(1, nsamples // batch)
        t0 = time.perf_counter()
        t_last = t0
        log = []
        loss_v = float("nan")
        for st in range(steps):
            # cosine decay to lr/20
            lr_t = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * st / steps)))
            img, label, _ = self.make_batch(rng, batch, optics, ri, seed * 1_000_003 + st,
                                            snr_range)
            loss = self.net.train_step(img, label, lr_t)
            # a line every log_every steps and at least every 10 s, so that a
            # slowed-down run (e.g. under hardware counters) is never silent
            now = time.perf_counter()
            if (st + 1) % log_every == 0 or st + 1 == steps or now - t_last >= 10.0:
                t_last = now
                loss_v = float(loss.mean())
                log.append((st + 1, loss_v))
                if progress:
                    progress(f"train step {st + 1}/{steps} loss {loss_v:.3f} "
                             f"({time.perf_counter() - t0:.1f}s)")
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        top1, topk = self.evaluate(optics, ri, heldout, seed + 1, snr_range)
        return TrainReport(steps, steps * batch, secs, loss_v, top1, topk, log)

    def evaluate(self, optics, ri, n: int, seed: int, snr_range=(0.03, 0.2)):
        """Held-out class accuracy on fresh synthetic images: top-1 and top-k."""
        rng = np.random.default_rng(seed)
        hit1 = hitk = 0
        for b0 in range(0, n, 256):
            m = min(256, n - b0)
            img, label, _ = self.make_batch(rng, m, optics, ri, seed * 7 + b0, snr_range)
            top, _ = self.ctx.cnn_propose(self.net.infer(img), self.cfg.topk)
            t = top.cpu().numpy()
            y = label.cpu().numpy()
            hit1 += int((t[:, 0] == y).sum())
            hitk += int((t == y[:, None]).any(axis=1).sum())
        return hit1 / n, hitk / n

    # ------------------------------------------------------------ search
    def propose(self, datc: torch.Tensor, ctff: torch.Tensor, ctx: api.Context | None = None):
        c = ctx or self.ctx
        img = c.cnn_images(self.search.nsam, self.cfg.S, datc, ctff)
        return c.cnn_propose(self.net.infer(img, c), self.cfg.topk)

    def launch(self, outc, ctff, datc, fsct, ring, topk: int = 32,
               ctx: api.Context | None = None):
        """Enqueue the learned search of one batch on the current stream, from
        images to winners: proposal -> ragged candidate lists (the proposed
        classes' cells, or the whole bank for a particle below the confidence
        threshold) -> one GEMM search launch -> FP32 CC3M rescoring of the
        top-k.  The results are copied into a pinned host buffer
        asynchronously; `finish` waits for them.  A caller may launch the next
        batch before finishing this one (two batches can be in flight); with
        a second context (its own workspace) on a second stream, the two
        batches also run on the GPU at the same time."""
        s = self.search
        c = ctx or self.ctx
        top, prob = self.propose(datc, ctff, c)
        loff, grp, cand, info = c.build_lists(
            top, prob, self.cfg.threshold, self.cell_off_dev, self.cells_dev, self.cfg.group,
            s.norient, s.group, self.maxlen)
        o, r, sg = c.search_ragged(s.nsam, s.maxr1, s.maxr2, s.rbfact, s.search_bank,
                                   outc, ctff, loff, grp, cand, topk, info=info)
        sc, best = c.rescore(s.nsam, s.a3df, outc, ctff, o, r, s.dm, s.nsamr, s.maxr1,
                             s.maxr2, s.rbfact, s.fsct_device(fsct), ring)
        npart = o.shape[0]
        f32 = torch.float32
        pack = torch.cat([o.view(f32).reshape(-1), r.view(f32).reshape(-1), sg.reshape(-1),
                          sc.reshape(-1), best.view(f32), info.view(f32).reshape(-1)])
        # two pinned buffers, used in turn, so that one batch can be unpacked
        # while the next one is in flight
        self._flip = getattr(self, "_flip", 0) ^ 1
        bufs = getattr(self, "_pinned", None)
        if bufs is None or bufs[0].numel() < pack.numel():
            bufs = [torch.empty(pack.numel(), dtype=f32, pin_memory=True) for _ in range(2)]
            self._pinned = bufs
        host = bufs[self._flip][:pack.numel()]
        host.copy_(pack, non_blocking=True)
        done = torch.cuda.Event()
        done.record()
        return (host, done, npart, topk, pack)

    def finish(self, pending):
        """Wait for a launched batch and return (SearchResult, confident)."""
        host, done, npart, topk, _ = pending
        done.synchronize()
        s = self.search
        pack = host.numpy().copy()
        n = npart * topk
        o_h = pack[:n].view(np.int32).reshape(npart, topk)
        r_h = pack[n:2 * n].view(np.int32).reshape(npart, topk)
        sg_h = pack[2 * n:3 * n].reshape(npart, topk)
        sc_h = pack[3 * n:4 * n].reshape(npart, topk)
        best_h = pack[4 * n:4 * n + npart].view(np.int32)
        info_h = pack[4 * n + npart:].view(np.int32).reshape(npart, 2)
        confident = info_h[:, 0] == 1
        self.last_valid = int(info_h[confident, 1].sum())
        ar = np.arange(npart)
        bo = o_h[ar, best_h]
        shifts = s.row_to_shift(np.where(o_h >= 0, r_h, 7 * 15 + 7).reshape(-1)).reshape(
            npart, topk, 2)
        res = SearchResult(s.euler[np.maximum(bo, 0)], shifts[ar, best_h], sc_h[ar, best_h], bo,
                           o_h, sg_h)
        return res, confident

    def run(self, outc, ctff, datc, fsct, ring, topk: int = 32, ctx: api.Context | None = None):
        """Learned search of one batch (launch + finish)."""
        return self.finish(self.launch(outc, ctff, datc, fsct, ring, topk, ctx))
