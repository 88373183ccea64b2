"""Global orientation search (the reference's mode 3/4 path) on the GPU.

PSEARCH (psearch.f:8-121) scans the SEARCHANG grid, each view through CCP
(ccp.f:8-275: NPSI in-plane steps x 4 quadrants, shifts on the NSAMR grid),
keeps the IPMAX best views, and LMAIN refines each of them with PREFINE
(lmain.f:615-642), whose score is CC3M.  Here:

  1. the orientation list is built on the host exactly as PSEARCH does
     (SEARCHANG grid + the RANDOM jitter of psearch.f:61-64 + the in-plane
     angles of ccp.f:144-150), one jitter per bank;
  2. the CTF-free template bank is built once per reference volume;
  3. every batch of particles is scored as an MFMA GEMM (BF16 or FP32) with
     the CCP peak/box-sum epilogue and a per-view top-k;
  4. the top-k candidates are rescored in FP32 with CC3M (the PREFINE score)
     at the candidate's angles and shift; the best rescored candidate wins
     (strict '>', earlier candidate on ties, lmain.f:634).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import api

PI_F = np.float32(3.1415926535897)


@dataclass
class SearchResult:
    euler: np.ndarray      # (np, 3) best (phi, theta, psi)
    shift: np.ndarray      # (np, 2) SHX, SHY (phase units)
    score: np.ndarray      # (np,) CC3M of the winner (FP32 rescoring)
    view: np.ndarray       # (np,) orientation index of the winner in the bank
    cand_o: np.ndarray     # (np, k) candidate orientations from the GEMM stage
    cand_score: np.ndarray # (np, k) their GEMM scores


def orientation_list(dang: float, asym: str, nasym: int, iran: int, pmask=(1, 1, 1, 1, 1),
                     start=(0.0, 0.0, 0.0)):
    """(euler (O,3), group, npsi, iran') in PSEARCH/CCP order."""
    sang, iquadmax = api.search_grid(dang, asym, nasym, pmask)
    mask = np.asarray(pmask, np.int32)
    # psearch.f:61-64
    d = np.float32(np.float32(dang) / np.float32(180.0) * PI_F)
    r = []
    s = int(iran)
    for _ in range(3):
        v, s = api.lcg_random(s)
        r.append(np.float32(v))
    phir = np.float32((r[0] - np.float32(0.5)) * d * np.float32(mask[0]))
    thetar = np.float32((r[1] - np.float32(0.5)) * d * np.float32(mask[1]))
    psir = np.float32((r[2] - np.float32(0.5)) * d * np.float32(mask[2]))
    # psearch.f:66-67, ccp.f:57-59
    dd = np.float32(np.float32(dang) / np.float32(5.0))
    if dd < 1.0:
        dd = np.float32(1.0)
    npsi = int(math.floor(float(np.float32(90.0) / dd) + 0.5))
    iq = iquadmax if mask[2] else 1
    npsi_eff = npsi if mask[2] else 1
    views = []
    for v in sang:
        x1 = np.float32(np.float32(mask[0]) * (v[0] + phir) + np.float32(start[0]))
        x2 = np.float32(np.float32(mask[1]) * (v[1] + thetar) + np.float32(start[1]))
        x3 = np.float32(np.float32(mask[2]) * (v[2] + psir) + np.float32(start[2]))
        x1 = np.float32(x1 - 2 * PI_F) if x1 > PI_F else x1
        x2 = np.float32(x2 - 2 * PI_F) if x2 > PI_F else x2
        x3 = np.float32(x3 - 2 * PI_F) if x3 > PI_F else x3
        views.append((x1, x2, x3))
    eul = np.zeros((len(views) * npsi_eff * iq, 3), np.float32)
    o = 0
    for (x1, x2, x3) in views:
        for ipsi in range(1, npsi_eff + 1):
            psis = np.float32(np.float32(ipsi - 1) * PI_F / np.float32(2.0) / np.float32(npsi))
            for iquad in range(1, iq + 1):
                psim = np.float32(x3 + np.float32(iquad - 1) * PI_F / np.float32(2.0) + psis)
                if psim > PI_F:  # psearch.f:88
                    psim = np.float32(psim - 2 * PI_F)
                eul[o] = (x1, x2, psim)
                o += 1
    return eul, npsi_eff * iq, npsi, s


class GlobalSearch:
    """Bank + batched search + FP32 CC3M rescoring for one reference volume."""

    def __init__(self, ctx: api.Context, a3df: torch.Tensor, *, dang: float, asym: str, nasym: int,
                 maxr1: int, maxr2: int, rbfact: float, dtype=torch.bfloat16, iran: int = -1,
                 pmask=(1, 1, 1, 1, 1)):
        self.ctx = ctx
        self.a3df = a3df
        self.nsam = a3df.shape[0]
        self.maxr1, self.maxr2, self.rbfact = maxr1, maxr2, rbfact
        self.euler, self.group, self.npsi, self.iran = orientation_list(dang, asym, nasym, iran,
                                                                        pmask)
        dm = api.rotation(self.euler, amag=1.0)
        self.dm_host = dm
        self.dm = torch.from_numpy(dm).to(a3df.device)
        self.dtype = dtype
        self.bank = ctx.search_bank(self.nsam, maxr1, maxr2, a3df, self.dm, dtype=dtype)
        # BF16 searches run on the separable bank (two-stage translations)
        # where the band fits it; the annulus bank stays for the GEMM form
        self.bank_sep = None
        if dtype == torch.bfloat16 and maxr2 <= ctx.SEP_MAXR2:
            self.bank_sep = ctx.search_bank(self.nsam, maxr1, maxr2, a3df, self.dm, dtype=dtype,
                                            separable=True)
        nsamh = self.nsam // 2
        nsamrh = nsamh // 2 if (nsamh // 2) * 2 == nsamh else (nsamh + 1) // 2
        self.nsamr = 2 * nsamrh
        self._fsct = (None, None)

    def fsct_device(self, fsct) -> torch.Tensor:
        """FSCT as a device tensor, uploaded once per distinct host array."""
        if torch.is_tensor(fsct):
            return fsct
        key, dev = self._fsct
        host = np.ascontiguousarray(fsct, np.float32)
        if key is None or key.shape != host.shape or not np.array_equal(key, host):
            dev = torch.from_numpy(host.copy()).to(self.a3df.device)
            self._fsct = (host.copy(), dev)
        return dev

    @property
    def norient(self) -> int:
        return self.euler.shape[0]

    @property
    def search_bank(self) -> torch.Tensor:
        """The bank the searches read: separable where built, else the annulus bank."""
        return self.bank_sep if self.bank_sep is not None else self.bank

    def candidates(self, outc: torch.Tensor, ctff: torch.Tensor, topk: int = 32):
        return self.ctx.search(self.nsam, self.maxr1, self.maxr2, self.rbfact, self.search_bank,
                               outc, ctff, topk=topk, group=self.group)

    def row_to_shift(self, row: np.ndarray) -> np.ndarray:
        """ccp.f:254-256: SX = IXM/NSAMR*2pi on the NSAMR grid."""
        sx = row // 15 - 7
        sy = row % 15 - 7
        f = np.float32(PI_F * np.float32(2.0))
        return np.stack([(sx.astype(np.float32) / np.float32(self.nsamr)) * f,
                         (sy.astype(np.float32) / np.float32(self.nsamr)) * f], -1).astype(np.float32)

    def rescore(self, outc, ctff, top_o, top_r, fsct, ring, keep_positive: bool = True):
        """FP32 CC3M at each candidate (the score PREFINE maximises)."""
        npart, k = top_o.shape
        o = top_o.cpu().numpy()
        r = top_r.cpu().numpy()
        valid = o >= 0
        oc = np.where(valid, o, 0)
        dm = self.dm_host[oc.reshape(-1)]
        shift = self.row_to_shift(np.where(valid, r, 7 * 15 + 7).reshape(-1))
        dev = outc.device
        pidx = torch.arange(npart, device=dev, dtype=torch.int32).repeat_interleave(k)
        amag = torch.ones(npart * k, dtype=torch.float32, device=dev)
        sc = self.ctx.cc3m(self.nsam, self.a3df, outc, ctff, pidx,
                           torch.from_numpy(np.ascontiguousarray(dm)).to(dev),
                           torch.from_numpy(np.ascontiguousarray(shift)).to(dev), amag, self.maxr1,
                           self.maxr2, self.rbfact,
                           torch.as_tensor(fsct, dtype=torch.float32, device=dev), ring)
        sc = sc.view(npart, k).cpu().numpy()
        sc = np.where(valid, sc, -np.inf)
        return sc, shift.reshape(npart, k, 2)

    def run(self, outc, ctff, fsct, ring, topk: int = 32) -> SearchResult:
        top_o, top_r, top_s = self.candidates(outc, ctff, topk)
        sc, shifts = self.rescore(outc, ctff, top_o, top_r, fsct, ring)
        best = np.argmax(sc, axis=1)  # first maximum: earlier candidate on ties
        npart = sc.shape[0]
        o = top_o.cpu().numpy()
        bo = o[np.arange(npart), best]
        return SearchResult(self.euler[bo], shifts[np.arange(npart), best],
                            sc[np.arange(npart), best], bo, o, top_s.cpu().numpy())


def angular_error_deg(e1: np.ndarray, e2: np.ndarray, sym_mats=None) -> np.ndarray:
    """Rotation angle between two ZYZ orientations, minimised over symmetry."""
    def mats(e):
        dm = api.rotation(np.asarray(e, np.float32), amag=0.0, sm=None)
        return dm.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)
    a, b = mats(e1), mats(e2)
    ops = sym_mats if sym_mats is not None else [np.eye(3)]
    best = np.full(len(a), 180.0)
    for s in ops:
        rel = np.einsum("nji,kj,nkl->nil", a, s, b)  # a^T S^T b
        tr = np.clip((np.trace(rel, axis1=1, axis2=2) - 1.0) / 2.0, -1.0, 1.0)
        best = np.minimum(best, np.degrees(np.arccos(tr)))
    return best


class LocalRefinement:
    """Mode-1 refinement (lmain.f:743-828) of particles near known parameters:
    the local grid search (every orientation of a (2 nang + 1)^3 grid DANG
    apart around the start, each at every shift of a (2 nsh + 1)^2 window DSH
    apart, scored by CC3M: fr_local_search_batch), then PREFINE (prefine.f,
    VA04A over CALCFX = -CC3M - FSH; fr_refine_batch) from the best grid
    point.  Shifts are in phase units (SHX = 2 pi dx / NSAM for dx pixels).

    prior: (XM - SHX, YM - SHY, SX, SY, SIG2) of FSH (fsh.f), one for all
    particles (the dataset's shift statistics, card10.f:541-566)."""

    def __init__(self, ctx: api.Context, a3df: torch.Tensor, *, maxr1: int, maxr2: int,
                 rbfact: float, nang: int = 6, dang_deg: float = 1.0, nsh: int = 4,
                 dsh: float | None = None, prior=(0.0, 0.0, 0.1, 0.1, 0.01),
                 mask=(1, 1, 1, 1, 1), maxit: int = 50):
        self.ctx = ctx
        self.a3df = a3df
        self.nsam = a3df.shape[0]
        self.maxr1, self.maxr2, self.rbfact = maxr1, maxr2, rbfact
        self.nang, self.nsh = nang, nsh
        self.dang = float(np.float32(np.deg2rad(dang_deg)))
        self.dsh = float(np.float32(2 * math.pi / self.nsam)) if dsh is None else float(dsh)
        self.prior = np.asarray(prior, np.float32)
        self.mask = np.asarray(mask, np.int32)
        self.maxit = maxit

    @property
    def norient(self) -> int:
        return (2 * self.nang + 1) ** 3

    def run(self, outc, ctff, start, fsct, ring):
        """start (np, 5) host PHI THETA PSI SHX SHY.  Returns (grid (np, 6):
        the best grid point and its CC3M, refined (np, 6): PREFINE's
        parameters and PRESA, CC3M evaluations per particle)."""
        fs = fsct if torch.is_tensor(fsct) else torch.as_tensor(
            np.ascontiguousarray(fsct, np.float32), device=outc.device)
        grid = self.ctx.local_search(self.nsam, self.a3df, outc, ctff, start, nang=self.nang,
                                     dang=self.dang, nsh=self.nsh, dsh=self.dsh,
                                     maxr1=self.maxr1, maxr2=self.maxr2, rbfact=self.rbfact,
                                     fsct=fs, ring=ring)
        n = grid.shape[0]
        prior = np.repeat(self.prior[None], n, axis=0)
        ref, nev = self.ctx.refine(self.nsam, self.a3df, outc, ctff, np.arange(n, dtype=np.int32),
                                   np.ones(n, np.float32), grid, self.mask, self.maxr1,
                                   self.maxr2, self.rbfact, fs, ring, prior, maxit=self.maxit)
        return grid, ref, nev
