"""Host-side mirror of Frealign's hot-path operators over the gfx950 library.

The functions keep the reference's names and argument meaning (radians,
phase-unit shifts, the RECB record of A3D3, PMASK, ...) and operate on
device tensors (torch is used only for device memory and streams).  Every
call goes through libfrealign_hip.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N

IDENTITY_SM = np.eye(3, dtype=np.float32).reshape(-1, order="F").copy()


def _p(t) -> int:
    """Raw address of a tensor / ndarray (or 0 for None)."""
    if t is None:
        return 0
    if isinstance(t, torch.Tensor):
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return t.data_ptr()
    if isinstance(t, np.ndarray):
        if not t.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return t.ctypes.data
    raise TypeError(type(t))


def _dev(t: torch.Tensor, dtype, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


class Context:
    """One fr_ctx on one device; launches go to torch's current stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        N.check(N.lib().fr_create(device, C.byref(h)), "fr_create")
        self.h = h
        self.device = device

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        N.check(N.lib().fr_set_stream(self.h, C.c_void_p(s)), "fr_set_stream")

    def close(self):
        if self.h:
            N.lib().fr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ profiling
    PROF = {"search_gemm": 0, "insert": 1, "cc3m": 2, "prep": 3, "search_list": 4, "cnn": 5, "plan": 6}

    def profile(self, on: bool = True):
        N.check(N.lib().fr_profile_enable(self.h, int(on)), "fr_profile_enable")

    def profile_read(self, name: str):
        """(total_ms, launches) of one kernel family since the last read."""
        ms, n = C.c_double(0), C.c_longlong(0)
        N.check(N.lib().fr_profile_read(self.h, self.PROF[name], C.byref(ms), C.byref(n)),
                "fr_profile_read")
        return ms.value, n.value

    # ---------------------------------------------------------------- FFT
    def fft_r2c(self, nsam: int, data: torch.Tensor):
        _dev(data, torch.float32, "data")
        nimg = data.numel() // (nsam * (nsam + 2))
        self._sync_stream()
        N.check(N.lib().fr_fft2d_r2c(self.h, nsam, nimg, _p(data)), "fr_fft2d_r2c")

    def fft_c2r(self, nsam: int, data: torch.Tensor):
        _dev(data, torch.float32, "data")
        nimg = data.numel() // (nsam * (nsam + 2))
        self._sync_stream()
        N.check(N.lib().fr_fft2d_c2r(self.h, nsam, nimg, _p(data)), "fr_fft2d_c2r")

    # ------------------------------------------------------------- CTFAPPLY
    def ctfapply(self, nsam: int, images: torch.Tensor, ppar: torch.Tensor, pssnr: torch.Tensor,
                 gpar, ictf: int = 1):
        """LMAIN's r2c + CTFAPPLY for a batch.  images: (nb, NSAM, NSAM+2) float32,
        destroyed.  ppar: (nb, 8) = SHX, SHY, DFMID1, DFMID2, ANGAST, AMAGP, 0, 0.
        gpar: CS, WL, WGH1, WGH2, THETATR, RIH, HALFW, RI2, RI3, TX, TY.
        Returns CTFF (nb, 2, NSAM, JC) c64, OUTC (nb, NSAM, JC), DATC (nb, NSAM, JC)."""
        jc = nsam // 2 + 1
        _dev(images, torch.float32, "images")
        nb = images.numel() // (nsam * (nsam + 2))
        _dev(ppar, torch.float32, "ppar")
        _dev(pssnr, torch.float32, "pssnr")
        if ppar.numel() != 8 * nb:
            raise ValueError("ppar must be (nb, 8)")
        if pssnr.numel() < jc + 1:
            raise ValueError("pssnr needs JC+1 entries")
        g = np.asarray(gpar, dtype=np.float32)
        if g.size != 11:
            raise ValueError("gpar needs 11 values")
        dev = images.device
        ctff = torch.empty((nb, 2, nsam, jc), dtype=torch.complex64, device=dev)
        outc = torch.empty((nb, nsam, jc), dtype=torch.complex64, device=dev)
        datc = torch.empty((nb, nsam, jc), dtype=torch.complex64, device=dev)
        self._sync_stream()
        N.check(N.lib().fr_prep_batch(self.h, nsam, nb, _p(images), _p(ppar), _p(pssnr), _p(g),
                                      ictf, _p(ctff), _p(outc), _p(datc)), "fr_prep_batch")
        return ctff, outc, datc

    # ------------------------------------------------------------- PEXTRACT
    def pextract(self, nsam: int, a3df: torch.Tensor, dm: torch.Tensor, ctff: torch.Tensor,
                 pidx: torch.Tensor | None = None) -> torch.Tensor:
        jc = nsam // 2 + 1
        _dev(a3df, torch.complex64, "a3df")
        _dev(dm, torch.float32, "dm")
        _dev(ctff, torch.complex64, "ctff")
        n = dm.numel() // 9
        npart = ctff.numel() // (2 * nsam * jc)
        if pidx is not None:
            _dev(pidx, torch.int32, "pidx")
            if pidx.numel() != n or (n and int(pidx.max()) >= npart) or (n and int(pidx.min()) < 0):
                raise ValueError("pidx out of range")
        elif n > npart:
            raise ValueError("more sections than CTF tables")
        if a3df.numel() != jc * nsam * nsam:
            raise ValueError("a3df must be a JC x NSAM x NSAM half-volume")
        out = torch.empty((n, nsam, jc), dtype=torch.complex64, device=a3df.device)
        self._sync_stream()
        N.check(N.lib().fr_pextract_batch(self.h, nsam, n, _p(a3df), _p(dm), _p(pidx), _p(ctff),
                                          _p(out)), "fr_pextract_batch")
        return out

    # ---------------------------------------------------------------- CC3M
    def cc3m(self, nsam: int, a3df, spec, ctff, pidx, dm, shift, amag, maxr1: int, maxr2: int,
             rbfact: float, fsct, ring) -> torch.Tensor:
        jc = nsam // 2 + 1
        for t, nm, dt in ((a3df, "a3df", torch.complex64), (spec, "spec", torch.complex64),
                          (ctff, "ctff", torch.complex64), (pidx, "pidx", torch.int32),
                          (dm, "dm", torch.float32), (shift, "shift", torch.float32),
                          (amag, "amag", torch.float32), (fsct, "fsct", torch.float32)):
            _dev(t, dt, nm)
        n = pidx.numel()
        npart = spec.numel() // (nsam * jc)
        if dm.numel() != 9 * n or shift.numel() != 2 * n or amag.numel() != n:
            raise ValueError("dm/shift/amag sizes do not match pidx")
        if n and (int(pidx.max()) >= npart or int(pidx.min()) < 0):
            raise ValueError("pidx out of range")
        if ctff.numel() != 2 * spec.numel() or fsct.numel() < nsam // 2 + 1:
            raise ValueError("ctff/fsct sizes")
        if a3df.numel() != nsam * nsam * jc:
            raise ValueError("a3df must be the NSAM x NSAM x JC half-volume")
        r = np.asarray(ring, dtype=np.float32)
        out = torch.empty(n, dtype=torch.float32, device=a3df.device)
        self._sync_stream()
        N.check(N.lib().fr_cc3m_batch(self.h, nsam, n, _p(a3df), _p(spec), _p(ctff), _p(pidx),
                                      _p(dm), _p(shift), _p(amag), maxr1, maxr2, float(rbfact),
                                      _p(fsct), _p(r), _p(out)), "fr_cc3m_batch")
        return out

    def presb(self, nsam: int, a3df, spec, ctff, pidx, dm, shift, amag, ring,
              rbin=None) -> torch.Tensor:
        """PRESB (presb.f:8-160) for n items; returns the n x NSAM RBIN array,
        updated in place when `rbin` is given (shells with no point keep their value)."""
        jc = nsam // 2 + 1
        for t, nm, dt in ((a3df, "a3df", torch.complex64), (spec, "spec", torch.complex64),
                          (ctff, "ctff", torch.complex64), (pidx, "pidx", torch.int32),
                          (dm, "dm", torch.float32), (shift, "shift", torch.float32),
                          (amag, "amag", torch.float32)):
            _dev(t, dt, nm)
        n = pidx.numel()
        npart = spec.numel() // (nsam * jc)
        if dm.numel() != 9 * n or shift.numel() != 2 * n or amag.numel() != n:
            raise ValueError("dm/shift/amag sizes do not match pidx")
        if n and (int(pidx.max()) >= npart or int(pidx.min()) < 0):
            raise ValueError("pidx out of range")
        if ctff.numel() != 2 * spec.numel() or a3df.numel() != nsam * nsam * jc:
            raise ValueError("ctff/a3df sizes")
        if rbin is None:
            rbin = torch.zeros((n, nsam), dtype=torch.float32, device=a3df.device)
        else:
            _dev(rbin, torch.float32, "rbin")
            if rbin.numel() != n * nsam:
                raise ValueError("rbin must hold n x NSAM floats")
        r = np.asarray(ring, dtype=np.float32)
        self._sync_stream()
        N.check(N.lib().fr_presb_batch(self.h, nsam, n, _p(a3df), _p(spec), _p(ctff), _p(pidx),
                                       _p(dm), _p(shift), _p(amag), _p(r), _p(rbin)),
                "fr_presb_batch")
        return rbin

    # ------------------------------------------------------------ insertion
    def insert(self, nsam: int, datc, ctff, recb: np.ndarray, half: np.ndarray, sym, pbc: float,
               boff: float, psize: float, vols: "HalfMaps", stats: "VoxelStats | None" = None):
        """The A3D3 flush for nb particles (fr_insert_batch); with `stats`
        the _SE insertion (fr_insert_batch_se) that also accumulates the
        per-voxel statistics."""
        _dev(datc, torch.complex64, "datc")
        _dev(ctff, torch.complex64, "ctff")
        recb = np.ascontiguousarray(recb, dtype=np.float32)
        half = np.ascontiguousarray(half, dtype=np.int32)
        nb = recb.shape[0]
        jc = nsam // 2 + 1
        if datc.numel() < nb * nsam * jc or ctff.numel() < 2 * nb * nsam * jc:
            raise ValueError("datc/ctff hold fewer particles than recb")
        if half.shape != (nb,):
            raise ValueError("half must be (nb,)")
        nsym, symop, jsym, _ = sym
        self._sync_stream()
        args = (self.h, nsam, nb, _p(datc), _p(ctff), _p(recb), _p(half), nsym, _p(jsym),
                _p(symop), float(pbc), float(boff), float(psize), _p(vols.num[0]),
                _p(vols.ctf2[0]), _p(vols.num[1]), _p(vols.ctf2[1]), _p(vols.nksum), _p(vols.ttd))
        if stats is None:
            N.check(N.lib().fr_insert_batch(*args), "fr_insert_batch")
            return
        if stats.nsam != nsam:
            raise ValueError("stats are for another box size")
        N.check(N.lib().fr_insert_batch_se(*args, _p(stats.ref), _p(stats.ksum), _p(stats.vsum),
                                           _p(stats.psum), _p(stats.asum)), "fr_insert_batch_se")

    # --------------------------------------------------------------- search
    def search_layout(self, nsam: int, maxr1: int, maxr2: int):
        kc, kpad = C.c_int(0), C.c_int(0)
        N.check(N.lib().fr_search_layout(self.h, nsam, maxr1, maxr2, C.byref(kc), C.byref(kpad)),
                "fr_search_layout")
        return kc.value, kpad.value

    SEP_ROW = 4096      # bf16 values per template of the separable bank
    SEP_MAXR2 = 32      # the separable bank holds MAXR2 <= 32

    def search_bank(self, nsam: int, maxr1: int, maxr2: int, a3df, dm: torch.Tensor,
                    dtype=torch.bfloat16, separable: bool = False) -> torch.Tensor:
        """CTF-free template bank for the orientations dm (norient, 9): the
        annulus in order (norient, kpad), or with `separable` (bf16, MAXR2 <= 32)
        the layout of the two-stage search (norient, 4096)."""
        _dev(a3df, torch.complex64, "a3df")
        _dev(dm, torch.float32, "dm")
        if dtype not in (torch.bfloat16, torch.float32):
            raise ValueError("bank dtype must be bfloat16 or float32")
        _, kpad = self.search_layout(nsam, maxr1, maxr2)
        n = dm.numel() // 9
        if separable:
            if dtype != torch.bfloat16 or maxr2 > self.SEP_MAXR2:
                raise ValueError("the separable bank is bfloat16 with MAXR2 <= 32")
            row, code = self.SEP_ROW, 2
        else:
            row, code = kpad, 1 if dtype == torch.bfloat16 else 0
        bank = torch.empty((n, row), dtype=dtype, device=a3df.device)
        self._sync_stream()
        N.check(N.lib().fr_search_bank(self.h, nsam, maxr1, maxr2, _p(a3df), n, _p(dm), code,
                                       _p(bank)), "fr_search_bank")
        return bank

    def _bank_code(self, nsam: int, maxr1: int, maxr2: int, bank: torch.Tensor) -> int:
        """0: FP32 annulus bank, 1: BF16 annulus bank, 2: BF16 separable bank."""
        if bank.dtype not in (torch.bfloat16, torch.float32) or not bank.is_cuda:
            raise ValueError("bank must be a device bfloat16/float32 tensor")
        _, kpad = self.search_layout(nsam, maxr1, maxr2)
        if bank.dim() == 2 and bank.shape[1] == kpad:
            return 1 if bank.dtype == torch.bfloat16 else 0
        if (bank.dim() == 2 and bank.dtype == torch.bfloat16 and bank.shape[1] == self.SEP_ROW
                and maxr2 <= self.SEP_MAXR2):
            return 2
        raise ValueError(f"bank rows must be {kpad} long (or {self.SEP_ROW}, separable bf16)")

    def psearch(self, nsam: int, a3df, outc, ctff, *, dangin: float, maxr1: int, maxr2: int,
                rbfact: float, sang: np.ndarray, iquadmax: int, par=None, amag=None,
                mask=(1, 1, 1, 1, 1), ipmax: int = 32, iran: int = -1, sm=None, asym: str = "C",
                testpar=None):
        """PSEARCH (psearch.f:8-121) over CCP with PAD and ROTATE2D, reference-exact,
        for every particle in call order.  par: (np, 5) PHI THETA PSI DSHX DSHY (the
        starting values; default zero).  Returns (par out = the best view, testpar
        (np, ipmax, 6), iran after the calls)."""
        _dev(a3df, torch.complex64, "a3df")
        _dev(outc, torch.complex64, "outc")
        _dev(ctff, torch.complex64, "ctff")
        jc = nsam // 2 + 1
        np_ = outc.numel() // (nsam * jc)
        if ctff.numel() != 2 * outc.numel():
            raise ValueError("ctff must hold two half-planes per particle")
        p = np.zeros((np_, 5), np.float32) if par is None else \
            np.ascontiguousarray(np.asarray(par, np.float32).reshape(np_, 5)).copy()
        am = np.ones(np_, np.float32) if amag is None else \
            np.ascontiguousarray(np.asarray(amag, np.float32).reshape(np_))
        sg = np.ascontiguousarray(np.asarray(sang, np.float32).reshape(-1, 3))
        tp = np.zeros((np_, ipmax, 6), np.float32) if testpar is None else \
            np.ascontiguousarray(np.asarray(testpar, np.float32).reshape(np_, ipmax, 6)).copy()
        m = np.ascontiguousarray(np.asarray(mask, np.int32))
        smv = None if sm is None else np.ascontiguousarray(sm, dtype=np.float32)
        ir = C.c_int(int(iran))
        self._sync_stream()
        N.check(N.lib().fr_psearch(self.h, nsam, np_, _p(am), _p(a3df), _p(outc), _p(ctff),
                                   float(dangin), int(maxr1), int(maxr2), _p(p), _p(sg),
                                   sg.shape[0], int(iquadmax), _p(m), float(rbfact), _p(tp),
                                   int(ipmax), C.byref(ir), _p(smv), asym.encode()[:1]),
                "fr_psearch")
        return p, tp, ir.value

    def ccp_views(self, nsam: int, a3df, outc, ctff, dm, *, maxr1: int, maxr2: int,
                  rbfact: float, dang: float, iquadmax: int, mask=(1, 1, 1, 1, 1)):
        """fr_ccp_view_batch: CCP of every (particle, view) item; dm (np, nview, 9)
        device.  Returns (score, code) device tensors (np, nview)."""
        _dev(dm, torch.float32, "dm")
        jc = nsam // 2 + 1
        np_ = outc.numel() // (nsam * jc)
        nview = dm.numel() // (9 * max(np_, 1))
        if dm.numel() != 9 * np_ * nview:
            raise ValueError("dm must be (particles, views, 9)")
        score = torch.empty((np_, nview), dtype=torch.float32, device=outc.device)
        code = torch.empty((np_, nview), dtype=torch.int32, device=outc.device)
        m = np.ascontiguousarray(np.asarray(mask, np.int32))
        self._sync_stream()
        N.check(N.lib().fr_ccp_view_batch(self.h, nsam, np_, nview, _p(a3df), _p(outc), _p(ctff),
                                          _p(dm), int(maxr1), int(maxr2), float(rbfact),
                                          float(dang), int(iquadmax), _p(m), _p(score), _p(code)),
                "fr_ccp_view_batch")
        return score, code

    def search(self, nsam: int, maxr1: int, maxr2: int, rbfact: float, bank: torch.Tensor,
               outc, ctff, topk: int = 32, group: int = 1):
        """Top-k views (orientation, shift row, score) per particle against the bank;
        `group` consecutive orientations form one view (its in-plane angles)."""
        _dev(outc, torch.complex64, "outc")
        _dev(ctff, torch.complex64, "ctff")
        code = self._bank_code(nsam, maxr1, maxr2, bank)
        jc = nsam // 2 + 1
        np_ = outc.numel() // (nsam * jc)
        if ctff.numel() != 2 * outc.numel():
            raise ValueError("ctff must hold two half-planes per particle")
        dev = outc.device
        top_o = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_r = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_s = torch.empty((np_, topk), dtype=torch.float32, device=dev)
        self._sync_stream()
        N.check(N.lib().fr_search_batch(self.h, nsam, maxr1, maxr2, float(rbfact), bank.shape[0],
                                        _p(bank), code, int(group), np_,
                                        _p(outc), _p(ctff), topk, _p(top_o), _p(top_r),
                                        _p(top_s)), "fr_search_batch")
        return top_o, top_r, top_s

    def search_list(self, nsam: int, maxr1: int, maxr2: int, rbfact: float, bank: torch.Tensor,
                    outc, ctff, cand: torch.Tensor, group: int, topk: int = 32):
        """fr_search_list_batch: particle p against bank rows cand[p] (-1 = empty);
        returns (bank orientation, shift row, score) of the topk best groups."""
        _dev(outc, torch.complex64, "outc")
        _dev(ctff, torch.complex64, "ctff")
        _dev(cand, torch.int32, "cand")
        code = self._bank_code(nsam, maxr1, maxr2, bank)
        jc = nsam // 2 + 1
        np_ = outc.numel() // (nsam * jc)
        if ctff.numel() != 2 * outc.numel():
            raise ValueError("ctff must hold two half-planes per particle")
        if cand.dim() != 2 or cand.shape[0] != np_:
            raise ValueError("cand must be (particles, ncand)")
        if np_ and int(cand.max()) >= bank.shape[0]:
            raise ValueError("candidate index beyond the bank")
        dev = outc.device
        top_o = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_r = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_s = torch.empty((np_, topk), dtype=torch.float32, device=dev)
        self._sync_stream()
        N.check(N.lib().fr_search_list_batch(
            self.h, nsam, maxr1, maxr2, float(rbfact), bank.shape[0], _p(bank),
            code, np_, _p(outc), _p(ctff), cand.shape[1],
            _p(cand), int(group), topk, _p(top_o), _p(top_r), _p(top_s)), "fr_search_list_batch")
        return top_o, top_r, top_s

    def build_lists(self, top: torch.Tensor, prob: torch.Tensor, threshold: float,
                    cell_off: torch.Tensor, cells: torch.Tensor, group: int, nbank: int,
                    bank_group: int, maxlen: int):
        """fr_build_lists: ragged candidate lists from proposed classes (device
        tensors in and out, no host synchronisation).  Returns (loff (np+1),
        grp (np), cand (np*maxlen), info (np, 2): confident, valid entries)."""
        _dev(top, torch.int32, "top")
        _dev(prob, torch.float32, "prob")
        _dev(cell_off, torch.int32, "cell_off")
        _dev(cells, torch.int32, "cells")
        if top.dim() != 2 or prob.shape != top.shape:
            raise ValueError("top and prob must be (particles, ktop)")
        if cell_off.numel() < 2:
            raise ValueError("cell_off must hold ncls+1 offsets")
        np_, ktop = top.shape
        dev = top.device
        loff = torch.empty(np_ + 1, dtype=torch.int32, device=dev)
        grp = torch.empty(np_, dtype=torch.int32, device=dev)
        cand = torch.empty(max(np_ * maxlen, 1), dtype=torch.int32, device=dev)
        info = torch.empty((np_, 2), dtype=torch.int32, device=dev)
        self._sync_stream()
        N.check(N.lib().fr_build_lists(
            self.h, np_, ktop, _p(top), _p(prob), float(threshold), cell_off.numel() - 1,
            _p(cell_off), _p(cells), int(group), int(nbank), int(bank_group), int(maxlen),
            _p(loff), _p(grp), _p(cand), _p(info)), "fr_build_lists")
        return loff, grp, cand, info

    def search_ragged(self, nsam: int, maxr1: int, maxr2: int, rbfact: float,
                      bank: torch.Tensor, outc, ctff, loff: torch.Tensor, grp: torch.Tensor,
                      cand: torch.Tensor, topk: int = 32, info: torch.Tensor | None = None):
        """fr_search_ragged_batch: particle p against bank rows
        cand[loff[p]:loff[p+1]] in view groups of grp[p]; returns (bank
        orientation, shift row, score) of the topk best groups.  `info`
        (from build_lists) lets the particles that fall back to the whole
        bank share its templates."""
        _dev(outc, torch.complex64, "outc")
        _dev(ctff, torch.complex64, "ctff")
        for t, nm in ((loff, "loff"), (grp, "grp"), (cand, "cand")):
            _dev(t, torch.int32, nm)
        code = self._bank_code(nsam, maxr1, maxr2, bank)
        jc = nsam // 2 + 1
        np_ = outc.numel() // (nsam * jc)
        if ctff.numel() != 2 * outc.numel():
            raise ValueError("ctff must hold two half-planes per particle")
        if loff.numel() != np_ + 1 or grp.numel() != np_:
            raise ValueError("loff must hold np+1 offsets and grp np group sizes")
        dev = outc.device
        top_o = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_r = torch.empty((np_, topk), dtype=torch.int32, device=dev)
        top_s = torch.empty((np_, topk), dtype=torch.float32, device=dev)
        self._sync_stream()
        N.check(N.lib().fr_search_ragged_batch(
            self.h, nsam, maxr1, maxr2, float(rbfact), bank.shape[0], _p(bank),
            code, np_, _p(outc), _p(ctff), _p(loff), _p(grp),
            _p(cand), None if info is None else _p(info), cand.numel(), topk, _p(top_o),
            _p(top_r), _p(top_s)), "fr_search_ragged_batch")
        return top_o, top_r, top_s

    def rescore(self, nsam: int, a3df, outc, ctff, top_o: torch.Tensor, top_r: torch.Tensor,
                dm_bank: torch.Tensor, nsamr: int, maxr1: int, maxr2: int, rbfact: float, fsct,
                ring):
        """fr_rescore_batch: FP32 CC3M of each particle's candidates and the
        winner's position; returns (scores (np, k), -inf for empty candidates,
        best (np))."""
        for t, nm, dt in ((a3df, "a3df", torch.complex64), (outc, "outc", torch.complex64),
                          (ctff, "ctff", torch.complex64), (top_o, "top_o", torch.int32),
                          (top_r, "top_r", torch.int32), (dm_bank, "dm_bank", torch.float32),
                          (fsct, "fsct", torch.float32)):
            _dev(t, dt, nm)
        if top_o.dim() != 2 or top_r.shape != top_o.shape:
            raise ValueError("top_o and top_r must be (particles, k)")
        jc = nsam // 2 + 1
        if a3df.numel() != nsam * nsam * jc:
            raise ValueError("a3df must be the NSAM x NSAM x JC half-volume")
        if outc.numel() // (nsam * jc) != top_o.shape[0] or ctff.numel() != 2 * outc.numel():
            raise ValueError("outc/ctff must hold one particle per candidate row")
        if dm_bank.dim() != 2 or dm_bank.shape[1] != 9 or fsct.numel() < nsam // 2 + 1:
            raise ValueError("dm_bank must be (nbank, 9); fsct NSAM/2+1 long")
        np_, k = top_o.shape
        r = np.asarray(ring, dtype=np.float32)
        score = torch.empty((np_, k), dtype=torch.float32, device=outc.device)
        best = torch.empty(np_, dtype=torch.int32, device=outc.device)
        self._sync_stream()
        N.check(N.lib().fr_rescore_batch(
            self.h, nsam, np_, k, _p(a3df), _p(outc), _p(ctff), _p(top_o), _p(top_r),
            dm_bank.shape[0], _p(dm_bank), int(nsamr), maxr1, maxr2, float(rbfact), _p(fsct),
            _p(r), _p(score), _p(best)), "fr_rescore_batch")
        return score, best

    # ---------------------------------------------------------------- SIGMA
    def sigma(self, nsam: int, a3df, datc, ctff, pidx, par, amag, maxr1: int, rclas: float,
              dmask, sigp: float, signset=None) -> np.ndarray:
        """SIGMA for n items (fr_sigma_batch): (n, 3) SIGMA2N, SIGMA2, LOGPC."""
        _dev(a3df, torch.complex64, "a3df")
        _dev(datc, torch.complex64, "datc")
        _dev(ctff, torch.complex64, "ctff")
        jc = nsam // 2 + 1
        npart = datc.numel() // (nsam * jc)
        p = np.ascontiguousarray(par, np.float32).reshape(-1, 5)
        n = p.shape[0]
        pi = np.ascontiguousarray(pidx, np.int32)
        am = np.ascontiguousarray(amag, np.float32)
        if pi.size != n or am.size != n:
            raise ValueError("one pidx and amag per item")
        if n and (pi.max() >= npart or pi.min() < 0):
            raise ValueError("pidx out of range")
        if a3df.numel() != nsam * nsam * jc or ctff.numel() != 2 * datc.numel():
            raise ValueError("a3df/ctff sizes")
        dm = np.ascontiguousarray(dmask, np.float32)
        ss = None if signset is None else np.ascontiguousarray(signset, np.float32)
        out = np.zeros((n, 3), np.float32)
        self._sync_stream()
        N.check(N.lib().fr_sigma_batch(self.h, nsam, n, _p(a3df), _p(datc), _p(ctff), _p(pi),
                                       _p(p), _p(am), int(maxr1), float(rclas), _p(dm),
                                       float(sigp), _p(ss), None, _p(out)), "fr_sigma_batch")
        return out

    def set_cc3m_variant(self, nsam: int, fused: bool, workgroups: int = 0):
        N.check(N.lib().fr_set_cc3m_variant(self.h, nsam, int(bool(fused)), int(workgroups)),
                "fr_set_cc3m_variant")

    def autotune_cc3m(self, nsam: int, a3df, spec, ctff, pidx, dm, shift, amag, maxr1: int,
                      maxr2: int, rbfact: float, fsct, ring) -> dict:
        """fr_autotune_cc3m on these items; returns the timings and the choice."""
        import json
        n = pidx.numel()
        out = torch.empty(n, dtype=torch.float32, device=a3df.device)
        buf = C.create_string_buffer(4096)
        r = np.asarray(ring, dtype=np.float32)
        self._sync_stream()
        N.check(N.lib().fr_autotune_cc3m(self.h, nsam, n, _p(a3df), _p(spec), _p(ctff), _p(pidx),
                                         _p(dm), _p(shift), _p(amag), maxr1, maxr2, float(rbfact),
                                         _p(fsct), _p(r), _p(out), buf, 4096),
                "fr_autotune_cc3m")
        return json.loads(buf.value.decode())

    # -------------------------------------------------------------- PREFINE
    def refine(self, nsam: int, a3df, spec, ctff, pidx, amag, par, mask, maxr1: int, maxr2: int,
               rbfact: float, fsct, ring, prior, maxit: int = 50):
        """PREFINE for n items (fr_refine_batch).  par (n, 6) host float32:
        PHI THETA PSI DSHX DSHY in; returns (refined par with PRESA in column 5,
        evaluations per item)."""
        _dev(a3df, torch.complex64, "a3df")
        _dev(spec, torch.complex64, "spec")
        _dev(ctff, torch.complex64, "ctff")
        _dev(fsct, torch.float32, "fsct")
        jc = nsam // 2 + 1
        npart = spec.numel() // (nsam * jc)
        p = np.array(par, np.float32).reshape(-1, 6).copy()
        n = p.shape[0]
        pi = np.ascontiguousarray(pidx, np.int32)
        am = np.ascontiguousarray(amag, np.float32)
        pr = np.ascontiguousarray(prior, np.float32).reshape(n, 5)
        if pi.size != n or am.size != n:
            raise ValueError("one pidx and amag per item")
        if n and (pi.max() >= npart or pi.min() < 0):
            raise ValueError("pidx out of range")
        if a3df.numel() != nsam * nsam * jc or ctff.numel() != 2 * spec.numel():
            raise ValueError("a3df/ctff sizes")
        m = np.ascontiguousarray(mask, np.int32)
        r = np.asarray(ring, np.float32)
        nev = np.zeros(n, np.int32)
        self._sync_stream()
        N.check(N.lib().fr_refine_batch(self.h, nsam, n, _p(a3df), _p(spec), _p(ctff), _p(pi),
                                        _p(am), _p(p), _p(m), maxr1, maxr2, float(rbfact),
                                        _p(fsct), _p(r), _p(pr), int(maxit), _p(nev)),
                "fr_refine_batch")
        return p, nev

    def ctfrefine(self, nsam: int, spectra, ppar, euler, groups, a3df, pssnr, gpar, *,
                  rmax1: float, rmax2: float, rbfact: float, ring, fastig: bool = False,
                  dfstd: float = 100.0, sig2=None, magnification: bool = False):
        """CTFREFINE (or MAGREFINE with magnification=True) for groups of
        consecutive particles (fr_ctfrefine_batch / fr_magrefine_batch).
        spectra: the (np, NSAM, NSAM+2) float32 r2c spectra fr_prep_batch leaves
        in its image buffer; ppar (np, 8) and euler (np, 3) host; groups: the
        group sizes.  Returns (np.ndarray (ng, 4): dDF1 dDF2 dANGAST or AMAG 0 0,
        and the final objective; evaluations per group)."""
        _dev(spectra, torch.float32, "spectra")
        _dev(a3df, torch.complex64, "a3df")
        _dev(pssnr, torch.float32, "pssnr")
        np_ = spectra.numel() // (nsam * (nsam + 2))
        pp = np.ascontiguousarray(np.asarray(ppar, np.float32).reshape(np_, 8))
        eu = np.ascontiguousarray(np.asarray(euler, np.float32).reshape(np_, 3))
        sizes = np.asarray(groups, np.int64)
        gs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        if gs[-1] != np_:
            raise ValueError("groups must cover the particles")
        ng = len(sizes)
        out = np.zeros((ng, 4), np.float32)
        nev = np.zeros(ng, np.int32)
        g = np.asarray(gpar, np.float32)
        r = np.asarray(ring, np.float32)
        self._sync_stream()
        if magnification:
            N.check(N.lib().fr_magrefine_batch(self.h, nsam, np_, _p(spectra), _p(pp), _p(eu), ng,
                                               _p(gs), _p(a3df), _p(pssnr), _p(g), float(rmax1),
                                               float(rmax2), float(rbfact), _p(r), _p(out),
                                               _p(nev)), "fr_magrefine_batch")
        else:
            s2 = np.zeros(ng, np.float32) if sig2 is None else \
                np.ascontiguousarray(np.asarray(sig2, np.float32).reshape(ng))
            N.check(N.lib().fr_ctfrefine_batch(self.h, nsam, np_, _p(spectra), _p(pp), _p(eu), ng,
                                               _p(gs), int(bool(fastig)), float(dfstd), _p(s2),
                                               _p(a3df), _p(pssnr), _p(g), float(rmax1),
                                               float(rmax2), float(rbfact), _p(r), _p(out),
                                               _p(nev)), "fr_ctfrefine_batch")
        return out, nev

    def local_search(self, nsam: int, a3df, spec, ctff, par, *, nang: int, dang: float,
                     nsh: int, dsh: float, maxr1: int, maxr2: int, rbfact: float, fsct, ring,
                     amag=None, sm=None, all_scores: bool = False):
        """The local grid search before PREFINE (fr_local_search_batch): per
        particle, (2 nang + 1)^3 orientations DANG apart (radians) around par's
        PHI THETA PSI, each scored by CC3M at the (2 nsh + 1)^2 shifts DSH apart
        (phase units) around its SHX SHY.  par (np, 5 or 6) host.  Returns the
        (np, 6) best grid points (PHI THETA PSI SHX SHY CC3M) and, with
        all_scores, the device (np, norient) best score and shift index of
        every orientation."""
        _dev(a3df, torch.complex64, "a3df")
        _dev(spec, torch.complex64, "spec")
        _dev(ctff, torch.complex64, "ctff")
        _dev(fsct, torch.float32, "fsct")
        jc = nsam // 2 + 1
        np_ = spec.numel() // (nsam * jc)
        pin = np.asarray(par, np.float32).reshape(np_, -1)
        p = np.zeros((np_, 6), np.float32)
        p[:, :5] = pin[:, :5]
        if a3df.numel() != nsam * nsam * jc or ctff.numel() != 2 * spec.numel():
            raise ValueError("a3df/ctff sizes")
        if fsct.numel() < jc:
            raise ValueError("fsct needs NSAM/2+1 values")
        am = np.ones(np_, np.float32) if amag is None else \
            np.ascontiguousarray(np.asarray(amag, np.float32).reshape(np_))
        r = np.asarray(ring, np.float32)
        smv = None if sm is None else np.ascontiguousarray(sm, dtype=np.float32)
        norient = (2 * nang + 1) ** 3
        sc = cd = None
        if all_scores:
            sc = torch.empty((np_, norient), dtype=torch.float32, device=spec.device)
            cd = torch.empty((np_, norient), dtype=torch.int32, device=spec.device)
        self._sync_stream()
        N.check(N.lib().fr_local_search_batch(self.h, nsam, np_, _p(a3df), _p(spec), _p(ctff),
                                              _p(am), _p(p), int(nang), float(dang), int(nsh),
                                              float(dsh), int(maxr1), int(maxr2), float(rbfact),
                                              _p(fsct), _p(r), _p(smv), _p(sc), _p(cd)),
                "fr_local_search_batch")
        if all_scores:
            return p, sc, cd
        return p

    # ------------------------------------------------ proposal network
    def cnn_init(self, shape, seed: int, master: torch.Tensor, wq: torch.Tensor):
        _dev(master, torch.float32, "master")
        _dev(wq, torch.float32, "wq")
        self._sync_stream()
        N.check(N.lib().fr_cnn_init(self.h, *shape, C.c_ulonglong(seed), _p(master), _p(wq)),
                "fr_cnn_init")

    def cnn_forward(self, shape, wq: torch.Tensor, img: torch.Tensor) -> torch.Tensor:
        _dev(wq, torch.float32, "wq")
        _dev(img, torch.float32, "img")
        S, K = shape[0], shape[3]
        nb = img.numel() // (S * S)
        if img.numel() != nb * S * S:
            raise ValueError("img must be nb x S x S")
        out = torch.empty((nb, K), dtype=torch.float32, device=img.device)
        self._sync_stream()
        N.check(N.lib().fr_cnn_forward(self.h, *shape, _p(wq), nb, _p(img), _p(out)),
                "fr_cnn_forward")
        return out

    def cnn_forward_bf16(self, shape, wq: torch.Tensor, img: torch.Tensor) -> torch.Tensor:
        """fr_cnn_forward_bf16: the logits with BF16 activations (C = 64)."""
        _dev(wq, torch.float32, "wq")
        _dev(img, torch.float32, "img")
        S, K = shape[0], shape[3]
        nb = img.numel() // (S * S)
        if img.numel() != nb * S * S:
            raise ValueError("img must be nb x S x S")
        out = torch.empty((nb, K), dtype=torch.float32, device=img.device)
        self._sync_stream()
        N.check(N.lib().fr_cnn_forward_bf16(self.h, *shape, _p(wq), nb, _p(img), _p(out)),
                "fr_cnn_forward_bf16")
        return out

    def cnn_gradients(self, shape, wq: torch.Tensor, img: torch.Tensor, label: torch.Tensor,
                      grad: torch.Tensor, bf16: bool = False) -> torch.Tensor:
        """Per-image losses and the gradient of their mean (fr_cnn_gradients);
        with `bf16` the mixed-precision step (fr_cnn_gradients_bf16: BF16 GEMM
        operands, FP32 accumulation)."""
        _dev(wq, torch.float32, "wq")
        _dev(img, torch.float32, "img")
        _dev(label, torch.int32, "label")
        _dev(grad, torch.float32, "grad")
        S, K = shape[0], shape[3]
        nb = img.numel() // (S * S)
        if label.numel() != nb or (nb and (int(label.min()) < 0 or int(label.max()) >= K)):
            raise ValueError("one label in 0..K-1 per image")
        loss = torch.empty(nb, dtype=torch.float32, device=img.device)
        self._sync_stream()
        fn = "fr_cnn_gradients_bf16" if bf16 else "fr_cnn_gradients"
        N.check(getattr(N.lib(), fn)(self.h, *shape, _p(wq), nb, _p(img), _p(label), _p(loss),
                                     _p(grad)), fn)
        return loss

    def cnn_adam(self, shape, master, wq, grad, m, v, step: int, lr: float):
        for t, nm in ((master, "master"), (wq, "wq"), (grad, "grad"), (m, "m"), (v, "v")):
            _dev(t, torch.float32, nm)
        self._sync_stream()
        N.check(N.lib().fr_cnn_adam(self.h, *shape, _p(master), _p(wq), _p(grad), _p(m), _p(v),
                                    int(step), float(lr)), "fr_cnn_adam")

    def cnn_propose(self, logits: torch.Tensor, topk: int):
        _dev(logits, torch.float32, "logits")
        nb, K = logits.shape
        top = torch.empty((nb, topk), dtype=torch.int32, device=logits.device)
        prob = torch.empty((nb, topk), dtype=torch.float32, device=logits.device)
        self._sync_stream()
        N.check(N.lib().fr_cnn_propose(self.h, K, nb, _p(logits), topk, _p(top), _p(prob)),
                "fr_cnn_propose")
        return top, prob

    def cnn_images(self, nsam: int, S: int, datc, ctff) -> torch.Tensor:
        _dev(datc, torch.complex64, "datc")
        _dev(ctff, torch.complex64, "ctff")
        jc = nsam // 2 + 1
        nb = datc.numel() // (nsam * jc)
        if ctff.numel() != 2 * datc.numel():
            raise ValueError("ctff must hold two half-planes per particle")
        img = torch.empty((nb, S, S), dtype=torch.float32, device=datc.device)
        self._sync_stream()
        N.check(N.lib().fr_cnn_images(self.h, nsam, S, nb, _p(datc), _p(ctff), _p(img)),
                "fr_cnn_images")
        return img

    def cnn_synth(self, nsam: int, S: int, a3df, dm, spar, gpar, disc: float, seed: int):
        _dev(a3df, torch.complex64, "a3df")
        _dev(dm, torch.float32, "dm")
        _dev(spar, torch.float32, "spar")
        nb = dm.numel() // 9
        if spar.numel() != 8 * nb or a3df.numel() != nsam * nsam * (nsam // 2 + 1):
            raise ValueError("spar must be nb x 8 and a3df the half-volume")
        g = np.ascontiguousarray(gpar, dtype=np.float32)
        if g.size != 11:
            raise ValueError("gpar has 11 entries")
        img = torch.empty((nb, S, S), dtype=torch.float32, device=dm.device)
        self._sync_stream()
        N.check(N.lib().fr_cnn_synth(self.h, nsam, S, nb, _p(a3df), _p(dm), _p(spar), _p(g),
                                     float(disc), C.c_ulonglong(seed), _p(img)), "fr_cnn_synth")
        return img

    def orient_nearest(self, dm, dmc, sym) -> torch.Tensor:
        for t, nm in ((dm, "dm"), (dmc, "dmc"), (sym, "sym")):
            _dev(t, torch.float32, nm)
        n = dm.numel() // 9
        lab = torch.empty(n, dtype=torch.int32, device=dm.device)
        self._sync_stream()
        N.check(N.lib().fr_orient_nearest(self.h, n, _p(dm), dmc.numel() // 9, _p(dmc),
                                          sym.numel() // 9, _p(sym), _p(lab)), "fr_orient_nearest")
        return lab

    def orient_within(self, dma, dmb, sym, max_deg: float) -> torch.Tensor:
        for t, nm in ((dma, "dma"), (dmb, "dmb"), (sym, "sym")):
            _dev(t, torch.float32, nm)
        na, nb = dma.numel() // 9, dmb.numel() // 9
        w = (na + 31) // 32
        mask = torch.zeros((nb, w), dtype=torch.int32, device=dma.device)
        self._sync_stream()
        N.check(N.lib().fr_orient_within(self.h, na, _p(dma), nb, _p(dmb), sym.numel() // 9,
                                         _p(sym), float(max_deg), _p(mask)), "fr_orient_within")
        return mask

    # ----------------------------------------------------------- statistics
    def sheltest(self, nsam: int, a, d, s, v, istp: int = 1, fmask: float = 1.0):
        jc = nsam // 2 + 1
        cap = jc + 2
        npix = np.zeros(cap, np.int32)
        npixt = np.zeros(cap, np.int32)
        dfsc = np.zeros(cap, np.float64)
        pssnr = np.zeros(cap, np.float32)
        nstp = C.c_int(0)
        self._sync_stream()
        N.check(N.lib().fr_sheltest(self.h, nsam, istp, _p(a), _p(d), _p(s), _p(v), float(fmask),
                                    _p(npix), _p(npixt), _p(dfsc), _p(pssnr), C.byref(nstp)),
                "fr_sheltest")
        k = nstp.value
        return dict(npix=npix[:k], npixt=npixt[:k], fsc=dfsc[:k], pssnr=pssnr[:k])

    def storeshift(self, nsam: int, std: float, vtd: float, iredun: int, a, s, d, v):
        b = torch.empty_like(a)
        c = torch.empty_like(d)
        self._sync_stream()
        N.check(N.lib().fr_storeshift(self.h, nsam, float(std), float(vtd), int(iredun), _p(a),
                                      _p(s), _p(d), _p(v), _p(b), _p(c)), "fr_storeshift")
        return b, c

    def combine_arrays(self, nsam: int, nn3: int, nvol1: int, nvol2: int, ifsc: int, a, s, d, v,
                       imp: int, nksum1, ttd1, ttd2, nksum: int, std: float, vtd: float):
        """COMBINE_ARRAYS (combine_arrays.f): copies summed into copy 1 in place
        (a, s, d, v hold the copies nn3 elements apart); returns the updated
        (NKSUM, STD, VTD)."""
        k1 = np.ascontiguousarray(nksum1, dtype=np.int64)
        t1 = np.ascontiguousarray(ttd1, dtype=np.float64)
        t2 = np.ascontiguousarray(ttd2, dtype=np.float64)
        nk, sd, vd = C.c_longlong(int(nksum)), C.c_double(std), C.c_double(vtd)
        self._sync_stream()
        N.check(N.lib().fr_combine_arrays(self.h, nsam, int(nn3), nvol1, nvol2, ifsc, _p(a), _p(s),
                                          _p(d), _p(v), imp, _p(k1), _p(t1), _p(t2), C.byref(nk),
                                          C.byref(sd), C.byref(vd)), "fr_combine_arrays")
        return nk.value, sd.value, vd.value

    def shiftvol(self, nsam: int, std: float, vtd: float, rlim: float, iredun: int, a, d, s, v,
                 ifsc: int, pssnr, fpart: float, fsc, fmask: float):
        """SHIFTVOL (shiftvol.f) in place on a (and s for ifsc 0); returns STD+VTD."""
        ps = np.ascontiguousarray(pssnr, dtype=np.float32)
        fs = np.ascontiguousarray(fsc, dtype=np.float32)
        sd = C.c_double(std)
        self._sync_stream()
        N.check(N.lib().fr_shiftvol(self.h, nsam, C.byref(sd), float(vtd), float(rlim), int(iredun),
                                    _p(a), _p(d), _p(s), _p(v), ifsc, _p(ps), float(fpart),
                                    _p(fs), float(fmask), min(ps.size, fs.size)), "fr_shiftvol")
        return sd.value

    def correct3d(self, nsam: int, a3dv, interp: int, ipad: int, sinclut):
        """CORRECT3D (correct3d.f) in place on the padded real volume."""
        lut = np.ascontiguousarray(sinclut, dtype=np.float32)
        self._sync_stream()
        N.check(N.lib().fr_correct3d(self.h, nsam, _p(a3dv), interp, ipad, _p(lut), lut.size),
                "fr_correct3d")

    def fft3d(self, nsam: int, vol, forward: bool):
        """3D r2c (forward) or c2r transform of one padded volume, in place."""
        self._sync_stream()
        fn = N.lib().fr_fft3d_r2c if forward else N.lib().fr_fft3d_c2r
        N.check(fn(self.h, nsam, _p(vol)), "fr_fft3d")

    def find_fpart(self, nsam: int, std: float, vtd: float, a, d, s, v, iredun: int, ri: float,
                   ric: float, psize: float, mw: float, dalt: float, fsc, pssnr):
        """FIND_FPART (fr_find_fpart): the particle fraction of the Wiener
        filter from the half-maps (the MW = 0 search, or the MW closed form);
        returns (FPART, PSSNR / FPART).  a, d: complex half-volumes; s, v:
        their CTF^2 sums; fsc, pssnr: SHELTEST's shells."""
        for t, nm, dt in ((a, "a", torch.complex64), (d, "d", torch.complex64),
                          (s, "s", torch.float32), (v, "v", torch.float32)):
            _dev(t, dt, nm)
        jc = nsam // 2 + 1
        n = max(len(fsc), len(pssnr), jc)
        fs = np.zeros(n, np.float32)
        fs[:len(fsc)] = fsc
        ps = np.zeros(n, np.float32)
        ps[:len(pssnr)] = pssnr
        fp = C.c_float(0.0)
        self._sync_stream()
        N.check(N.lib().fr_find_fpart(self.h, nsam, float(std), float(vtd), _p(a), _p(d), _p(s),
                                      _p(v), int(iredun), float(ri), float(ric), float(psize),
                                      float(mw), float(dalt), fs.ctypes.data, ps.ctypes.data, n,
                                      C.byref(fp)), "fr_find_fpart")
        return fp.value, ps[:len(pssnr)]

    def finalize(self, nsam: int, vols: "HalfMaps", iredun: int, psize: float, ri: float,
                 ric: float, mw: float, dalt: float, ffilt: bool, rlim: float):
        """fr_finalize: the end of a refinement cycle (frealign_v9.f:790-950) in
        one call.  vols.num[0] and vols.ctf2[0] are overwritten (the final
        real-space map ends in vols.num[0], padded layout); returns a dict with
        the final map as an (NSAM, NSAM, NSAM) view, the masked half-maps'
        transforms (b, c) and the FSC table's numbers."""
        jc = nsam // 2 + 1
        cap = jc + 2
        b = torch.empty_like(vols.num[0])
        c = torch.empty_like(vols.num[1])
        fsc = np.zeros(cap, np.float32)
        pssnr = np.zeros(cap, np.float32)
        npix = np.zeros(cap, np.int32)
        npixt = np.zeros(cap, np.int32)
        nk = np.ascontiguousarray(vols.nksum.cpu().numpy(), np.int64)
        td = np.ascontiguousarray(vols.ttd.cpu().numpy(), np.float64)
        nstp, fm, fp = C.c_int(0), C.c_float(0.0), C.c_float(0.0)
        self._sync_stream()
        N.check(N.lib().fr_finalize(
            self.h, nsam, _p(vols.num[0]), _p(vols.ctf2[0]), _p(vols.num[1]), _p(vols.ctf2[1]),
            nk.ctypes.data, td.ctypes.data, int(iredun), float(psize), float(ri), float(ric),
            float(mw), float(dalt), int(bool(ffilt)), float(rlim), _p(b), _p(c), cap,
            fsc.ctypes.data, pssnr.ctypes.data, npix.ctypes.data, npixt.ctypes.data,
            C.byref(nstp), C.byref(fm), C.byref(fp)), "fr_finalize")
        k = nstp.value
        final = vols.num[0].view(torch.float32).reshape(nsam, nsam, nsam + 2)[..., :nsam]
        return dict(final=final, b=b, c=c, fsc=fsc[:k], pssnr=pssnr[:k], npix=npix[:k],
                    npixt=npixt[:k], fmask=fm.value, fpart=fp.value)

    def mask3d(self, nsam: int, a3dv, rio: float, ric: float, halfw: float):
        """MASK3D (mask3d.f) in place on the padded real volume; returns
        (FMASK, AVE, STD)."""
        fm, av, sd = C.c_float(), C.c_float(), C.c_float()
        self._sync_stream()
        N.check(N.lib().fr_mask3d(self.h, nsam, _p(a3dv), float(rio), float(ric), float(halfw),
                                  C.byref(fm), C.byref(av), C.byref(sd)), "fr_mask3d")
        return fm.value, av.value, sd.value


@dataclass
class HalfMaps:
    """The two half-map accumulators of A3D3 (A3DF/S3DF and D3DF/V3DF) with
    their NKSUM1 / TTD1 scalars, resident on the device."""
    nsam: int
    num: tuple
    ctf2: tuple
    nksum: torch.Tensor
    ttd: torch.Tensor

    @staticmethod
    def zeros(nsam: int, device="cuda") -> "HalfMaps":
        jc = nsam // 2 + 1
        shp = (nsam, nsam, jc)
        num = tuple(torch.zeros(shp, dtype=torch.complex64, device=device) for _ in range(2))
        ctf2 = tuple(torch.zeros(shp, dtype=torch.float32, device=device) for _ in range(2))
        return HalfMaps(nsam, num, ctf2, torch.zeros(2, dtype=torch.int64, device=device),
                        torch.zeros(2, dtype=torch.float64, device=device))


@dataclass
class VoxelStats:
    """The per-voxel statistics of the _SE insertion (NNSTAT = 1,
    pinsert_trilin_se.f:186-194) over both half-maps: KSUM (terms received),
    VSUM (sum |term|^2), PSUM (sum of phase differences to the reference
    transform `ref`), ASUM (sum |term|), resident on the device."""
    nsam: int
    ref: torch.Tensor
    ksum: torch.Tensor
    vsum: torch.Tensor
    psum: torch.Tensor
    asum: torch.Tensor

    @staticmethod
    def zeros(nsam: int, ref: torch.Tensor, device="cuda") -> "VoxelStats":
        _dev(ref, torch.complex64, "ref")
        jc = nsam // 2 + 1
        if ref.numel() != nsam * nsam * jc:
            raise ValueError("ref must be the (NSAM, NSAM, NSAM/2+1) half-volume")
        shp = (nsam, nsam, jc)
        z = lambda dt: torch.zeros(shp, dtype=dt, device=device)  # noqa: E731
        return VoxelStats(nsam, ref, z(torch.int32), z(torch.float32), z(torch.float32),
                          z(torch.float32))


class Plan:
    """A graph-captured score-and-insert pipeline for batches of one box size
    (fr_plan_*): CTFAPPLY -> CC3M at each particle's parameters -> insertion
    into `vols`."""

    def __init__(self, ctx: "Context", nsam: int, nb: int, a3df, pssnr, gpar, maxr1: int,
                 maxr2: int, rbfact: float, fsct, ring, sym, pbc: float, boff: float,
                 psize: float, vols: "HalfMaps", ictf: int = 1):
        _dev(a3df, torch.complex64, "a3df")
        _dev(pssnr, torch.float32, "pssnr")
        _dev(fsct, torch.float32, "fsct")
        if vols.nsam != nsam or a3df.numel() != nsam * nsam * (nsam // 2 + 1):
            raise ValueError("a3df and half-maps must match NSAM")
        nsym, symop, jsym, _ = sym
        self.ctx, self.nsam, self.nb, self.vols = ctx, nsam, nb, vols
        self._keep = (a3df, pssnr, fsct)
        g = np.ascontiguousarray(gpar, np.float32)
        r = np.ascontiguousarray(ring, np.float32)
        h = C.c_void_p()
        ctx._sync_stream()
        N.check(N.lib().fr_plan_create(ctx.h, nsam, nb, _p(a3df), _p(pssnr), _p(g), ictf, maxr1,
                                       maxr2, float(rbfact), _p(fsct), _p(r), nsym, _p(jsym),
                                       _p(symop), float(pbc), float(boff), float(psize),
                                       _p(vols.num[0]), _p(vols.ctf2[0]), _p(vols.num[1]),
                                       _p(vols.ctf2[1]), _p(vols.nksum), _p(vols.ttd),
                                       C.byref(h)), "fr_plan_create")
        self.h = h

    def run(self, images: torch.Tensor, ppar: np.ndarray, recb: np.ndarray,
            half: np.ndarray) -> np.ndarray:
        _dev(images, torch.float32, "images")
        if images.numel() != self.nb * self.nsam * (self.nsam + 2):
            raise ValueError("images must be nb x NSAM x (NSAM+2)")
        pp = np.ascontiguousarray(ppar, np.float32).reshape(self.nb, 8)
        rb = np.ascontiguousarray(recb, np.float32).reshape(self.nb, 10)
        hf = np.ascontiguousarray(half, np.int32).reshape(self.nb)
        score = np.zeros(self.nb, np.float32)
        self.ctx._sync_stream()
        N.check(N.lib().fr_plan_load(self.h, _p(images)), "fr_plan_load")
        N.check(N.lib().fr_plan_run(self.h, _p(pp), _p(rb), _p(hf), _p(score)), "fr_plan_run")
        return score

    def close(self):
        if self.h:
            N.lib().fr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- geometry
POWELL_FN = C.CFUNCTYPE(C.c_float, C.c_int, C.POINTER(C.c_float), C.c_void_p)


def powell_minimize(fn, x, e, escale: float = 100.0, icon: int = 1, maxit: int = 50,
                    user=None):
    """VA04A (fr_powell_minimize) on a host objective.  fn: a POWELL_FN, or a
    Python callable f(x: np.ndarray) -> float.  Returns (x, F, evaluations)."""
    x = np.array(x, np.float32).copy()
    e = np.ascontiguousarray(e, np.float32)
    if not isinstance(fn, (POWELL_FN, C._CFuncPtr)):
        py = fn

        def tramp(n, px, _u):
            return float(py(np.ctypeslib.as_array(px, shape=(n,)).copy()))
        fn = POWELL_FN(tramp)
    f = C.c_float(0)
    nev = C.c_int(0)
    N.check(N.lib().fr_powell_minimize(x.size, _p(x), _p(e), float(escale), int(icon), int(maxit),
                                       C.cast(fn, C.c_void_p), user, C.byref(f), C.byref(nev)),
            "fr_powell_minimize")
    return x, np.float32(f.value), nev.value


def cnn_param_count(S: int, C_: int, nconv: int, K: int) -> int:
    n = C.c_longlong(0)
    N.check(N.lib().fr_cnn_param_count(S, C_, nconv, K, C.byref(n)), "fr_cnn_param_count")
    return n.value


def search_grid(dang: float, asym: str = "C", nasym: int = 1, pmask=(1, 1, 1, 1, 1),
                rmm: float = 0.1, ri: float = 50.0, asymstore: bool = True):
    """SEARCHANG (searchang.f): returns (sang (n,3) radians float32, iquadmax)."""
    m = np.asarray(pmask, dtype=np.int32)
    n, iq = C.c_int(0), C.c_int(0)
    N.check(N.lib().fr_search_grid(rmm, ri, dang, asym.encode(), nasym, _p(m), int(asymstore),
                                   None, 0, C.byref(n), C.byref(iq)), "fr_search_grid")
    sang = np.zeros((max(n.value, 1), 3), np.float32)
    N.check(N.lib().fr_search_grid(rmm, ri, dang, asym.encode(), nasym, _p(m), int(asymstore),
                                   _p(sang), n.value, C.byref(n), C.byref(iq)), "fr_search_grid")
    return sang[: n.value], iq.value


def point_group(asym: str, nasym: int):
    """GETSYMMAT: (nsym, symop (72,) float32, jsym (8,) int32, iredun)."""
    symop = np.zeros(72, np.float32)
    jsym = np.zeros(8, np.int32)
    nsym, ired = C.c_int(0), C.c_int(0)
    N.check(N.lib().fr_point_group(asym.encode(), nasym, _p(symop), _p(jsym), C.byref(nsym),
                                   C.byref(ired)), "fr_point_group")
    return nsym.value, symop, jsym, ired.value


def rotation(euler, amag: float = 1.0, sm=IDENTITY_SM) -> np.ndarray:
    """DM matrices (n, 9) of cc3m.f:48-65 (amag=0: pextract.f form)."""
    e = np.ascontiguousarray(np.asarray(euler, dtype=np.float32).reshape(-1, 3))
    dm = np.zeros((e.shape[0], 9), np.float32)
    smv = None if sm is None else np.ascontiguousarray(sm, dtype=np.float32)
    N.check(N.lib().fr_rotation(e.shape[0], _p(e), float(amag), _p(smv) if smv is not None else None,
                                _p(dm)), "fr_rotation")
    return dm


def lcg_random(iran: int):
    s = C.c_int(iran)
    v = N.lib().fr_lcg_random(C.byref(s))
    return v, s.value
