"""The reference's files on the host: MRC/CCP4 stacks and maps (iof.f) and
parameter files (card10.f reads them, lmain.f writes them), over the native
readers and writers of libfrealign_hip (fr_mrc_*, fr_par_*).

Units.  A par file holds angles in degrees, shifts in Angstroms, OCC and
SCORE in percent (lmain.f:913-923).  In memory the hot path uses radians,
shifts in phase units (SHX = shift / PSX / NSAM * 2 pi, PSX = RELMAG * DSTEP
* 10^4 / MAG, card10.f:269-295) and fractions; `ParFile.to_internal` and
`par_line` convert as the reference does.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

PI_F = np.float32(3.1415926535897)


# ------------------------------------------------------------------- MRC
@dataclass
class MrcInfo:
    nx: int
    ny: int
    nz: int
    mode: int
    psize: float
    dmin: float
    dmax: float
    dmean: float
    rms: float
    nsymbt: int


def mrc_info(path: str) -> MrcInfo:
    nxyz = np.zeros(3, np.int32)
    mode, nsym = C.c_int(0), C.c_int(0)
    ps = C.c_float(0.0)
    st = np.zeros(4, np.float32)
    N.check(N.lib().fr_mrc_info(path.encode(), nxyz.ctypes.data, C.addressof(mode),
                                C.addressof(ps), st.ctypes.data, C.addressof(nsym)),
            "fr_mrc_info")
    return MrcInfo(int(nxyz[0]), int(nxyz[1]), int(nxyz[2]), mode.value, ps.value,
                   float(st[0]), float(st[1]), float(st[2]), float(st[3]), nsym.value)


def mrc_read(path: str, first: int = 0, count: int | None = None) -> np.ndarray:
    """Sections first .. first+count-1 as float32 (count, NY, NX)."""
    info = mrc_info(path)
    if count is None:
        count = info.nz - first
    out = np.empty((count, info.ny, info.nx), np.float32)
    N.check(N.lib().fr_mrc_read(path.encode(), int(first), int(count), out.ctypes.data),
            "fr_mrc_read")
    return out


def mrc_write(path: str, data: np.ndarray, psize: float, label: str | None = None) -> None:
    """A mode-2 file of data (NZ, NY, NX) (2D arrays: one section)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    lab = None if label is None else label.encode()[:80]
    N.check(N.lib().fr_mrc_write(path.encode(), nx, ny, nz, float(psize), a.ctypes.data, lab),
            "fr_mrc_write")


def title(asym: str) -> str:
    """IOPEN's title line (iof.f:198-205) for a map of the given symmetry."""
    import time
    t = list("Frealign vX.XX - XX.XX.XX, output symmetry ".ljust(80))
    vx = "frealign_amd".ljust(15)[:15]
    t[10:25] = list(vx)
    t[43:46] = list(asym.ljust(3)[:3])
    t[46:48] = list("  ")
    t[48:72] = list(time.strftime("%a %b %d %H:%M:%S %Y").ljust(24)[:24])
    return "".join(t)


# -------------------------------------------------------------- par files
PAR_HEADER = ("C           PSI   THETA     PHI       SHX       SHY     MAG  FILM      DF1"
              "      DF2  ANGAST     OCC      LogP      SIGMA   SCORE  CHANGE")


@dataclass
class ParRecord:
    c: int
    psi: float
    theta: float
    phi: float
    shx: float      # Angstrom
    shy: float
    mag: float
    film: int
    df1: float
    df2: float
    angast: float   # degrees
    occ: float      # percent
    logp: float
    sigma: float
    score: float    # percent
    change: float = 0.0


@dataclass
class ParFile:
    records: list
    comments: list = field(default_factory=list)
    part_ssnr: np.ndarray | None = None   # sqrt(Part_SSNR) per shell, from the table
    formats: set = field(default_factory=set)

    def to_internal(self, nsam: int, relmag: float, dstep: float):
        """card10.f:269-295 for every record: (C, PSI, THETA, PHI, SHX, SHY,
        AMAGP-free magnification, FILM, DF1, DF2, ANGAST, OCC, LogP, SIGMA,
        PRESA) with angles in radians, shifts in phase units, OCC and PRESA as
        fractions; angles outside +-180 wrapped first."""
        out = np.zeros((len(self.records), 15), np.float64)
        for k, r in enumerate(self.records):
            ang = []
            for a in (r.psi, r.theta, r.phi):
                a = np.float32(a)
                if a > 180.0:
                    a = np.float32(a - np.float32(360.0))
                if a < -180.0:
                    a = np.float32(a + np.float32(360.0))
                ang.append(np.float32(np.float32(a / np.float32(180.0)) * PI_F))
            psx = np.float32(np.float32(np.float32(relmag) * np.float32(dstep) * np.float32(10000.0))
                             / np.float32(abs(r.mag)))
            f2pi = np.float32(PI_F * np.float32(2.0))
            shx = np.float32(np.float32(np.float32(np.float32(r.shx) / psx) / np.float32(nsam)) * f2pi)
            shy = np.float32(np.float32(np.float32(np.float32(r.shy) / psx) / np.float32(nsam)) * f2pi)
            if abs(shx) > PI_F:
                shx = np.float32(0.0)
            if abs(shy) > PI_F:
                shy = np.float32(0.0)
            out[k] = (r.c, ang[0], ang[1], ang[2], shx, shy, r.mag, r.film, r.df1, r.df2,
                      np.float32(np.float32(r.angast) / np.float32(180.0)) * PI_F,
                      np.float32(r.occ) / np.float32(100.0), r.logp, r.sigma,
                      np.float32(r.score) / np.float32(100.0))
        return out


def par_parse(line: str):
    """One line through fr_par_parse: (None for a comment | values, format)."""
    rec = np.zeros(15, np.float64)
    fmt = C.c_int(0)
    r = N.lib().fr_par_parse(line.encode(), rec.ctypes.data, C.addressof(fmt))
    if r == 1:
        return None, -1
    N.check(r, "fr_par_parse")
    return rec, fmt.value


def par_line(r: ParRecord) -> str:
    buf = C.create_string_buffer(160)
    N.check(N.lib().fr_par_format(int(r.c), float(r.psi), float(r.theta), float(r.phi),
                                  float(r.shx), float(r.shy), float(r.mag), int(r.film),
                                  float(r.df1), float(r.df2), float(r.angast), float(r.occ),
                                  float(r.logp), float(r.sigma), float(r.score),
                                  float(r.change), buf, 160), "fr_par_format")
    return buf.value.decode()


def par_read(path: str) -> ParFile:
    """card10.f:92-210 for one dataset's file: comment lines kept, the
    Part_SSNR table (card10.f:104-126: shells 2 .. NSAM/2, column 3 the FSC,
    column 5 sqrt(Part_SSNR)) picked up, blank or non-positive records
    skipped, repeated particle numbers collapsed to the first."""
    recs, comments, fmts = [], [], set()
    ssnr = []
    in_table = False
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if line[:1] == "C":
                comments.append(line)
                if "Part_SSNR" in line:
                    in_table = True
                    continue
                if in_table:
                    # FORMAT(1X,I4,25X,F7.3,10X,F11.4): shell (columns 2-5), FSC
                    # (31-37), sqrt(Part_SSNR) (48-58)
                    try:
                        shell = int(line[1:5])
                        ssnr.append((shell, float(line[47:58].strip() or 0.0)))
                    except ValueError:
                        in_table = False
                continue
            in_table = False
            if not line.strip():
                continue
            vals, fmt = par_parse(line)
            if vals is None or vals[0] <= 0:
                continue
            if recs and int(abs(vals[0])) == abs(recs[-1].c):
                continue
            fmts.add(fmt)
            recs.append(ParRecord(int(vals[0]), *[float(v) for v in vals[1:7]], int(vals[7]),
                                  *[float(v) for v in vals[8:15]]))
    table = None
    if ssnr:
        table = np.array([v for _, v in sorted(ssnr)], np.float32)
    return ParFile(recs, comments, table, fmts)


def par_write(path: str, records, trailer=None) -> None:
    with open(path, "w") as f:
        f.write(PAR_HEADER + "\n")
        for r in records:
            f.write(par_line(r) + "\n")
        for line in trailer or []:
            f.write(line + "\n")


def fsc_table(nsam: int, psize: float, fsc, npixt, pssnr=None, npix=None, fspr=None,
              rlim: float = 0.5, fmask: float = 0.0, fpart: float = 0.0, version: str = ""):
    """The 'C' lines OPRESSTATHALF appends to the output par file
    (opresstathalf.f:46-168): the mask fractions, then for shells 2 .. NSTP
    with RING RAD <= RLIM the row NO, RESOL, RING RAD, FSPR, FSC, Part_FSC,
    sqrt(Part_SSNR), sqrt(Rec_SSNR), CC, EXP. C, SIG C, ERFC, TOTVOX
    (FORMAT 7014), the NPIX-weighted 'Average' row (FORMAT 7015) and the date
    line.  Rec_SSNR = 2 FSC / (1 - FSC + 1e-4), scaled by FMASK / FPART when
    FPART is set, which also gives Part_FSC = SSNR / (|SSNR| + 2); FSPR is
    written when given; CC, EXP. C, SIG C and ERFC (the PRESB and variance
    statistics of the refinement) are written as zeros.  npix (the shell's
    voxel count with both halves non-zero) weights the averages; npixt is
    the shell's voxel count."""
    import time
    lines = ["C", f"C  Fraction_mask, Fraction_particle = {fmask:11.5f}{fpart:11.5f}", "C",
             "C" + " " * 49 + "sqrt" + " " * 7 + "sqrt",
             "C  NO.  RESOL  RING RAD   FSPR    FSC  Part_FSC  Part_SSNR  Rec_SSNR       CC "
             "  EXP. C    SIG C  ERFC  TOTVOX"]
    f32 = np.float32
    npix = npixt if npix is None else npix
    av_pr = av_fsc = av_pfsc = av_ssnr = av_pssnr = f32(0.0)
    isumv = 0
    for i in range(2, len(fsc) + 1):
        k = i - 1
        ring = f32(f32(k) / f32(nsam))
        if ring > rlim:
            break
        fi = f32(fsc[k])
        ssnr = f32(f32(2.0) * fi / (f32(1.0) - fi + f32(0.0001)))
        pf = f32(0.0)
        if fi >= 1.0:
            pf = f32(1.0)
        elif fpart != 0.0:
            ssnr = f32(f32(f32(fmask) / f32(fpart)) * ssnr)
            pf = f32(ssnr / (abs(ssnr) + f32(2.0)))
        sq = f32(math.sqrt(abs(float(ssnr))))
        psn = f32(0.0) if pssnr is None else f32(max(float(pssnr[k]), 0.0))
        ps = f32(math.sqrt(float(psn)))
        pr = f32(0.0) if fspr is None else f32(fspr[k])
        lines.append("C" + f"{i:4d}{f32(psize) / ring:8.2f}{ring:10.4f}{pr:7.2f}{fi:7.3f}"
                     f"{pf:10.3f}{ps:11.4f}{sq:10.2f}"
                     f"{0.0:9.4f}{0.0:9.4f}{0.0:9.4f}{0.0:6.2f}{int(npixt[k]):8d}")
        w = f32(int(npix[k]))
        av_pr = f32(av_pr + w * pr)
        av_fsc = f32(av_fsc + w * fi)
        av_pfsc = f32(av_pfsc + w * pf)
        av_ssnr = f32(av_ssnr + w * f32(sq * sq))
        av_pssnr = f32(av_pssnr + w * psn)
        isumv += int(npix[k])
    den = f32(isumv) if isumv else f32(np.nan)
    lines.append("C  Average             "
                 f"{av_pr / den:7.2f}{av_fsc / den:7.3f}{av_pfsc / den:10.3f}"
                 f"{math.sqrt(float(av_pssnr / den)):11.4f}{math.sqrt(float(av_ssnr / den)):10.2f}"
                 f"{0.0:9.4f}{0.0:9.4f}{0.0:9.4f}{0.0:6.2f}")
    lines.append("C")
    lines.append("C Date and time      " + time.strftime("%d-%m-%Y,  %H:%M")
                 + "    Frealign V" + (version or "frealign_amd").ljust(15)[:15])
    return lines
