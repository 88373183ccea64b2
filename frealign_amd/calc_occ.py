"""CALC_OCC (calc_occ.f): occupancies of a multi-class run from its classes'
parameter files, `python -m frealign_amd.calc_occ < answers`, with the
reference program's dialogue on standard input: the number of files N, the
OCC change multiplier F, the N input parameter files, then the N output
files.

Each class's file is read as calc_occ.f:93-156 reads it: lines starting with
'C' are kept in order (those before the last record are the header, the rest
the footer), every other line is a record in the v9 fixed format
FORMAT(I7,3F8.2,2F10.2,I8,I6,2F9.1,2F8.2,I10,F11.4,2F8.2) (calc_occ.f:145),
or in FrealignX's, with a PSHFT column after ANGAST (calc_occ.f:133), once a
record longer than 140 characters is seen.  A(I) is the class's mean OCC.
Each particle's OCC and SIGMA are updated by fr_calc_occ
(calc_occ.f:179-217), and each output file is the header, the records in the
input's format with the new OCC and SIGMA, and the footer, in which the
Part_SSNR table's column becomes the A-weighted root mean square of the
classes' columns (calc_occ.f:275-302).  Output files are byte-identical to
the reference program's.
"""
from __future__ import annotations

import math
import sys

import numpy as np

from . import _native as N
from . import fileio

F32 = np.float32

# calc_occ.f:145 (v9) and :133 (FrealignX): (width, decimals or -1 for I)
_V9 = [(7, -1), (8, 2), (8, 2), (8, 2), (10, 2), (10, 2), (8, -1), (6, -1), (9, 1), (9, 1),
       (8, 2), (8, 2), (10, -1), (11, 4), (8, 2), (8, 2)]
_FX = _V9[:11] + [(8, 2)] + _V9[11:]
_OLD = ("Phase res. / B factor", "Phase residual target", "Phase residual threshold")


class CalcOccError(ValueError):
    pass


def _field(text: str, d: int):
    """Fortran Iw / Fw.d input editing of one field: blanks ignored, a blank
    field is 0, an F field with no decimal point or exponent carries d
    implied decimals."""
    t = text.replace(" ", "")
    if not t:
        return 0 if d < 0 else F32(0.0)
    try:
        if d < 0:
            return int(t)
        if "." in t or any(c in t for c in "eEdD"):
            return F32(float(t.replace("d", "e").replace("D", "e")))
        return F32(int(t) / 10.0 ** d)
    except ValueError as e:
        raise CalcOccError(f"bad field {text!r}") from e


def _fixed(line: str, spec):
    line = line.ljust(200)
    out, pos = [], 0
    for w, d in spec:
        out.append(_field(line[pos:pos + w], d))
        pos += w
    return out


def _fmt_f(v, w: int, d: int) -> str:
    """Fortran Fw.d output editing (NaN and infinities right-justified,
    asterisks for a value too wide)."""
    v = float(v)
    if math.isnan(v):
        s = "NaN".rjust(w)
    elif math.isinf(v):
        s = ("Inf" if v > 0 else "-Inf").rjust(w)
    else:
        s = f"{v:{w}.{d}f}"
    return "*" * w if len(s) > w else s


class _ClassFile:
    def __init__(self):
        self.clines = []   # every 'C' line, in order (LINE(1..) of calc_occ.f)
        self.nh = 0        # C lines before the last record
        self.recs = []
        self.a = F32(0.0)

    @property
    def footer(self):
        return self.clines[self.nh:]


def _read(path: str, state: dict) -> _ClassFile:
    cf = _ClassFile()
    with open(path) as f:
        for raw in f:
            ln = raw.rstrip("\n")[:200]
            if ln[:1] == "C":
                if ("PRES" in ln and "DPRES" in ln) or any(x in ln for x in _OLD):
                    raise CalcOccError("Old par file. Convert to new format.")
                cf.clines.append(ln)
                continue
            n = len(ln.rstrip(" "))
            if n < 120:
                raise CalcOccError("Old par file. Convert to new format.")
            if n > 140 and not state["frealignx"]:
                print("  FrealignX parameter file\n")
                state["frealignx"] = True
            cf.nh = len(cf.clines)
            v = _fixed(ln, _FX if state["frealignx"] else _V9)
            if not state["frealignx"]:
                v = v[:11] + [F32(0.0)] + v[11:]   # PSHFT = 0
            if v[0] <= 0:
                print("  ***WARNING*** blank line in parameter file")
                continue
            cf.a = F32(cf.a + v[12])
            cf.recs.append(v)
    cf.a = F32(cf.a / F32(len(cf.recs))) if cf.recs else F32(np.nan)
    return cf


def _record(v, occ, sig, frealignx: bool) -> str:
    line = fileio.par_line(fileio.ParRecord(
        v[0], v[1], v[2], v[3], v[4], v[5], float(v[6]), v[7], v[8], v[9], v[10], float(occ),
        float(v[13]), float(sig), v[15], v[16]))
    if frealignx:
        line = line[:91] + _fmt_f(v[11], 8, 2) + line[91:]
    return line


class CalcOcc:
    """The program: read the N class files, update the occupancies, write
    the N output files (calc_occ.f:36-307)."""

    def __init__(self, f: float):
        self.f = F32(f)
        self.state = {"frealignx": False}
        self.files = []

    def read(self, paths):
        for i, p in enumerate(paths):
            cf = _read(p, self.state)
            self.files.append(cf)
            print(f" No. of header lines read = {cf.nh:10d}\n"
                  f" No. of footer lines read = {len(cf.footer):10d}\n"
                  f" No. of particles read    = {len(cf.recs):10d}\n"
                  f" Average occupancy in %   = {_fmt_f(cf.a, 10, 4)}\n")
        first = self.files[0]
        for i, cf in enumerate(self.files[1:], start=2):
            if len(cf.recs) != len(first.recs):
                raise CalcOccError("No. of particles varies")
        for k in range(len(first.recs)):
            for i, cf in enumerate(self.files[1:], start=2):
                if cf.recs[k][0] != first.recs[k][0]:
                    raise CalcOccError(f"Particle sequences differ (files 1 and {i}, "
                                       f"particle {k + 1})")
                if cf.recs[k][7] != first.recs[k][7]:
                    raise CalcOccError(f"Film no. differs (files 1 and {i}, particle {k + 1})")

    def update(self):
        n, np_ = len(self.files), len(self.files[0].recs)
        self.avg = np.array([cf.a for cf in self.files], np.float32)
        self.occ = np.array([[r[12] for r in cf.recs] for cf in self.files],
                            np.float32).reshape(n, np_)
        algp = np.array([[r[13] for r in cf.recs] for cf in self.files],
                        np.float32).reshape(n, np_)
        self.sig = np.array([[r[14] for r in cf.recs] for cf in self.files],
                            np.float32).reshape(n, np_)
        N.check(N.lib().fr_calc_occ(n, np_, float(self.f), self.avg.ctypes.data,
                                    self.occ.ctypes.data, algp.ctypes.data,
                                    self.sig.ctypes.data), "fr_calc_occ")

    def write(self, i: int, path: str):
        cf = self.files[i]
        fx = self.state["frealignx"]
        with open(path, "w") as fo:
            for ln in cf.clines[:cf.nh]:
                fo.write(ln.rstrip(" ") + "\n")
            for k, v in enumerate(cf.recs):
                fo.write(_record(v, self.occ[i, k], self.sig[i, k], fx) + "\n")
            for k in range(cf.nh, len(cf.clines)):
                ln = cf.clines[k]
                if not self.state.get("lpssnr"):
                    fo.write(ln.rstrip(" ") + "\n")
                    if "Part_SSNR" in ln:
                        self.state["lpssnr"] = True
                    continue
                if "Average" in ln:
                    fo.write(ln.rstrip(" ") + "\n")
                    self.state["lpssnr"] = False
                    continue
                # FORMAT(A47,F11.4,A51) in every class's file, line k
                sump, suma = F32(0.0), F32(0.0)
                for ii, other in enumerate(self.files):
                    o = other.clines[k] if k < len(other.clines) else ""
                    ps = _field(o.ljust(200)[47:58], 4)
                    sump = F32(sump + F32(F32(ps * ps) * self.avg[ii]))
                    suma = F32(suma + self.avg[ii])
                sump = F32(sump / suma)
                own = ln.ljust(200)
                fo.write(own[:47] + _fmt_f(np.sqrt(sump), 11, 4) + own[58:109] + "\n")


def _next_nonblank(stream):
    for line in stream:
        if line.strip():
            return line
    raise CalcOccError("end of input")


def _list_value(line: str, ty):
    tok = line.replace(",", " ").split()[0]
    return ty(float(tok)) if ty is int and "." in tok else ty(tok)


def main(stream=None) -> int:
    stream = stream if stream is not None else sys.stdin
    print("\n  CALC_OCC (frealign_amd)\n")
    try:
        print(" No. of input files?")
        n = _list_value(_next_nonblank(stream), int)
        print(f"{n:6d}\n")
        if n == 0:
            print("\n  N = 0: Nothing to do. Exiting...")
            return 0
        print(" OCC change multiplier?")
        f = _list_value(_next_nonblank(stream), float)
        print(f"{_fmt_f(F32(f), 6, 2)}\n")
        prog = CalcOcc(f)
        paths = []
        for i in range(n):
            paths.append(stream.readline().rstrip("\n")[:200].rstrip(" "))
            print("*************************************************\n\n"
                  f" Input file name for file {i + 1:6d}\n{paths[-1]}")
        prog.read(paths)
        prog.update()
        print("*************************************************\n"
              "*************************************************\n")
        for i in range(n):
            out = stream.readline().rstrip("\n")[:200].rstrip(" ")
            print(f" Output file name for file {i + 1:6d}\n{out}")
            prog.write(i, out)
    except (CalcOccError, OSError, ValueError) as e:
        print(f" ****ERROR**** {e}")
        return 1
    print("\n  Normal termination of calc_occ")
    return 0


if __name__ == "__main__":
    sys.exit(main())
