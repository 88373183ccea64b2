"""The reference's file formats on the host (fr_io.cpp through fileio.py) and
the card deck (cards.py), without a GPU.

MRC (iof.f): a written file reads back bit for bit with the header SET_MHDR
and ICLOSE write (dimensions, CELLA = PSIZE * N, mode 2, the statistics from
FP64 sums); modes 0 (unsigned bytes) and 1 (INTEGER*2), byte-swapped files
and an extended header are read as IREAD does.  Par records (lmain.f:922-923)
are compared with a flang-compiled Fortran program's F/I editing of the same
values, byte for byte, and read back by both sides; the card deck of a
mult_refine_n.com heredoc parses into the values card1.f ... card7.f read.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from frealign_amd import cards, fileio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _hdr(nx, ny, nz, mode, psize, stamp=b"\x44\x41", nsymbt=0, swap=False):
    h = np.zeros(256, np.int32)
    hf = h.view(np.float32)
    h[0:3] = (nx, ny, nz)
    h[3] = mode
    h[7:10] = (nx, ny, nz)
    hf[10:13] = (psize * nx, psize * ny, psize * nz)
    hf[13:16] = 90.0
    h[16:19] = (1, 2, 3)
    h[23] = nsymbt
    b = bytearray(h.byteswap().tobytes() if swap else h.tobytes())
    b[208:212] = b"MAP "
    b[212:214] = stamp
    return bytes(b)


def test_mrc_roundtrip_and_header(tmp_path):
    r = np.random.default_rng(3)
    a = r.standard_normal((3, 16, 24)).astype(np.float32) * 5 + 1
    p = str(tmp_path / "x.mrc")
    fileio.mrc_write(p, a, 1.25, fileio.title("D7 "))
    info = fileio.mrc_info(p)
    assert (info.nx, info.ny, info.nz, info.mode, info.nsymbt) == (24, 16, 3, 2, 0)
    assert abs(info.psize - 1.25) < 1e-7
    assert info.dmin == a.min() and info.dmax == a.max()
    mean = a.astype(np.float64).sum() / a.size
    rms = np.sqrt((a * a).astype(np.float64).sum() / a.size - mean * mean)
    assert abs(info.dmean - mean) < 1e-6 and abs(info.rms - rms) < 1e-5
    back = fileio.mrc_read(p)
    assert np.array_equal(back, a)
    assert np.array_equal(fileio.mrc_read(p, 1, 2), a[1:3])
    raw = open(p, "rb").read(1024)
    assert raw[208:212] == b"MAP " and raw[212:214] == b"\x44\x41"
    assert np.frombuffer(raw[220:224], np.int32)[0] == 1 and raw[224:232] == b"Frealign"
    assert os.path.getsize(p) == 1024 + a.nbytes


def test_mrc_modes_swap_and_extended_header(tmp_path):
    r = np.random.default_rng(4)
    u8 = r.integers(0, 256, (2, 4, 6), dtype=np.uint8)
    p0 = str(tmp_path / "m0.mrc")
    open(p0, "wb").write(_hdr(6, 4, 2, 0, 2.0) + u8.tobytes())
    assert np.array_equal(fileio.mrc_read(p0), u8.astype(np.float32))
    i16 = r.integers(-30000, 30000, (2, 4, 6), dtype=np.int16)
    p1 = str(tmp_path / "m1.mrc")
    open(p1, "wb").write(_hdr(6, 4, 2, 1, 2.0, nsymbt=80) + bytes(80) + i16.tobytes())
    assert np.array_equal(fileio.mrc_read(p1), i16.astype(np.float32))
    assert fileio.mrc_info(p1).nsymbt == 80
    f32 = r.standard_normal((2, 4, 6)).astype(np.float32)
    p2 = str(tmp_path / "be.mrc")
    open(p2, "wb").write(_hdr(6, 4, 2, 2, 2.0, stamp=b"\x11\x11", swap=True)
                         + f32.byteswap().tobytes())
    assert np.array_equal(fileio.mrc_read(p2), f32)
    assert abs(fileio.mrc_info(p2).psize - 2.0) < 1e-7
    with pytest.raises(Exception):
        fileio.mrc_read(p2, 1, 5)


def _records(n, seed):
    r = np.random.default_rng(seed)
    v = np.zeros((n, 16))
    v[:, 0] = r.integers(1, 999999, n)
    v[:, 1:4] = r.uniform(-10, 360, (n, 3))
    v[:, 4:6] = r.uniform(-300, 300, (n, 2))
    v[:, 6] = r.uniform(1000, 99999, n)
    v[:, 7] = r.integers(0, 9999, n)
    v[:, 8:10] = r.uniform(5000, 40000, (n, 2))
    v[:, 10] = r.uniform(-90, 90, n)
    v[:, 11] = r.uniform(0, 100, n)
    v[:, 12] = r.uniform(-1e6, 1e6, n)
    v[:, 13] = r.uniform(0, 20, n)
    v[:, 14:16] = r.uniform(-50, 100, (n, 2))
    v[:4, 1] = [-0.001, 1234567.0, 99999.994, -9999.996]  # -0.00, overflow, rounding at the edge
    v[4, 12] = 3.0e9                                      # I10 overflow
    return v.astype(np.float32)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_par_records_match_fortran_editing(tmp_path):
    exe = str(tmp_path / "fr_par_check")
    rr = subprocess.run([FLANG, "-O1", os.path.join(ROOT, "tests", "fortran", "fr_par_check.f90"),
                         "-o", exe], capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    v = _records(300, 7)
    fin = tmp_path / "vals.txt"
    with open(fin, "w") as f:
        for row in v:
            f.write(" ".join(repr(float(x)) for x in row) + "\n")
    rr = subprocess.run([exe, "write", str(fin), str(tmp_path / "f.par")], capture_output=True,
                        text=True)
    assert rr.returncode == 0, rr.stderr
    want = open(tmp_path / "f.par").read().splitlines()
    got = [fileio.par_line(fileio.ParRecord(int(round(float(x[0]))), *[float(y) for y in x[1:7]],
                                            int(round(float(x[7]))), *[float(y) for y in x[8:16]]))
           for x in v]
    bad = [(k, w, g) for k, (w, g) in enumerate(zip(want, got)) if w != g]
    assert not bad, bad[:3]
    assert all(len(g) == 136 for g in got)
    # both sides read the records back the same way
    with open(tmp_path / "g.par", "w") as f:
        f.write("\n".join(got[5:]) + "\n")
    rr = subprocess.run([exe, "read", str(tmp_path / "g.par"), str(tmp_path / "r.txt")],
                        capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    fort = np.loadtxt(tmp_path / "r.txt")
    for k, line in enumerate(got[5:]):
        rec, fmt = fileio.par_parse(line)
        assert fmt == 2
        assert np.allclose(rec, fort[k, :15], rtol=0, atol=1e-4 * np.maximum(1, np.abs(rec)))


def test_par_parse_formats_and_file(tmp_path):
    v9 = fileio.par_line(fileio.ParRecord(12, 10.5, 80.25, -3.0, 1.5, -2.25, 50000, 3, 15000.0,
                                          14500.0, 45.0, 100.0, -1234, 0.75, 55.5, 1.25))
    rec, fmt = fileio.par_parse(v9)
    assert fmt == 2 and rec[0] == 12 and rec[12] == -1234 and abs(rec[14] - 55.5) < 1e-9
    # the fixed v9 columns without SIGMA (card10.f:151-156): I13 LogP, F8.2 SCORE
    fixed = v9[:99] + f"{-1234:13d}{55.5:8.2f}"
    rec, fmt = fileio.par_parse(fixed)
    assert fmt == 1 and rec[12] == -1234 and rec[13] == 0.0 and abs(rec[14] - 55.5) < 1e-9
    # pre-v9 (card10.f:140-142): shifts in pixels, the phase residual last
    old = f"{7:7d}{10.0:8.2f}{20.0:8.2f}{30.0:8.2f}{1.5:8.2f}{-0.5:8.2f}{50000.0:8.0f}{2:6d}" \
          f"{15000.0:9.1f}{14000.0:9.1f}{12.0:8.2f}{35.0:7.2f}"
    rec, fmt = fileio.par_parse(old)
    assert fmt == 0 and rec[6] == 50000.0 and rec[11] == 100.0 and abs(rec[14] - 35.0) < 1e-9
    # Fw.d input without a decimal point carries d implied decimals
    rec, _ = fileio.par_parse(old.replace("   10.00", "    1000", 1))
    assert abs(rec[1] - 10.0) < 1e-12
    assert fileio.par_parse("C comment")[0] is None
    recs = [fileio.ParRecord(k + 1, 1.0 * k, 2.0, 3.0, 0.5, -0.5, 50000, 1, 15000.0, 15000.0,
                             0.0, 100.0, 0, 1.0, 50.0) for k in range(3)]
    p = str(tmp_path / "a.par")
    table = fileio.fsc_table(64, 2.0, np.linspace(1, 0, 33), np.arange(33) * 10,
                             np.linspace(4, 0, 33))
    fileio.par_write(p, recs, table)
    pf = fileio.par_read(p)
    assert [r.c for r in pf.records] == [1, 2, 3] and pf.formats == {2}
    assert pf.part_ssnr is not None and len(pf.part_ssnr) >= 30
    assert abs(pf.part_ssnr[0] - np.sqrt(np.linspace(4, 0, 33)[1])) < 1e-3  # shell 2


HEREDOC = """M,1,F,T,F,F,0,F,F,F,F,0,F,1,1	!CFORM,IFLAG,FMAG,FDEF,FASTIG,FPART,IEWALD,FBEAUT,FFILT,FBFACT,FMATCH,IFSC,FDUMP,IMEM,INTERP
108.0,0.0,1.5,500.0,0.07,0.0,5.0,0.0,5.0,10,20	!RO,RI,PSIZE,WGH,XSTD,PBC,BOFF,DANG,ITMAX,IPMAX
1,1,1,1,1				!MASK
1,2000								!IFIRST,ILAST
D7									!ASYM symmetry card (I=icosahedral)
1.0,7.5,90.0,0.0,2.7,300.0,0.0,0.0			!RELMAG,DSTEP,TARGET,THRESH,CS,AKV,TX,TY
6.0,   100.0,   6.0, 8.0, 100.0, 0.0		!RREC,RMIN,RMAX,RCLAS,DFSTD,RBFACT
stack.mrc
stack_reproject.mrc
in.par
out.par
out.shft
-100., 0., 0., 0., 0., 0., 0., 0.					! terminator with RELMAG=0.0
ref.mrc
weights
map1.mrc
map2.mrc
phasediffs
pointspread
"""


def test_card_deck_of_a_refinement_heredoc():
    d = cards.parse_deck(HEREDOC)
    assert (d.cform, d.iflag, d.fmag, d.fdef, d.ifsc, d.imem, d.interp) == ("M", 1, False, True,
                                                                            0, 1, 1)
    assert (d.ro, d.psize, d.mw, d.wgh, d.pbc, d.dang, d.itmax, d.ipmax) == (108.0, 1.5, 500.0,
                                                                            0.07, 5.0, 5.0, 10,
                                                                            20)
    assert d.pmask == [1, 1, 1, 1, 1] and d.dmask[0] == -1.0
    assert (d.ifirst, d.ilast, d.asym) == (1, 2000, "D7") and d.point_group() == ("D", 7)
    assert len(d.datasets) == 1 and not d.reconstruct  # RELMAG < 0: refinement only
    ds = d.datasets[0]
    assert (ds.dstep, ds.cs, ds.akv, ds.rrec, ds.rmax1, ds.rmax2, ds.rclas, ds.dfstd) == \
        (7.5, 2.7, 300.0, 6.0, 100.0, 6.0, 8.0, 100.0)
    assert (ds.finpat1, ds.finpar, ds.foutpar, ds.foutsh) == ("stack.mrc", "in.par", "out.par",
                                                             "out.shft")
    assert (d.f3d, d.map1, d.map2) == ("ref.mrc", "map1.mrc", "map2.mrc")
    assert abs(ds.wl - 0.019687) < 1e-5  # 300 kV (card6.f:51-52)
    with pytest.raises(cards.CardError):
        cards.parse_deck(HEREDOC.replace("1,1,1,1,1", "1,1,2,1,1"))
