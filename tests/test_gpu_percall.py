"""A Fortran host calling CTFAPPLY, PEXTRACT, CC3M, PRESB, SIGMA and
PINSERT_TRILIN with the reference's own symbols and argument lists, linked
against libfrealign_hip.so in place of the reference routines, against the
FP32 oracle on the same inputs.

Bars as the batched parity tests: CTFAPPLY 1e-5 relative L2 per output,
PEXTRACT 1e-5 relative L2, CC3M 1e-4 relative, PRESB 2e-5 absolute per shell
(unwritten shells untouched), SIGMA 1e-3 relative (SIGMA2 carries a
cancellation) and LOGPC 1e-4, PINSERT_TRILIN (one stripe, accumulated onto
non-zero volumes) NKSUM exact and the added volumes 1e-5 relative L2."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
LIB = os.path.join(ROOT, "frealign_amd", "libfrealign_hip.so")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_percall_vs_oracle(tmp_path):
    L = O.oracle()
    nsam, maxr1, maxr2 = 64, 2, 20
    jc = nsam // 2 + 1
    ic = nsam * jc
    r = np.random.default_rng(11)
    vol = np.ascontiguousarray(np.fft.rfftn(r.standard_normal((nsam,) * 3)).astype(np.complex64))
    spec = (r.standard_normal((nsam, jc)) + 1j * r.standard_normal((nsam, jc))).astype(np.complex64)
    datc = (r.standard_normal((nsam, jc)) + 1j * r.standard_normal((nsam, jc))).astype(np.complex64)
    ctff = (0.5 + 0.2 * r.standard_normal((2, nsam, jc))).astype(np.complex64)
    fsct = np.zeros(jc, np.float32)
    fsct[: jc // 2] = 1.0
    rbin0 = np.full(nsam, 7.0, np.float32)
    phi, theta, psi, shx, shy, amag = 0.3, 1.1, -0.7, 0.05, -0.08, 1.01
    ri = 0.4 * nsam
    ri2, ri3, rih, halfw = (ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0
    rbfact, rclas, sigp, signset = 0.001, 0.3, 0.01, 0.0
    dmask = np.array([nsam / 2 + 2, nsam / 2 - 1, nsam / 2 + 1, 0.35 * nsam], np.float32)
    sm = O.IDENTITY.astype(np.float32)
    # CTFAPPLY: 300 kV, CS 2.7 mm, 7% amplitude contrast, 1.5 A pixels
    wl = 12.26 / np.sqrt(300000.0 + 0.9785 * 300000.0 ** 2 / 1e6)
    cpar = np.array([2.7e7, wl, np.sqrt(1 - 0.07 ** 2), 0.07, 15000.0, 15600.0, 0.4,
                     wl / (1.5 * nsam), 1.0, rih, halfw, rih * rih, (ri + 3) ** 2], np.float32)
    pssnr = np.exp(-((np.arange(jc + 1) / nsam) * 4.0) ** 2).astype(np.float32)
    # PINSERT_TRILIN: stripe 2 of 3 (ST = (2JC-1)/3 as pinsert.f:95), onto
    # non-zero volumes and counters.  The starting volumes are of the size of
    # one call's increments: the library adds a voxel's terms first and the
    # sum once, the reference each term in turn, so on a much larger base the
    # two differ by the base's rounding, not by the insertion.
    irrec2, ip = int(0.4 * nsam) ** 2, 2
    occ, tmp, dx, dy = 0.8, -0.004, 0.02, 0.05
    stw = np.float32(2 * jc - 1) / np.float32(3)
    dmi = np.zeros(9, np.float32)
    L.or_dm(0.7, 1.9, -2.3, 1.0, None, dmi)
    b3 = (0.002 * vol).astype(np.complex64).reshape(-1)
    t3 = r.uniform(0, 1, vol.size).astype(np.float32)
    nk0, ttd0 = 5, 0.25
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([nsam, maxr1, maxr2], np.int32).tofile(f)
        np.array([phi, theta, psi, shx, shy, amag, ri2, ri3, rih, halfw, rbfact, rclas, sigp,
                  signset], np.float32).tofile(f)
        dmask.tofile(f)
        sm.tofile(f)
        for a in (vol, spec, datc, ctff, fsct, rbin0):
            a.tofile(f)
        cpar.tofile(f)
        pssnr.tofile(f)
        np.array([irrec2, ip], np.int32).tofile(f)
        np.array([occ, tmp, dx, dy, stw], np.float32).tofile(f)
        dmi.tofile(f)
        b3.tofile(f)
        t3.tofile(f)
        np.array([nk0], np.int64).tofile(f)
        np.array([ttd0], np.float64).tofile(f)
    exe = tmp_path / "fr_percall_check"
    rr = subprocess.run([FLANG, "-O1", os.path.join(ROOT, "tests", "fortran",
                                                   "fr_percall_check.f90"), LIB,
                         f"-Wl,-rpath,{os.path.dirname(LIB)}", "-o", str(exe)],
                        capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    print("running the Fortran host", flush=True)
    rr = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True,
                        timeout=120)
    assert rr.returncode == 0, rr.stdout + rr.stderr
    assert "FR_PERCALL_CHECK done" in rr.stdout
    raw = np.fromfile(fout, dtype=np.uint8)
    off = 0
    outc = raw[off:off + 8 * ic].view(np.complex64)
    off += 8 * ic
    ccv = raw[off:off + 4].view(np.float32)[0]
    off += 4
    rbin = raw[off:off + 4 * nsam].view(np.float32)
    off += 4 * nsam
    s2n, s2, lpc = raw[off:off + 12].view(np.float32)
    off += 12
    nv = vol.size
    views = {}
    for name, nbytes, dt in (("ctff", 16 * ic, np.complex64), ("outd", 8 * ic, np.complex64),
                             ("datc", 8 * ic, np.complex64), ("b3df", 8 * nv, np.complex64),
                             ("t3df", 4 * nv, np.float32), ("nksum", 8, np.int64),
                             ("ttd", 8, np.float64)):
        views[name] = raw[off:off + nbytes].view(dt)
        off += nbytes
    assert off == raw.size
    # CTFAPPLY vs the oracle
    cf_ = np.zeros(2 * ic, np.complex64)
    oc = np.zeros(ic, np.complex64)
    dc = np.zeros(ic, np.complex64)
    L.or32_ctfapply(nsam, spec.reshape(-1), shx, shy, *[float(x) for x in cpar], pssnr, 1, cf_,
                    oc, dc)
    for name, want in (("ctff", cf_), ("outd", oc), ("datc", dc)):
        e = np.linalg.norm(views[name] - want) / np.linalg.norm(want)
        print(f"CTFAPPLY {name} rel L2 {e:.1e}")
        assert e < 1e-5, name
    # PINSERT_TRILIN vs the stripe oracle, on the same starting volumes
    bo, to = b3.copy(), t3.copy()
    ttd = C.c_double(ttd0)
    nk = C.c_int64(nk0)
    L.or32_pinsert_trilin(nsam, irrec2, spec.reshape(-1), ctff.reshape(-1), occ, bo, to,
                          C.byref(ttd), shx, shy, C.byref(nk), dmi, tmp, dx, dy, float(stw), ip)
    assert nk.value > nk0
    assert int(views["nksum"][0]) == nk.value
    db_gpu, db_or = views["b3df"] - b3, bo - b3
    dt_gpu, dt_or = views["t3df"] - t3, to - t3
    eb = np.linalg.norm(db_gpu - db_or) / np.linalg.norm(db_or)
    et = np.linalg.norm(dt_gpu - dt_or) / np.linalg.norm(dt_or)
    ett = abs(views["ttd"][0] - ttd.value) / abs(ttd.value - ttd0)
    print(f"PINSERT_TRILIN NKSUM {nk.value - nk0} B3DF {eb:.1e} T3DF {et:.1e} TTD {ett:.1e}")
    assert eb < 1e-5 and et < 1e-5 and ett < 1e-5
    # oracle
    w = np.zeros(ic, np.complex64)
    L.or32_pextract(nsam, w, vol.reshape(-1), phi, theta, psi, ctff.reshape(-1), sm)
    e = np.linalg.norm(outc - w) / np.linalg.norm(w)
    print(f"PEXTRACT rel L2 {e:.1e}")
    assert e < 1e-5
    c32 = L.or32_cc3m(nsam, amag, spec.reshape(-1), vol.reshape(-1), maxr1, maxr2, phi, theta,
                      psi, shx, shy, rbfact, fsct, ctff.reshape(-1), ri2, ri3, rih, halfw, sm)
    print(f"CC3M {ccv:.7f} oracle {c32:.7f}")
    assert abs(ccv - c32) <= 1e-4 * abs(c32)
    rb = rbin0.copy()
    L.or32_presb(nsam, amag, spec.reshape(-1), vol.reshape(-1), phi, theta, psi, shx, shy, rb,
                 ctff.reshape(-1), ri2, ri3, rih, halfw, sm)
    assert np.array_equal(rbin[1::2], rb[1::2])
    assert np.max(np.abs(rbin[0::2] - rb[0::2])) < 2e-5
    out = np.zeros(3, np.float32)
    L.or32_sigma(nsam, amag, vol.reshape(-1), datc.reshape(-1), phi, theta, psi, shx, shy,
                 ctff.reshape(-1), maxr1, rclas, signset, dmask, sigp, sm, out)
    got = np.array([s2n, s2, lpc])
    rel = np.abs(got - out) / np.abs(out)
    print(f"SIGMA {got} oracle {out} rel {rel}")
    assert rel[0] < 1e-3 and rel[1] < 1e-3 and rel[2] < 1e-4
    assert math.isfinite(ccv)


def test_a3d3_percall_vs_oracle():
    """a3d3_ (the reference's A3D3 symbol and argument list, NVOL = 1) called
    for four particles in turn, IFSC = 0 and ISM = 1 as the reference's FSC
    run: the half-maps alternate with ISWITCH, a particle with OCC = 0 is not
    recorded but still switches, and every call flushes.  Against the FP32
    oracle's PINSERT into the same halves: NKSUM1 exact, volumes 1e-5 relative
    L2, TTD1/TTD2 1e-6 relative; the bookkeeping (RECB, ISW, IBUF, ISWITCH)
    as a3d3.f:39-85,145 leave it."""
    from frealign_amd import api

    L = O.oracle()
    lib = C.CDLL(LIB)
    nsam, nn1 = 32, 40
    jc = nsam // 2 + 1
    ic = nsam * jc
    nv = nsam * nsam * jc
    nn3 = nn1 * nn1 * (nn1 // 2 + 1)
    r = np.random.default_rng(23)
    npart = 4
    datc = (r.standard_normal((npart, ic)) + 1j * r.standard_normal((npart, ic))).astype(np.complex64)
    ctff = (0.6 + 0.2 * (r.standard_normal((npart, 2 * ic))
                         + 1j * r.standard_normal((npart, 2 * ic)))).astype(np.complex64)
    eul = r.uniform(-np.pi, np.pi, (npart, 3)).astype(np.float32)
    shf = r.uniform(-0.3, 0.3, (npart, 2)).astype(np.float32)
    occ = np.array([1.0, 0.7, 0.0, 0.9], np.float32)
    sig2n = r.uniform(0.5, 2.0, npart).astype(np.float32)
    presa = r.uniform(0.2, 0.8, npart).astype(np.float32)
    rrec = np.array([0.0, 0.45], np.float32)  # RREC(NSET), NSET = 2
    nsym, symop, jsym, _ = api.point_group("D", 7)
    pbc, boff, psize = 5.0, 40.0, 1.5
    a3df, d3df = np.zeros(nn3, np.complex64), np.zeros(nn3, np.complex64)
    s3df, v3df = np.zeros(nn3, np.float32), np.zeros(nn3, np.float32)
    recbf = np.zeros(nn1 // 2 * nn1 + ic, np.complex64)  # the copy runs past NN1/2*NN1
    ctfbf = np.zeros(nn1 * nn1 + 2 * ic, np.complex64)
    recb = np.zeros(10, np.float32)
    isw = np.full(2, -1, np.int32)
    nksum1 = np.zeros(2, np.int64)
    ttd1, ttd2 = np.zeros(2, np.float64), np.zeros(2, np.float64)
    iswitch, ibuf = C.c_int(0), C.c_int(0)
    dummy = np.zeros(16, np.float32)
    idummy = np.zeros(64, np.int32)
    f, i = O.byref_f, O.byref_i
    expect_half = []
    for k in range(npart):
        before = iswitch.value
        args = [i(2), C.byref(iswitch), i(nsam), i(1), f(0.4 * nsam), rrec.ctypes, i(k + 1),
                i(npart), f(1.0), np.ascontiguousarray(datc[k]).ctypes,
                np.ascontiguousarray(ctff[k]).ctypes, dummy.ctypes, dummy.ctypes, dummy.ctypes,
                dummy.ctypes, idummy.ctypes, f(float(eul[k, 0])), f(float(eul[k, 1])),
                f(float(eul[k, 2])), f(float(shf[k, 0])), f(float(shf[k, 1])),
                f(float(presa[k])), f(pbc), f(boff), a3df.ctypes, s3df.ctypes,
                C.byref(C.c_double(0.0)), d3df.ctypes, v3df.ctypes, C.byref(C.c_double(0.0)),
                i(nn1), i(1), dummy.ctypes, i(1), i(1), f(0.0), i(0), i(nsym), i(8),
                jsym.ctypes, symop.ctypes, i(0), C.byref(C.c_int64(0)), i(0), i(1), i(1), i(1),
                i(1), f(float(occ[k])), f(float(sig2n[k])), f(psize), C.c_char_p(b"D7 "),
                f(0.0), f(0.0), i(0), i(0), nksum1.ctypes, ttd1.ctypes, ttd2.ctypes,
                idummy.ctypes, idummy.ctypes, i(1), recbf.ctypes, recb.ctypes, isw.ctypes,
                C.byref(ibuf), ctfbf.ctypes, i(nn1), i(nn3), O.IDENTITY.ctypes, i(0), i(1),
                C.c_size_t(3)]
        assert len(args) == 73
        print(f"A3D3 call K={k + 1} ISWITCH={before} OCC={occ[k]}", flush=True)
        lib.a3d3_(*args)
        assert ibuf.value == 0
        assert iswitch.value == 1 - before
        expect_half.append(before if occ[k] != 0.0 else -1)
        if occ[k] != 0.0:
            assert isw[0] == before
            want = [*eul[k], *shf[k], 1.0, occ[k], sig2n[k], presa[k], rrec[1]]
            assert np.array_equal(recb, np.array(want, np.float32))
            assert np.array_equal(recbf[:ic], datc[k])
            assert np.array_equal(ctfbf[:2 * ic], ctff[k])
    assert expect_half == [0, 1, -1, 1]
    # the oracle's PINSERT of the recorded particles into their halves
    num = [np.zeros(nv, np.complex64) for _ in range(2)]
    c2 = [np.zeros(nv, np.float32) for _ in range(2)]
    ttd = [C.c_double(0.0), C.c_double(0.0)]
    nk = [C.c_int64(0), C.c_int64(0)]
    for k, h in enumerate(expect_half):
        if h < 0:
            continue
        on = np.float32(occ[k]) / np.float32(sig2n[k])
        L.or32_pinsert(nsam, float(rrec[1]), 1.0, datc[k], ctff[k], num[h], c2[h],
                       C.byref(ttd[h]), float(eul[k, 0]), float(eul[k, 1]), float(eul[k, 2]),
                       float(shf[k, 0]), float(shf[k, 1]), float(presa[k]), pbc, boff, nsym,
                       jsym, symop, C.byref(nk[h]), psize, float(on), O.IDENTITY, 0)
    assert int(nksum1[0]) == nk[0].value + nk[1].value and nksum1[1] == 0
    for h, (gn, gc, gt) in enumerate(((a3df, s3df, ttd1), (d3df, v3df, ttd2))):
        e_num = np.linalg.norm(gn[:nv] - num[h]) / np.linalg.norm(num[h])
        e_c2 = np.linalg.norm(gc[:nv] - c2[h]) / np.linalg.norm(c2[h])
        e_t = abs(gt[0] - ttd[h].value) / ttd[h].value
        print(f"A3D3 half {h}: numerator {e_num:.1e} CTF^2 {e_c2:.1e} TTD {e_t:.1e}")
        assert e_num < 1e-5 and e_c2 < 1e-5 and e_t < 1e-6
        assert not gn[nv:].any() and not gc[nv:].any() and gt[1] == 0.0


def _a3d3_args(k, npart, nsam, nn1, nn3, iswitch, ibuf, datc, ctff, eul, shf, presa, occ, sig2n,
               rrec, maps, sums, bufs, sym, nvol, ifsc, pbc=5.0, boff=40.0, psize=1.5,
               stats=None):
    """a3d3_'s 73 arguments; stats = (ASUM, VSUM, PSUM, C3DF, KSUM) for NNSTAT = 1."""
    f, i = O.byref_f, O.byref_i
    a3df, s3df, d3df, v3df = maps
    nksum1, ttd1, ttd2 = sums
    recbf, recb, isw, ctfbf = bufs
    nsym, symop, jsym = sym
    dummy = np.zeros(16, np.float32)
    idummy = np.zeros(64, np.int32)
    nv1 = nv2 = max(1, (nvol + 1) // 2) if ifsc == 0 else nvol
    sa = [dummy.ctypes] * 4 + [idummy.ctypes] if stats is None else [a.ctypes for a in stats]
    args = [i(2), C.byref(iswitch), i(nsam), i(1), f(0.4 * nsam), rrec.ctypes, i(k + 1),
            i(npart), f(1.0), np.ascontiguousarray(datc).ctypes,
            np.ascontiguousarray(ctff).ctypes] + sa + [
            f(float(eul[0])), f(float(eul[1])), f(float(eul[2])),
            f(float(shf[0])), f(float(shf[1])), f(float(presa)), f(pbc), f(boff), a3df.ctypes,
            s3df.ctypes, C.byref(C.c_double(0.0)), d3df.ctypes, v3df.ctypes,
            C.byref(C.c_double(0.0)), i(nn1), i(1), dummy.ctypes, i(1), i(1), f(0.0), i(0),
            i(nsym), i(8), jsym.ctypes, symop.ctypes, i(0 if stats is None else 1),
            C.byref(C.c_int64(0)), i(ifsc), i(1),
            i(nvol), i(nv1), i(nv2), f(float(occ)), f(float(sig2n)), f(psize),
            C.c_char_p(b"D7 "), f(0.0), f(0.0), i(0), i(0), nksum1.ctypes, ttd1.ctypes,
            ttd2.ctypes, idummy.ctypes, idummy.ctypes, i(1), recbf.ctypes, recb.ctypes,
            isw.ctypes, C.byref(ibuf), ctfbf.ctypes, i(nn1), i(nn3), O.IDENTITY.ctypes, i(0),
            i(1), C.c_size_t(3)]
    assert len(args) == 73
    return args


def test_a3d3_private_volumes_vs_oracle():
    """a3d3_ with NVOL = 4 (IFSC = 0): six particles, a flush of four slots at
    K = 4 and of two at K = ILAST = 6.  Slot I goes to private volume
    IS = (I-1)/2+1 of its half (a3d3_s.f:39-40), NKSUM1(I) and TTD1/TTD2(IS)
    per slot and volume.  The half-maps stay on the device until the last
    flush: the host arrays are unchanged after K = 4 until fr_percall_flush,
    which then shows the first flush's volumes.  Bars as the NVOL = 1 test."""
    from frealign_amd import api

    L = O.oracle()
    lib = C.CDLL(LIB)
    nsam, nn1, nvol = 32, 36, 4
    jc = nsam // 2 + 1
    ic = nsam * jc
    nv = nsam * nsam * jc
    nn3 = nn1 * nn1 * (nn1 // 2 + 1)
    r = np.random.default_rng(29)
    npart = 6
    datc = (r.standard_normal((npart, ic)) + 1j * r.standard_normal((npart, ic))).astype(np.complex64)
    ctff = (0.6 + 0.2 * (r.standard_normal((npart, 2 * ic))
                         + 1j * r.standard_normal((npart, 2 * ic)))).astype(np.complex64)
    eul = r.uniform(-np.pi, np.pi, (npart, 3)).astype(np.float32)
    shf = r.uniform(-0.3, 0.3, (npart, 2)).astype(np.float32)
    occ = r.uniform(0.5, 1.0, npart).astype(np.float32)
    sig2n = r.uniform(0.5, 2.0, npart).astype(np.float32)
    presa = r.uniform(0.2, 0.8, npart).astype(np.float32)
    rrec = np.array([0.0, 0.45], np.float32)
    nsym, symop, jsym, _ = api.point_group("D", 7)
    nvh = 2  # NVOL1 = NVOL2 = NVOL/2 private volumes per half
    maps = (np.zeros(nn3 * nvh, np.complex64), np.zeros(nn3 * nvh, np.float32),
            np.zeros(nn3 * nvh, np.complex64), np.zeros(nn3 * nvh, np.float32))
    sums = (np.zeros(nvol, np.int64), np.zeros(nvh, np.float64), np.zeros(nvh, np.float64))
    bufs = (np.zeros(nn1 // 2 * nn1 * nvol + ic, np.complex64), np.zeros(10 * nvol, np.float32),
            np.full(nvol, -1, np.int32), np.zeros(nn1 * nn1 * nvol + 2 * ic, np.complex64))
    iswitch, ibuf = C.c_int(0), C.c_int(0)
    slot_of = []  # (half, IS, slot) of each particle
    for k in range(npart):
        print(f"A3D3 NVOL={nvol} call K={k + 1}", flush=True)
        h = iswitch.value
        slot = ibuf.value
        slot_of.append((h, slot // 2, slot))
        lib.a3d3_(*_a3d3_args(k, npart, nsam, nn1, nn3, iswitch, ibuf, datc[k], ctff[k], eul[k],
                              shf[k], presa[k], occ[k], sig2n[k], rrec, maps, sums, bufs,
                              (nsym, symop, jsym), nvol, 0))
        if k == 3:
            assert ibuf.value == 0
            assert not maps[0].any() and not maps[2].any()  # still on the device
            lib.fr_percall_flush()
            snap = [m.copy() for m in maps]
    assert ibuf.value == 0
    assert [s[:2] for s in slot_of] == [(0, 0), (1, 0), (0, 1), (1, 1), (0, 0), (1, 0)]

    def oracle_into(parts):
        num = [[np.zeros(nv, np.complex64) for _ in range(nvh)] for _ in range(2)]
        c2 = [[np.zeros(nv, np.float32) for _ in range(nvh)] for _ in range(2)]
        ttd = [[C.c_double(0.0) for _ in range(nvh)] for _ in range(2)]
        nk = [C.c_int64(0) for _ in range(nvol)]
        for k in parts:
            h, is_, slot = slot_of[k]
            on = np.float32(occ[k]) / np.float32(sig2n[k])
            L.or32_pinsert(nsam, float(rrec[1]), 1.0, datc[k], ctff[k], num[h][is_], c2[h][is_],
                           C.byref(ttd[h][is_]), float(eul[k, 0]), float(eul[k, 1]),
                           float(eul[k, 2]), float(shf[k, 0]), float(shf[k, 1]), float(presa[k]),
                           5.0, 40.0, nsym, jsym, symop, C.byref(nk[slot]), 1.5, float(on),
                           O.IDENTITY, 0)
        return num, c2, ttd, nk

    def compare(got_maps, want, label):
        num, c2, _, _ = want
        for h in range(2):
            gn, gc = got_maps[2 * h], got_maps[2 * h + 1]
            for is_ in range(nvh):
                a = gn[is_ * nn3:is_ * nn3 + nv]
                b = gc[is_ * nn3:is_ * nn3 + nv]
                e_num = np.linalg.norm(a - num[h][is_]) / np.linalg.norm(num[h][is_])
                e_c2 = np.linalg.norm(b - c2[h][is_]) / np.linalg.norm(c2[h][is_])
                print(f"A3D3 {label} half {h} volume {is_ + 1}: numerator {e_num:.1e} "
                      f"CTF^2 {e_c2:.1e}")
                assert e_num < 1e-5 and e_c2 < 1e-5
                assert not gn[is_ * nn3 + nv:(is_ + 1) * nn3].any()

    compare(snap, oracle_into(range(4)), "after the first flush")
    want = oracle_into(range(npart))
    compare(maps, want, "at K = ILAST")
    num, c2, ttd, nk = want
    assert [int(v) for v in sums[0]] == [v.value for v in nk]
    for h, got in ((0, sums[1]), (1, sums[2])):
        for is_ in range(nvh):
            assert abs(got[is_] - ttd[h][is_].value) <= 1e-6 * ttd[h][is_].value


def test_a3d3_statistics_vs_oracle():
    """a3d3_ with NNSTAT = 1 (NVOL = 2, IFSC = 0, four particles): slot I's
    per-voxel KSUM/VSUM/PSUM/ASUM go to column I of the (NN3+1, NVOL) arrays
    (a3d3_s.f:70-71 -> PINSERT_TRILIN_SE) against C3DF, copied back with the
    half-maps at K = ILAST.  KSUM exact against the oracle's PINSERT_SE, the
    sums within 1e-5 relative L2, the half-maps as without statistics."""
    from frealign_amd import api

    L = O.oracle()
    lib = C.CDLL(LIB)
    nsam, nn1, nvol = 32, 36, 2   # NN1 > NSAM: the RECBF/CTFBF columns do not overlap
    jc = nsam // 2 + 1
    ic = nsam * jc
    nv = nsam * nsam * jc
    nn3 = nn1 * nn1 * (nn1 // 2 + 1)
    r = np.random.default_rng(31)
    npart = 4
    datc = (r.standard_normal((npart, ic)) + 1j * r.standard_normal((npart, ic))).astype(np.complex64)
    ctff = (0.6 + 0.2 * (r.standard_normal((npart, 2 * ic))
                         + 1j * r.standard_normal((npart, 2 * ic)))).astype(np.complex64)
    ref = (r.standard_normal(nv) + 1j * r.standard_normal(nv)).astype(np.complex64)
    eul = r.uniform(-np.pi, np.pi, (npart, 3)).astype(np.float32)
    shf = r.uniform(-0.3, 0.3, (npart, 2)).astype(np.float32)
    occ = r.uniform(0.5, 1.0, npart).astype(np.float32)
    sig2n = r.uniform(0.5, 2.0, npart).astype(np.float32)
    presa = r.uniform(0.2, 0.8, npart).astype(np.float32)
    rrec = np.array([0.0, 0.45], np.float32)
    nsym, symop, jsym, _ = api.point_group("C", 1)
    maps = (np.zeros(nn3, np.complex64), np.zeros(nn3, np.float32),
            np.zeros(nn3, np.complex64), np.zeros(nn3, np.float32))
    sums = (np.zeros(nvol, np.int64), np.zeros(1, np.float64), np.zeros(1, np.float64))
    bufs = (np.zeros(nn1 // 2 * nn1 * nvol + ic, np.complex64), np.zeros(10 * nvol, np.float32),
            np.full(nvol, -1, np.int32), np.zeros(nn1 * nn1 * nvol + 2 * ic, np.complex64))
    col = nn3 + 1
    stats = (np.zeros(col * nvol, np.float32), np.zeros(col * nvol, np.float32),
             np.zeros(col * nvol, np.float32), ref, np.zeros(col * nvol, np.int32))
    iswitch, ibuf = C.c_int(0), C.c_int(0)
    slot_of = []
    for k in range(npart):
        print(f"A3D3 NNSTAT=1 call K={k + 1}", flush=True)
        slot_of.append(ibuf.value)
        lib.a3d3_(*_a3d3_args(k, npart, nsam, nn1, nn3, iswitch, ibuf, datc[k], ctff[k], eul[k],
                              shf[k], presa[k], occ[k], sig2n[k], rrec, maps, sums, bufs,
                              (nsym, symop, jsym), nvol, 0, stats=stats))
    assert slot_of == [0, 1, 0, 1]
    ks = [np.zeros(nv, np.int32) for _ in range(nvol)]
    vs, ps, as_ = ([np.zeros(nv, np.float32) for _ in range(nvol)] for _ in range(3))
    for k in range(npart):
        sl = slot_of[k]
        on = np.float32(occ[k]) / np.float32(sig2n[k])
        L.or32_pinsert_se(nsam, float(rrec[1]), 1.0, datc[k], ctff[k], np.zeros(nv, np.complex64),
                          np.zeros(nv, np.float32), C.byref(C.c_double(0.0)), float(eul[k, 0]),
                          float(eul[k, 1]), float(eul[k, 2]), float(shf[k, 0]), float(shf[k, 1]),
                          float(presa[k]), 5.0, 40.0, nsym, jsym, symop, C.byref(C.c_int64(0)),
                          1.5, float(on), O.IDENTITY, 0, ref, ks[sl], vs[sl], ps[sl], as_[sl])
    asum, vsum, psum, _, ksum = stats
    for sl in range(nvol):
        o = col * sl
        assert np.array_equal(ksum[o:o + nv], ks[sl]) and not ksum[o + nv:o + col].any()
        assert int(ksum[o:o + nv].sum()) == int(sums[0][sl])
        e = [np.linalg.norm(g[o:o + nv] - w) / np.linalg.norm(w)
             for g, w in ((vsum, vs[sl]), (psum, ps[sl]), (asum, as_[sl]))]
        print(f"A3D3 NNSTAT slot {sl + 1}: KSUM {ks[sl].sum()} VSUM {e[0]:.1e} PSUM {e[1]:.1e} "
              f"ASUM {e[2]:.1e}")
        assert max(e) < 1e-5


def test_pinsert_trilin_se_percall_vs_oracle():
    """pinsert_trilin_se_ with NNSTAT = 1 on one stripe: B3DF/T3DF/NKSUM/TTD
    and the statistics against the oracle's PINSERT_TRILIN_SE."""
    L = O.oracle()
    lib = C.CDLL(LIB)
    f, i = O.byref_f, O.byref_i
    nsam, imp, ip = 64, 3, 2
    jc = nsam // 2 + 1
    nv = nsam * nsam * jc
    r = np.random.default_rng(37)
    spec = (r.standard_normal(nsam * jc) + 1j * r.standard_normal(nsam * jc)).astype(np.complex64)
    ctff = (0.5 + 0.2 * (r.standard_normal(2 * nsam * jc)
                         + 1j * r.standard_normal(2 * nsam * jc))).astype(np.complex64)
    ref = (r.standard_normal(nv) + 1j * r.standard_normal(nv)).astype(np.complex64)
    dm = np.zeros(9, np.float32)
    L.or_dm(0.4, 2.2, -1.3, 1.0, None, dm)
    irrec2 = (nsam * 2 // 5) ** 2
    tmp, occ, shx, shy, dx, dy = -0.002, 0.7, 0.09, -0.04, 0.0, 0.0
    st = float(np.float32(2 * jc - 1) / np.float32(imp))
    out = []
    for which in ("gpu", "oracle"):
        b3 = np.zeros(nv, np.complex64)
        t3 = np.zeros(nv, np.float32)
        ttd, nk = C.c_double(0.0), C.c_int64(0)
        ksum = np.zeros(nv, np.int32)
        vsum, psum, asum = (np.zeros(nv, np.float32) for _ in range(3))
        if which == "gpu":
            print("[pinsert_trilin_se_] NNSTAT=1", flush=True)
            ns1 = np.zeros(4 * nsam, np.int32)
            lib.pinsert_trilin_se_(i(nsam), f(1.0), i(irrec2), spec.ctypes, ctff.ctypes, f(occ),
                                   b3.ctypes, t3.ctypes, C.byref(ttd), f(shx), f(shy),
                                   asum.ctypes, vsum.ctypes, psum.ctypes, ref.ctypes, ksum.ctypes,
                                   np.zeros(4, np.float32).ctypes, i(1), i(1), f(0.0), i(0), i(1),
                                   C.byref(nk), i(jc), dm.ctypes, f(tmp), i(nsam // 2),
                                   i(nsam * jc), f(dx), f(dy), f(st), i(ip), ns1.ctypes)
        else:
            L.or32_pinsert_trilin_se(nsam, irrec2, spec, ctff, occ, b3, t3, C.byref(ttd), shx, shy,
                                     C.byref(nk), dm, tmp, dx, dy, st, ip, ref, ksum, vsum, psum,
                                     asum)
        out.append((b3, t3, ttd.value, nk.value, ksum, vsum, psum, asum))
    g, o = out
    assert g[3] == o[3] > 0 and np.array_equal(g[4], o[4])
    e = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in
         ((g[0], o[0]), (g[1], o[1]), (g[5], o[5]), (g[6], o[6]), (g[7], o[7]))]
    print(f"[pinsert_trilin_se_] NKSUM {g[3]}; rel L2 B3DF {e[0]:.1e} T3DF {e[1]:.1e} "
          f"VSUM {e[2]:.1e} PSUM {e[3]:.1e} ASUM {e[4]:.1e}")
    assert max(e) < 1e-5 and abs(g[2] - o[2]) <= 1e-6 * abs(o[2])


def test_psearch_and_resident_reference_vs_oracle():
    """psearch_ (the reference's PSEARCH symbol and argument list) for three
    particles in turn, IRAN carried from call to call, against or32_psearch:
    the best view bit for bit, TESTPAR as fr_psearch's test.  Then cc3m_ on
    the resident reference: the same value on a repeat call, the new
    reference's value after the host array is rewritten in place (its
    fingerprint changes) and after fr_percall_invalidate."""
    from frealign_amd import api, synth

    L = O.oracle()
    lib = C.CDLL(LIB)
    ctx = api.Context(0)
    nsam, psize, ri = 64, 2.0, 26.0
    optics = synth.Optics(psize=psize)
    sym = api.point_group("D", 7)
    vol = synth.blob_phantom(nsam, 20, 24.0, seed=51, sym=sym, device="cuda")
    a3df_t = synth.half_volume(vol)
    npart = 3
    parts = synth.make_particles(ctx, a3df_t, npart, seed=52, optics=optics, snr=2.0, ri=ri,
                                 shift_a=4.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to("cuda")
    ctff_t, outc_t, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    a3df = np.ascontiguousarray(a3df_t.cpu().numpy().reshape(-1))
    outc = outc_t.cpu().numpy().reshape(npart, -1)
    ctff = ctff_t.cpu().numpy().reshape(npart, -1)
    ctx.close()
    dang, maxr1, maxr2, rbfact, ipmax = 15.0, 1, 14, 0.0005, 8
    sang, iq = api.search_grid(dang, "D", 7)
    flat = np.ascontiguousarray(sang.reshape(-1))
    mask = np.ones(5, np.int32)
    f, i = O.byref_f, O.byref_i
    iran = C.c_int(-23)
    iran_o = C.c_int(-23)
    dummy = np.zeros(64, np.float32)
    slots = agree = 0
    for p in range(npart):
        print(f"[psearch_] particle {p + 1}", flush=True)
        par = [C.c_float(0.0) for _ in range(6)]
        tp = np.zeros(6 * ipmax, np.float32)
        lib.psearch_(i(nsam), i(1), f(1.0), outc[p].ctypes, a3df.ctypes, f(dang), i(maxr1),
                     i(maxr2), *[C.byref(v) for v in par], flat.ctypes, i(len(sang)),
                     dummy.ctypes, dummy.ctypes, dummy.ctypes, i(iq), mask.ctypes, f(rbfact),
                     dummy.ctypes, i(1), dummy.ctypes, i(0), f(0.0), ctff[p].ctypes, f(0.0),
                     f(0.0), f(0.0), f(0.0), tp.ctypes, i(ipmax), C.c_char_p(b"D7 "),
                     C.byref(iran), dummy.ctypes, O.IDENTITY.ctypes, C.c_size_t(3))
        ot = np.zeros(6 * ipmax, np.float32)
        ph, th, ps, dx, dy = (C.c_float(0.0) for _ in range(5))
        L.or32_psearch(nsam, 1.0, outc[p], a3df, dang, maxr1, maxr2, C.byref(ph), C.byref(th),
                       C.byref(ps), C.byref(dx), C.byref(dy), flat, len(sang), iq, mask, rbfact,
                       ctff[p], ot, ipmax, C.byref(iran_o), O.IDENTITY)
        assert iran.value == iran_o.value
        assert [par[j].value for j in range(3)] == [ph.value, th.value, ps.value]
        tp, ot = tp.reshape(ipmax, 6), ot.reshape(ipmax, 6)
        for q in range(ipmax):
            if ot[q, 5] == 0.0:
                assert tp[q, 5] == 0.0
                continue
            slots += 1
            near = [abs(ot[q, 5] - ot[s, 5]) <= 1e-4 * abs(ot[q, 5]) for s in (q - 1, q + 1)
                    if 0 <= s < ipmax]
            if any(near):
                continue
            assert abs(tp[q, 5] - ot[q, 5]) < 1e-4 * abs(ot[q, 5]), (p, q, tp[q], ot[q])
            agree += np.array_equal(tp[q, :5], ot[q, :5])
    print(f"[psearch_] TESTPAR slots with identical angles and shifts: {agree}/{slots}")
    assert agree >= 0.95 * slots

    # cc3m_ on the resident reference
    jc = nsam // 2 + 1
    fsct = np.ones(jc, np.float32)
    ri2, ri3, rih, halfw = (ri - 3) ** 2, (ri + 3) ** 2, ri - 3, 6.0
    ang = (0.4, 1.2, -0.5, 0.03, -0.02)

    def cc3m_gpu():
        return lib.cc3m_(i(nsam), i(1), f(1.0), outc[0].ctypes, a3df.ctypes, i(maxr1), i(maxr2),
                         *[f(v) for v in ang], i(0), dummy.ctypes, f(rbfact), dummy.ctypes, i(1),
                         dummy.ctypes, fsct.ctypes, i(0), f(0.0), ctff[0].ctypes, f(ri2), f(ri3),
                         f(rih), f(halfw), dummy.ctypes, dummy.ctypes, O.IDENTITY.ctypes)

    lib.cc3m_.restype = C.c_float

    def cc3m_oracle():
        return L.or32_cc3m(nsam, 1.0, outc[0], a3df, maxr1, maxr2, *ang, rbfact, fsct, ctff[0],
                           ri2, ri3, rih, halfw, O.IDENTITY)

    first, again = cc3m_gpu(), cc3m_gpu()
    want = cc3m_oracle()
    assert first == again and abs(first - want) <= 1e-4 * abs(want)
    a3df *= np.complex64(0.5 + 0.25j)  # rewritten in place: the fingerprint changes
    a3df[: nsam * jc] *= np.complex64(1.5)
    changed, want2 = cc3m_gpu(), cc3m_oracle()
    print(f"[cc3m_] resident {first:.6f} / {again:.6f} (oracle {want:.6f}); "
          f"after the rewrite {changed:.6f} (oracle {want2:.6f})")
    assert abs(changed - want2) <= 1e-4 * abs(want2) and abs(changed - first) > 1e-3
    lib.fr_percall_invalidate()
    assert abs(cc3m_gpu() - want2) <= 1e-4 * abs(want2)
