"""CALC_OCC (frealign_amd.calc_occ over fr_calc_occ) against the reference
program calc_occ.f, built from its source by oracle/Makefile into
oracle/_ref/calc_occ: the same answers on standard input, the output
parameter files compared byte for byte.  Cases: three classes in the v9
format with a Part_SSNR table in every footer, particles with all-zero
occupancies, LogP spreads beyond the 10-unit cut-off and a record with
particle number 0; FrealignX records (a PSHFT column); and, without the
reference program, the closed form of F = 1 (the new occupancies are the
shares EXP(LogP - max) A / sum * 100 and sum to 100 per particle)."""
import io
import os
import subprocess

import numpy as np
import pytest

from frealign_amd import calc_occ, fileio

REF = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "calc_occ")
NP = 40


def _class_files(tmp, ncls, seed, frealignx=False, footer=True):
    rng = np.random.default_rng(seed)
    occ = rng.uniform(0, 100, (ncls, NP)).round(2)
    occ[:, 3] = 0.0                                     # all-zero occupancies
    logp = rng.integers(-3000, -2900, (ncls, NP))
    logp[0, 5] = -2800                                  # one class far ahead
    sig = rng.uniform(0.5, 2.0, (ncls, NP)).round(4)
    names = []
    for c in range(ncls):
        lines = [fileio.PAR_HEADER, "C  class " + str(c + 1)]
        for k in range(NP):
            ptcl = 0 if k == 7 else k + 1               # a blank record, skipped
            ln = fileio.par_line(fileio.ParRecord(
                ptcl, float(rng.uniform(0, 360)), float(rng.uniform(0, 180)),
                float(rng.uniform(0, 360)), float(rng.normal(0, 5)), float(rng.normal(0, 5)),
                50000.0, k // 10, 15000.0 + k, 14800.0 - k, 30.0, float(occ[c, k]),
                float(logp[c, k]), float(sig[c, k]), float(rng.uniform(0, 80)),
                float(rng.uniform(0, 5))))
            if frealignx:
                ln = ln[:91] + f"{rng.uniform(-3, 3):8.2f}" + ln[91:]
            lines.append(ln)
        if footer:
            lines += fileio.fsc_table(64, 2.0, np.linspace(1, 0, 33) ** (c + 1),
                                      np.arange(33) * 10,
                                      np.linspace(4 + c, 0, 33) ** 2, rlim=0.45, version="test")
            lines.append("C  trailing note")
        p = os.path.join(tmp, f"in{c + 1}.par")
        with open(p, "w") as f:
            f.write("\n".join(lines) + "\n")
        names.append(p)
    return names


def _answers(ins, outs, f):
    return "\n".join([str(len(ins)), f"{f}"] + ins + outs) + "\n"


def _ours(ins, outdir, f):
    outs = [os.path.join(outdir, os.path.basename(p).replace("in", "out")) for p in ins]
    assert calc_occ.main(io.StringIO(_answers(ins, outs, f))) == 0
    return outs


def _reference(ins, outdir, f):
    outs = [os.path.join(outdir, os.path.basename(p).replace("in", "out")) for p in ins]
    r = subprocess.run([REF], input=_answers(ins, outs, f), capture_output=True, text=True,
                       timeout=60, cwd=outdir)
    assert r.returncode == 0 and "Normal termination" in r.stdout, r.stdout + r.stderr
    return outs


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/calc_occ not built")
@pytest.mark.parametrize("frealignx,f", [(False, 0.5), (False, 1.0), (True, 0.3)])
def test_calc_occ_matches_reference_program(tmp_path, frealignx, f):
    ins = _class_files(str(tmp_path), 3, seed=11 + int(frealignx), frealignx=frealignx)
    (tmp_path / "ref").mkdir()
    (tmp_path / "ours").mkdir()
    want = _reference(ins, str(tmp_path / "ref"), f)
    got = _ours(ins, str(tmp_path / "ours"), f)
    for a, b in zip(want, got):
        wa, gb = open(a, "rb").read(), open(b, "rb").read()
        if wa != gb:
            wl, gl = wa.decode().splitlines(), gb.decode().splitlines()
            diff = [(i, x, y) for i, (x, y) in enumerate(zip(wl, gl)) if x != y][:3]
            pytest.fail(f"{os.path.basename(b)} differs ({len(wl)} vs {len(gl)} lines): {diff}")


def test_calc_occ_closed_form(tmp_path):
    ins = _class_files(str(tmp_path), 4, seed=5, footer=False)
    outs = _ours(ins, str(tmp_path), 1.0)
    inp = [fileio.par_read(p).records for p in ins]
    out = [fileio.par_read(p).records for p in outs]
    a = np.array([np.mean([r.occ for r in rs]) for rs in inp])
    for k in range(len(out[0])):
        lp = np.array([rs[k].logp for rs in inp], np.float64)
        share = np.where(lp.max() - lp < 10, np.exp(lp - lp.max()) * a, 0.0)
        share = share / share.sum() * 100
        got = np.array([rs[k].occ for rs in out])
        assert np.max(np.abs(got - share)) <= 0.006, (k, got, share)
        # SIGMA: the occupancy-weighted mean of the classes' SIGMA, in every class
        assert abs(out[0][k].sigma - np.sum([rs[k].sigma for rs in inp] * share / 100)) < 2e-4
        assert len({rs[k].sigma for rs in out}) == 1


def test_calc_occ_rejects_inconsistent_files(tmp_path):
    ins = _class_files(str(tmp_path), 2, seed=3, footer=False)
    lines = open(ins[1]).read().splitlines()
    lines = lines[:-1]                                   # one particle fewer
    open(ins[1], "w").write("\n".join(lines) + "\n")
    outs = [str(tmp_path / "o1.par"), str(tmp_path / "o2.par")]
    assert calc_occ.main(io.StringIO(_answers(ins, outs, 0.5))) == 1
