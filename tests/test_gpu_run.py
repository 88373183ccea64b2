"""A Frealign run from its card deck (frealign_amd.run) on files written the
way the reference writes them: an MRC particle stack, a v9 parameter file and
an MRC reference map, at 64^2, C1.

  * IFLAG 0 at the true parameters: the output maps (cards 13, 15, 16) exist
    with the stack's dimensions and pixel size, the FSC table appended to the
    output parameter file parses (it is what the next run reads as its
    Part_SSNR table, card10.f:104-126), and the half-map FSC of the run agrees
    on the low shells (within 0.05; the run's is taken after CORRECT3D and
    MASK3D) with the one of a reconstruction built directly through the
    library from the same particles.
  * IFLAG 1 from parameters perturbed by 3 degrees: the refined orientations
    written to the output file are closer to the truth than the input ones,
    and the shift file holds the differences.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, cards, fileio, run, search, synth  # noqa: E402

NSAM, PSIZE, N = 64, 2.0, 96
DSTEP = 5.0                     # um: MAG = DSTEP * 1e4 / PSIZE = 25,000 gives AMAGP = 1


def _dataset(ctx, tmp, seed=5):
    optics = synth.Optics(psize=PSIZE)
    vol = synth.blob_phantom(NSAM, 40, 24.0, seed=seed, device="cuda")
    a3df = synth.half_volume(vol)
    parts = synth.make_particles(ctx, a3df, N, seed=seed + 1, optics=optics, snr=2.0, ri=26.0,
                                 shift_a=2.0)
    fileio.mrc_write(os.path.join(tmp, "stack.mrc"),
                     parts.images[:, :, :NSAM].cpu().numpy(), PSIZE, fileio.title("C1 "))
    fileio.mrc_write(os.path.join(tmp, "ref.mrc"), vol.cpu().numpy(), PSIZE, fileio.title("C1 "))
    return parts, a3df


def _par(tmp, name, parts, euler):
    mag = DSTEP * 1e4 / PSIZE
    recs = []
    for k in range(N):
        phi, theta, psi = np.degrees(euler[k])
        sx, sy = parts.shift[k] * NSAM * PSIZE / (2 * np.pi)  # phase units -> A
        recs.append(fileio.ParRecord(k + 1, float(psi) % 360, float(theta) % 360,
                                     float(phi) % 360, float(sx), float(sy), mag, k // 16,
                                     float(parts.df1[k]), float(parts.df2[k]),
                                     float(np.degrees(parts.angast[k])), 100.0, 0, 0.05, 50.0))
    fileio.par_write(os.path.join(tmp, name), recs)


def _deck(tmp, iflag, finpar, rec=True, f3d="map.mrc"):
    j = lambda n: os.path.join(tmp, n)  # noqa: E731
    relmag = "0." if rec else "-100."
    return f"""M,{iflag},F,F,F,F,0,F,F,F,F,0,F,0,1	!CFORM,IFLAG,...
52.0,0.0,{PSIZE},0.0,0.07,0.0,100.0,0.0,15.0,10,4	!RO,RI,PSIZE,MW,WGH,XSTD,PBC,BOFF,DANG,ITMAX,IPMAX
1,1,1,1,1
1,{N}
C1
1.0,{DSTEP},90.0,0.0,2.7,300.0,0.0,0.0
4.0,100.0,10.0,10.0,100.0,0.0
{j('stack.mrc')}
{j('proj.mrc')}
{j(finpar)}
{j('out.par')}
{j('out.shft')}
{relmag}, 0., 0., 0., 0., 0., 0., 0.
{j(f3d)}
{j('weights')}
{j('map1.mrc')}
{j('map2.mrc')}
{j('pha')}
{j('poi')}
"""


def test_run_reconstruction_from_files(tmp_path):
    ctx = api.Context(0)
    tmp = str(tmp_path)
    parts, _ = _dataset(ctx, tmp)
    _par(tmp, "in.par", parts, parts.euler)
    print("[run] IFLAG 0 from the files", flush=True)
    r = run.Run(cards.parse_deck(_deck(tmp, 0, "in.par")), ctx=ctx)
    r.run()
    for name in ("map.mrc", "map1.mrc", "map2.mrc"):
        info = fileio.mrc_info(os.path.join(tmp, name))
        assert (info.nx, info.ny, info.nz) == (NSAM, NSAM, NSAM)
        assert abs(info.psize - PSIZE) < 1e-6 and info.rms > 0
    table = fileio.par_read(os.path.join(tmp, "out.par"))
    assert table.part_ssnr is not None and len(table.part_ssnr) > 10
    # the same reconstruction through the library directly
    optics = synth.Optics(psize=PSIZE)
    gpar = optics.gpar(NSAM, 26.0)
    pp = parts.ppar(with_shift=True)
    pssnr = torch.from_numpy(r.pssnr).cuda()
    ctff, outc, datc = ctx.ctfapply(NSAM, parts.images.clone(), pp, pssnr, r.gpar)
    vols = api.HalfMaps.zeros(NSAM)
    recb = np.zeros((N, 10), np.float32)
    recb[:, 0:3] = r.internal[:, [3, 2, 1]]
    recb[:, 5:8] = 1.0
    recb[:, 8] = 0.5
    recb[:, 9] = r.rrec
    ctx.insert(NSAM, datc, ctff, recb, (np.arange(N) % 2).astype(np.int32),
               api.point_group("C", 1), 100.0, 0.0, PSIZE, vols)
    nk = float(vols.nksum.sum().item())
    ttd = vols.ttd.cpu().numpy()
    b, c = ctx.storeshift(NSAM, ttd[0] / nk * 0.1, ttd[1] / nk * 0.1, 1, vols.num[0],
                          vols.ctf2[0], vols.num[1], vols.ctf2[1])
    fsc = ctx.sheltest(NSAM, b, c, vols.ctf2[0], vols.ctf2[1])["fsc"]
    # the run's FSC is taken after CORRECT3D and MASK3D: compare the low shells,
    # where the mask changes little, and require a correlated reconstruction
    print(f"[run] FSC shells 2-8: run {np.round(r.fsc[1:8], 3)}, direct {np.round(fsc[1:8], 3)}")
    assert np.all(r.fsc[1:6] > 0.8) and np.all(fsc[1:6] > 0.8)
    assert np.max(np.abs(r.fsc[1:4] - fsc[1:4])) < 0.05


def test_run_refinement_from_files(tmp_path):
    ctx = api.Context(0)
    tmp = str(tmp_path)
    parts, _ = _dataset(ctx, tmp, seed=9)
    rng = np.random.default_rng(3)
    start = (parts.euler + np.radians(rng.normal(0, 3.0, (N, 3)))).astype(np.float32)
    _par(tmp, "start.par", parts, start)
    print("[run] IFLAG 1 from perturbed parameters", flush=True)
    r = run.Run(cards.parse_deck(_deck(tmp, 1, "start.par", rec=False, f3d="ref.mrc")), ctx=ctx)
    r.run()
    out = fileio.par_read(os.path.join(tmp, "out.par"))
    assert [x.c for x in out.records] == list(range(1, N + 1))
    got = np.radians(np.array([[x.phi, x.theta, x.psi] for x in out.records], np.float32))
    e0 = search.angular_error_deg(start, parts.euler)
    e1 = search.angular_error_deg(got, parts.euler)
    print(f"[run] median angular error to the truth: input {np.median(e0):.2f}, "
          f"refined {np.median(e1):.2f} deg")
    assert np.median(e1) < 0.5 * np.median(e0)
    sh = fileio.par_read(os.path.join(tmp, "out.shft"))
    assert len(sh.records) == N
    dpsi = np.array([x.psi for x in sh.records])
    want = np.array([(a.psi - b.psi + 180) % 360 - 180 for a, b in
                     zip(out.records, fileio.par_read(os.path.join(tmp, "start.par")).records)])
    assert np.max(np.abs(dpsi - want)) < 0.02
