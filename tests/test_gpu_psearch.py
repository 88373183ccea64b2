"""The GPU global search against the oracle's PSEARCH (psearch.f:8-121 ->
ccp.f:8-275 with PAD and ROTATE2D) on the same prepared particles, the same
SEARCHANG grid and the same RANDOM jitter.

The GEMM searches score each in-plane angle from a section extracted at that
angle; CCP rotates one padded section by nearest-neighbour lookup.  The two
therefore agree on the selected orientation, not bit for bit on the scores:
the bar is the same best view within one grid step (modulo the point group)
for every particle, and the gaps between the two scores are printed.
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import oracle_lib as O  # noqa: E402

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import api, search, synth  # noqa: E402

DEV = "cuda"


@pytest.fixture(scope="module")
def ctx():
    c = api.Context(0)
    yield c
    c.close()


def oracle_psearch(nsam, outc, ctff, a3df, dang, maxr1, maxr2, rbfact, asym, nasym, iran,
                   ipmax):
    """or32_psearch per particle (threads: the oracle releases the GIL)."""
    L = O.oracle()
    sang, iquadmax = api.search_grid(dang, asym, nasym)
    flat = np.ascontiguousarray(sang.reshape(-1))
    mask = np.ones(5, np.int32)

    def one(p):
        tp = np.zeros(6 * ipmax, np.float32)
        ir = C.c_int(iran)
        ph, th, ps, dx, dy = (C.c_float(0.0) for _ in range(5))
        L.or32_psearch(nsam, 1.0, outc[p].reshape(-1), a3df, dang, maxr1, maxr2, C.byref(ph),
                       C.byref(th), C.byref(ps), C.byref(dx), C.byref(dy), flat, len(sang),
                       iquadmax, mask, rbfact, ctff[p].reshape(-1), tp, ipmax, C.byref(ir),
                       O.IDENTITY)
        return tp.reshape(ipmax, 6)

    with ThreadPoolExecutor(8) as ex:
        return np.stack(list(ex.map(one, range(outc.shape[0]))))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_search_best_view_matches_oracle_psearch(ctx, dtype):
    nsam, psize, ri = 64, 2.0, 26.0
    dang, maxr1, maxr2, rbfact, iran = 15.0, 1, 14, 0.0005, -5
    optics = synth.Optics(psize=psize)
    sym = api.point_group("D", 7)
    vol = synth.blob_phantom(nsam, 12, 24.0, seed=7, sym=sym, device=DEV)
    a3df = synth.half_volume(vol)
    npart = 16
    parts = synth.make_particles(ctx, a3df, npart, seed=8, optics=optics, snr=1.0, ri=ri,
                                 shift_a=2.0)
    gpar = optics.gpar(nsam, ri)
    pssnr = torch.from_numpy(synth.default_pssnr(nsam, ri)).to(DEV)
    ctff, outc, _ = ctx.ctfapply(nsam, parts.images.clone(), parts.ppar(), pssnr, gpar)
    print(f"[psearch] gpu search {dtype}, {npart} particles at {nsam}^2, D7 {dang} deg",
          flush=True)
    gs = search.GlobalSearch(ctx, a3df, dang=dang, asym="D", nasym=7, maxr1=maxr1, maxr2=maxr2,
                             rbfact=rbfact, dtype=dtype, iran=iran)
    top_o, top_r, top_s = gs.candidates(outc, ctff, topk=8)
    top_o, top_s = top_o.cpu().numpy(), top_s.cpu().numpy()
    print("[psearch] oracle PSEARCH", flush=True)
    tp = oracle_psearch(nsam, outc.cpu().numpy(), ctff.cpu().numpy(),
                        a3df.cpu().numpy().reshape(-1), dang, maxr1, maxr2, rbfact, "D", 7, iran,
                        8)
    assert (top_o[:, 0] >= 0).all() and (tp[:, 0, 5] > 0).all()
    gpu_best = gs.euler[top_o[:, 0]]
    ora_best = tp[:, 0, :3]
    ops = synth.symmetry_matrices(sym)
    err = search.angular_error_deg(gpu_best, ora_best, ops)
    gap = np.abs(top_s[:, 0] - tp[:, 0, 5]) / np.abs(tp[:, 0, 5])
    print(f"[psearch] {dtype}: best-view distance to PSEARCH (deg) {np.round(err, 1)}")
    print(f"[psearch] {dtype}: relative score gap GEMM vs CCP {np.round(gap, 3)}")
    # one grid step: the view spacing (15 deg) and CCP's in-plane step (3 deg)
    assert np.mean(err <= dang + 3.0) >= 0.99
