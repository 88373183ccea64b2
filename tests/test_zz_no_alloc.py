"""fr_sheltest_noalloc (no allocation: the shell accumulators live in a fixed
device array) against fr_sheltest (context scratch), both through the C-ABI,
on two calls in a row, the second with a smaller box.

SHELTEST's sums (sheltest.f:48-88) are FP64 atomics, so their summation order
varies from run to run: NPIX/NPIXT and the shell count must be identical, the
FSC within 1e-12 absolute and PSSNR within 1e-6 relative (FP32 output)."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no HIP device", allow_module_level=True)

from frealign_amd import _native as N  # noqa: E402
from frealign_amd import api  # noqa: E402

DEV = "cuda"


def _call(fn, ctx, nsam, istp, a, d, s, v):
    jc = nsam // 2 + 1
    cap = jc + 2
    npix = np.zeros(cap, np.int32)
    npixt = np.zeros(cap, np.int32)
    dfsc = np.zeros(cap, np.float64)
    pssnr = np.zeros(cap, np.float32)
    nstp = C.c_int(0)
    torch.cuda.synchronize()
    N.check(getattr(N.lib(), fn)(ctx.h, nsam, istp, a.data_ptr(), d.data_ptr(), s.data_ptr(),
                                 v.data_ptr(), 1.0, npix.ctypes.data, npixt.ctypes.data,
                                 dfsc.ctypes.data, pssnr.ctypes.data, C.byref(nstp)), fn)
    k = nstp.value
    return k, npix[:k], npixt[:k], dfsc[:k], pssnr[:k]


def _maps(rng, nsam):
    jc = nsam // 2 + 1
    n = nsam * nsam * jc
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    d = (a + 0.7 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    s = rng.uniform(0.1, 3, n).astype(np.float32)
    v = rng.uniform(0.1, 3, n).astype(np.float32)
    t = lambda x: torch.from_numpy(x).to(DEV)  # noqa: E731
    return t(a), t(d), t(s), t(v)


def test_sheltest_noalloc_equals_sheltest():
    ctx = api.Context(0)
    rng = np.random.default_rng(2026)
    try:
        # the second call is smaller: stale sums of the first must not leak in
        for nsam, istp in ((128, 1), (48, 2)):
            print(f"[no-alloc] fr_sheltest_noalloc vs fr_sheltest, nsam={nsam} istp={istp}",
                  flush=True)
            maps = _maps(rng, nsam)
            ref = _call("fr_sheltest", ctx, nsam, istp, *maps)
            got = _call("fr_sheltest_noalloc", ctx, nsam, istp, *maps)
            assert got[0] == ref[0]
            assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
            assert np.max(np.abs(got[3] - ref[3])) < 1e-12
            assert np.all(np.abs(got[4] - ref[4]) <= 1e-6 * np.abs(ref[4]) + 1e-30)
            assert got[1].sum() > 0
    finally:
        ctx.close()
