"""GPU: updec_segsum_compact_noalloc, the segmented SUM that allocates
nothing (its accumulators live in a fixed device array of the library),
against updec_segsum_compact on the same inputs, byte for byte, called
through the C ABI.  Two calls in a row, the second smaller, so a state left
by the first call (the fixed array keeps it) must not leak into the second;
then a call past the array's limit must fail with UPDEC_EINVAL and leave the
output untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UPDEC_EINVAL = 1


def _err(U):
    return U.lib().updec_last_error().decode(errors="replace")


def _noalloc(U, col, P, S, rows, assoc, groups, ngroups, out):
    fn = U.lib().updec_segsum_compact_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return U.lib().updec_segsum_compact_noalloc(U._ptr(col), P, S, rows, U._ptr(assoc), U._ptr(groups), ngroups,
                                                U._ptr(out), U.MODE_AUTO, None)


def _case(U, P, rows, nseg, seed):
    rng = np.random.default_rng(seed)
    host = U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, P, seed, 3, digits=P - 2, neg_permille=500), rows)
    assoc = rng.permutation(rows).astype(np.int32)
    cuts = np.sort(rng.choice(np.arange(1, rows), size=nseg - 1, replace=False))
    groups = np.concatenate([[0], cuts]).astype(np.int32)
    return host, assoc, groups


@pytest.mark.parametrize("P,S", [(15, 2), (76, 20)])
def test_noalloc_equals_allocating_entry_point(U, O, gpu, P, S):
    # first a large call (many segments, some tiles crossing starts), then a
    # smaller one with fewer segments
    for rows, nseg, seed in ((400_000, 5_000, 1), (70_000, 7, 2)):
        host, assoc, groups = _case(U, P, rows, nseg, seed + P)
        col = U.DeviceBuffer.from_numpy(host)
        da = U.DeviceBuffer.from_numpy(assoc)
        dg = U.DeviceBuffer.from_numpy(groups)
        want = U.DeviceBuffer(nseg * 133).zero()
        got = U.DeviceBuffer(nseg * 133).zero()
        U.segsum_compact(col, P, S, rows, da, dg, nseg, want)
        rc = _noalloc(U, col, P, S, rows, da, dg, nseg, got)
        assert rc == 0, _err(U)
        U.device_sync()
        w, g = want.numpy(), got.numpy()
        assert np.array_equal(w, g), (rows, nseg)
        # and both equal the oracle
        ref = O.segsum_compact(host, P, S, rows, assoc, groups)
        assert np.array_equal(g.reshape(-1, 133), ref)


def test_noalloc_rejects_more_segments_than_its_scratch(U, gpu):
    lim = ctypes.c_int32()
    assert U.lib().updec_segsum_noalloc_max_segments(15, ctypes.byref(lim)) == 0
    n = lim.value + 1
    host = U.gen_compact_host(U.gen_desc(U.GEN_DIGITS, 15, 5, 1, digits=10), 16)
    col = U.DeviceBuffer.from_numpy(host)
    groups = U.DeviceBuffer.from_numpy(np.zeros(n, dtype=np.int32))
    out = U.DeviceBuffer(133).zero()
    rc = _noalloc(U, col, 15, 2, 16, None, groups, n, out)
    assert rc == UPDEC_EINVAL
    assert "scratch" in _err(U)
    U.device_sync()
    assert not out.numpy().any()
