"""The allocation-free entry point sc_groundtruth_nn1_noalloc against
sc_groundtruth_nn1, through the C ABI: two calls in a row, the second with a
smaller input, must give the same ids and distance bits as the allocating
entry point, and a batch too large for the fixed scratch is refused."""
import ctypes as C

import numpy as np
import pytest

import sc_amd as S
import sco

pytestmark = pytest.mark.gpu


def _noalloc(ctx, base, q):
    nq, dim = q.shape
    ids = S.Dev(ctx, (nq,), np.int32)
    dists = S.Dev(ctx, (nq,), np.float64)
    S._check("sc_groundtruth_nn1_noalloc", S.lib().sc_groundtruth_nn1_noalloc(
        ctx.h, base.ptr, base.shape[0], q.ptr, nq, dim, ids.ptr, dists.ptr))
    return ids.numpy(), dists.numpy()


@pytest.mark.parametrize("dim", [128, 960])
def test_noalloc_groundtruth_matches_allocating_entry(ctx, dim):
    mix = sco.mix_d960() if dim == 960 else sco.mix_d128(ncomp=9)
    for n, nq in ((60000, 3000), (20011, 257)):       # second call smaller
        print("noalloc groundtruth dim=%d n=%d nq=%d" % (dim, n, nq), flush=True)
        base = sco.gen(mix, sco.STREAM_BASE, 0, n, dim)
        q = sco.gen(mix, sco.STREAM_QUERY, 0, nq, dim)
        q[:7] = base[[3, 3, 11, 500, 9000, 9000, n - 1]]        # exact hits (distance 0)
        base[n - 2] = base[3]                                   # a duplicated row: lowest id
        db, dq = S.Dev.from_numpy(ctx, base), S.Dev.from_numpy(ctx, q)
        want_ids, want_d = S.groundtruth_nn1(ctx, db, dq)
        got_ids, got_d = _noalloc(ctx, db, dq)
        assert np.array_equal(got_ids, want_ids.numpy())
        assert np.array_equal(got_d.view(np.uint64), want_d.numpy().view(np.uint64))
        assert got_ids[0] == 3 and got_d[0] == 0.0


def test_noalloc_groundtruth_refuses_oversized_batch(ctx):
    base = S.Dev(ctx, (64, 16), np.float32)
    nq = (16 << 20) // 12 + 1
    q = S.Dev(ctx, (nq, 16), np.float32)
    ids = S.Dev(ctx, (nq,), np.int32)
    rc = S.lib().sc_groundtruth_nn1_noalloc(ctx.h, base.ptr, 64, q.ptr, nq, 16, ids.ptr, None)
    assert rc == S.SC_ERR_ARG
    assert b"fixed scratch" in S.lib().sc_last_error()
