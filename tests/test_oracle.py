"""Pins the CPU oracle (oracle/sc_oracle.c) before it is trusted as the checker.

Each test checks the oracle against something independent of its own code:
published known-answer vectors, the reference's own sources compiled here
(oracle/_ref, marked `ref`), brute force, closed forms, or FP64 invariants.
"""
import heapq
import os
import tempfile

import numpy as np
import pytest

import sco

ref_missing = pytest.mark.skipif(sco.ref() is None, reason="oracle/_ref not built")


# ------------------------------------------------------------------ Philox
def test_philox_random123_kat():
    # Random123 kat_vectors, philox4x32_10
    assert list(sco.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(sco.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(sco.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                           [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_generator_chunks_and_grid():
    m = sco.mix_d128()
    a = sco.gen(m, sco.STREAM_BASE, 0, 300, 128)
    b = sco.gen(m, sco.STREAM_BASE, 100, 200, 128)
    assert np.array_equal(a[100:], b)                      # any chunk regenerates
    assert np.array_equal(a, sco.gen(m, sco.STREAM_BASE, 0, 300, 128))
    assert not np.array_equal(a, sco.gen(m, sco.STREAM_QUERY, 0, 300, 128))
    assert a.min() >= 0 and a.max() <= 255 and np.all(a == np.round(a))
    # Irwin-Hall noise has unit variance: the within-row spread is sigma-sized
    big = sco.gen(m, sco.STREAM_TRAIN, 0, 4000, 128)
    assert 60 < big.mean() < 200
    m960 = sco.mix_d960()
    c = sco.gen(m960, sco.STREAM_BASE, 0, 50, 960)
    assert c.min() >= 0 and c.max() <= 1
    assert np.all(c * 65536 == np.round(c * 65536))        # 2^-16 grid
    assert 0.05 < (c - c.mean(axis=0)).std() < 0.2         # ~0.1 noise


# ------------------------------------------------------------------ S1
def _s1_closed(x, y):
    t = (x.astype(np.float32) - y.astype(np.float32)).astype(np.float64)
    return np.cumsum(t * t, axis=-1)[..., -1]               # sequential FP64


def test_distance_closed_form():
    rng = np.random.default_rng(1)
    for d in (1, 16, 60, 64, 128, 480):
        x = rng.standard_normal(d).astype(np.float32) * 100
        y = rng.standard_normal(d).astype(np.float32) * 100
        assert sco.dist(x, y) == _s1_closed(x, y)
    # integer data: exact integer distance
    x = rng.integers(0, 256, 128).astype(np.float32)
    y = rng.integers(0, 256, 128).astype(np.float32)
    assert sco.dist(x, y) == float(((x.astype(np.int64) - y.astype(np.int64)) ** 2).sum())


@ref_missing
@pytest.mark.ref
def test_distance_matches_reference():
    rng = np.random.default_rng(2)
    for d in (16, 64, 128, 960):
        for _ in range(20):
            x = (rng.standard_normal(d) * rng.uniform(0.01, 300)).astype(np.float32)
            y = (rng.standard_normal(d) * rng.uniform(0.01, 300)).astype(np.float32)
            assert sco.dist(x, y) == sco.ref().scref_dist(x, y, d)


# ------------------------------------------------------------------ encoders
def _codebooks(rng, x, mc, kc, mp, kp, dup=True):
    n, dim = x.shape
    bsc, bsp = dim // mc, dim // mp
    cq = np.empty((mc, kc, bsc), np.float32)
    for j in range(mc):
        cq[j] = x[rng.choice(n, kc, replace=False), j * bsc:(j + 1) * bsc]
        cq[j] += rng.uniform(-0.5, 0.5, (kc, bsc)).astype(np.float32)
        if dup:  # exact duplicates force distance ties
            cq[j, kc - 1] = cq[j, 1]
            cq[j, kc - 2] = cq[j, 0]
    pq = (rng.standard_normal((mp, kp, bsp)) * 8).astype(np.float32)
    if dup:
        pq[:, kp - 1] = pq[:, 3]
    return cq, pq


def _bf_encode(x, cq, pq, mc, kc, mp, kp, nc):
    """Brute force: exact S1 distances, then the tie rules by explicit sorting."""
    n, dim = x.shape
    bsc, bsp = dim // mc, dim // mp
    cid = np.zeros((n, mc * nc), np.uint16)
    codes = np.zeros((n, mp), np.uint8)
    for i in range(n):
        r = np.empty(dim, np.float32)
        for j in range(mc):
            xb = x[i, j * bsc:(j + 1) * bsc]
            d = _s1_closed(xb[None, :], cq[j])
            if nc == 1:
                order = np.lexsort((np.arange(kc), d))       # (dist, idx ASC)
                cid[i, j] = order[0]
            else:
                order = np.lexsort((-np.arange(kc), d))      # (dist, idx DESC)
                cid[i, j], cid[i, mc + j] = order[0], order[1]
            r[j * bsc:(j + 1) * bsc] = xb - cq[j, order[0]]
        for s in range(mp):
            d = _s1_closed(r[None, s * bsp:(s + 1) * bsp], pq[s])
            codes[i, s] = np.lexsort((np.arange(kp), d))[0]
    return cid, codes


@pytest.mark.parametrize("nc", [1, 2])
@pytest.mark.parametrize("integer", [True, False])
def test_encode_bruteforce(nc, integer):
    rng = np.random.default_rng(10 + nc)
    mc, kc, mp, kp, dim, n = 2, 12, 4, 16, 32, 300
    x = sco.gen(sco.mix_d128(ncomp=5), sco.STREAM_TRAIN, 0, n, dim)
    if not integer:
        x = x + rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp)
    # plant exact ties: some vectors equal to a duplicated centroid
    x[:4, :dim // mc] = cq[0, 1]
    x[4:8, dim // mc:] = cq[1, 0]
    cid, codes = sco.encode(x, cq, pq, mc, kc, mp, kp, nc)
    bcid, bcodes = _bf_encode(x, cq, pq, mc, kc, mp, kp, nc)
    assert np.array_equal(cid, bcid)
    assert np.array_equal(codes, bcodes)
    if nc == 2:
        assert np.all(cid[:, :mc] != cid[:, mc:])
        # the tie at centroid 1 / kc-1 goes to the later index
        assert np.all(cid[:4, 0] == kc - 1) and np.all(cid[:4, mc] == 1)


def test_encode_given_pins():
    """Encoder::rencode's restatement: with each vector's own nearest cells
    it reproduces Encoder::encode's codes (pinned above against brute force);
    with permuted rows and other cells it equals brute force on the explicit
    residual (ties planted in the PQ codebook)."""
    rng = np.random.default_rng(77)
    mc, kc, mp, kp, dim, n = 2, 12, 4, 16, 32, 300
    x = sco.gen(sco.mix_d128(ncomp=5), sco.STREAM_TRAIN, 0, n, dim)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp)
    cid, codes = sco.encode(x, cq, pq, mc, kc, mp, kp, 1)
    assert np.array_equal(sco.encode_given(x, np.arange(n), cid, cq, pq, mc, kc, mp, kp), codes)
    rows = rng.integers(0, n, 500)
    gcid = rng.integers(0, kc, (500, mc))
    got = sco.encode_given(x, rows, gcid, cq, pq, mc, kc, mp, kp)
    bsc, bsp = dim // mc, dim // mp
    for i in range(500):
        r = x[rows[i]].copy()
        for j in range(mc):
            r[j * bsc:(j + 1) * bsc] = r[j * bsc:(j + 1) * bsc] - cq[j, gcid[i, j]]
        for s in range(mp):
            d = _s1_closed(r[None, s * bsp:(s + 1) * bsp], pq[s])
            assert got[i, s] == np.lexsort((np.arange(kp), d))[0]


@ref_missing
@pytest.mark.ref
@pytest.mark.parametrize("mc,kc", [(1, 32), (2, 16)])
def test_encode_and_buckets_match_reference(mc, kc):
    rng = np.random.default_rng(20 + mc)
    mp, kp, dim, n = 8, 32, 64, 1500
    x = sco.gen(sco.mix_d128(ncomp=7), sco.STREAM_BASE, 0, n, dim)
    x[:6] = np.tile(np.concatenate([c[1] for c in _codebooks(rng, x, mc, kc, mp, kp)[0]]), (6, 1))
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp)
    with tempfile.TemporaryDirectory() as t:
        fv, cp, pp = (os.path.join(t, f) for f in ("x.fvecs", "cq.ctr_", "pq.ctr_"))
        sco.write_fvecs(fv, x)
        sco.write_codebook(cp, cq, kc, mc, dim)
        sco.write_codebook(pp, pq, kp, mp, dim)
        ncells = kc ** mc
        rcid = np.zeros((n, mc), np.uint16)
        rcodes = np.zeros((n, mp), np.uint8)
        rL = np.zeros(ncells, np.int32)
        rpid = np.zeros(n, np.int32)
        rcs = np.zeros((n, mp), np.uint8)
        got = sco.ref().scref_encode_imi(fv.encode(), cp.encode(), pp.encode(), rcid,
                                         rcodes, rL, rpid, rcs, n, ncells)
    assert got == n
    cid, codes = sco.encode(x, cq, pq, mc, kc, mp, kp, 1)
    assert np.array_equal(cid, rcid)
    assert np.array_equal(codes, rcodes)
    keys = sco.cell_keys(cid, mc, kc, 1)
    for i in range(0, n, 97):
        assert keys[i] == sco.ref().scref_vector_base(np.ascontiguousarray(cid[i]), mc, kc)
    L, pid, cs = sco.build_dense(keys, ncells, codes)
    assert np.array_equal(L, rL) and np.array_equal(pid, rpid) and np.array_equal(cs, rcs)


def test_keys_and_dense_cells():
    rng = np.random.default_rng(3)
    n, mc, kc = 500, 2, 6
    cid = rng.integers(0, kc, (n, 2 * mc)).astype(np.uint16)
    k2 = sco.cell_keys(cid, mc, kc, 2)
    uid = cid[:, :mc].astype(np.uint64) * kc + cid[:, mc:]
    assert np.array_equal(k2, uid[:, 0] * kc * kc + uid[:, 1])
    k1 = sco.cell_keys(np.ascontiguousarray(cid[:, :mc]), mc, kc, 1)
    assert np.array_equal(k1, cid[:, 0].astype(np.uint64) * kc + cid[:, 1])
    codes = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    L, pid, cs = sco.build_dense(k1, kc * kc, codes)
    assert L[-1] == n and np.all(np.diff(np.concatenate([[0], L])) >= 0)
    start = 0
    for c in range(kc * kc):
        members = pid[start:L[c]]
        assert np.array_equal(members, np.nonzero(k1 == c)[0])   # ids ascending
        assert np.array_equal(cs[start:L[c]], codes[members])
        start = L[c]


# ------------------------------------------------------------------ heaps
@ref_missing
@pytest.mark.ref
def test_heap4_matches_reference_heap():
    import ctypes as C
    rng = np.random.default_rng(4)
    L = sco.lib()
    for trial in range(40):
        cap = int(rng.integers(3, 40))
        # few distinct scores -> many ties; inserts past capacity are dropped
        ov = np.zeros(cap, np.float32); oi = [np.zeros(cap, np.int32) for _ in range(4)]
        rv = np.zeros(cap, np.float32); ri = [np.zeros(cap, np.int32) for _ in range(4)]
        h = sco.Heap4(ov.ctypes.data, *(a.ctypes.data for a in oi), 0, cap)
        rpos = C.c_int32(0)
        for step in range(300):
            if rng.random() < 0.55:
                v = float(rng.integers(0, 6)) if trial % 2 else float(rng.standard_normal())
                ids = [int(a) for a in rng.integers(0, 1000, 4)]
                L.sco_heap4_insert(C.byref(h), v, *ids)
                sco.ref().scref_heap4_insert(rv, *ri, v, *ids, C.byref(rpos), cap)
            else:
                a, b, c, d = (C.c_int32() for _ in range(4))
                got = L.sco_heap4_pop(C.byref(h), C.byref(a), C.byref(b), C.byref(c), C.byref(d))
                out = np.zeros(4, np.int32)
                rgot = sco.ref().scref_heap4_pop(rv, *ri, out, C.byref(rpos))
                assert got == rgot
                if got:
                    assert [a.value, b.value, c.value, d.value] == list(out)
            assert h.pos == rpos.value
            n = h.pos
            assert np.array_equal(ov[:n], rv[:n])
            for x, y in zip(oi, ri):
                assert np.array_equal(x[:n], y[:n])


@ref_missing
@pytest.mark.ref
def test_heap1_matches_reference_heap():
    """The single-id heap of search_ivfadc (pq_insert/pq_pop,
    include/sc_utilities.h:229-300) against the reference's compiled heap:
    random insert/pop sequences with many ties, drops past capacity and pops
    of an emptied heap; the arrays must agree after every step."""
    import ctypes as C
    rng = np.random.default_rng(14)
    L, R = sco.lib(), sco.ref()
    for trial in range(60):
        cap = int(rng.integers(2, 48))
        ov = np.zeros(cap, np.float32); oi = np.zeros(cap, np.int32)
        rv = np.zeros(cap, np.float32); ri = np.zeros(cap, np.int32)
        h = sco.Heap1(ov.ctypes.data, oi.ctypes.data, 0, cap)
        rpos = C.c_int32(0)
        for step in range(400):
            if rng.random() < 0.55:
                v = float(rng.integers(0, 5)) if trial % 2 else float(rng.standard_normal())
                i = int(rng.integers(0, 10 ** 6))
                L.sco_heap1_insert(C.byref(h), v, i)
                R.scref_heap1_insert(rv, ri, v, i, C.byref(rpos), cap)
            else:
                assert L.sco_heap1_pop(C.byref(h)) == R.scref_heap1_pop(rv, ri, C.byref(rpos))
            assert h.pos == rpos.value
            n = h.pos
            assert np.array_equal(ov[:n], rv[:n]) and np.array_equal(oi[:n], ri[:n])


def _replay_ivfadc(ix, q, w, T):
    """search_ivfadc steps 1-2 (include/query.h:225-259) driven by the
    reference's compiled pq_insert/pq_pop: every coarse score inserted in
    centroid order into a heap of capacity kc, then up to w pops, stopping
    once the budget T is reached."""
    import ctypes as C
    R = sco.ref()
    dqc, _ = sco.lut(q, ix.cq, ix.pq, ix.norm_c, ix.norm_r, ix.dim, ix.mc, ix.kc, ix.mp, ix.kp)
    kc = ix.kc
    acc = np.float32(0)
    for t in q:
        acc = np.float32(acc + np.float32(t * t))
    hv = np.zeros(kc, np.float32); hid = np.zeros(kc, np.int32)
    pos = C.c_int32(0)
    for i in range(kc):
        R.scref_heap1_insert(hv, hid, np.float32(acc + dqc[i]), i, C.byref(pos), kc)
    cells, total = [], 0
    for _ in range(w):
        bid = R.scref_heap1_pop(hv, hid, C.byref(pos))
        cells.append(bid)
        total += int(ix.L[bid] - ix.L[bid - 1]) if bid > 0 else int(ix.L[0])
        if total >= T:
            break
    return cells, total


@ref_missing
@pytest.mark.ref
def test_ivfadc_traversal_replay_with_reference_heap():
    """The oracle's IVFADC traversal equals the replay on the reference's own
    single-id heap, over budgets that stop early, late and at w."""
    for kc in (8, 37):
        ix, x, _, _ = _make_index(sco.IVFADC, n=3000, kc=kc)
        q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 10, ix.dim)
        for qi in range(10):
            for w, T in ((1, 10 ** 9), (kc // 2, 700), (kc, 10 ** 9), (kc, 1)):
                ids, dists, st, ex = ix.search(sco.IVFADC, q[qi], R=15, w=w, T=T, trace=True)
                cells, total = _replay_ivfadc(ix, q[qi], w, T)
                assert list(ex["cells"]) == cells
                assert st.sum == total


def test_heap_lone_child_is_not_compared():
    """Known answer of the array heap (children of i are {2i, 2i+1}; the
    sift-down stops while j + 1 < size, so a lone last child is never
    compared): insert 4,1,0,6; pop; insert 7; pop; pop returns 6 while 4 is
    still queued.  The sequence was found by searching small op sequences
    against the reference's compiled pq_insert4/pq_pop4."""
    import ctypes as C
    cap = 8
    v = np.zeros(cap, np.float32); ids = [np.zeros(cap, np.int32) for _ in range(4)]
    h = sco.Heap4(v.ctypes.data, *(a.ctypes.data for a in ids), 0, cap)
    ins = lambda val: sco.lib().sco_heap4_insert(C.byref(h), float(val), val, 0, 0, 0)
    a, b, c, d = (C.c_int32() for _ in range(4))
    pop = lambda: (sco.lib().sco_heap4_pop(C.byref(h), C.byref(a), C.byref(b), C.byref(c),
                                           C.byref(d)), a.value)[1]
    for val in (4, 1, 0, 6):
        ins(val)
    assert pop() == 0
    ins(7)
    assert pop() == 1
    assert pop() == 6          # out of order: 4 is still in the heap
    assert pop() == 4
    assert pop() == 7


# ------------------------------------------------------------------ top-R
def test_topr_reference_kat_and_ties():
    # test/test_algorithm.cpp:72-79: sort_id({1.0,3.2,2.4,2.3}) -> 1.0, 2.3 (id 3)
    pos = sco.topr(np.array([1.0, 3.2, 2.4, 2.3], np.float32), 4)
    assert list(pos) == [0, 3, 2, 1]
    d = np.array([2, 1, 2, 1, 0, 2], np.float32)
    assert list(sco.topr(d, 4)) == [4, 1, 3, 0]            # ties by position
    rng = np.random.default_rng(5)
    d = rng.integers(0, 50, 1000).astype(np.float32)
    assert np.array_equal(sco.topr(d, 100), np.lexsort((np.arange(1000), d))[:100])


def test_recall_definition():
    gt = np.array([5, 7, 9], np.int32)
    ids = np.array([[1, 5, 2], [7, 0, 0], [1, 2, 3]], np.int32)
    assert sco.recall(gt, ids, 3) == pytest.approx(2 / 3)
    assert sco.recall(gt, ids, 1) == pytest.approx(1 / 3)


# ------------------------------------------------------------------ search
def _make_index(kind, n=3000, dim=32, kc=8, mp=4, kp=16, seed=7):
    rng = np.random.default_rng(seed)
    mc = 2 if kind in (sco.IMI, sco.SCIMI) else 1
    x = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_BASE, 0, n, dim)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp, dup=False)
    nc = {sco.SC: 2, sco.SC3: 3, sco.SCIMI: 2}.get(kind, 1)
    if kind == sco.SC3:
        cid, codes = sco.encode_general(x, cq, pq, mc, kc, mp, kp, nc)
    else:
        cid, codes = sco.encode(x, cq, pq, mc, kc, mp, kp, nc)
    keys = sco.cell_keys(cid, mc, kc, nc)
    if kind == sco.SCIMI:
        ck, L, pid, cs = sco.build_sparse(keys, codes)
        return sco.OracleIndex(dim, mc, kc, mp, kp, cq, pq, L, pid, cs, ckeys=ck), x, cid, codes
    ncells = {sco.IVFADC: kc, sco.SC3: kc ** 3}.get(kind, kc * kc)
    L, pid, cs = sco.build_dense(keys, ncells, codes)
    ix = sco.OracleIndex(dim, mc, kc, mp, kp, cq, pq, L, pid, cs)
    return ix, x, cid, codes


def test_adc_matches_fp64_reconstruction():
    """ADC value = ||q - c - r||^2 for the candidate's centroid c and PQ
    reconstruction r (FP64), within FP32 rounding; q = c + r gives ~0."""
    for kind in (sco.IVFADC, sco.IMI, sco.SC, sco.SC3, sco.SCIMI):
        ix, x, cid, codes = _make_index(kind)
        q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 3, ix.dim)
        for qi in range(3):
            ids, dists, st, ex = ix.search(kind, q[qi], R=10, w=ix.kc, T=10 ** 9, trace=True)
            for p in range(0, ex["all_ids"].size, 37):
                vid = ex["all_ids"][p]
                bsc, bsp = ix.dim // ix.mc, ix.dim // ix.mp
                cvec = np.concatenate([ix.cq.reshape(ix.mc, ix.kc, bsc)[j, cid[vid, j]]
                                       for j in range(ix.mc)])
                rvec = np.concatenate([ix.pq.reshape(ix.mp, ix.kp, bsp)[s, codes[vid, s]]
                                       for s in range(ix.mp)])
                exact = np.sum((q[qi].astype(np.float64) - cvec - rvec) ** 2)
                scale = np.sum(q[qi].astype(np.float64) ** 2) + np.sum(cvec.astype(np.float64) ** 2) * 2 + 1e4
                assert abs(ex["all_dists"][p] - exact) <= 4e-6 * scale
        # q = c + r exactly reconstructs a database vector's code -> ADC ~ 0
        vid = 11
        bsc, bsp = ix.dim // ix.mc, ix.dim // ix.mp
        cvec = np.concatenate([ix.cq.reshape(ix.mc, ix.kc, bsc)[j, cid[vid, j]] for j in range(ix.mc)])
        rvec = np.concatenate([ix.pq.reshape(ix.mp, ix.kp, bsp)[s, codes[vid, s]] for s in range(ix.mp)])
        qq = (cvec + rvec).astype(np.float32)
        ids, dists, st, ex = ix.search(kind, qq, R=1, w=ix.kc, T=10 ** 9, trace=True)
        pos = int(np.nonzero(ex["all_ids"] == vid)[0][0])
        assert abs(ex["all_dists"][pos]) <= 1e-5 * np.sum(qq.astype(np.float64) ** 2)


# ------------------------------------------------------------------ two halves
def test_sparse_directory_matches_dense_cells():
    """The sparse directory lists exactly the non-empty dense cells: keys
    ascending, the same sizes, members and codes in the same order."""
    ix, x, cid, codes = _make_index(sco.SCIMI, n=2500, kc=5)
    keys = sco.cell_keys(cid, 2, 5, 2)
    assert keys.max() < 5 ** 4
    L, pid, cs = sco.build_dense(keys, 5 ** 4, codes)
    sizes = np.diff(np.concatenate([[0], L]))
    assert np.array_equal(ix.ckeys, np.nonzero(sizes)[0].astype(np.uint64))
    assert np.array_equal(np.diff(np.concatenate([[0], ix.L])), sizes[sizes > 0])
    assert np.array_equal(ix.pid, pid) and np.array_equal(ix.codes, cs)
    # the key is vector_base over the per-half (nearest, second) uids
    # (src/sc_encoder.cpp:39-50): uid_h = n_h*kc + s_h, key = uid_0*kc^2 + uid_1
    want = ((cid[:, 0].astype(np.uint64) * 5 + cid[:, 2]) * 25 +
            cid[:, 1].astype(np.uint64) * 5 + cid[:, 3])
    assert np.array_equal(keys, want)


def _scimi_grid_order(ix, q):
    """Brute force over the whole (n0, s0, n1, s1) grid: every cell's score
    from the halves' SC pair scores in FP32, the grid sorted by (score, key),
    the empty cells dropped."""
    kc = ix.kc
    dqc, _ = sco.lut(q, ix.cq, ix.pq, ix.norm_c, ix.norm_r, ix.dim, 2, kc, ix.mp, ix.kp)
    bsc = ix.dim // 2
    v = []
    for h in range(2):
        acc = np.float32(0)
        for t in q[h * bsc:(h + 1) * bsc]:
            acc = np.float32(acc + np.float32(t * t))
        v.append((acc + dqc[h * kc:(h + 1) * kc]).astype(np.float32))
    g = np.indices((kc, kc, kc, kc)).reshape(4, -1).astype(np.int64)
    n0, s0, n1, s1 = g
    score = ((v[0][n0] + v[0][s0]).astype(np.float32) +
             (v[1][n1] + v[1][s1]).astype(np.float32)).astype(np.float32)
    key = (n0 * kc + s0) * kc * kc + n1 * kc + s1
    order = np.lexsort((key, score))
    pos = {int(k): c for c, k in enumerate(ix.ckeys)}
    return [pos[int(key[o])] for o in order if int(key[o]) in pos]


@pytest.fixture
def scimi_select(request):
    """Ranks every two-halves directory by selection (histogram buckets up to
    the budget, then a sort of those cells only) or by sorting every cell."""
    sco.lib().sco_scimi_set_select_min(1 if request.param else 1 << 20)
    yield request.param
    sco.lib().sco_scimi_set_select_min(1 << 20)


@pytest.mark.parametrize("scimi_select", [False, True], indirect=True)
def test_scimi_ranking_matches_grid_enumeration(scimi_select):
    """Two-halves search: the cells taken are the prefix of the brute-force
    grid order that stops at w cells or T candidates, whatever the budgets
    (including every cell of the index and T = 0), by either ranking path."""
    ix, x, cid, codes = _make_index(sco.SCIMI, n=3000, kc=6)
    sizes = np.diff(np.concatenate([[0], ix.L]))
    q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 8, ix.dim)
    for qi in range(8):
        order = _scimi_grid_order(ix, q[qi])
        assert len(order) == ix.ckeys.size
        for w, T in ((3, 10 ** 9), (50, 400), (10 ** 5, 1500), (10 ** 5, 10 ** 9), (9, 0)):
            ids, dists, st, ex = ix.search(sco.SCIMI, q[qi], R=30, w=w, T=T, trace=True)
            want, tot = [], 0
            for c in order:
                if len(want) >= w or tot >= T:
                    break
                want.append(c)
                tot += int(sizes[c])
            assert list(ex["cells"]) == want
            assert st.sum == tot and st.ncells == len(want)
            assert st.exhausted == (len(want) == len(order) and len(want) < w and tot < T)
            # the scan's coarse terms use the nearest centroid of each half
            k2 = ix.kc ** 2
            assert np.array_equal(ex["s1"], (ix.ckeys[want] // k2 // ix.kc).astype(np.int32))
            assert np.array_equal(ex["s2"], (ix.ckeys[want] % k2 // ix.kc).astype(np.int32))


def test_scimi_refuses_dense_index_and_other_kinds_refuse_sparse():
    ix, x, _, _ = _make_index(sco.SCIMI, n=500, kc=4)
    with pytest.raises(ValueError):
        ix.search(sco.IMI, x[0], R=5, w=10, T=10)
    ixd, xd, _, _ = _make_index(sco.IMI, n=500, kc=4)
    with pytest.raises(ValueError):
        ixd.search(sco.SCIMI, xd[0], R=5, w=10, T=10)


def _replay_traversal(ix, kind, q, w, T):
    """Direct replay of the multi-sequence step (multi_query.h:133-205,
    sc_query.h:111-186) driven by the reference's compiled heap."""
    import ctypes as C
    R = sco.ref()
    dqc, dqr = sco.lut(q, ix.cq, ix.pq, ix.norm_c, ix.norm_r, ix.dim, ix.mc, ix.kc, ix.mp, ix.kp)
    kc = ix.kc
    Lc = ix.L
    length = lambda b: (Lc[b] - Lc[b - 1]) if b > 0 else Lc[0]
    hv = np.zeros(w, np.float32); hid = [np.zeros(w, np.int32) for _ in range(4)]
    pos = C.c_int32(0)
    ins = lambda s, a, b, c, d: R.scref_heap4_insert(hv, *hid, s, a, b, c, d, C.byref(pos), w)
    out = np.zeros(4, np.int32)
    trav = np.zeros((kc, kc), bool)
    cells, total = [], 0
    if kind == sco.IMI:
        qs = [np.float32(0), np.float32(0)]
        bsc = ix.dim // 2
        for h in range(2):
            acc = np.float32(0)
            for t in q[h * bsc:(h + 1) * bsc]:
                acc = np.float32(acc + np.float32(t * t))
            qs[h] = acc
        v = [np.float32(qs[h] + dqc[h * kc:(h + 1) * kc]) for h in range(2)]
        o = [np.lexsort((np.arange(kc), v[h])) for h in range(2)]
        vs = [v[h][o[h]] for h in range(2)]
        ins(np.float32(vs[0][0] + vs[1][0]), 0, 0, o[0][0], o[1][0])
        visited = 0
        while visited < w and total < T:
            if not R.scref_heap4_pop(hv, *hid, out, C.byref(pos)):
                break                       # whole grid visited
            h1, h2, h3, h4 = (int(z) for z in out)
            bid = h3 * kc + h4
            if length(bid) > 0:
                cells.append(bid); total += length(bid)
            trav[h1, h2] = True
            visited += 1
            if total < T:
                if h1 < kc - 1 and (h2 == 0 or trav[h1 + 1, h2 - 1]):
                    ins(np.float32(vs[0][h1 + 1] + vs[1][h2]), h1 + 1, h2, o[0][h1 + 1], o[1][h2])
                if h2 < kc - 1 and (h1 == 0 or trav[h1 - 1, h2 + 1]):
                    ins(np.float32(vs[0][h1] + vs[1][h2 + 1]), h1, h2 + 1, o[0][h1], o[1][h2 + 1])
    else:
        acc = np.float32(0)
        for t in q:
            acc = np.float32(acc + np.float32(t * t))
        v = np.float32(acc + dqc)
        o = np.lexsort((np.arange(kc), v))
        vs = v[o]
        d0 = np.float32(vs[0] + vs[1])
        ins(d0, 0, 1, o[0], o[1]); ins(d0, 1, 0, o[1], o[0])
        while len(cells) < w and total < T:
            if not R.scref_heap4_pop(hv, *hid, out, C.byref(pos)):
                break
            h1, h2, h3, h4 = (int(z) for z in out)
            bid = h3 * kc + h4
            l = length(bid) if bid > 0 else 0
            if l > 0:
                cells.append(bid); total += l
            trav[h1, h2] = True
            if total < T:
                if h1 < kc - 1 and (h2 == 0 or trav[h1 + 1, h2 - 1]):
                    ins(np.float32(vs[h1 + 1] + vs[h2]), h1 + 1, h2, o[h1 + 1], o[h2])
                if h2 < kc - 1 and (h1 == 0 or trav[h1 - 1, h2 + 1]):
                    ins(np.float32(vs[h1] + vs[h2 + 1]), h1, h2 + 1, o[h1], o[h2 + 1])
    # the popped set is a down-set of the rank grid
    for a in range(kc):
        row = trav[a]
        if row.any():
            assert row[:np.nonzero(row)[0].max() + 1].all() or (kind == sco.SC and a == 0)
    return cells, total


@ref_missing
@pytest.mark.ref
@pytest.mark.parametrize("kind", [sco.IMI, sco.SC])
def test_traversal_replay_with_reference_heap(kind):
    ix, x, _, _ = _make_index(kind, n=4000, kc=10)
    q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 12, ix.dim)
    for qi in range(12):
        for w, T in ((6, 10 ** 9), (40, 500), (100, 2500), (200, 10 ** 9)):
            ids, dists, st, ex = ix.search(kind, q[qi], R=20, w=w, T=T, trace=True)
            cells, total = _replay_traversal(ix, kind, q[qi], w, T)
            assert list(ex["cells"]) == cells
            assert st.sum == total


def _replay_traversal3(ix, q, w, T):
    """Direct replay of search_mr_ivf3's multi-sequence step
    (sc_query.h:296-402) over rank triples, driven by the reference's
    compiled pq_insert6/pq_pop6; returns the non-empty cells in order, the
    candidate total and the number of triples popped."""
    import ctypes as C
    R = sco.ref()
    dqc, dqr = sco.lut(q, ix.cq, ix.pq, ix.norm_c, ix.norm_r, ix.dim, ix.mc, ix.kc, ix.mp, ix.kp)
    kc = ix.kc
    Lc = ix.L
    length = lambda b: int(Lc[b] - Lc[b - 1]) if b > 0 else 0
    hv = np.zeros(w, np.float32)
    hid = [np.zeros(w, np.int32) for _ in range(6)]
    pos = C.c_int32(0)
    acc = np.float32(0)
    for t in q:
        acc = np.float32(acc + np.float32(t * t))
    v = np.float32(acc + dqc)
    o = np.lexsort((np.arange(kc), v))
    vs = v[o]
    ins = lambda s, a, b, c: R.scref_heap6_insert(hv, *hid, s, a, b, c, int(o[a]), int(o[b]),
                                                  int(o[c]), C.byref(pos), w)
    out = np.zeros(6, np.int32)
    trav = np.zeros((kc, kc, kc), bool)
    ins(np.float32(np.float32(3.0) * vs[0]), 0, 0, 0)
    cells, total, visited = [], 0, 0
    while len(cells) < w and total < T and visited < kc ** 3:
        if not R.scref_heap6_pop(hv, *hid, out, C.byref(pos)):
            break
        h1, h2, h3, h4, h5, h6 = (int(z) for z in out)
        bid = (h4 * kc + h5) * kc + h6
        if length(bid) > 0:
            cells.append(bid)
            total += length(bid)
        trav[h1, h2, h3] = True
        visited += 1
        if total < T:
            if h1 < kc - 1 and (h2 == 0 or trav[h1 + 1, h2 - 1, h3]) and \
                    (h3 == 0 or trav[h1 + 1, h2, h3 - 1]):
                ins(np.float32(np.float32(vs[h1 + 1] + vs[h2]) + vs[h3]), h1 + 1, h2, h3)
            if h2 < kc - 1 and (h3 == 0 or trav[h1, h2 + 1, h3 - 1]) and \
                    (h1 == 0 or trav[h1 - 1, h2 + 1, h3]):
                ins(np.float32(np.float32(vs[h1] + vs[h2 + 1]) + vs[h3]), h1, h2 + 1, h3)
            if h3 < kc - 1 and (h2 == 0 or trav[h1, h2 - 1, h3 + 1]) and \
                    (h1 == 0 or trav[h1 - 1, h2, h3 + 1]):
                ins(np.float32(np.float32(vs[h1] + vs[h2]) + vs[h3 + 1]), h1, h2, h3 + 1)
    # the popped set is a down-set of the rank grid (closed under lowering
    # any coordinate), which is what lets a depth map stand for it
    a, b, c = np.nonzero(trav)
    for x, y, z in zip(a, b, c):
        assert trav[:x + 1, :y + 1, :z + 1].all()
    return cells, total, visited


@ref_missing
@pytest.mark.ref
def test_traversal3_replay_with_reference_heap():
    """The three-candidate traversal of the oracle equals the replay driven by
    the reference's own heap, including a grid walked to exhaustion (kc=4:
    every one of the 64 triples popped) and heap drops (w small)."""
    for kc, n in ((8, 4000), (4, 1500)):
        ix, x, _, _ = _make_index(sco.SC3, n=n, kc=kc)
        q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 8, ix.dim)
        for qi in range(8):
            for w, T in ((3, 10 ** 9), (40, 300), (200, 1200), (5000, 10 ** 9)):
                ids, dists, st, ex = ix.search(sco.SC3, q[qi], R=20, w=w, T=T, trace=True)
                cells, total, visited = _replay_traversal3(ix, q[qi], w, T)
                assert list(ex["cells"]) == cells
                assert st.sum == total and st.nvisited == visited
                if kc == 4 and w == 5000 and T == 10 ** 9:
                    assert visited == 64 and total == n


def test_encode_general_bruteforce():
    """SCEncoder::encode (general nc): the nc nearest by float-rounded S1
    distance, ties to the lower index; residual against the nearest; PQ
    argmin of float-rounded distances, first index on ties."""
    rng = np.random.default_rng(31)
    for mc, nc in ((1, 3), (2, 3), (1, 1), (2, 2)):
        kc, mp, kp, dim, n = 12, 4, 16, 32, 400
        x = sco.gen(sco.mix_d128(ncomp=5), sco.STREAM_TRAIN, 0, n, dim)
        x[::3] += rng.uniform(-0.5, 0.5, x[::3].shape).astype(np.float32)
        cq, pq = _codebooks(rng, x, mc, kc, mp, kp)      # duplicated centroids: exact ties
        cid, codes = sco.encode_general(x, cq, pq, mc, kc, mp, kp, nc)
        bsc, bsp = dim // mc, dim // mp
        cqb = cq.reshape(mc, kc, bsc)
        pqb = pq.reshape(mp, kp, bsp)
        for i in range(n):
            r = np.empty(dim, np.float32)
            for j in range(mc):
                xb = x[i, j * bsc:(j + 1) * bsc]
                d = _s1_sequential(xb, cqb[j])
                top = np.lexsort((np.arange(kc), d))[:nc]
                assert list(cid[i, j * nc:(j + 1) * nc]) == list(top)
                r[j * bsc:(j + 1) * bsc] = xb - cqb[j, top[0]]
            for s_ in range(mp):
                d = _s1_sequential(r[s_ * bsp:(s_ + 1) * bsp], pqb[s_])
                assert codes[i, s_] == int(np.flatnonzero(d == d.min())[0])
        assert (cid.reshape(n, mc, nc)[:, :, 0] != cid.reshape(n, mc, nc)[:, :, -1]).all() or nc == 1


def test_search_topr_is_canonical_selection():
    for kind in (sco.IVFADC, sco.IMI, sco.SC, sco.SC3):
        ix, x, _, _ = _make_index(kind)
        q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 5, ix.dim)
        for qi in range(5):
            ids, dists, st, ex = ix.search(kind, q[qi], R=30, w=min(ix.kc, 20), T=800, trace=True)
            w = min(ix.kc, 20)   # IVF/IMI bound popped cells, SC non-empty cells
            assert st.sum >= 800 or w in (st.ncells, st.nvisited) or st.exhausted
            sel = np.lexsort((np.arange(ex["all_dists"].size), ex["all_dists"]))[:30]
            assert np.array_equal(ids, ex["all_ids"][sel])
            assert np.array_equal(dists, ex["all_dists"][sel])


def _s1_sequential(q, y):
    """S1 restated independently: float differences, squared in float64 and
    summed in dimension order (cumsum is sequential), stored as float32."""
    t = (q.astype(np.float32)[None, :] - y.astype(np.float32)).astype(np.float64)
    return np.cumsum(t * t, axis=1)[:, -1].astype(np.float32)


def test_search_real_distances():
    """real_dist = true: the traced search's candidates, in traversal order,
    ranked by their exact distance to the raw vectors; R above the candidate
    count keeps traversal order; a query equal to a gathered vector finds it
    at distance 0 first."""
    rng = np.random.default_rng(5)
    for kind in (sco.IVFADC, sco.IMI, sco.SC):
        ix, x, _, _ = _make_index(kind)
        raw = x + rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)   # non-integer rows
        q = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_QUERY, 0, 4, ix.dim)
        for qi in range(4):
            for R, T in ((30, 800), (4000, 300)):
                w = min(ix.kc, 20)
                ids, dists, st = ix.search_real(kind, q[qi], raw, R, w, T)
                _, _, st2, ex = ix.search(kind, q[qi], R=R, w=w, T=T, trace=True)
                cand = ex["all_ids"]
                assert st.sum == st2.sum == cand.size
                d = _s1_sequential(q[qi], raw[cand])
                if cand.size >= R:
                    sel = np.lexsort((np.arange(cand.size), d))[:R]
                else:
                    sel = np.arange(cand.size)
                m = min(R, cand.size)
                assert np.array_equal(ids[:m], cand[sel]) and np.array_equal(dists[:m], d[sel])
                assert (ids[m:] == -1).all() and np.isinf(dists[m:]).all()
        # a gathered vector used as the query is found first, at distance 0
        _, _, _, ex = ix.search(kind, q[0], R=30, w=min(ix.kc, 20), T=800, trace=True)
        v = int(ex["all_ids"][7])
        ids, dists, _ = ix.search_real(kind, raw[v], raw, 30, min(ix.kc, 20), 800)
        assert ids[0] == v and dists[0] == 0.0


def test_bruteforce_nn1_and_recall():
    """Ground truth: the first strict minimum of the float64 S1 distance in id
    order (duplicated rows resolve to the lower id; a query equal to a row
    has distance 0); recall@R counts queries whose ground truth is among
    their first R results."""
    rng = np.random.default_rng(21)
    base = sco.gen(sco.mix_d128(ncomp=6), sco.STREAM_BASE, 0, 900, 48)
    base += rng.uniform(-0.5, 0.5, base.shape).astype(np.float32)
    base[700] = base[123]                      # exact duplicate, later id
    base[800] = base[50]
    q = sco.gen(sco.mix_d128(ncomp=6), sco.STREAM_QUERY, 0, 40, 48)
    q[3] = base[700]                           # ties at distance 0: 123 wins
    q[4] = base[50]
    gt = sco.bruteforce_nn1(base, q)
    for i in range(q.shape[0]):
        t = (q[i][None, :] - base).astype(np.float64)
        d = np.cumsum(t * t, axis=1)[:, -1]
        assert gt[i] == int(np.flatnonzero(d == d.min())[0])
    assert gt[3] == 123 and gt[4] == 50
    ids = np.tile(np.arange(10, dtype=np.int32), (40, 1))
    ids[:, 2] = gt
    ids[::2, 2] = -5                           # half the queries miss
    assert sco.recall(gt, ids, 3) == 0.5
    assert sco.recall(gt, ids, 2) == float(np.mean(gt < 2))


def test_exhaustive_ivf_is_full_adc_scan():
    """Config-1 mapping: kc=1 coarse cell, w=1, T >= N scans every vector."""
    rng = np.random.default_rng(8)
    n, dim, mp, kp = 2000, 32, 4, 16
    x = sco.gen(sco.mix_d128(ncomp=4), sco.STREAM_BASE, 0, n, dim)
    cq = np.zeros((1, 1, dim), np.float32)
    pq = x[rng.choice(n, kp, replace=False)].reshape(kp, mp, dim // mp).transpose(1, 0, 2).copy()
    cid, codes = sco.encode(x, cq, pq, 1, 1, mp, kp, 1)
    L, pid, cs = sco.build_dense(sco.cell_keys(cid, 1, 1, 1), 1, codes)
    ix = sco.OracleIndex(dim, 1, 1, mp, kp, cq, pq, L, pid, cs)
    q = sco.gen(sco.mix_d128(ncomp=4), sco.STREAM_QUERY, 0, 1, dim)[0]
    ids, dists, st, ex = ix.search(sco.IVFADC, q, R=50, w=1, T=n, trace=True)
    assert st.sum == n and np.array_equal(np.sort(ex["all_ids"]), np.arange(n))
    recon = np.concatenate([pq[s][codes[:, s]] for s in range(mp)], axis=1).astype(np.float64)
    exact = ((q.astype(np.float64) - recon) ** 2).sum(1)
    assert np.max(np.abs(ex["all_dists"] - exact[ex["all_ids"]])) <= 1e-5 * exact.max()


def test_error_paths():
    ix, x, _, _ = _make_index(sco.IVFADC)
    q = x[0]
    with pytest.raises(ValueError):
        ix.search(sco.IVFADC, q, R=10, w=ix.kc + 1, T=10)      # pq_pop on empty heap
    with pytest.raises(ValueError):
        ix.search(sco.IMI, q, R=10, w=4, T=10)                 # mc != 2
    ix2, x2, _, _ = _make_index(sco.SC, n=200, kc=4)
    ids, d, st, _ = ix2.search(sco.SC, x2[0], R=500, w=1000, T=10 ** 9)
    assert st.exhausted == 1 and st.sum == 200                 # every cell visited
    assert np.all(ids[200:] == -1)


# ------------------------------------------------------------------ k-means
def test_kmeans_pp_and_lloyd_properties():
    m = sco.mix_d128(ncomp=6)
    x = sco.gen(m, sco.STREAM_TRAIN, 0, 3000, 16)
    k = 24
    idx = sco.kmeans_pp(x, k, seed=99)
    assert len(set(map(tuple, x[idx]))) == k                    # D^2 > 0 for new seeds
    assert np.array_equal(idx, sco.kmeans_pp(x, k, seed=99))
    assert not np.array_equal(idx, sco.kmeans_pp(x, k, seed=98))
    assert not np.array_equal(idx, sco.kmeans_pp(x, k, seed=99, lagged=1))
    c0, _ = sco.kmeans(x, k, 0, seed=99)
    assert np.array_equal(c0, x[idx])
    prev = None
    for it in (1, 2, 5, 12):
        c, lab = sco.kmeans(x, k, it, seed=99)
        cprev, _ = sco.kmeans(x, k, it - 1, seed=99)
        d = np.stack([_s1_closed(x, cprev[j][None, :]) for j in range(k)], 1).astype(np.float32)
        want = k - 1 - np.argmin(d[:, ::-1], axis=1)           # ties -> later centroid
        assert np.array_equal(lab, want)
        for j in range(k):
            mem = x[lab == j].astype(np.float64)
            assert mem.shape[0] > 0
            assert np.array_equal(c[j], (mem.sum(0) / mem.shape[0]).astype(np.float32))
        obj = float(np.sum(_s1_closed(x, c[lab])))
        if prev is not None:
            assert obj <= prev + 1e-6 * prev
        prev = obj


def _kpp_replay(x, k, seed, lagged):
    """k-means++ seeding restated with numpy (the Philox stream checked by the
    Random123 vectors above): first seed u64(round 0) mod n; then weights
    w_i = floor(D2_i * 2^16), pivot p = u64(round c) mod W, pick = the first i
    whose inclusive prefix sum of w exceeds p (W = 0: u64 mod n); D2 folds in
    the newest seed, or the previous one when lagged (k-means.h:244-250)."""
    def u64(rnd, which):
        o = sco.philox([rnd, 0, 0xA5, which], [seed, 1])
        return (int(o[1]) << 32) | int(o[0])
    n = x.shape[0]
    idx = [u64(0, 0) % n]
    d2 = _s1_closed(x, x[idx[0]][None, :])
    for c in range(1, k):
        w = np.ldexp(d2, 16).astype(np.uint64)
        W = int(w.sum(dtype=np.uint64))
        r = u64(c, 1)
        if W == 0:
            pick = r % n
        else:
            pick = int(np.searchsorted(np.cumsum(w, dtype=np.uint64), np.uint64(r % W),
                                       side="right"))
        idx.append(pick)
        d2 = np.minimum(d2, _s1_closed(x, x[idx[c - 1] if lagged else idx[c]][None, :]))
    return np.array(idx, np.int64)


@pytest.mark.parametrize("lagged", [0, 1])
def test_kmeans_pp_sampler_exact_replay(lagged):
    """The oracle's seeds equal an independent numpy replay of the defined
    sampler, pick for pick, incl. duplicated points (D2 = 0 weights)."""
    m = sco.mix_d128(ncomp=5)
    x = sco.gen(m, sco.STREAM_TRAIN, 0, 700, 8)
    x[100:140] = x[5]                     # duplicates: zero weight once x[5] is a seed
    for seed in (1, 7, 20151001):
        for k in (2, 9, 40):
            want = _kpp_replay(x, k, seed, lagged)
            assert np.array_equal(sco.kmeans_pp(x, k, seed=seed, lagged=lagged), want)
    # all points identical: W = 0 at every round, picks fall back to u64 mod n
    z = np.full((13, 4), 3.0, np.float32)
    assert np.array_equal(sco.kmeans_pp(z, 6, seed=3, lagged=lagged), _kpp_replay(z, 6, 3, lagged))


def test_kmeans_pp_second_pick_is_d2_weighted():
    """Distribution of the second seed over many seeds: P(j | first f) must be
    D2(j, f) / sum D2 (k-means.h:216-230), not D / sum D: on the points
    {0, 1, 3, 7} the two rules differ by up to 0.19 in probability."""
    x = np.array([[0.0], [1.0], [3.0], [7.0]], np.float32)
    counts = np.zeros((4, 4))
    idx = np.zeros(2, np.int64)
    for seed in range(4000):
        assert sco.lib().sco_kmeans_pp(x, 4, 1, 2, seed + 11, 0, idx, 1) == 0   # one thread
        counts[idx[0], idx[1]] += 1
    for f in range(4):
        tot = counts[f].sum()
        assert tot > 600
        d2 = (x[:, 0].astype(np.float64) - x[f, 0]) ** 2
        p = d2 / d2.sum()
        assert counts[f, f] == 0
        # 5 sigma of a binomial count around the D^2 expectation
        sig = np.sqrt(tot * p * (1 - p)) + 1
        assert np.all(np.abs(counts[f] - tot * p) <= 5 * sig), (f, counts[f], tot * p)
        pd = np.abs(x[:, 0].astype(np.float64) - x[f, 0])
        pd = pd / pd.sum()
        assert np.max(np.abs(counts[f] / tot - pd)) > 0.05   # D weighting would be rejected


def test_residual_second_nearest():
    rng = np.random.default_rng(9)
    x = sco.gen(sco.mix_d128(ncomp=3), sco.STREAM_TRAIN, 0, 200, 16)
    cq = x[rng.choice(200, 10, replace=False)].copy()
    r0 = sco.residual(x, cq, 1, 10, 0)
    r1 = sco.residual(x, cq, 1, 10, 1)
    d = np.stack([_s1_closed(x, cq[j][None, :]).astype(np.float32) for j in range(10)], 1)
    o = np.lexsort((np.tile(np.arange(10), (200, 1)), d), axis=1)
    assert np.array_equal(r0, x - cq[o[:, 0]])
    assert np.array_equal(r1, x - cq[o[:, 1]])


@ref_missing
@pytest.mark.ref
@pytest.mark.parametrize("mc,part,second", [(1, 8, 1), (2, 4, 1), (1, 1, 1), (2, 8, 0),
                                            (1, 8, 0)])
def test_residual_matches_reference(tmp_path, mc, part, second):
    """PQQuantizer / SCQuantizer::calc_residual_vector as the reference runs
    them.  PQQuantizer walks the codebook's own mc blocks (which=0 with mc);
    SCQuantizer walks the quantizer's `part` blocks of dim/part over the same
    floats, reading the codebook as [part][kc][dim/part]
    (include/sc_quantizer.h:96-124), and subtracts the second nearest of each
    such block: which=1 with mc = part.  Exact ties are planted (duplicated
    centroids, vectors equal to a centroid)."""
    rng = np.random.default_rng(mc * 100 + part * 10 + second)
    n, dim, kc = 700, 64, 12
    x = sco.gen(sco.mix_d128(ncomp=4), sco.STREAM_TRAIN, 0, n, dim)
    x[n // 2:] += rng.uniform(-0.5, 0.5, (n - n // 2, dim)).astype(np.float32)
    bsc = dim // mc
    cq = np.stack([x[rng.choice(n, kc, replace=False), j * bsc:(j + 1) * bsc]
                   for j in range(mc)])                       # [mc][kc][bsc]
    cq[:, kc - 1] = cq[:, 2]
    flat = cq.reshape(-1)
    x[:6] = flat[:dim]                                        # ties in the part walk too
    path = str(tmp_path / "cq.ctr_")
    sco.write_codebook(path, cq, kc, mc, dim)
    want = x.copy()
    sco.ref().scref_residual(want, n, dim, part, path.encode(), second)
    got = sco.residual(x, flat, part if second else mc, kc, second)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    if second and part != mc:   # the codebook's own blocks give other residuals
        assert not np.array_equal(sco.residual(x, flat, mc, kc, 1), want)
