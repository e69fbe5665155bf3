"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer and index outputs (codes, cell ids, L/pid, neighbour ids) must match
bit for bit; FP32 outputs (tables, ADC distances) are computed with the same
operation sequence as the oracle and must match bit for bit as well.
"""
import os

import numpy as np
import pytest

import sc_amd as S
import sco

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# ------------------------------------------------------------------ generator
def test_generator_matches_oracle(ctx):
    for mix, omix, dim, row0, n in ((S.mix_d128(), sco.mix_d128(), 128, 0, 5000),
                                    (S.mix_d128(), sco.mix_d128(), 128, 123456789, 700),
                                    (S.mix_d960(), sco.mix_d960(), 960, 5, 200)):
        g = S.gen_rows(ctx, mix, S.STREAM_BASE, row0, n, dim).numpy()
        o = sco.gen(omix, sco.STREAM_BASE, row0, n, dim)
        assert np.array_equal(_bits(g), _bits(o))


# ------------------------------------------------------------------ encode
def _codebooks(rng, x, mc, kc, mp, kp, ties=True):
    n, dim = x.shape
    bsc, bsp = dim // mc, dim // mp
    cq = np.empty((mc, kc, bsc), np.float32)
    for j in range(mc):
        cq[j] = x[rng.choice(n, kc, replace=False), j * bsc:(j + 1) * bsc]
        cq[j] += rng.uniform(-0.5, 0.5, (kc, bsc)).astype(np.float32)
        if ties:
            cq[j, kc - 1] = cq[j, 1]
            cq[j, kc - 2] = cq[j, 0]
    pq = (rng.standard_normal((mp, kp, bsp)) * 8).astype(np.float32)
    if ties:
        pq[:, kp - 1] = pq[:, 3]
    return cq, pq


ENC_CASES = [  # dim, mc, kc, mp, kp, nc, n
    (128, 1, 256, 8, 256, 2, 3001),
    (128, 2, 64, 8, 256, 1, 2500),
    (128, 2, 32, 8, 256, 2, 2500),
    (128, 1, 1, 8, 256, 1, 777),
    (960, 2, 32, 16, 256, 1, 600),
    (64, 1, 300, 4, 100, 2, 1000),
]


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16])
@pytest.mark.parametrize("case", ENC_CASES)
def test_encode_matches_oracle(ctx, case, prec):
    dim, mc, kc, mp, kp, nc, n = case
    bf16_shape_ok = (dim // mc) % 16 == 0 and (dim // mp) % 16 == 0
    rng = np.random.default_rng(sum(case))
    mix = sco.mix_d960() if dim == 960 else sco.mix_d128(ncomp=7)
    x = sco.gen(mix, sco.STREAM_TRAIN, 0, n, dim)
    if dim != 960:
        x[n // 2:] += rng.uniform(-0.5, 0.5, (n - n // 2, dim)).astype(np.float32)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp, ties=kc > 2)
    bsc = dim // mc
    if kc > 1:
        x[:5, :bsc] = cq[0, 1]                    # exact ties on the coarse distance
    args = (ctx, S.Dev.from_numpy(ctx, x), S.Dev.from_numpy(ctx, cq),
            S.Dev.from_numpy(ctx, pq), mc, kc, mp, kp, nc, prec)
    if prec == S.PREC_BF16 and not bf16_shape_ok:
        # the BF16 form needs block lengths that are multiples of 16: refused
        # before any launch, with the reason
        with pytest.raises(S.ScError, match="multiples of 16"):
            S.encode(*args)
        return
    cid, codes, st = S.encode(*args)
    ocid, ocodes = sco.encode(x, cq, pq, mc, kc, mp, kp, nc)
    assert np.array_equal(cid.numpy(), ocid)
    assert np.array_equal(codes.numpy(), ocodes)
    if prec == S.PREC_BF16:
        print("bf16 encode: %d candidates re-checked, %d rows fully exact (n=%d)"
              % (st.rechecked, st.fallback, n))


GIVEN_CASES = [  # dim, mc, kc, mp, kp, n (raw rows), m (positions)
    (128, 1, 256, 8, 256, 3001, 4000),
    (128, 2, 16, 8, 256, 2000, 2000),
    (960, 2, 32, 16, 256, 300, 500),
    (64, 1, 300, 4, 100, 700, 1),
]


@pytest.mark.parametrize("case", GIVEN_CASES)
def test_encode_given_matches_oracle(ctx, case):
    """Encoder::rencode: codes of raw rows (permuted, repeated) against given
    cells, bit-exact; rows=None codes every row in order; a row or cell out
    of range is refused with the reason."""
    dim, mc, kc, mp, kp, n, m = case
    rng = np.random.default_rng(sum(case) + 5)
    mix = sco.mix_d960() if dim == 960 else sco.mix_d128(ncomp=7)
    x = sco.gen(mix, sco.STREAM_TRAIN, 0, n, dim)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp, ties=kc > 2)
    rows = rng.integers(0, n, m).astype(np.int32)
    cid = rng.integers(0, kc, (m, mc)).astype(np.uint16)
    dx, dcq, dpq = (S.Dev.from_numpy(ctx, a) for a in (x, cq, pq))
    got = S.encode_given(ctx, dx, S.Dev.from_numpy(ctx, cid), dcq, dpq, mc, kc, mp, kp,
                         rows=S.Dev.from_numpy(ctx, rows)).numpy()
    assert np.array_equal(got, sco.encode_given(x, rows, cid, cq, pq, mc, kc, mp, kp))
    own, codes, _ = S.encode(ctx, dx, dcq, dpq, mc, kc, mp, kp, 1)
    assert np.array_equal(S.encode_given(ctx, dx, own, dcq, dpq, mc, kc, mp, kp).numpy(),
                          codes.numpy())
    bad_rows = rows.copy()
    bad_rows[-1] = n
    with pytest.raises(S.ScError, match="row index"):
        S.encode_given(ctx, dx, S.Dev.from_numpy(ctx, cid), dcq, dpq, mc, kc, mp, kp,
                       rows=S.Dev.from_numpy(ctx, bad_rows))
    bad_cid = cid.copy()
    bad_cid[0, mc - 1] = kc
    with pytest.raises(S.ScError, match="cell id"):
        S.encode_given(ctx, dx, S.Dev.from_numpy(ctx, bad_cid), dcq, dpq, mc, kc, mp, kp,
                       rows=S.Dev.from_numpy(ctx, rows))


# ------------------------------------------------------------------ index + search
OKIND = {S.KIND_IVFADC: sco.IVFADC, S.KIND_IMI: sco.IMI, S.KIND_SC: sco.SC,
         S.KIND_SC3: sco.SC3, S.KIND_SCIMI: sco.SCIMI}


def _index_pair(ctx, kind, n=6000, dim=64, kc=16, mp=8, kp=64, seed=3):
    rng = np.random.default_rng(seed)
    mc = 2 if kind in (S.KIND_IMI, S.KIND_SCIMI) else 1
    nc = {S.KIND_SC: 2, S.KIND_SC3: 3, S.KIND_SCIMI: 2}.get(kind, 1)
    x = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_BASE, 0, n, dim)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp, ties=False)
    dx, dcq, dpq = (S.Dev.from_numpy(ctx, a) for a in (x, cq, pq))
    if kind == S.KIND_SC3:      # SCEncoder::encode, general nc
        cid, codes = S.encode_general(ctx, dx, dcq, dpq, mc, kc, mp, kp, nc)
        ocid, ocodes = sco.encode_general(x, cq, pq, mc, kc, mp, kp, nc)
    else:
        cid, codes, _ = S.encode(ctx, dx, dcq, dpq, mc, kc, mp, kp, nc)
        ocid, ocodes = sco.encode(x, cq, pq, mc, kc, mp, kp, nc)
    assert np.array_equal(cid.numpy(), ocid) and np.array_equal(codes.numpy(), ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, mc, kc, dpq, mp, kp, nc)
    if kind == S.KIND_SCIMI:
        ck, L, pid, cs = sco.build_sparse(sco.cell_keys(ocid, mc, kc, nc), ocodes)
        return ix, sco.OracleIndex(dim, mc, kc, mp, kp, cq, pq, L, pid, cs, ckeys=ck)
    ncells = kc ** (mc * nc)
    L, pid, cs = sco.build_dense(sco.cell_keys(ocid, mc, kc, nc), ncells, ocodes)
    oix = sco.OracleIndex(dim, mc, kc, mp, kp, cq, pq, L, pid, cs)
    return ix, oix


@pytest.mark.parametrize("kind", [S.KIND_IVFADC, S.KIND_IMI, S.KIND_SC, S.KIND_SC3,
                                  S.KIND_SCIMI])
def test_search_real_distances_match_oracle(ctx, kind):
    """SC_FORM_REAL (real_dist = true): candidates ranked by the exact S1
    distance to the raw vectors, bit-exact ids and distances against the
    oracle, including queries with fewer candidates than R."""
    rng = np.random.default_rng(11)
    ix, oix = _index_pair(ctx, kind)
    n = oix.pid.size
    raw = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_BASE, 0, n, ix.dim)
    raw += rng.uniform(-0.5, 0.5, raw.shape).astype(np.float32)
    draw = S.Dev.from_numpy(ctx, raw)
    q = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, 0, 24, ix.dim)
    dq = S.Dev.from_numpy(ctx, q)
    with pytest.raises(S.ScError):
        ix.search(dq, kind, 50, 8, 800, form=S.FORM_REAL)   # no raw vectors yet
    ix.attach_raw(draw)
    for R, w, T in ((50, 8, 800), (400, 3, 150)):
        ids, dists, sums, _ = ix.search(dq, kind, R, w, T, form=S.FORM_REAL)
        gi, gd, gs = ids.numpy(), dists.numpy(), sums.numpy()
        for i in range(q.shape[0]):
            oi, od, ost = oix.search_real(kind, q[i], raw, R, w, T)
            assert gs[i] == ost.sum
            assert np.array_equal(gi[i], oi), (R, i)
            assert np.array_equal(gd[i].view(np.uint32), od.view(np.uint32)), (R, i)


@pytest.mark.parametrize("dim,n,nq", [(128, 5000, 300), (128, 70, 1), (960, 1500, 37),
                                      (48, 20000, 129)])
def test_groundtruth_nn1_matches_oracle(ctx, dim, n, nq):
    """sc_groundtruth_nn1 against the oracle's brute force: ids bit-exact
    (duplicated rows resolve to the lower id), distances equal to the FP64
    S1 value; a query equal to a base row finds it at distance 0."""
    rng = np.random.default_rng(dim + n)
    mix = sco.mix_d960(ncomp=8) if dim == 960 else sco.mix_d128(ncomp=8)
    base = sco.gen(mix, sco.STREAM_BASE, 0, n, dim)
    base += rng.uniform(-0.5, 0.5, base.shape).astype(np.float32)
    q = sco.gen(mix, sco.STREAM_QUERY, 0, nq, dim)
    if n > 100:
        base[n - 1] = base[17]
        base[n // 2] = base[n // 3]
        q[0] = base[n - 1]                     # equal distance 0 at ids 17 and n-1
        if nq > 1:
            q[nq - 1] = base[n // 2]
    ids, dists = S.groundtruth_nn1(ctx, S.Dev.from_numpy(ctx, base), S.Dev.from_numpy(ctx, q))
    gi, gd = ids.numpy(), dists.numpy()
    assert np.array_equal(gi, sco.bruteforce_nn1(base, q))
    for i in range(nq):
        assert gd[i] == sco.dist(q[i], base[gi[i]])
    if n > 100:
        assert gi[0] == 17 and gd[0] == 0.0
        if nq > 1:
            assert gi[nq - 1] == n // 3


@pytest.mark.parametrize("kind", [S.KIND_IVFADC, S.KIND_IMI, S.KIND_SC])
def test_index_build_and_tables_match_oracle(ctx, kind):
    ix, oix = _index_pair(ctx, kind)
    e = ix.export()
    assert np.array_equal(e["L"], oix.L)
    assert np.array_equal(e["pid"], oix.pid)
    assert np.array_equal(e["codes"], oix.codes)
    assert np.array_equal(_bits(e["norm_c"]), _bits(oix.norm_c))
    assert np.array_equal(_bits(e["norm_r"]), _bits(oix.norm_r))
    assert np.array_equal(_bits(e["dot_cr"]), _bits(oix.dot_cr))
    ix.close()


SEARCH_CASES = [  # kind, R, w, T
    (S.KIND_IVFADC, 100, 16, 3000),
    (S.KIND_IVFADC, 10, 3, 100000),
    (S.KIND_IMI, 100, 200, 2000),
    (S.KIND_IMI, 50, 5, 100000),         # heap capacity 5: inserts dropped
    (S.KIND_SC, 100, 500, 3000),
    (S.KIND_SC, 100, 4, 100000),         # cell budget binds
    (S.KIND_SC, 300, 500, 40),           # fewer candidates than R
    (S.KIND_SC, 7, 500, 0),              # T = 0: the loop test fails first, no cell
]


def _compare_search(ctx, ix, oix, kind, R, w, T, nq=150, form=S.FORM_PLAIN):
    q = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, 0, nq, ix.dim)
    ids, dists, sums, st = ix.search(S.Dev.from_numpy(ctx, q), kind, R, w, T, form=form)
    oids, od, osum = oix.search_batch(OKIND[kind], q, R, w, T)
    assert np.array_equal(sums.numpy(), osum)
    assert np.array_equal(ids.numpy(), oids)
    assert np.array_equal(_bits(dists.numpy()), _bits(od))
    assert st.queries == nq and st.candidates == int(osum.sum())
    return st


@pytest.mark.parametrize("case", SEARCH_CASES)
def test_search_matches_oracle(ctx, case):
    kind, R, w, T = case
    ix, oix = _index_pair(ctx, kind)
    st = _compare_search(ctx, ix, oix, kind, R, w, T)
    if T == 0:
        # include/sc_query.h:130 tests `sum < T` before every pop: with T = 0
        # nothing is popped, no candidate is gathered and every result slot
        # is empty
        assert st.candidates == 0 and st.cells == 0 and st.visited == 0
    ix.close()


# ------------------------------------------------------------------ two halves (SC_KIND_SCIMI)
SCIMI_CASES = [  # kc, R, w, T
    (8, 100, 100000, 3000),
    (8, 100, 5, 100000),         # the cell budget binds
    (8, 300, 500, 40),           # fewer candidates than R
    (8, 7, 500, 0),              # T = 0: no cell
    (4, 50, 10 ** 6, 10 ** 6),   # every cell of the index taken (exhausted)
    (16, 100, 100000, 2500),
    (300, 100, 100000, 1500),    # kc > 256: 16-byte ranking records
]


@pytest.mark.parametrize("case", SCIMI_CASES[:3])
def test_search_scimi_record_ranking_matches_oracle(ctx, case, monkeypatch):
    """The ranking from the cells' records alone (no pair tables)."""
    monkeypatch.setenv("SC_RANK_PAIRS", "0")
    kc, R, w, T = case
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=kc)
    _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=100)
    ix.close()


@pytest.mark.parametrize("case", SCIMI_CASES)
def test_search_scimi_matches_oracle(ctx, case):
    """Two halves with (nearest, second) cells per half: the sparse
    directory built on the GPU equals the oracle's (keys, L, pid, codes), and
    ids, distances and candidate counts equal the oracle's search."""
    kc, R, w, T = case
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=kc)
    e = ix.export()
    assert np.array_equal(ix.cell_keys(), oix.ckeys)
    assert np.array_equal(e["L"], oix.L) and np.array_equal(e["pid"], oix.pid)
    assert np.array_equal(e["codes"], oix.codes)
    assert np.array_equal(_bits(e["dot_cr"]), _bits(oix.dot_cr))
    st = _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=100)
    _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=100, form=S.FORM_MIXED)
    if T == 0:
        assert st.candidates == 0 and st.cells == 0
    if kc == 4:
        assert st.exhausted == 100 and st.candidates == 100 * oix.pid.size
    with pytest.raises(S.ScError) as ex:
        ix.search(S.Dev.from_numpy(ctx, np.zeros((1, ix.dim), np.float32)), S.KIND_IMI, R, w, T)
    assert ex.value.rc == S.SC_ERR_CONFIG
    ix.close()


@pytest.mark.parametrize("big", ["auto", "0"])
@pytest.mark.parametrize("pairs", ["1", "0"])
@pytest.mark.parametrize("cap", ["16", "1"])
def test_search_scimi_small_rank_list(ctx, cap, pairs, big, monkeypatch):
    """An on-chip ranking list of 16 (or 1) cells: queries go through the
    bounded subset and the second histogram level, then the overflowing ones
    through the pruned ranking (or, without it, a longer list and the
    device-wide sorted fallback), with and without the pair tables; results
    must not change."""
    monkeypatch.setenv("SC_RANK_CAP", cap)
    monkeypatch.setenv("SC_RANK_PAIRS", pairs)
    if big == "0":
        monkeypatch.setenv("SC_RANK_BIG", "0")
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=8)
    for R, w, T in ((100, 100000, 3000), (50, 7, 10 ** 6), (20, 10 ** 6, 10 ** 6)):
        _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=60)
    ix.close()


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_search_scimi_rerank_past_capacity(ctx, form, monkeypatch):
    """An on-chip ranking list of one cell and 1200 queries in one batch:
    more queries outgrow it than the pruned re-ranking takes in its
    device-driven pass (512 workgroups that read the count), so the rest are
    ranked after the batch's one host synchronisation and scanned again
    exactly; ids, distances and counts must equal the oracle's."""
    monkeypatch.setenv("SC_RANK_CAP", "1")
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=8)
    f = S.FORM_MIXED if form == "mixed" else S.FORM_PLAIN
    st = _compare_search(ctx, ix, oix, S.KIND_SCIMI, 100, 100000, 3000, nq=1200, form=f)
    assert st.heap_fallback == 1200
    ix.close()


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_search_scimi_small_passes(ctx, form, monkeypatch):
    """1200 queries in passes of 256 with an on-chip ranking list of 16
    cells: every pass re-ranks its overflowing queries on the second stream
    while the others are scanned, and the passes reuse the same scratch and
    events; ids, distances and counts must equal the oracle's."""
    monkeypatch.setenv("SC_SEARCH_CHUNK", "256")
    monkeypatch.setenv("SC_RANK_CAP", "16")
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=8)
    f = S.FORM_MIXED if form == "mixed" else S.FORM_PLAIN
    for R, w, T in ((100, 100000, 3000), (20, 10 ** 6, 10 ** 6)):
        _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=1200, form=f)
    ix.close()


@pytest.mark.parametrize("parts", ["1", "3", "auto"])
@pytest.mark.parametrize("case", SCIMI_CASES)
def test_search_scimi_pruned_ranking_matches_oracle(ctx, case, parts, monkeypatch):
    """The pruned ranking of large indexes (centroids visited in order of
    their bound, a falling histogram cut, fine-bucket emission), forced onto
    the small indexes, with each query on one workgroup, on three, or on as
    many as the batch size gives: ids, distances and candidate counts equal
    the oracle's, in both forms."""
    monkeypatch.setenv("SC_RANK_BIG", "1")
    if parts != "auto":
        monkeypatch.setenv("SC_RANK_PARTS", parts)
    kc, R, w, T = case
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=kc)
    st = _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=100)
    _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=100, form=S.FORM_MIXED)
    # batches of a few queries (each query shared over several workgroups)
    for nq in (1, 4):
        _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=nq)
    if T == 0:
        assert st.candidates == 0 and st.cells == 0
    if kc == 4:
        assert st.exhausted == 100 and st.candidates == 100 * oix.pid.size
    ix.close()


@pytest.mark.parametrize("cap", ["64", "1"])
def test_search_scimi_pruned_ranking_overflow(ctx, cap, monkeypatch):
    """A pruned-ranking list of 64 (or 1) entries per query: queries are run
    again from their first run's cut, and those still over capacity go to
    the device-wide sort; results must not change.  Coarse codebooks with
    duplicated centroids give many equal cell scores."""
    monkeypatch.setenv("SC_RANK_BIG", "1")
    monkeypatch.setenv("SC_RANK_BIG_CAP", cap)
    rng = np.random.default_rng(9)
    n, dim, kc, mp, kp = 6000, 64, 16, 8, 64
    x = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_BASE, 0, n, dim)
    cq, pq = _codebooks(rng, x, 2, kc, mp, kp, ties=True)
    dx, dcq, dpq = (S.Dev.from_numpy(ctx, a) for a in (x, cq, pq))
    cid, codes, _ = S.encode(ctx, dx, dcq, dpq, 2, kc, mp, kp, 2)
    ocid, ocodes = sco.encode(x, cq, pq, 2, kc, mp, kp, 2)
    assert np.array_equal(cid.numpy(), ocid) and np.array_equal(codes.numpy(), ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    ck, L, pid, cs = sco.build_sparse(sco.cell_keys(ocid, 2, kc, 2), ocodes)
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, L, pid, cs, ckeys=ck)
    for R, w, T in ((100, 100000, 3000), (50, 7, 10 ** 6), (20, 10 ** 6, 10 ** 6), (100, 40, 200)):
        st = _compare_search(ctx, ix, oix, S.KIND_SCIMI, R, w, T, nq=60)
        if T == 3000:   # every query's first list outgrows the capacity
            assert st.heap_fallback == 60
    ix.close()


def test_search_scimi_load_sparse_and_refusals(ctx):
    """sc_index_load_sparse from the oracle's arrays gives the same search;
    a dense load of a two-halves index, unordered keys and an empty cell are
    refused."""
    ix, oix = _index_pair(ctx, S.KIND_SCIMI, kc=8)
    lx = S.Index.load_sparse(ctx, oix.ckeys, oix.L, oix.pid, oix.codes, oix.dim, oix.cq, oix.kc,
                             oix.pq, oix.mp, oix.kp)
    assert np.array_equal(lx.cell_keys(), oix.ckeys)
    _compare_search(ctx, lx, oix, S.KIND_SCIMI, 100, 100000, 2000, nq=40)
    bad = oix.ckeys.copy()
    bad[[1, 2]] = bad[[2, 1]]
    with pytest.raises(S.ScError, match="ascend"):
        S.Index.load_sparse(ctx, bad, oix.L, oix.pid, oix.codes, oix.dim, oix.cq, oix.kc, oix.pq,
                            oix.mp, oix.kp)
    L2 = oix.L.copy()
    L2[1] = L2[0]
    with pytest.raises(S.ScError, match="empty"):
        S.Index.load_sparse(ctx, oix.ckeys, L2, oix.pid, oix.codes, oix.dim, oix.cq, oix.kc,
                            oix.pq, oix.mp, oix.kp)
    with pytest.raises(S.ScError, match="load_sparse"):
        S.Index.load(ctx, oix.L, oix.pid, oix.codes, oix.dim, oix.cq, 2, oix.kc, oix.pq, oix.mp,
                     oix.kp, 2)
    lx.close()
    ix.close()


# ------------------------------------------------------------------ three candidates (nc = 3)
ENC_GENERAL_CASES = [  # dim, mc, kc, mp, kp, nc, n
    (128, 1, 256, 8, 256, 3, 3001),
    (128, 2, 32, 8, 256, 3, 1500),
    (64, 1, 300, 4, 100, 2, 1000),
    (960, 2, 32, 16, 256, 1, 500),
    (32, 1, 5, 4, 16, 3, 700),
]


@pytest.mark.parametrize("case", ENC_GENERAL_CASES)
def test_encode_general_matches_oracle(ctx, case):
    """SCEncoder::encode with general nc: the nc nearest by float-rounded
    distance (ties to the lower index) and the PQ codes, bit-exact."""
    dim, mc, kc, mp, kp, nc, n = case
    rng = np.random.default_rng(sum(case) + 1)
    mix = sco.mix_d960() if dim == 960 else sco.mix_d128(ncomp=7)
    x = sco.gen(mix, sco.STREAM_TRAIN, 0, n, dim)
    if dim != 960:
        x[n // 2:] += rng.uniform(-0.5, 0.5, (n - n // 2, dim)).astype(np.float32)
    cq, pq = _codebooks(rng, x, mc, kc, mp, kp, ties=kc > 2)
    x[:5, :dim // mc] = cq[0, 1]                  # exact ties on the coarse distance
    cid, codes = S.encode_general(ctx, S.Dev.from_numpy(ctx, x), S.Dev.from_numpy(ctx, cq),
                                  S.Dev.from_numpy(ctx, pq), mc, kc, mp, kp, nc)
    ocid, ocodes = sco.encode_general(x, cq, pq, mc, kc, mp, kp, nc)
    assert np.array_equal(cid.numpy(), ocid)
    assert np.array_equal(codes.numpy(), ocodes)


SC3_CASES = [  # kc, R, w, T
    (12, 100, 500, 3000),
    (12, 100, 4, 100000),        # cell budget (and heap capacity) binds: inserts dropped
    (12, 300, 500, 40),          # fewer candidates than R
    (12, 7, 500, 0),             # T = 0: no cell
    (4, 50, 5000, 100000),       # the whole 4^3 grid popped
    (32, 100, 2000, 2500),
]


@pytest.mark.parametrize("case", SC3_CASES)
def test_search_sc3_matches_oracle(ctx, case):
    """SCQuery::search_mr_ivf3: the (nearest, second, third) index built on
    the GPU equals the oracle's; ids, distances and candidate counts of the
    plain form equal the oracle's search; the mixed form is refused."""
    kc, R, w, T = case
    ix, oix = _index_pair(ctx, S.KIND_SC3, kc=kc)
    e = ix.export()
    assert np.array_equal(e["L"], oix.L) and np.array_equal(e["pid"], oix.pid)
    assert np.array_equal(e["codes"], oix.codes)
    st = _compare_search(ctx, ix, oix, S.KIND_SC3, R, w, T, nq=80)
    if kc == 4 and T == 100000:
        assert st.visited == 80 * 64
    # the three-candidate kind has no mixed form: refused, not run plain
    with pytest.raises(S.ScError, match="not available") as e:
        _compare_search(ctx, ix, oix, S.KIND_SC3, R, w, T, nq=80, form=S.FORM_MIXED)
    assert e.value.rc == S.SC_ERR_ARG
    with pytest.raises(S.ScError):
        ix.search(S.Dev.from_numpy(ctx, np.zeros((1, ix.dim), np.float32)), S.KIND_SC, R, w, T)
    ix.close()


@pytest.mark.parametrize("case", SEARCH_CASES)
@pytest.mark.parametrize("kc,kp", [(16, 64), (64, 64)])
def test_search_mixed_form_matches_oracle(ctx, case, kc, kp, monkeypatch):
    """SC_FORM_MIXED (FP16 table scan + exact re-rank) returns the plain form's
    ids and distances bit for bit; kc = kp = 64 takes the precomputed-LUT path.
    The pop budget is lifted so that long traversals stay in the fused kernel.
    The multi-index takes the staged mixed scan (per-vector dot_cr sums)."""
    monkeypatch.setenv("SC_POP_BUDGET", "1000000")
    kind, R, w, T = case
    ix, oix = _index_pair(ctx, kind, kc=kc, kp=kp)
    st = _compare_search(ctx, ix, oix, kind, R, w, T, form=S.FORM_MIXED)
    if kind != S.KIND_IMI:
        assert st.reranked > 0 or st.candidates == 0     # the exact re-rank ran
    ix.close()


@pytest.mark.parametrize("kind", [S.KIND_IMI, S.KIND_SCIMI])
@pytest.mark.parametrize("mp,kp", [(8, 64), (16, 32), (4, 64)])
def test_search_mixed_imi_scan_matches_oracle(ctx, kind, mp, kp):
    """The mixed IMI-sequence scan (FP16 diff_qr table + per-vector dot_cr
    sums, exact re-rank) equals the oracle's plain search for both
    two-block kinds, mp = 8 and 16 (and mp = 4, which scans plain)."""
    ix, oix = _index_pair(ctx, kind, kc=8 if kind == S.KIND_SCIMI else 16, mp=mp, kp=kp)
    for R, w, T in ((100, 100000, 3000), (300, 500, 40), (10, 100000, 10 ** 6)):
        _compare_search(ctx, ix, oix, kind, R, w, T, nq=80, form=S.FORM_MIXED)
    ix.close()


@pytest.mark.parametrize("kind", [S.KIND_IMI, S.KIND_SCIMI])
def test_search_mixed_imi_overflow_redo(ctx, kind, monkeypatch):
    """A margin widened a billion times keeps every candidate: the kept sets
    outgrow the buffer, those queries are redone exactly, and the results do
    not change."""
    monkeypatch.setenv("SC_MIXED_MARGIN", "1e9")
    ix, oix = _index_pair(ctx, kind, kc=8 if kind == S.KIND_SCIMI else 16)
    _compare_search(ctx, ix, oix, kind, 50, 100000, 4000, nq=40, form=S.FORM_MIXED)
    ix.close()


def test_search_mixed_form_overflow_redo(ctx, monkeypatch):
    """An inflated error margin makes the kept set outgrow the buffer; those
    queries are redone exactly and the results must not change."""
    ix, oix = _index_pair(ctx, S.KIND_SC, kc=64, kp=64)
    monkeypatch.setenv("SC_MIXED_MARGIN", "1e9")
    _compare_search(ctx, ix, oix, S.KIND_SC, 50, 300, 2500, nq=40, form=S.FORM_MIXED)
    monkeypatch.delenv("SC_MIXED_MARGIN")
    ix.close()


@pytest.mark.parametrize("kind,kc,kp,form", [(S.KIND_SC, 16, 64, S.FORM_PLAIN),
                                             (S.KIND_SC, 64, 64, S.FORM_PLAIN),
                                             (S.KIND_IVFADC, 64, 64, S.FORM_PLAIN),
                                             (S.KIND_IMI, 16, 64, S.FORM_PLAIN),
                                             (S.KIND_IMI, 16, 64, S.FORM_MIXED),
                                             (S.KIND_SC, 64, 64, S.FORM_MIXED)])
def test_search_plan_replays_match_oracle(ctx, kind, kc, kp, form):
    """A captured plan replayed several times, with eager searches and new
    query contents between the replays, and again after autotuning; every
    replay must equal the oracle (kc = kp = 64 takes the tiled LUT kernel
    inside the graph; IMI captures the staged kernels in its graph)."""
    ix, oix = _index_pair(ctx, kind, kc=kc, kp=kp)
    okind = {S.KIND_IVFADC: sco.IVFADC, S.KIND_IMI: sco.IMI, S.KIND_SC: sco.SC,
             S.KIND_SC3: sco.SC3}[kind]
    R, w, T, nq = 50, (12 if kind == S.KIND_IVFADC else 300), 2500, 130
    qs = [sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, i * nq, nq, ix.dim) for i in range(3)]
    dq = S.Dev.from_numpy(ctx, qs[0])
    plan = S.Plan(ix, dq, kind, R, w, T, form=form)
    other = S.Dev.from_numpy(ctx, qs[2][:40])
    for rep in range(5):
        if rep == 3:
            plan.autotune()
        qh = qs[rep % 2]
        dq.copy_from(qh)
        st = plan.execute(stats=True)
        oids, od, osum = oix.search_batch(okind, qh, R, w, T)
        assert np.array_equal(plan.sums.numpy(), osum), rep
        assert np.array_equal(plan.ids.numpy(), oids), rep
        assert np.array_equal(_bits(plan.dists.numpy()), _bits(od)), rep
        assert st.queries == nq and st.candidates == int(osum.sum())
        # eager work between replays
        ix.search(other, kind, R, w, T)
        ctx.sync()
    plan.close()
    ix.close()


@pytest.mark.parametrize("form", [S.FORM_PLAIN, S.FORM_MIXED])
def test_search_plan_staged_graph_fallback(ctx, monkeypatch, form):
    """IMI plan graph with a 2-entry on-chip heap: the graph flags the queries
    whose heap outgrows it and sc_plan_execute finishes them after the graph;
    every replay equals the oracle.  The mixed form's exact redo runs inside
    the graph (an inflated margin makes it happen)."""
    monkeypatch.setenv("SC_HEAP_ONCHIP", "2")
    if form == S.FORM_MIXED:
        monkeypatch.setenv("SC_MIXED_MARGIN", "1e9")
    ix, oix = _index_pair(ctx, S.KIND_IMI)
    R, w, T, nq = 50, 300, 2500, 70
    q = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, 0, nq, ix.dim)
    dq = S.Dev.from_numpy(ctx, q)
    plan = S.Plan(ix, dq, S.KIND_IMI, R, w, T, form=form)
    oids, od, osum = oix.search_batch(sco.IMI, q, R, w, T)
    for rep in range(3):
        st = plan.execute(stats=True)
        assert st.heap_fallback > 0
        assert np.array_equal(plan.sums.numpy(), osum), rep
        assert np.array_equal(plan.ids.numpy(), oids), rep
        assert np.array_equal(_bits(plan.dists.numpy()), _bits(od)), rep
        assert st.candidates == int(osum.sum())
    plan.close()
    ix.close()


@pytest.mark.parametrize("form", [S.FORM_PLAIN, S.FORM_MIXED])
def test_search_heap_fallback_path(ctx, monkeypatch, form):
    """A 2-entry on-chip heap forces every multi-sequence query onto the
    global-memory heap of capacity w; results must not change (the mixed
    form with an inflated margin, so that its scans overflow as well)."""
    if form == S.FORM_MIXED:
        monkeypatch.setenv("SC_MIXED_MARGIN", "1e9")
    for kind in (S.KIND_IMI, S.KIND_SC):
        ix, oix = _index_pair(ctx, kind)
        monkeypatch.setenv("SC_HEAP_ONCHIP", "2")
        st = _compare_search(ctx, ix, oix, kind, 50, 300, 2500, nq=60, form=form)
        assert st.heap_fallback > 0
        monkeypatch.delenv("SC_HEAP_ONCHIP")
        ix.close()


def test_search_cell_list_overflow_path(ctx, monkeypatch):
    """A 4-cell on-chip cell list flags most fused queries; the staged kernels
    finish them and the results must not change."""
    ix, oix = _index_pair(ctx, S.KIND_SC)
    monkeypatch.setenv("SC_FUSED_CELLS", "4")
    _compare_search(ctx, ix, oix, S.KIND_SC, 50, 300, 2500, nq=60)
    monkeypatch.delenv("SC_FUSED_CELLS")
    ix.close()


@pytest.mark.parametrize("kind", [S.KIND_IMI, S.KIND_SC])
def test_search_past_staged_ranks(ctx, kind, monkeypatch):
    """kc = 300: the one-workgroup traversal of the staged kernels stages 256
    ranks per list; the queries that walk the whole grid go past them and are
    redone by the global-heap kernel; every result must equal the oracle's."""
    monkeypatch.setenv("SC_NO_FUSED", "1")          # SC through the staged kernels too
    ix, oix = _index_pair(ctx, kind, n=4000, kc=300)
    _compare_search(ctx, ix, oix, kind, 50, 100000, 3000, nq=24)
    _compare_search(ctx, ix, oix, kind, 50, 100000, 100000, nq=6)
    ix.close()


@pytest.mark.parametrize("form", [S.FORM_PLAIN, S.FORM_MIXED])
def test_search_pop_budget_handoff(ctx, monkeypatch, form):
    """SC queries whose traversal pops more cells than the fused kernel's
    budget are finished by the staged kernels; a budget of 3 hands most of
    them on, and every result must still equal the oracle's."""
    monkeypatch.setenv("SC_POP_BUDGET", "3")
    ix, oix = _index_pair(ctx, S.KIND_SC)
    st = _compare_search(ctx, ix, oix, S.KIND_SC, 100, 500, 3000, form=form)
    assert st.visited > 3 * 150 * 0.5       # most queries popped past the budget
    ix.close()


@pytest.mark.parametrize("kind", [S.KIND_IVFADC, S.KIND_IMI, S.KIND_SC, S.KIND_SCIMI])
def test_search_scan_order_invariant(ctx, kind, monkeypatch):
    """The staged scan's workgroup order (launch order, sorted by first
    cell, sorted and split over the XCDs) changes no output bit, in the
    plain and the real-distance forms."""
    monkeypatch.setenv("SC_NO_FUSED", "1")
    ix, oix = _index_pair(ctx, kind)
    raw = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_BASE, 0, oix.pid.size, ix.dim)
    ix.attach_raw(S.Dev.from_numpy(ctx, raw))
    q = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, 0, 300, ix.dim)
    dq = S.Dev.from_numpy(ctx, q)
    for form in (S.FORM_PLAIN, S.FORM_REAL):
        out = {}
        for order in ("0", "1", "2"):
            monkeypatch.setenv("SC_SCAN_ORDER", order)
            ids, dists, sums, _ = ix.search(dq, kind, 50, 8, 800, form=form)
            out[order] = (ids.numpy(), dists.numpy().view(np.uint32), sums.numpy())
        for order in ("1", "2"):
            for a, b in zip(out["0"], out[order]):
                assert np.array_equal(a, b), (form, order)
    ix.close()


def test_search_exhausted_grid(ctx):
    ix, oix = _index_pair(ctx, S.KIND_SC, n=300, kc=4)
    st = _compare_search(ctx, ix, oix, S.KIND_SC, 400, 1000, 10 ** 6, nq=20)
    assert st.exhausted == 20
    ix.close()


def test_search_kind_mismatch_errors(ctx):
    ix, _ = _index_pair(ctx, S.KIND_SC, n=500)
    q = S.Dev.from_numpy(ctx, np.zeros((2, 64), np.float32))
    with pytest.raises(S.ScError) as e:
        ix.search(q, S.KIND_IMI, 10, 10, 10)
    assert e.value.rc == -3
    ix.close()
    ix, _ = _index_pair(ctx, S.KIND_IVFADC, n=500)
    with pytest.raises(S.ScError) as e:
        ix.search(q, S.KIND_IVFADC, 10, ix.kc + 1, 10)
    assert e.value.rc == -7
    ix.close()


def test_index_load_roundtrip(ctx):
    ix, oix = _index_pair(ctx, S.KIND_SC)
    ix2 = S.Index.load(ctx, oix.L, oix.pid, oix.codes, 64, oix.cq, 1, oix.kc, oix.pq, oix.mp,
                       oix.kp, 2)
    _compare_search(ctx, ix2, oix, S.KIND_SC, 100, 500, 3000, nq=40)
    ix.close()
    ix2.close()


# ------------------------------------------------------------------ k-means
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16])
@pytest.mark.parametrize("lagged", [0, 1])
def test_kmeans_matches_oracle_on_integer_data(ctx, lagged, prec):
    """Labels and centres equal the oracle's, for the exact form and for the
    mixed form (BF16 MFMA assignment distances, exact re-check)."""
    x = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_TRAIN, 0, 6000, 16)
    c, lab = S.kmeans(ctx, S.Dev.from_numpy(ctx, x), 40, 6, seed=11, lagged=lagged, prec=prec)
    oc, olab = sco.kmeans(x, 40, 6, seed=11, lagged=lagged)
    assert np.array_equal(lab.numpy(), olab)
    assert np.array_equal(_bits(c.numpy()), _bits(oc))     # integer data: exact FP64 sums


def test_kmeans_empty_cluster_repair(ctx):
    base = sco.gen(sco.mix_d128(ncomp=5), sco.STREAM_TRAIN, 0, 12, 8)
    x = np.repeat(base, 150, axis=0)                       # 12 distinct points, k = 16
    c, lab = S.kmeans(ctx, S.Dev.from_numpy(ctx, x), 16, 4, seed=5)
    oc, olab = sco.kmeans(x, 16, 4, seed=5)
    assert np.array_equal(lab.numpy(), olab)
    assert np.array_equal(_bits(c.numpy()), _bits(oc))


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16])
def test_kmeans_exact_on_fractional_data(ctx, prec):
    """Fractional data (PQ codebooks train on residuals): the Lloyd sums run
    over each centre's members in index order, the oracle's sequential order,
    so labels and centres equal the oracle's bit for bit, and two runs agree
    bit for bit."""
    rng = np.random.default_rng(4)
    x = sco.gen(sco.mix_d128(ncomp=9), sco.STREAM_TRAIN, 0, 20000, 16)
    x = x + rng.uniform(-0.5, 0.5, x.shape).astype(np.float32) / 3.0
    dx = S.Dev.from_numpy(ctx, x)
    c, lab = S.kmeans(ctx, dx, 32, 6, seed=2, prec=prec)
    c2, lab2 = S.kmeans(ctx, dx, 32, 6, seed=2, prec=prec)
    oc, olab = sco.kmeans(x, 32, 6, seed=2)
    assert np.array_equal(lab.numpy(), olab)
    assert np.array_equal(_bits(c.numpy()), _bits(oc))
    assert np.array_equal(_bits(c2.numpy()), _bits(c.numpy()))
    assert np.array_equal(lab2.numpy(), lab.numpy())


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dim,mc,kc,n", [(32, 2, 20, 3000), (128, 8, 256, 2100),
                                          (960, 2, 300, 700), (64, 1, 2, 500)])
def test_residual_matches_oracle(ctx, which, dim, mc, kc, n):
    """Residuals against the nearest / second-nearest centroid of each block
    (float-rounded distances, lowest index on ties), with duplicated
    centroids and vectors equal to a centroid planted."""
    x = sco.gen(sco.mix_d960() if dim == 960 else sco.mix_d128(ncomp=6), sco.STREAM_TRAIN, 0, n,
                dim)
    rng = np.random.default_rng(which + dim)
    bs = dim // mc
    cq = np.stack([x[rng.choice(n, kc, replace=False), j * bs:(j + 1) * bs] for j in range(mc)])
    if kc > 3:
        cq[:, kc - 1] = cq[:, 1]
    x[:3, :bs] = cq[0, 1]
    d = S.Dev.from_numpy(ctx, x)
    S.residual(ctx, d, S.Dev.from_numpy(ctx, cq), mc, kc, which)
    o = sco.residual(x, cq, mc, kc, which)
    assert np.array_equal(_bits(d.numpy()), _bits(o))


# ------------------------------------------------------------------ full size
def test_full_size_sc_index_properties(ctx):
    """1M x 128, SC kc=256 (the dense-partition benchmark shape): exact encode
    on sampled rows, cell invariants, and search on 64 queries vs the oracle."""
    n, dim, kc, mp, kp = 1_000_000, 128, 256, 8, 256
    mix = S.mix_d128()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 50_000, dim)
    cq, _ = S.kmeans(ctx, xt, kc, 4, seed=1, want_labels=False)
    res = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 50_000, dim)
    S.residual(ctx, res, cq, 1, kc, 0)
    rh = res.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, 4, seed=2 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, cq, dpq, 1, kc, mp, kp, 2)
    cidh, codesh = cid.numpy(), codes.numpy()
    bcid, bcodes, bst = S.encode(ctx, x, cq, dpq, 1, kc, mp, kp, 2, S.PREC_BF16)
    assert np.array_equal(bcid.numpy(), cidh)          # BF16 form == FP32 form, all rows
    assert np.array_equal(bcodes.numpy(), codesh)
    print("full size bf16: %d re-checked, %d exact rows" % (bst.rechecked, bst.fallback))
    del bcid, bcodes
    rows = np.random.default_rng(0).choice(n, 3000, replace=False)
    xr =np.concatenate([sco.gen(sco.mix_d128(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq.numpy(), pq, 1, kc, mp, kp, 2)
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, cq, 1, kc, dpq, mp, kp, 2)
    e = ix.export()
    keys = sco.cell_keys(cidh, 1, kc, 2)
    assert e["L"][-1] == n
    starts = np.concatenate([[0], e["L"][:-1]])
    assert np.all(np.diff(e["pid"])[np.diff(keys[e["pid"]]) == 0] > 0)   # ids ascend in a cell
    assert np.array_equal(e["codes"], codesh[e["pid"]])
    assert np.all(keys[e["pid"]][starts[e["L"] > starts]] == np.nonzero(e["L"] > starts)[0])
    oix = sco.OracleIndex(dim, 1, kc, mp, kp, cq.numpy(), pq, e["L"], e["pid"], e["codes"])
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 64, dim)
    ids, dists, sums, st = ix.search(S.Dev.from_numpy(ctx, q), S.KIND_SC, 100, 100000, 10000)
    oids, od, osum = oix.search_batch(sco.SC, q, 100, 100000, 10000)
    assert np.array_equal(sums.numpy(), osum)
    assert np.array_equal(ids.numpy(), oids)
    assert np.array_equal(_bits(dists.numpy()), _bits(od))
    ix.close()


def test_full_size_imi_search(ctx):
    """1M x 128 inverted multi-index, two 64-dim halves with K=1024 each, PQ
    8x256 (the multi-index benchmark shape): encode on sampled rows, index
    invariants, and search (w=100k, T=10k, top-100) on 48 queries vs the
    oracle."""
    n, dim, kc, mp, kp = 1_000_000, 128, 1024, 8, 256
    mix = S.mix_d128()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 50_000, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 64:(h + 1) * 64])),
                            kc, 3, seed=11 + h, want_labels=False)[0].numpy() for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, 3, seed=21 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 1)
    cidh, codesh = cid.numpy(), codes.numpy()
    rows = np.random.default_rng(1).choice(n, 2000, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d128(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq, pq, 2, kc, mp, kp, 1)
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 1)
    e = ix.export()
    assert e["L"].size == kc * kc and e["L"][-1] == n
    keys = sco.cell_keys(cidh, 2, kc, 1)
    assert np.array_equal(np.bincount(keys.astype(np.int64), minlength=kc * kc),
                          np.diff(np.concatenate([[0], e["L"]])))
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, e["L"], e["pid"], e["codes"])
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 48, dim)
    ids, dists, sums, st = ix.search(S.Dev.from_numpy(ctx, q), S.KIND_IMI, 100, 100000, 10000)
    oids, od, osum = oix.search_batch(sco.IMI, q, 100, 100000, 10000)
    assert np.array_equal(sums.numpy(), osum)
    assert np.array_equal(ids.numpy(), oids)
    assert np.array_equal(_bits(dists.numpy()), _bits(od))
    ix.close()


def test_large_k_sc_index_mixed_batch_sizes(ctx):
    """K=4096 (nearest, second) cells over D=128 (the large dense-partition
    shape, here with 1M base vectors): BF16 and FP32 encodes identical,
    encode on sampled rows vs the oracle, and searches served in batches of
    1, 64 and 4096 queries (eager calls and a hipGraph plan) vs the oracle."""
    n, dim, kc, mp, kp = 1_000_000, 128, 4096, 8, 256
    mix = S.mix_d128()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 100_000, dim)
    cq, _ = S.kmeans(ctx, xt, kc, 3, seed=5, want_labels=False)
    S.residual(ctx, xt, cq, 1, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, 3, seed=31 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, cq, dpq, 1, kc, mp, kp, 2, S.PREC_BF16)
    fcid, fcodes, _ = S.encode(ctx, x, cq, dpq, 1, kc, mp, kp, 2, S.PREC_FP32)
    cidh, codesh = cid.numpy(), codes.numpy()
    assert np.array_equal(cidh, fcid.numpy()) and np.array_equal(codesh, fcodes.numpy())
    del fcid, fcodes
    rows = np.random.default_rng(2).choice(n, 600, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d128(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq.numpy(), pq, 1, kc, mp, kp, 2)
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, cq, 1, kc, dpq, mp, kp, 2)
    e = ix.export()
    assert e["L"].size == kc * kc and e["L"][-1] == n
    oix = sco.OracleIndex(dim, 1, kc, mp, kp, cq.numpy(), pq, e["L"], e["pid"], e["codes"])
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 4096, dim)
    chk = np.random.default_rng(3).choice(4096, 40, replace=False)
    oids, od, osum = oix.search_batch(sco.SC, q[chk], 100, 100000, 10000)
    for nb in (1, 64, 4096):
        starts = sorted({int(c) // nb * nb for c in (chk[:8] if nb == 1 else chk)})
        for start in starts:
            dq = S.Dev.from_numpy(ctx, np.ascontiguousarray(q[start:start + nb]))
            ids, dists, sums, _ = ix.search(dq, S.KIND_SC, 100, 100000, 10000, stats=False)
            gi, gd, gs = ids.numpy(), dists.numpy(), sums.numpy()
            for k, c in enumerate(chk):
                if start <= c < start + nb:
                    assert gs[c - start] == osum[k]
                    assert np.array_equal(gi[c - start], oids[k])
                    assert np.array_equal(_bits(gd[c - start]), _bits(od[k]))
    plan = S.Plan(ix, S.Dev.from_numpy(ctx, q), S.KIND_SC, 100, 100000, 10000)
    plan.autotune()
    plan.execute(stats=False)
    pi, pd = plan.ids.numpy(), plan.dists.numpy()
    assert np.array_equal(pi[chk], oids) and np.array_equal(_bits(pd[chk]), _bits(od))
    plan.close()
    ix.close()


@pytest.mark.parametrize("d,k,n", [(64, 1024, 30000), (128, 256, 20000), (16, 256, 50000)])
def test_kmeans_mixed_form_equals_exact(ctx, d, k, n):
    """SC_PREC_BF16 k-means (BF16 MFMA Lloyd assignment with an exact
    re-check of every centre within the error bound) gives the exact form's
    labels and centres bit for bit on integer data, also with many exact
    distance ties (duplicated points), and the oracle's on a sample."""
    x = sco.gen(sco.mix_d128(ncomp=17), sco.STREAM_TRAIN, 0, n, 128)[:, :d].copy()
    x[n // 2:n // 2 + 500] = x[:500]            # duplicated points: equal distances
    dx = S.Dev.from_numpy(ctx, x)
    c1, l1 = S.kmeans(ctx, dx, k, 4, seed=7, prec=S.PREC_FP32)
    c2, l2 = S.kmeans(ctx, dx, k, 4, seed=7, prec=S.PREC_BF16)
    assert np.array_equal(l1.numpy(), l2.numpy())
    assert np.array_equal(_bits(c1.numpy()), _bits(c2.numpy()))
    if n <= 30000 and k <= 256:
        oc, olab = sco.kmeans(x, k, 4, seed=7)
        assert np.array_equal(l2.numpy(), olab)


def test_d960_imi_search(ctx):
    """D=960 inverted multi-index (two 480-dim halves with K=1024 each, PQ
    16x256, the high-dimensional shape) on 60k base vectors: encode on
    sampled rows and searches at budgets from 1k to 30k candidates (one call
    per budget) vs the oracle."""
    n, dim, kc, mp, kp = 60_000, 960, 1024, 16, 256
    mix = S.mix_d960()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 20_000, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 480:(h + 1) * 480])),
                            kc, 2, seed=41 + h, want_labels=False)[0].numpy() for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 60:(s + 1) * 60])),
                            kp, 2, seed=51 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 1)
    cidh, codesh = cid.numpy(), codes.numpy()
    rows = np.random.default_rng(4).choice(n, 300, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d960(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq, pq, 2, kc, mp, kp, 1)
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 1)
    e = ix.export()
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, e["L"], e["pid"], e["codes"])
    q = sco.gen(sco.mix_d960(), sco.STREAM_QUERY, 0, 24, dim)
    dq = S.Dev.from_numpy(ctx, q)
    for T in (1000, 12000, 30000):
        ids, dists, sums, _ = ix.search(dq, S.KIND_IMI, 100, 200000, T)
        oids, od, osum = oix.search_batch(sco.IMI, q, 100, 200000, T)
        assert np.array_equal(sums.numpy(), osum), T
        assert np.array_equal(ids.numpy(), oids), T
        assert np.array_equal(_bits(dists.numpy()), _bits(od)), T
    ix.close()


def test_d960_imi_full_size_budget_plans(ctx):
    """High-dimensional batched search at its full base size: D=960, IMI with
    two 480-dim halves of K=1024, PQ 16x256, 1M base vectors.  Encode on
    sampled rows vs the oracle; 2000 queries with budgets drawn from
    U[1k, 30k] run as one hipGraph plan per form (identical ids, sums and
    distance bits between the forms), and 24 sampled queries equal the
    oracle's search with their own budget."""
    n, dim, kc, mp, kp, R, w = 1_000_000, 960, 1024, 16, 256, 100, 200_000
    mix = S.mix_d960()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 30_000, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 480:(h + 1) * 480])),
                            kc, 2, seed=41 + h, prec=S.PREC_BF16, want_labels=False)[0].numpy()
                   for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 60:(s + 1) * 60])),
                            kp, 2, seed=51 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 1)
    rows = np.random.default_rng(7).choice(n, 200, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d960(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq, pq, 2, kc, mp, kp, 1)
    assert np.array_equal(cid.numpy()[rows], ocid) and np.array_equal(codes.numpy()[rows], ocodes)
    del x
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 1)
    del cid, codes
    nq = 2000
    q = S.gen_rows(ctx, mix, S.STREAM_QUERY, 0, nq, dim)
    T = np.random.default_rng(960).integers(1000, 30001, nq).astype(np.int32)
    dT = S.Dev.from_numpy(ctx, T)
    got = {}
    for form in (S.FORM_PLAIN, S.FORM_MIXED):
        plan = S.Plan(ix, q, S.KIND_IMI, R, w, 30000, form=form, budgets=dT)
        for _ in range(2):
            plan.execute(stats=False)
        got[form] = (plan.ids.numpy(), plan.dists.numpy(), plan.sums.numpy())
        plan.close()
    for a, b in zip(got[S.FORM_PLAIN], got[S.FORM_MIXED]):
        assert np.array_equal(_bits(a) if a.dtype == np.float32 else a,
                              _bits(b) if b.dtype == np.float32 else b)
    e = ix.export()
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, e["L"], e["pid"], e["codes"])
    qh = q.numpy()
    gi, gd, gs = got[S.FORM_PLAIN]
    for i in np.random.default_rng(3).choice(nq, 24, replace=False):
        oi, od, st, _ = oix.search(sco.IMI, qh[i], R, w, int(T[i]))
        assert gs[i] == st.sum, i
        assert np.array_equal(gi[i], oi), i
        assert np.array_equal(_bits(gd[i]), _bits(od)), i
    ix.close()


def test_exhaustive_pq_config(ctx):
    """Exhaustive PQ ADC (the reference's CPU configuration: 100k base vectors,
    PQ 8x256 trained on 50k, top-100) mapped onto search_ivfadc with one
    zero coarse centroid, w=1 and T >= N: every vector is a candidate; ids
    and distances vs the oracle."""
    n, dim, mp, kp = 100_000, 128, 8, 256
    mix = S.mix_d128()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 50_000, dim).numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(xt[:, s * 16:(s + 1) * 16])),
                            kp, 5, seed=61 + s, want_labels=False)[0].numpy() for s in range(mp)])
    cq = np.zeros((1, 1, dim), np.float32)
    dcq, dpq = S.Dev.from_numpy(ctx, cq), S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 1, 1, mp, kp, 1)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 1, 1, dpq, mp, kp, 1)
    e = ix.export()
    assert e["L"][0] == n and np.array_equal(e["pid"], np.arange(n))
    oix = sco.OracleIndex(dim, 1, 1, mp, kp, cq, pq, e["L"], e["pid"], e["codes"])
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 64, dim)
    ids, dists, sums, _ = ix.search(S.Dev.from_numpy(ctx, q), S.KIND_IVFADC, 100, 1, n)
    oids, od, osum = oix.search_batch(sco.IVFADC, q, 100, 1, n)
    assert (sums.numpy() == n).all() and np.array_equal(sums.numpy(), osum)
    assert np.array_equal(ids.numpy(), oids)
    assert np.array_equal(_bits(dists.numpy()), _bits(od))
    ix.close()


def test_full_size_scimi_index_and_search(ctx):
    """1M x 128, two 64-dim halves with K=256 each and (nearest, second)
    cells per half, PQ 8x256 (the dense-partition benchmark shape): the BF16
    encode equals the FP32 encode on every row, sampled rows equal the
    oracle's encode, the sparse directory holds exactly the vectors' keys,
    and search (w=100k, T=10k, top-100) on 64 queries equals the oracle."""
    n, dim, kc, mp, kp = 1_000_000, 128, 256, 8, 256
    mix = S.mix_d128()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 50_000, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 64:(h + 1) * 64])),
                            kc, 3, seed=31 + h, want_labels=False)[0].numpy() for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, 3, seed=41 + s, want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 2)
    cidh, codesh = cid.numpy(), codes.numpy()
    bcid, bcodes, bst = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 2, S.PREC_BF16)
    assert np.array_equal(bcid.numpy(), cidh) and np.array_equal(bcodes.numpy(), codesh)
    print("two halves bf16: %d re-checked, %d exact rows" % (bst.rechecked, bst.fallback))
    del bcid, bcodes
    rows = np.random.default_rng(2).choice(n, 2000, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d128(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq, pq, 2, kc, mp, kp, 2)
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    e = ix.export()
    ck = ix.cell_keys()
    vkeys = sco.cell_keys(cidh, 2, kc, 2)
    uk, cnt = np.unique(vkeys, return_counts=True)
    assert np.array_equal(ck, uk)                                  # exactly the vectors' cells
    assert np.array_equal(np.diff(np.concatenate([[0], e["L"]])), cnt)
    assert np.array_equal(vkeys[e["pid"]], np.repeat(uk, cnt))     # members in key order
    assert np.all(np.diff(e["pid"])[np.diff(vkeys[e["pid"]]) == 0] > 0)   # ids ascend in a cell
    assert np.array_equal(e["codes"], codesh[e["pid"]])
    print("two halves: %d non-empty cells of %d" % (ck.size, kc ** 4))
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, e["L"], e["pid"], e["codes"], ckeys=ck)
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 64, dim)
    oids, od, osum = oix.search_batch(sco.SCIMI, q, 100, 100000, 10000)
    for form in (S.FORM_PLAIN, S.FORM_MIXED):
        ids, dists, sums, st = ix.search(S.Dev.from_numpy(ctx, q), S.KIND_SCIMI, 100, 100000,
                                         10000, form=form)
        assert np.array_equal(sums.numpy(), osum)
        assert np.array_equal(ids.numpy(), oids)
        assert np.array_equal(_bits(dists.numpy()), _bits(od))
    ix.close()


@pytest.mark.parametrize("kind,kc,nc_ok", [(S.KIND_SC, 16, True), (S.KIND_IMI, 16, True),
                                           (S.KIND_IVFADC, 16, True), (S.KIND_SC3, 12, True),
                                           (S.KIND_SCIMI, 8, True)])
def test_search_per_query_budgets(ctx, kind, kc, nc_ok):
    """A batch whose queries carry their own candidate budgets (drawn from a
    wide range, some above the batch maximum and some negative, which clamp)
    returns for every query what a search of that query alone with its budget
    returns in the oracle; eager calls and a plan of every kind."""
    ix, oix = _index_pair(ctx, kind, kc=kc)
    nq = 90
    rng = np.random.default_rng(int(kind) + 17)
    q = sco.gen(sco.mix_d128(ncomp=11), sco.STREAM_QUERY, 0, nq, ix.dim)
    Tmax = 3000
    T = rng.integers(0, Tmax + 1, nq).astype(np.int32)
    T[:3] = [-5, Tmax + 999, 0]
    w = 16 if kind == S.KIND_IVFADC else 400
    dq, dT = S.Dev.from_numpy(ctx, q), S.Dev.from_numpy(ctx, T)
    okind = OKIND[kind]
    want = [oix.search(okind, q[i], 100, w, int(min(max(T[i], 0), Tmax))) for i in range(nq)]
    ids, dists, sums, _ = ix.search(dq, kind, 100, w, Tmax, budgets=dT)
    gi, gd, gs = ids.numpy(), dists.numpy(), sums.numpy()
    for i in range(nq):
        oi, od, st, _ = want[i]
        assert gs[i] == st.sum, i
        assert np.array_equal(gi[i], oi), i
        assert np.array_equal(_bits(gd[i]), _bits(od)), i
    # plans: fused graph (SC, IVFADC), staged graph (IMI), staged calls
    plan = S.Plan(ix, dq, kind, 100, w, Tmax, budgets=dT)
    plan.execute(stats=False)
    assert np.array_equal(plan.ids.numpy(), gi) and np.array_equal(plan.sums.numpy(), gs)
    T2 = T[::-1].copy()                        # new budgets, same buffer
    dT.copy_from(T2)
    plan.execute(stats=False)
    for i in range(0, nq, 7):
        oi, od, st, _ = oix.search(okind, q[i], 100, w, int(min(max(T2[i], 0), Tmax)))
        assert plan.sums.numpy()[i] == st.sum and np.array_equal(plan.ids.numpy()[i], oi)
    plan.close()
    ix.close()


def test_two_halves_k4096_100m_base(ctx):
    """The large dense-partition shape: two 64-dim halves with K=4096
    (nearest, second) cells each, PQ 8x256, a 100M-vector base generated and
    encoded a 1M-vector chunk at a time (BF16 MFMA distances with the exact
    re-check), 8e8 bytes of codes.  Sampled rows' cell ids and codes equal
    the oracle's encode and sit in their key's cell; search (w=100k, T=10k,
    top-100) of 32 queries served in batches of 1, 4 and 27 equals the
    oracle (the pruned ranking on the GPU against the oracle's selection
    ranking), in both forms."""
    n, dim, kc, mp, kp = 100_000_000, 128, 4096, 8, 256
    mix = S.mix_d128()
    # codebooks from 1M training vectors, 4 Lloyd iterations each
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 1_000_000, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 64:(h + 1) * 64])),
                            kc, 4, seed=51 + h, prec=S.PREC_BF16, want_labels=False)[0].numpy()
                   for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, 4, seed=61 + s, prec=S.PREC_BF16, want_labels=False)[0].numpy()
                   for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    xt.free()
    del th, rh

    def fill(r0, m, buf):
        S.gen_rows(ctx, mix, S.STREAM_BASE, r0, m, dim, out=buf)

    cid, codes, rech, exact = S.encode_stream(ctx, fill, n, dim, dcq, dpq, 2, kc, mp, kp, 2,
                                              prec=S.PREC_BF16)
    print("100M encode: %d codewords re-checked, %d exact (row, block) pairs" % (rech, exact))
    rows = np.random.default_rng(5).choice(n, 1500, replace=False)
    xr = np.concatenate([sco.gen(sco.mix_d128(), sco.STREAM_BASE, int(r), 1, dim) for r in rows])
    ocid, ocodes = sco.encode(xr, cq, pq, 2, kc, mp, kp, 2)
    cidh, codesh = cid.numpy(), codes.numpy()
    assert np.array_equal(cidh[rows], ocid) and np.array_equal(codesh[rows], ocodes)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    cid.free()
    codes.free()
    e = ix.export()
    ck = ix.cell_keys()
    L, pid = e["L"], e["pid"]
    assert ck.size > 0 and np.all(np.diff(ck) > 0) and int(L[-1]) == n
    assert np.all(np.diff(L) > 0) and L[0] > 0
    keys = sco.cell_keys(np.ascontiguousarray(cidh[rows]), 2, kc, 2)
    c = np.searchsorted(ck, keys)
    assert np.array_equal(ck[c], keys)
    for r, ci in zip(rows, c):
        lo = int(L[ci - 1]) if ci > 0 else 0
        assert r in pid[lo:int(L[ci])]
        assert np.array_equal(e["codes"][lo + int(np.nonzero(pid[lo:int(L[ci])] == r)[0][0])],
                              codesh[r])
    del cidh, codesh
    print("100M base: %d non-empty cells" % ck.size)
    oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, L, pid, e["codes"], ckeys=ck)
    del e
    q = sco.gen(sco.mix_d128(), sco.STREAM_QUERY, 0, 32, dim)
    oids, od, osum = oix.search_batch(sco.SCIMI, q, 100, 100000, 10000)
    dq = S.Dev.from_numpy(ctx, q)
    for form in (S.FORM_PLAIN, S.FORM_MIXED):
        out = [ix.search(dq.rows(a, b - a), S.KIND_SCIMI, 100, 100000, 10000, form=form)
               for a, b in ((0, 1), (1, 5), (5, 32))]
        ids = np.concatenate([o[0].numpy() for o in out])
        dists = np.concatenate([o[1].numpy() for o in out])
        sums = np.concatenate([o[2].numpy() for o in out])
        assert np.array_equal(sums, osum)
        assert np.array_equal(ids, oids)
        assert np.array_equal(_bits(dists), _bits(od))
    ix.close()
