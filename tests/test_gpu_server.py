"""Serving form on the GPU: every query's result (row digests and nnz) is the same alone (the
per-chain path, one query at a time) and in the mix (concurrent slots, prefix cache, product
plans, admission waits, streamed roots); admission with a tiny budget serialises queries."""
import numpy as np
import pytest

from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    gb = configs.BibSpec(nA=1500, nP=2500, nV=15, nT=150, lam_auth=2.0, terms_per_paper=6,
                         seed=configs.SEED_BASE + 91).build()
    gl = configs.LocSpec(nU=800, nL=1500, nC=20, nY=30, nS=24, geo_p=0.1, seed=configs.SEED_BASE + 92).build()
    G = [smcm.Graph(ctx, gb), smcm.Graph(ctx, gl)]
    stream = configs.cfg5_stream(gb.dimensions, gl.dimensions, n_queries=80, qps=1e6, lmax=7, smax=1500,
                                 seed=configs.SEED_BASE + 93)
    hgs = [gb, gl]
    queries = [(q.graph, q.path, (q.src_rows + hgs[q.graph].type_offset[q.path.vertexTypes[0]]).astype(np.int32))
               for q in stream]
    alone = []
    for gi, mp, src in queries:
        r, _ = G[gi].run(mp, src)
        alone.append((r.nnz, r.row_digest()))
        r.free()
    return G, queries, [q.arrival for q in stream], alone


def compare(outs, digs, alone):
    for i, (o, d, (nnz, dig)) in enumerate(zip(outs, digs, alone)):
        assert o["status"] == 0, (i, o)
        assert o["nnz"] == nnz, i
        np.testing.assert_array_equal(d, dig, err_msg=f"query {i}")


@pytest.mark.parametrize("slots,plan,cache,block,large", [(4, 1, 1 << 30, 32 << 30, 4 << 30),
                                                          (1, 0, 0, 32 << 30, 4 << 30),
                                                          (8, 1, 1 << 20, 1 << 20, 4 << 30),
                                                          (3, 0, 1 << 30, 32 << 30, 4 << 30),
                                                          (4, 1, 1 << 30, 32 << 30, 1 << 10)])
def test_alone_equals_mix(ctx, served, slots, plan, cache, block, large):
    """large = 1 KiB sends most queries through the hand-over to the large-lane workers."""
    G, queries, arrivals, alone = served
    srv = smcm.Server(ctx, G, slots=slots, plan_cost=plan, cache=cache, block=block, large_bytes=large)
    try:
        outs, digs = srv.replay(queries, np.zeros(len(queries)))  # everything arrives at once
        compare(outs, digs, alone)
        st = srv.stats()
        if cache >= (1 << 30):
            assert st["cache_hits"] > 0 and st["cache_inserts"] > 0, st
            assert any(o["cache_k"] > 0 for o in outs)
    finally:
        srv.close()


def test_paced_arrivals(ctx, served):
    G, queries, arrivals, alone = served
    srv = smcm.Server(ctx, G, slots=4)
    try:
        arr = np.linspace(0.0, 0.2, len(queries))
        outs, digs = srv.replay(queries, arr)
        compare(outs, digs, alone)
        for o, a in zip(outs, arr):
            assert o["t_submit"] >= a - 1e-3
            assert o["t_done"] >= o["t_start"] >= o["t_submit"] - 1e-6
    finally:
        srv.close()


def test_tiny_budget_serialises(ctx, served):
    """One-row queries cannot be split: with a 1-byte budget they run one at a time."""
    G, queries, arrivals, alone = served
    one = [(gi, mp, src[:1]) for gi, mp, src in queries[:30]]
    ref = []
    for gi, mp, src in one:
        r, _ = G[gi].run(mp, src)
        ref.append((r.nnz, r.row_digest()))
        r.free()
    srv = smcm.Server(ctx, G, slots=4, budget=1)
    try:
        outs, digs = srv.replay(one, np.zeros(len(one)))
        compare(outs, digs, ref)
        iv = sorted((o["t_start"], o["t_done"]) for o in outs)
        for (s0, e0), (s1, e1) in zip(iv, iv[1:]):
            assert s1 >= e0 - 1e-4, (s0, e0, s1, e1)
    finally:
        srv.close()


def test_oversized_queries_split(ctx, served):
    """Queries predicted above half the budget run as blocks of source rows: same results."""
    G, queries, arrivals, alone = served
    srv = smcm.Server(ctx, G, slots=4, budget=4 << 20, block=1 << 20)
    try:
        outs, digs = srv.replay(queries, np.zeros(len(queries)))
        compare(outs, digs, alone)
        assert any(o["blocks"] > 1 for o in outs)
    finally:
        srv.close()


def test_bad_query_rejected(ctx, served):
    G, queries, arrivals, alone = served
    srv = smcm.Server(ctx, G, slots=2)
    try:
        mp = configs.metapath_from_letters("A-P-V", "bib")
        with pytest.raises(smcm.SmcmError):
            srv.replay([(0, mp, np.array([0], np.int32))], [0.0])  # global id 0 is a paper, not an author
    finally:
        srv.close()


def test_cancel(ctx, served):
    """Cancelled queries finish with SMCM_ECANCELED, whether queued, waiting for admission (a
    one-byte budget admits one query at a time) or running; a query past its last check finishes
    normally with its usual result.  The other queries' results are unchanged, and the server
    keeps serving afterwards."""
    import time
    G, queries, arrivals, alone = served
    srv = smcm.Server(ctx, G, slots=3, budget=1, cache=0)
    try:
        tickets = [srv.submit(gi, mp, src) for gi, mp, src in queries[:40]]
        time.sleep(0.002)  # some queries have started
        cancelled = set(range(0, 40, 3))
        for i in sorted(cancelled, reverse=True):
            srv.cancel(tickets[i][0])
        n_cancelled = 0
        for i, (t, dig) in enumerate(tickets):
            o = srv.poll(t)
            nnz, ref = alone[i]
            if i in cancelled and o["status"] == smcm.SMCM_ECANCELED:
                n_cancelled += 1
                assert "cancelled" in o["error"], o
                continue
            assert o["status"] == 0, (i, o)
            assert o["nnz"] == nnz, i
            np.testing.assert_array_equal(dig, ref, err_msg=f"query {i}")
        assert n_cancelled >= len(cancelled) // 2, n_cancelled
        with pytest.raises(smcm.SmcmError) as ei:
            srv.cancel(tickets[0][0])  # polled: the ticket is released
        assert ei.value.code == smcm.SMCM_EINVAL
        with pytest.raises(smcm.SmcmError):
            srv.cancel(1 << 40)
        outs, digs = srv.replay(queries[:20], np.zeros(20))
        compare(outs, digs, alone[:20])
    finally:
        srv.close()
