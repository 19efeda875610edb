"""The benchmark workloads on the GPU at their own sizes, each checked against the oracle or, for
the batched and serving forms, against the per-chain path (itself checked against the oracle):

* config 2 (A=500k, P=800k, V=5k, T=20k): every one of the 500k A-P-V-P-A rows and 5100
  A-P-T-P-A rows (random rows plus the 100 authors with the most papers) against the oracle's
  row digests; plans equal to the oracle DP's; the executed product count equal to the oracle's
  count for that plan over all rows;
* config 3 (U=100k location graph): the L = 8-12 meta-paths from 256 seeded users, reference
  plan and product plan, against the oracle's left-to-right evaluation (every order gives the
  same exact product);
* config 4: the first 512 chains of the 4096-chain batch over the config-1 graph (with terms),
  persistent executor forced on and forced off, against the per-chain path;
* config 5: the first 48 queries of the Poisson stream over the config-1 and config-3 graphs,
  served in the mix against each query run alone.
The row digest is mix(width) + sum over the row of mix(mix(column) ^ count) mod 2^64 on both
sides (smcm_result_row_digest / orc_row_digest)."""
import numpy as np
import pytest

import oracle as O
from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu

C_RTOL = 1e-12  # tree-ordered c_ij sums vs the one-thread row-order sum


def first_mismatch(got, want):
    bad = np.flatnonzero(got != want)
    return f"{bad.size} rows differ, first {bad[:5].tolist()}"


@pytest.fixture(scope="module")
def cfg2(ctx):
    hg = configs.CFG2.build()
    G = smcm.Graph(ctx, hg)
    yield hg, G
    G.close()


def test_config2_apvpa_every_row(ctx, cfg2):
    hg, G = cfg2
    mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
    nnz, dig, _ = G.run_blocked(mp, 48 << 30, digest=True)
    st = ctx.last_stats()
    mats = O.chain_matrices(hg, mp)
    c_ref, _ = O.estimate(mats, O.coldims(hg, mp))
    _, _, _, pairs = O.dp(c_ref)
    assert st["pairs"] == pairs.ravel().tolist()
    c, _, _, gp = G.plan(mp)
    np.testing.assert_allclose(c, c_ref, rtol=C_RTOL, atol=0)
    np.testing.assert_array_equal(gp, pairs)
    want, wnnz, wprods = O.plan_digest(mats, pairs, 1 << 15)
    assert nnz == wnnz
    assert st["products"] == wprods
    assert np.array_equal(dig, want), first_mismatch(dig, want)


def test_config2_aptpa_rows(ctx, cfg2):
    hg, G = cfg2
    mp = configs.metapath_from_letters("A-P-T-P-A", "bib")
    nnz, dig, _ = G.run_blocked(mp, 48 << 30, digest=True)
    st = ctx.last_stats()
    mats = O.chain_matrices(hg, mp)
    c_ref, _ = O.estimate(mats, O.coldims(hg, mp))
    _, _, _, pairs = O.dp(c_ref)
    assert st["pairs"] == pairs.ravel().tolist()
    # products of the plan over all rows (the root's operands only: its 4.5e10 entries stay unbuilt)
    assert st["products"] == O.plan_products(mats, pairs)
    deg = np.diff(mats[0].ptr)
    rng = np.random.default_rng(7)
    rows = np.unique(np.concatenate([rng.choice(mats[0].nrows, 5000, replace=False),
                                     np.argsort(deg, kind="stable")[-100:]])).astype(np.int32)
    sub = O.chain_matrices(hg, mp, rows)
    cs, _ = O.estimate(sub, O.coldims(hg, mp))
    _, _, _, spairs = O.dp(cs)
    want, _, _ = O.plan_digest(sub, spairs, 1 << 12)
    assert np.array_equal(dig[rows], want), first_mismatch(dig[rows], want)
    # the same rows as a source subset give the same digests
    src = (rows + hg.type_offset[configs.A]).astype(np.int32)
    _, dsub, _ = G.run_blocked(mp, 48 << 30, src_rows=src, digest=True)
    assert np.array_equal(dsub, want), first_mismatch(dsub, want)


@pytest.fixture(scope="module")
def cfg3(ctx):
    hg = configs.CFG3.build()
    G = smcm.Graph(ctx, hg)
    yield hg, G
    G.close()


@pytest.mark.parametrize("p", configs.CFG3_PATHS)
def test_config3_long_chains(ctx, cfg3, p):
    hg, G = cfg3
    mp = configs.metapath_from_letters(p, "loc")
    assert 8 <= mp.pathLen <= 12
    rng = np.random.default_rng(3)
    users = np.sort(rng.choice(int(hg.dimensions[configs.LU]), 256, replace=False)).astype(np.int32)
    mats = O.chain_matrices(hg, mp, users)
    want, wnnz, _ = O.plan_digest(mats, O.left_to_right(mp.pathLen), 256)
    src = (users + hg.type_offset[configs.LU]).astype(np.int32)
    try:
        for pc in (0, 1):
            ctx.set_option("plan_cost", pc)
            nnz, dig, _ = G.run_blocked(mp, 8 << 30, src_rows=src, digest=True)
            assert nnz == wnnz, (p, pc)
            assert np.array_equal(dig, want), f"{p} plan_cost={pc}: " + first_mismatch(dig, want)
    finally:
        ctx.set_option("plan_cost", 0)


@pytest.fixture(scope="module")
def cfg1(ctx):
    hg = configs.CFG1.build()
    G = smcm.Graph(ctx, hg)
    yield hg, G
    G.close()


def test_config4_batch(ctx, cfg1):
    hg, G = cfg1
    batch = configs.cfg4_batch(hg.dimensions, 512)
    off = hg.type_offset
    paths = [q.path for q in batch]
    srcs = [(q.src_rows + off[q.path.vertexTypes[0]]).astype(np.int32) for q in batch]
    ref = []
    for mp, s in zip(paths, srcs):
        r, _ = G.run(mp, s)
        ref.append((r.nnz, r.row_digest()))
        r.free()
    saved = ctx.get_option("batch_small_products")
    try:
        for thr in (0, 1 << 30):  # executor off / on for every job below 2^30 estimated products
            ctx.set_option("batch_small_products", thr)
            nnz, dig, _, st = G.run_batch(paths, srcs, digests=True)
            for i, (n, d) in enumerate(ref):
                what = f"threshold {thr}, chain {i} {paths[i].toString()}"
                assert nnz[i] == n, what
                assert np.array_equal(dig[i], d), what
            if thr == 0:
                assert st["small_jobs"] == 0, st
            else:
                assert st["small_jobs"] > st["jobs"] // 2, st
    finally:
        ctx.set_option("batch_small_products", saved)


def test_config5_stream(ctx, cfg1, cfg3):
    hg1, G1 = cfg1
    hg3, G3 = cfg3
    stream = configs.cfg5_stream(hg1.dimensions, hg3.dimensions, n_queries=48, qps=5.0)
    hgs, Gs = [hg1, hg3], [G1, G3]
    queries = [(q.graph, q.path, (q.src_rows + hgs[q.graph].type_offset[q.path.vertexTypes[0]]).astype(np.int32))
               for q in stream]
    assert {q.graph for q in stream} == {0, 1}
    alone = []
    for gi, mp, src in queries:
        n, d, _ = Gs[gi].run_blocked(mp, 32 << 30, src_rows=src, digest=True)
        alone.append((n, d))
    srv = smcm.Server(ctx, Gs, plan_cost=1)  # the default serving configuration
    try:
        outs, digs = srv.replay(queries, np.array([q.arrival for q in stream]))
        hits = srv.stats()["cache_hits"]
    finally:
        srv.close()
    for i, (o, d, (n, want)) in enumerate(zip(outs, digs, alone)):
        assert o["status"] == 0, (i, o)
        assert o["nnz"] == n, i
        assert np.array_equal(d, want), f"query {i}"
    assert hits > 0
