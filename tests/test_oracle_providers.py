"""The cost-provider oracle (oracle/providers.py) against the reference's own programs and
against what the mathematics fixes.

Fixtures (tests/golden/golden_<graph>_providers.npz, golden_<graph>_<prov>_acc.txt) come from
the reference's planners efficient/{meta,mnc,dm}.cpp (DP tables m_ and s_, via
oracle/provider_harness.cpp) and its estimate programs accuracy/{meta,mnc,dm}.cpp, compiled from
their sources and run with one thread by tests/golden/make_golden_providers.py.  LGraph's ranks
are unseeded in the reference (lg.cpp setN), so it is pinned by properties of the estimator.
"""
import os

import numpy as np
import pytest

import oracle as O
import providers as P
from resmcm_ari_amd.hin import MetaPath

from test_oracle_golden import load as load_graph, make_golden

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRAPHS = ["bib_tiny", "bib_tiny2", "loc_tiny", "edge", "cfg1"]


def acc_lines(name, prov):
    with open(os.path.join(HERE, f"golden_{name}_{prov}_acc.txt")) as f:
        return [line.split() for line in f if line.strip()]


def chain(hg, mp):
    mats = O.chain_matrices(hg, mp)
    gm, bins = P.global_mats(hg, mp, mats)
    return gm, bins


def rel_close(a, b, rtol, what):
    a, b = float(a), float(b)
    assert abs(a - b) <= rtol * max(abs(a), abs(b), 1e-300), f"{what}: {a!r} vs {b!r}"


def check_tables(res, z, pre, L, p, rtol):
    ref_m, ref_s = z[pre + "_m"], z[pre + "_s"]
    for i in range(L):
        for j in range(i + 1, L):
            rel_close(res["m"][i, j], ref_m[i, j], rtol, f"{p}: m[{i},{j}]")
            assert int(res["s"][i, j]) == int(ref_s[i, j]), f"{p}: s[{i},{j}]"


@pytest.mark.parametrize("name", GRAPHS)
def test_metaac_matches_reference(name):
    z = np.load(os.path.join(HERE, f"golden_{name}_providers.npz"))
    hg = make_golden.GRAPHS[name][0]()
    acc = acc_lines(name, "meta")
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        gm, _ = chain(hg, mp)
        res = P.metaac(gm, hg.n)
        if mp.pathLen > 1:
            # densities summed in the same row order: the DP is bit-exact
            check_tables(res, z, f"meta_p{k}", mp.pathLen, p, 0.0)
            assert float(z[f"meta_p{k}_ret"]) == float(np.int64(res["m"][0, mp.pathLen - 1]))
        est = P.metaac(gm, hg.n, accuracy=True)["estimate"]
        assert acc[k][0] == p
        rel_close(est, float(acc[k][1]), 5e-6, f"{p}: MetaAC estimate")


@pytest.mark.parametrize("name", GRAPHS)
def test_mnc_matches_reference(name):
    z = np.load(os.path.join(HERE, f"golden_{name}_providers.npz"))
    hg = make_golden.GRAPHS[name][0]()
    acc = acc_lines(name, "mnc")
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        gm, _ = chain(hg, mp)
        if mp.pathLen > 1:
            res = P.mnc(gm, hg.n)
            check_tables(res, z, f"mnc_p{k}", mp.pathLen, p, 1e-15)
            # the planner returns es_nnz truncated to an integer (edge)
            assert float(z[f"mnc_p{k}_ret"]) == float(int(res["estimate"]))
        est = P.mnc(gm, hg.n, accuracy=True)["estimate"]
        assert acc[k][0] == p
        rel_close(est, float(acc[k][1]), 5e-6, f"{p}: MNC estimate")


@pytest.mark.parametrize("name", GRAPHS)
def test_dmap_matches_reference(name):
    z = np.load(os.path.join(HERE, f"golden_{name}_providers.npz"))
    hg = make_golden.GRAPHS[name][0]()
    acc = acc_lines(name, "dm")
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        gm, _ = chain(hg, mp)
        res = P.dmap(gm, hg.n)
        if mp.pathLen > 1:
            # the density sums are OpenMP reductions in the reference: tolerance
            check_tables(res, z, f"dm_p{k}", mp.pathLen, p, 1e-12)
            rel_close(res["c"][0, mp.pathLen - 1], float(z[f"dm_p{k}_ret"]), 1e-12, f"{p}: DMap c[0][L-1]")
        assert acc[k][0] == p
        rel_close(res["estimate"], float(acc[k][1]), 5e-6, f"{p}: DMap estimate")


# ---- what the mathematics fixes --------------------------------------------------------------

def _bib(seed=3):
    return make_golden.GRAPHS["bib_tiny"][0]()


def test_metaac_density_is_nnz_over_n2():
    hg = _bib()
    mp = MetaPath("1-3-0-2-4")  # A-P-V
    gm, _ = chain(hg, mp)
    for G in gm:
        rel_close(P.metaac_density(G, hg.n), len(G.col_g) / hg.n / hg.n, 1e-14, "density")


def test_mnc_exact_when_left_rows_have_one_entry():
    """MNC's first branch is the exact nnz when every row of A has at most one entry (each
    row of A*B copies one row of B) or every column of B has at most one entry."""
    hg = make_golden.GRAPHS["bib_tiny2"][0]()
    for path in ("0-1-2-4-0", "0-2-3-5-0", "1-3-0-1-2"):  # P-V-P, P-T-P, A-P-V
        mp = MetaPath(path)
        mats = O.chain_matrices(hg, mp)
        gm, _ = P.global_mats(hg, mp, mats)
        a, b = P.mnc_leaf(gm[0], hg.n), P.mnc_leaf(gm[1], hg.n)
        exact = O.spgemm(mats[0], mats[1]).nnz
        if a.max_h_r <= 1 or b.max_h_c <= 1:
            est = P.mnc_sparsity(a, b, hg.n) * P.LD(hg.n) * P.LD(hg.n)
            rel_close(est, exact, 1e-15, f"{path}: MNC exact case")


def test_mnc_propagation_preserves_mass():
    hg = make_golden.GRAPHS["bib_tiny2"][0]()
    mp = MetaPath("1-3-0-1-2-4-0-0-1")  # A-P-V-P-A
    gm, _ = chain(hg, mp)
    a, b = P.mnc_leaf(gm[0], hg.n), P.mnc_leaf(gm[1], hg.n)
    out = P.mnc_propagate(a, b, hg.n)
    # each propagated count is a rounded share of es_nnz: the sums differ by at most one per row
    assert abs(float(out.h_r.sum()) - float(out.es_nnz)) <= np.count_nonzero(a.h_r) + 1
    assert abs(float(out.h_c.sum()) - float(out.es_nnz)) <= np.count_nonzero(b.h_c) + 1


def test_dmap_leaf_and_product_closed_forms():
    n = 600  # three 256-blocks
    # one entry in block (0, 1) of A, one in block (1, 2) of B
    class G:
        pass
    A, B = G(), G()
    A.rows_of_entries, A.col_g = np.array([3]), np.array([300])
    B.rows_of_entries, B.col_g = np.array([260, 261]), np.array([520, 530])
    da, db = P.dmap_leaf(A, n), P.dmap_leaf(B, n)
    assert da[0, 1] == 1 / 256 / 256 and db[1, 2] == 2 / 256 / 256
    prod = P.dmap_product(da, db)
    expect = 1 - (1 - da[0, 1] * db[1, 2]) ** 256
    rel_close(prod[0, 2], expect, 1e-15, "single pair")
    assert np.count_nonzero(prod) == 1
    rel_close(P.dmap_density(da), 1 / 256 / 256 / 9, 1e-15, "density")


def test_lgraph_ranks_are_exponential_and_seeded():
    u = np.arange(20000)
    r = P.lgraph_ranks(11, 0, u)
    assert r.shape == (20000, 32) and (r > 0).all()
    assert abs(r.mean() - 1.0) < 0.01 and abs(r.var() - 1.0) < 0.03
    np.testing.assert_array_equal(r, P.lgraph_ranks(11, 0, u))
    assert not np.array_equal(r, P.lgraph_ranks(12, 0, u))
    assert not np.array_equal(r, P.lgraph_ranks(11, 1, u))


def test_lgraph_estimates_distinct_reachable_rows():
    """31 / (sum of 32 min-ranks) estimates how many rows reach a column (min of k Exp(1)
    ranks is Exp(k)); the accuracy variant sums it over the column type, so on A-P-V-P-A it
    estimates nnz of the product, and the efficient variant weights it by the column's list
    multiplicity (lg.cpp:150-166)."""
    hg = make_golden.GRAPHS["bib_tiny2"][0]()
    mp = MetaPath("1-3-0-1-2-4-0-0-1")  # A-P-V-P-A
    mats = O.chain_matrices(hg, mp)
    gm, bins = P.global_mats(hg, mp, mats)
    types = list(mp.vertexTypes)
    exact = O.run_chain(hg, mp)[0].nnz
    ests = [float(P.lgraph(gm, hg.n, seed, accuracy=True, bins=bins, types=types)["estimate"]) for seed in range(8)]
    assert abs(np.mean(ests) / exact - 1.0) < 0.1
    res = P.lgraph(gm[:1], hg.n, 5)
    G = gm[0]
    # each A-P column (a paper) is reached by its authors: the leaf estimate is close to
    # sum over entries of the in-degree of the entry's column
    indeg = np.bincount(G.col_g, minlength=hg.n)
    target = float((indeg[G.col_g]).sum())
    assert abs(float(res["c"][0, 0]) / target - 1.0) < 0.15
