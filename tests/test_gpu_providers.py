"""The reference's other DP cost providers on the GPU (smcm_cost_estimate, option plan_cost 2-5)
against the oracle restatement (oracle/providers.py), itself pinned to the reference's programs.

Tolerances: MetaAC and DMap compute in binary64 like the reference, but their densities are tree
sums here (the reference's are sequential or OpenMP reductions): rel 1e-12.  MNC and LGraph are
long double in the reference (and the oracle) and binary64 here: rel 1e-9.  Split tables and
chain orders must be identical; every plan's product is bit-identical to the reference plan's.
"""
import numpy as np
import pytest

import oracle as O
import providers as P
from resmcm_ari_amd import smcm
from resmcm_ari_amd.hin import MetaPath

from test_oracle_golden import make_golden

pytestmark = pytest.mark.gpu

GRAPHS = ["bib_tiny", "bib_tiny2", "loc_tiny", "edge", "cfg1"]
RTOL = {"metaac": 1e-12, "dmap": 1e-12, "mnc": 1e-9, "lgraph": 1e-9}
SEED = 0x5EED0000


def _graphs():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = make_golden.GRAPHS[name][0]()
        return cache[name]

    return get


graph_of = _graphs()


def paths_of(name):
    return [MetaPath(p.toString()) for p in make_golden.letters(make_golden.GRAPHS[name][2], make_golden.GRAPHS[name][1])]


def oracle_run(prov, hg, mp, accuracy):
    mats = O.chain_matrices(hg, mp)
    gm, bins = P.global_mats(hg, mp, mats)
    if prov == "metaac":
        return P.metaac(gm, hg.n, accuracy)
    if prov == "mnc":
        return P.mnc(gm, hg.n, accuracy)
    if prov == "dmap":
        return P.dmap(gm, hg.n)
    return P.lgraph(gm, hg.n, SEED, accuracy=accuracy, bins=bins, types=list(mp.vertexTypes))


def close(a, b, rtol, what):
    a, b = float(a), float(b)
    assert abs(a - b) <= rtol * max(abs(a), abs(b), 1e-300), f"{what}: {a!r} vs {b!r}"


@pytest.fixture(scope="module")
def graphs(ctx):
    out = {}
    yield lambda name: out.setdefault(name, smcm.Graph(ctx, graph_of(name)))
    for g in out.values():
        g.close()


@pytest.mark.parametrize("accuracy", [False, True], ids=["planner", "estimate"])
@pytest.mark.parametrize("prov", list(RTOL))
@pytest.mark.parametrize("name", GRAPHS)
def test_provider_matches_oracle(ctx, graphs, name, prov, accuracy):
    hg = graph_of(name)
    G = graphs(name)
    ctx.set_option("lgraph_seed", SEED)
    for mp in paths_of(name):
        L = mp.pathLen
        est, c, m, s = G.cost_estimate(mp, prov, accuracy=accuracy)
        ref = oracle_run(prov, hg, mp, accuracy)
        what = f"{name} {mp.toString()} {prov} accuracy={accuracy}"
        close(est, ref["estimate"], RTOL[prov], what + ": estimate")
        if accuracy and prov in ("lgraph",):
            continue
        for i in range(L):
            for j in range(i + 1, L):
                close(m[i, j], ref["m"][i, j], RTOL[prov], what + f": m[{i},{j}]")
                assert int(s[i, j]) == int(ref["s"][i, j]), what + f": s[{i},{j}]"


@pytest.mark.parametrize("prov", list(RTOL))
def test_plan_cost_runs_provider_plans(ctx, graphs, prov):
    """plan_cost = 2..5: the chain runs the provider's plan; results bit-identical to the
    reference plan's (counts are exact integers mod 2^64)."""
    name = "bib_tiny2"
    hg = graph_of(name)
    G = graphs(name)
    ctx.set_option("lgraph_seed", SEED)
    for mp in paths_of(name):
        if mp.pathLen < 2:
            continue
        ref_nnz, ref_dig, _ = G.run_blocked(mp, 1 << 30, digest=True)
        ctx.set_option("plan_cost", P.PROVIDERS[prov])
        try:
            _, _, st, pairs = G.plan(mp)
            nnz, dig, _ = G.run_blocked(mp, 1 << 30, digest=True)
        finally:
            ctx.set_option("plan_cost", 0)
        ref = oracle_run(prov, hg, mp, False)
        np.testing.assert_array_equal(pairs, np.asarray(ref["pairs"], np.int32).reshape(-1, 2),
                                      err_msg=f"{mp.toString()} {prov}: plan")
        assert nnz == ref_nnz
        np.testing.assert_array_equal(dig, ref_dig, err_msg=f"{mp.toString()} {prov}: result")


def test_plan_cost_rejects_unknown(ctx):
    with pytest.raises(smcm.SmcmError):
        ctx.set_option("plan_cost", 6)


@pytest.mark.parametrize("prov", list(RTOL))
def test_provider_edge_cases(ctx, graphs, prov):
    """One-matrix chains, a relation with no entries (edge graph: V-P-T has no P-T entries for
    some papers; P-A on the isolated paper), and the estimate-only variants agree with the oracle."""
    name = "edge"
    hg = graph_of(name)
    G = graphs(name)
    ctx.set_option("lgraph_seed", SEED)
    for path in ("0-0-1", "2-4-0-2-3", "0-2-3-5-0-0-1", "1-3-0-0-1-3-0-0-1"):
        mp = MetaPath(path)
        for accuracy in (False, True):
            est, c, m, s = G.cost_estimate(mp, prov, accuracy=accuracy)
            ref = oracle_run(prov, hg, mp, accuracy)
            close(est, ref["estimate"], RTOL[prov], f"{path} {prov} accuracy={accuracy}")
            L = mp.pathLen
            if accuracy and prov == "lgraph":
                continue
            for i in range(L):
                for j in range(i + 1, L):
                    assert int(s[i, j]) == int(ref["s"][i, j]), f"{path} {prov}: s[{i},{j}]"
