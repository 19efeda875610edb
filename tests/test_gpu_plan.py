"""Product-aware planning (option plan_cost = 1) on the GPU: the product-cost table against the
oracle's restatement, the DP against the oracle's product DP on the GPU's tables, and results
bit-identical to the reference plan's (every plan gives the same exact integer product)."""
import numpy as np
import pytest

import oracle as O
from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu

P_RTOL = 1e-7  # logarithms summed in 2^-32 fixed point on the GPU vs float64 on the CPU

BIB = ["A-P-V-P-A", "P-A-P-T-P", "T-P-A-P-V-P", "A-P-T-P-A-P-V", "V-P-T-P-T-P-A-P-V", "P-T-P"]
LOC = ["U-L-C-L-Y-L-U-L-S-L-U", "U-L-Y-L-U-L-U", "L-C-L-S-L"]


@pytest.fixture(scope="module")
def graphs(ctx):
    gb = configs.tiny_bib().build()
    gl = configs.tiny_loc().build()
    return {"bib": (gb, smcm.Graph(ctx, gb)), "loc": (gl, smcm.Graph(ctx, gl))}


@pytest.fixture()
def product_plans(ctx):
    ctx.set_option("plan_cost", 1)
    yield
    ctx.set_option("plan_cost", 0)


def cases():
    for p in BIB:
        yield "bib", p, None
        yield "bib", p, [1, 4, 5]
    for p in LOC:
        yield "loc", p, None
        yield "loc", p, [0, 2, 3, 7]


@pytest.mark.parametrize("schema,p,sub", list(cases()))
def test_products_table_vs_oracle(graphs, schema, p, sub):
    hg, G = graphs[schema]
    mp = configs.metapath_from_letters(p, schema)
    t0 = mp.vertexTypes[0]
    src = None if sub is None else (hg.type_offset[t0] + np.asarray(sub)).astype(np.int32)
    mats = O.chain_matrices(hg, mp, None if sub is None else np.asarray(sub, np.int32))
    _, R = O.estimate(mats, O.coldims(hg, mp))
    want = O.products_estimate(mats, R, O.coldims(hg, mp))
    got = G.products_estimate(mp, src)
    np.testing.assert_allclose(got, want, rtol=P_RTOL, atol=1e-9)


@pytest.mark.parametrize("schema,p,sub", list(cases()))
def test_product_plan_same_result(ctx, graphs, product_plans, schema, p, sub):
    hg, G = graphs[schema]
    mp = configs.metapath_from_letters(p, schema)
    t0 = mp.vertexTypes[0]
    src = None if sub is None else (hg.type_offset[t0] + np.asarray(sub)).astype(np.int32)
    c, m, s, pairs = G.plan(mp, src)
    P = G.products_estimate(mp, src)
    _, _, want_pairs = O.dp_products(c, P)
    np.testing.assert_array_equal(pairs, want_pairs)
    r1, _ = G.run(mp, src)
    got = r1.to_host()
    r1.free()
    ctx.set_option("plan_cost", 0)
    r0, _ = G.run(mp, src)
    ref = r0.to_host()
    r0.free()
    ctx.set_option("plan_cost", 1)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def plan_splits(pairs, L):
    """(i, k, j, is_root) of every product of a plan."""
    st, out = [], []
    for k, l in np.asarray(pairs).reshape(-1, 2).tolist():
        a = (k, k) if k >= 0 else st.pop(-2 if l < 0 else -1)
        b = (l, l) if l >= 0 else st.pop()
        st.append((a[0], b[1]))
        out.append((a[0], a[1], b[1], a[0] == 0 and b[1] == L - 1))
    return out


@pytest.mark.parametrize("schema,p", [("bib", p) for p in BIB if p.count("-") >= 3] + [("loc", p) for p in LOC])
def test_product_plan_memory_bound(ctx, graphs, product_plans, schema, p):
    """A budget just below the largest product of the unbounded product plan: the plan is the
    oracle's bounded product DP on the GPU's tables, and the result is unchanged."""
    hg, G = graphs[schema]
    mp = configs.metapath_from_letters(p, schema)
    L = mp.pathLen
    c, _, _, pairs = G.plan(mp)
    P = G.products_estimate(mp)
    need = max(O.split_bytes(c, i, k, j, r) for i, k, j, r in plan_splits(pairs, L))
    if need < 2:
        pytest.skip("the unbounded plan needs no intermediate")
    budget = int(need) - 1
    old = ctx.get_option("budget")
    ctx.set_option("budget", budget)
    try:
        _, _, _, got = G.plan(mp)
        r, _ = G.run(mp)
        res = r.to_host()
        r.free()
    finally:
        ctx.set_option("budget", old)
    _, _, want = O.dp_products(c, P, float(budget))
    np.testing.assert_array_equal(got, want)
    ctx.set_option("plan_cost", 0)
    r0, _ = G.run(mp)
    ref = r0.to_host()
    r0.free()
    ctx.set_option("plan_cost", 1)
    for a, b in zip(res, ref):
        np.testing.assert_array_equal(a, b)
