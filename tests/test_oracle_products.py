"""The product-cost estimate (column densities of X_ik against the estimator's row densities of
X_{k+1,j}) and the product-aware DP, checked on the CPU: exact cases of the estimate (two leaves;
a bijective middle factor) and the DP against brute-force enumeration of every parenthesisation."""
import numpy as np
import pytest

import oracle as O
from helpers import random_csr
from resmcm_ari_amd import configs


@pytest.mark.parametrize("p", ["A-P-V-P-A", "P-A-P-T-P", "T-P-A-P-V", "A-P-T-P-V-P"])
def test_two_leaves_exact(p):
    """P[i, i, i+1] = sum_v colcount_i[v] * deg_{i+1}[v] = products(M_i, M_{i+1}) (no row holds more
    entries than its column type has vertices in these graphs)."""
    hg = configs.tiny_bib(seed=4).build()
    mp = configs.metapath_from_letters(p, "bib")
    for rows in (None, np.array([0, 3, 5, 17, 29], np.int32)):
        mats = O.chain_matrices(hg, mp, rows)
        _, R = O.estimate(mats, O.coldims(hg, mp))
        P = O.products_estimate(mats, R, O.coldims(hg, mp))
        for i in range(mp.pathLen - 1):
            want = O.products(mats[i], mats[i + 1])
            assert P[i, i, i + 1] == pytest.approx(want, rel=1e-12, abs=1e-9), (p, i)


def test_bijective_middle_factor():
    """With M_1 a permutation, X_01 = M_0 M_1 has M_0's column counts permuted, so the estimate of
    X_01 * M_2 is the exact product count."""
    rng = np.random.default_rng(3)
    M0 = random_csr(rng, 30, 20, row_len=lambda r: r.integers(0, 8))
    perm = rng.permutation(20).astype(np.int32)
    M1 = O.CSR(20, 20, np.arange(21, dtype=np.int64), perm, None)
    M2 = random_csr(rng, 20, 25, row_len=lambda r: r.integers(0, 10))
    mats = [M0, M1, M2]
    coldims = np.array([20, 20, 25], np.int32)
    _, R = O.estimate(mats, coldims)
    P = O.products_estimate(mats, R, coldims)
    want = O.products(O.spgemm(M0, M1), M2)
    assert P[0, 1, 2] == pytest.approx(want, rel=1e-9)


def trees(i, j):
    if i == j:
        yield i
        return
    for k in range(i, j):
        for a in trees(i, k):
            for b in trees(k + 1, j):
                yield (a, b, i, k, j)


def tree_cost(t, c, P, root=True):
    if isinstance(t, int):
        return 0.0
    a, b, i, k, j = t
    return tree_cost(a, c, P, False) + tree_cost(b, c, P, False) + P[i, k, j] + (0.0 if root else c[i, j])


def plan_cost(pairs, c, P, L):
    st = []
    total = 0.0
    for q, (k, l) in enumerate(pairs):
        a = (k, k) if k >= 0 else st.pop(-2 if l < 0 else -1)
        b = (l, l) if l >= 0 else st.pop()
        i, kk, j = a[0], a[1], b[1]
        total += P[i, kk, j] + (0.0 if (i == 0 and j == L - 1) else c[i, j])
        st.append((i, j))
    return total


@pytest.mark.parametrize("seed", range(10))
def test_product_dp_vs_bruteforce(seed):
    rng = np.random.default_rng(seed)
    L = int(rng.integers(2, 7))
    c = rng.integers(0, 5, size=(L, L)).astype(np.float64) * 10.0 ** rng.integers(0, 3)
    P = rng.integers(0, 5, size=(L, L, L)).astype(np.float64)  # small ints: many ties
    m, s, pairs = O.dp_products(c, P)
    best = min(tree_cost(t, c, P) for t in trees(0, L - 1))
    assert m[0, L - 1] == best
    assert plan_cost(pairs.tolist(), c, P, L) == best


def tree_splits(t, root=True):
    """(i, k, j, is_root) of every product of tree t."""
    if isinstance(t, int):
        return []
    a, b, i, k, j = t
    return tree_splits(a, False) + tree_splits(b, False) + [(i, k, j, root)]


@pytest.mark.parametrize("seed", range(40))
def test_product_dp_memory_bound_vs_bruteforce(seed):
    """Products whose estimated bytes exceed the budget are never planned when some plan avoids
    them all; when none does, the plan is the unbounded optimum."""
    rng = np.random.default_rng(100 + seed)
    L = int(rng.integers(3, 7))
    c = rng.integers(0, 50, size=(L, L)).astype(np.float64)
    P = rng.integers(0, 20, size=(L, L, L)).astype(np.float64)
    all_trees = list(trees(0, L - 1))

    def need(t):
        return max(O.split_bytes(c, i, k, j, r) for i, k, j, r in tree_splits(t))

    # the budget just below the largest product of an unbounded optimum (that plan is then
    # excluded), or at a quantile of every tree's largest product
    opt = min(all_trees, key=lambda t: tree_cost(t, c, P))
    budget = need(opt) - 0.5 if seed % 2 == 0 else float(np.quantile([need(t) for t in all_trees], 0.4))
    ok = [t for t in all_trees if need(t) <= budget]
    pool = ok if ok else all_trees
    best = min(tree_cost(t, c, P) for t in pool)
    m, s, pairs = O.dp_products(c, P, budget)
    assert m[0, L - 1] == best
    assert plan_cost(pairs.tolist(), c, P, L) == best
    if ok:
        st = []
        for k, l in pairs.tolist():
            a = (k, k) if k >= 0 else st.pop(-2 if l < 0 else -1)
            b = (l, l) if l >= 0 else st.pop()
            st.append((a[0], b[1]))
            assert O.split_bytes(c, a[0], a[1], b[1], a[0] == 0 and b[1] == L - 1) <= budget
    # no bound: the plain product DP
    m0, _, _ = O.dp_products(c, P)
    assert m0[0, L - 1] == min(tree_cost(t, c, P) for t in all_trees)
