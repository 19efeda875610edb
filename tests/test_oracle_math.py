"""Checks of the oracle that do not come from the reference (its product is Boolean):
uint64 path counts against dense integer products, closed forms of X * X^T, integer SpMV chains,
order independence; the DP against brute-force enumeration including ties; estimator special
cases fixed by the formulas."""
import itertools

import numpy as np
import pytest

import oracle as O
from helpers import random_csr
from resmcm_ari_amd import configs
from resmcm_ari_amd.hin import MetaPath


def dense(csr):
    return csr.to_dense()


def trees(i, j):
    if i == j:
        yield i
        return
    for k in range(i, j):
        for a in trees(i, k):
            for b in trees(k + 1, j):
                yield (a, b, k)


def plan_of(t):
    pairs = []

    def emit(n):
        if isinstance(n, int):
            return n
        a, b, _ = n
        x, y = emit(a), emit(b)
        pairs.append((x, y))
        return -1

    emit(t)
    return pairs


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_spgemm_counts_vs_dense(seed):
    rng = np.random.default_rng(seed)
    A = random_csr(rng, 40, 30, row_len=lambda r: r.integers(0, 12), with_vals=True)
    B = random_csr(rng, 30, 50, row_len=lambda r: r.integers(0, 15), with_vals=True)
    C = O.spgemm(A, B)
    with np.errstate(over="ignore"):
        want = dense(A) @ dense(B)  # uint64 matmul wraps mod 2^64
    got = dense(C)
    np.testing.assert_array_equal(got, want)
    # structure = Boolean product
    bool_want = (dense(A) != 0).astype(np.int64) @ (dense(B) != 0).astype(np.int64) > 0
    np.testing.assert_array_equal(np.diff(C.ptr), bool_want.sum(1))


def test_duplicates_count_as_paths():
    rng = np.random.default_rng(9)
    A = random_csr(rng, 30, 20, row_len=lambda r: r.integers(0, 15), dup=True, sorted_rows=False)
    B = random_csr(rng, 20, 25, row_len=lambda r: r.integers(0, 6))
    C = O.spgemm(A, B)
    np.testing.assert_array_equal(dense(C), dense(A) @ dense(B))


@pytest.mark.parametrize("p", ["A-P-V-P-A", "P-A-P-T-P", "T-P-A-P-V", "A-P-T-P-V-P"])
def test_chain_vs_dense_all_orders(p):
    hg = configs.tiny_bib(seed=5).build()
    mp = configs.metapath_from_letters(p, "bib")
    mats = O.chain_matrices(hg, mp)
    D = [dense(m) for m in mats]
    want = D[0]
    for d in D[1:]:
        want = want @ d
    for t in trees(0, mp.pathLen - 1):
        res = O.execute_plan(mats, plan_of(t))
        np.testing.assert_array_equal(dense(res), want, err_msg=f"{p} plan {plan_of(t)}")


def test_palindrome_closed_forms():
    """A-P-V-P-A = Y Y^T with Y = A-P-V: symmetric, diagonal = squared row norms of Y, total =
    ||Y^T 1||^2."""
    hg = configs.CFG1.build()
    mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
    res, _, _, _ = O.run_chain(hg, mp)
    Y = O.run_chain(hg, configs.metapath_from_letters("A-P-V", "bib"))[0]
    yv = Y.values().astype(np.uint64)
    rows = np.repeat(np.arange(Y.nrows), np.diff(Y.ptr))
    diag_want = np.zeros(Y.nrows, np.uint64)
    np.add.at(diag_want, rows, yv * yv)
    crow = np.repeat(np.arange(res.nrows), np.diff(res.ptr))
    on_diag = res.col == crow
    diag = np.zeros(res.nrows, np.uint64)
    diag[crow[on_diag]] = res.val[on_diag]
    np.testing.assert_array_equal(diag, diag_want)
    colsum = np.zeros(Y.ncols, np.uint64)
    np.add.at(colsum, Y.col, yv)
    assert int(res.val.sum(dtype=np.uint64)) == int((colsum * colsum).sum(dtype=np.uint64))
    # symmetry: (u, v, x) present iff (v, u, x) present
    key = crow.astype(np.int64) * res.ncols + res.col
    tkey = res.col.astype(np.int64) * res.ncols + crow
    o1, o2 = np.argsort(key), np.argsort(tkey)
    np.testing.assert_array_equal(key[o1], tkey[o2])
    np.testing.assert_array_equal(res.val[o1], res.val[o2])


@pytest.mark.parametrize("p", ["A-P-V-P-A", "P-V-P-A", "A-P-T-P-V"])
def test_row_sums_equal_spmv_chain(p):
    hg = configs.CFG1.build()
    mp = configs.metapath_from_letters(p, "bib")
    mats = O.chain_matrices(hg, mp)
    res, _, _, _ = O.run_chain(hg, mp)
    x = np.ones(mats[-1].ncols, np.uint64)
    for m in reversed(mats):
        rows = np.repeat(np.arange(m.nrows), np.diff(m.ptr))
        y = np.zeros(m.nrows, np.uint64)
        np.add.at(y, rows, x[m.col] * m.values())
        x = y
    got = np.zeros(res.nrows, np.uint64)
    np.add.at(got, np.repeat(np.arange(res.nrows), np.diff(res.ptr)), res.val)
    np.testing.assert_array_equal(got, x)


def brute_dp(c):
    """Minimum of sum over internal nodes of (c_ik + c_{k+1,j} + c_ij) over all trees, and the
    lowest optimal root split (exact arithmetic on integer-valued tables)."""
    L = c.shape[0]

    def cost(t):
        if isinstance(t, int):
            return 0, t, t
        a, b, k = t
        ca, i, _ = cost(a)
        cb, _, j = cost(b)
        return ca + cb + c[i, k] + c[k + 1, j] + c[i, j], i, j

    best, best_k = None, None
    for t in trees(0, L - 1):
        v = cost(t)[0]
        k = t[2] if not isinstance(t, int) else None
        if best is None or v < best or (v == best and k < best_k):
            best, best_k = v, k
    return best, best_k


@pytest.mark.parametrize("seed", range(12))
def test_dp_vs_bruteforce(seed):
    rng = np.random.default_rng(seed)
    L = int(rng.integers(2, 8))
    # small integers: exact sums, frequent ties
    c = np.triu(rng.integers(0, 4, size=(L, L)).astype(np.float64), 1)
    cost, m, s, pairs = O.dp(c)
    best, best_k = brute_dp(c)
    assert cost == best
    if L > 1:
        assert s[0, L - 1] == best_k


def test_dp_all_zero_picks_lowest_k():
    L = 6
    c = np.zeros((L, L))
    _, _, s, pairs = O.dp(c)
    # lowest k (= i) for every interval: M0 (M1 (M2 (M3 (M4 M5))))
    assert pairs.tolist() == [[4, 5], [3, -1], [2, -1], [1, -1], [0, -1]]


def test_estimator_special_cases():
    """Leaf = min(1, deg/dim); a row with one neighbour gives 1-(1-r); all-zero right densities
    give 0; a right density of 1 gives 1."""
    A = O.CSR(4, 3, np.array([0, 0, 1, 3, 6], np.int64), np.array([2, 0, 1, 0, 1, 2], np.int32))
    B = O.CSR(3, 2, np.array([0, 2, 2, 3], np.int64), np.array([0, 1, 1], np.int32))
    c, R = O.estimate([A, B], [3, 2])
    np.testing.assert_array_equal(R[0][:, 0], np.minimum(1.0, np.diff(A.ptr) / 3.0))
    rB = np.minimum(1.0, np.diff(B.ptr) / 2.0)  # [1, 0, 0.5]
    np.testing.assert_array_equal(R[1][:, 0], rB)
    assert R[0][0, 1] == 0.0                     # empty row
    assert R[0][1, 1] == 1 - (1 - rB[2])         # one neighbour
    assert R[0][2, 1] == 1.0                     # neighbour with density 1
    want = [1 - np.prod([1 - rB[v] for v in A.col[A.ptr[u]:A.ptr[u + 1]]]) for u in range(4)]
    np.testing.assert_array_equal(R[0][:, 1], want)
    assert c[0, 1] == sum(float(x) * 2.0 for x in R[0][:, 1])


def test_estimator_exact_on_functional_relations():
    """When every row of the right factor has exactly one entry (P-V), r for P-V-P equals
    |P_v| / |P| per paper, and A-P-V estimates 1-(1-1/|V|)^deg exactly in binary64."""
    hg = configs.tiny_bib(seed=2).build()
    mp = configs.metapath_from_letters("A-P-V", "bib")
    mats = O.chain_matrices(hg, mp)
    c, R = O.estimate(mats, O.coldims(hg, mp))
    nV = hg.dimensions[configs.V]
    for u in range(mats[0].nrows):
        deg = mats[0].ptr[u + 1] - mats[0].ptr[u]
        rho = 1.0
        for _ in range(deg):
            rho *= (1 - 1.0 / nV)
        assert R[0][u, 1] == 1 - rho


@pytest.mark.parametrize("name", ["bib_tiny", "loc_tiny", "edge"])
def test_plan_digest_equals_full_execution(name):
    """orc_plan_digest (row blocks of M_0, the root reduced to digests without materialising it)
    equals the digests of the fully executed plan, for the DP plan and left to right, at block
    sizes 1, 7 and all rows; its product count equals the count over a full execution, which is
    checked against a brute-force count from dense operand structures."""
    import test_oracle_golden as TG
    from resmcm_ari_amd.hin import MetaPath

    z, hg = TG.load(name)
    for p in z["paths"]:
        mp = MetaPath(str(p))
        if mp.pathLen < 2:
            continue
        mats = O.chain_matrices(hg, mp)
        c, _ = O.estimate(mats, O.coldims(hg, mp))
        for plan in (O.dp(c)[3], O.left_to_right(mp.pathLen)):
            full = O.execute_plan(mats, plan)
            want = O.row_digest(full)
            count = O.plan_products(mats, plan)
            # brute force: products of a step = sum over A's entries of B's row lengths; a base
            # matrix's entries are counted with their duplicates, a product's are distinct
            mult = [m.to_dense().astype(np.int64) for m in mats]  # values 1: entry multiplicities
            stack, brute = [], 0
            for k, l in plan:
                A = mult[k] if k >= 0 else stack.pop(-2 if l < 0 else -1)
                B = mult[l] if l >= 0 else stack.pop()
                brute += int((A @ B.sum(axis=1)).sum())
                stack.append(((A @ B) != 0).astype(np.int64))
            assert count == brute, p
            for block in (1, 7, mats[0].nrows + 1):
                d, nnz, prods = O.plan_digest(mats, plan, block)
                np.testing.assert_array_equal(d, want, err_msg=f"{p} block {block}")
                assert nnz == full.nnz and prods == count, p
