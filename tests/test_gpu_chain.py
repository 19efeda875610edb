"""Whole chains on the GPU vs the oracle: typed matrices, estimator (bit-exact per row), DP plan
and the product (bit-exact rowptr / sorted global columns / uint64 counts)."""
import numpy as np
import pytest

import oracle as O
from helpers import assert_same_csr
from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu

C_RTOL = 1e-12  # default tree-ordered c_ij sums vs the reference's one-thread row-order sum


def all_paths(schema, L):
    """Every schema-valid meta-path with L matrices (letters joined by '-')."""
    sch = configs.BIB_SCHEMA if schema == "bib" else configs.LOC_SCHEMA
    names = {v: k for k, v in (configs.BIB_NAMES if schema == "bib" else configs.LOC_NAMES).items()}
    out = []
    stack = [[s] for s in sch]
    while stack:
        p = stack.pop()
        if len(p) == L + 1:
            out.append("-".join(names[t] for t in p))
            continue
        for t in sch[p[-1]]:
            stack.append(p + [t])
    return sorted(out)


@pytest.fixture(scope="module")
def tiny_graphs(ctx):
    gb = configs.tiny_bib().build()
    gl = configs.tiny_loc().build()
    return {"bib": (gb, smcm.Graph(ctx, gb)), "loc": (gl, smcm.Graph(ctx, gl))}


def check_chain(hg, G, mp, src=None, exact_cost=False):
    local = hg.local_ids()
    src_local = None if src is None else local[np.asarray(src)]
    ref, c_ref, R_ref, pairs_ref = O.run_chain(hg, mp, src_rows=src_local)
    r, times = G.run(mp, src_rows=src)
    ptr, col, val = r.to_host()
    bins = hg.bins
    assert_same_csr(ptr, col, val, ref, col_map=bins[mp.vertexTypes[-1]], what=mp.toString())
    rows = r.row_ids()
    expect_rows = bins[mp.vertexTypes[0]] if src is None else np.asarray(src)
    np.testing.assert_array_equal(rows, expect_rows)
    r.free()
    if mp.pathLen > 1:
        c, m, s, pairs = G.plan(mp, src_rows=src)
        if exact_cost:
            np.testing.assert_array_equal(c, c_ref)
        else:
            np.testing.assert_allclose(c, c_ref, rtol=C_RTOL, atol=0)
            np.testing.assert_array_equal(pairs, pairs_ref)
        for i in range(mp.pathLen):
            Ri = G.estimator_rows(mp, i, src_rows=src)
            np.testing.assert_array_equal(Ri, R_ref[i], err_msg=f"{mp} r_{i},* differs")
    return times


@pytest.mark.parametrize("schema,L", [("bib", 1), ("bib", 2), ("bib", 3), ("bib", 4), ("loc", 2), ("loc", 4)])
def test_all_paths_tiny(tiny_graphs, schema, L):
    hg, G = tiny_graphs[schema]
    for p in all_paths(schema, L):
        check_chain(hg, G, configs.metapath_from_letters(p, schema))


def test_long_chain_tiny(tiny_graphs):
    hg, G = tiny_graphs["loc"]
    for p in ["U-L-C-L-Y-L-U-L-S-L-U", "U-L-C-L-U-L-Y-L-S-L-U-L-U", "L-Y-L-C-L-S-L-U-L"]:
        check_chain(hg, G, configs.metapath_from_letters(p, "loc"))


def test_exact_cost_sum(ctx, tiny_graphs):
    """exact_cost_sum reproduces the reference's one-thread c_ij bit for bit."""
    hg, G = tiny_graphs["bib"]
    ctx.set_option("exact_cost_sum", 1)
    try:
        for p in ["A-P-V-P-A", "P-A-P-T-P", "T-P-A-P-V"]:
            check_chain(hg, G, configs.metapath_from_letters(p, "bib"), exact_cost=True)
    finally:
        ctx.set_option("exact_cost_sum", 0)


def test_source_subset_is_slice(tiny_graphs):
    hg, G = tiny_graphs["bib"]
    mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
    full, _ = G.run(mp)
    fptr, fcol, fval = full.to_host()
    rng = np.random.default_rng(5)
    authors = hg.bins[configs.A]
    src = rng.choice(authors, size=17, replace=False)
    check_chain(hg, G, mp, src=src)
    sub, _ = G.run(mp, src_rows=src)
    sptr, scol, sval = sub.to_host()
    local = hg.local_ids()
    for q, v in enumerate(src):
        u = local[v]
        np.testing.assert_array_equal(scol[sptr[q]:sptr[q + 1]], fcol[fptr[u]:fptr[u + 1]])
        np.testing.assert_array_equal(sval[sptr[q]:sptr[q + 1]], fval[fptr[u]:fptr[u + 1]])


def test_every_plan_same_result(tiny_graphs):
    """Order independence: all parenthesisations give identical bits."""
    hg, G = tiny_graphs["bib"]
    mp = configs.metapath_from_letters("A-P-T-P-V-P", "bib")
    base, _ = G.run(mp)
    b = base.to_host()
    L = mp.pathLen
    for plan in enumerate_plans(L):
        r, _ = G.run(mp, plan=plan)
        o = r.to_host()
        for x, y in zip(b, o):
            np.testing.assert_array_equal(x, y)


def enumerate_plans(L):
    """All post-order plans (get_optimal_chain_order format) for L matrices."""

    def trees(i, j):
        if i == j:
            yield ("leaf", i)
            return
        for k in range(i, j):
            for lt in trees(i, k):
                for rt in trees(k + 1, j):
                    yield ("node", lt, rt)

    out = []
    for t in trees(0, L - 1):
        pairs = []

        def emit(n):
            if n[0] == "leaf":
                return n[1]
            a = emit(n[1])
            b = emit(n[2])
            pairs.append((a, b))
            return -1

        emit(t)
        out.append(pairs)
    return out


def test_bad_inputs(ctx, tiny_graphs):
    hg, G = tiny_graphs["bib"]
    from resmcm_ari_amd.hin import MetaPath
    with pytest.raises(smcm.SmcmError) as e:
        G.run(MetaPath(vertexTypes=[1, 9], edgeTypes=[3]))
    assert e.value.code == smcm.SMCM_EINVAL
    with pytest.raises(smcm.SmcmError):
        G.run(configs.metapath_from_letters("A-P-A", "bib"), src_rows=[0])  # vertex 0 is a paper
    with pytest.raises(smcm.SmcmError):
        G.run(configs.metapath_from_letters("A-P-A", "bib"), plan=[[0, 0]])


@pytest.fixture(scope="module")
def cfg1(ctx):
    hg = configs.CFG1.build()
    return hg, smcm.Graph(ctx, hg)


@pytest.mark.parametrize("p", ["A-P-V-P-A", "P-T-P-V", "P-V-P-A", "V-P-A-P-V"])
def test_config1(cfg1, p):
    hg, G = cfg1
    check_chain(hg, G, configs.metapath_from_letters(p, "bib"))


def test_cli_matches_reference_lines(tmp_path):
    """The drop-in program writes the reference's line format; its nnz column equals the
    reference program's on the same inputs (fixtures from the reference build)."""
    import os
    import subprocess

    from resmcm_ari_amd import _build
    import test_oracle_golden as TG

    for name in ["edge", "bib_tiny", "cfg1"]:
        z, hg = TG.load(name)
        files = [str(tmp_path / f"{name}_{x}.txt") for x in ("graph", "vertex", "edge")]
        hg.write_text(*files)
        mpf = tmp_path / f"{name}_mp.txt"
        mpf.write_text("".join(str(p) + "\n" for p in z["paths"]))
        out = tmp_path / f"{name}_out.txt"
        subprocess.run([os.path.join(_build.LIBDIR, "smcm_resmcm"), "8", "1", *files, str(mpf), str(out)],
                       check=True, stdout=subprocess.DEVNULL)
        got = [line.split() for line in out.read_text().splitlines()]
        want = TG.ref_lines(name, "resmcm")
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert len(g) == 6
            assert g[0] == w[0] and g[1] == w[1], (name, g, w)


@pytest.mark.parametrize("prov", ["meta", "mnc", "dm", "lg"])
def test_cli_cost_providers(tmp_path, prov):
    """SMCM_PLAN_COST makes the drop-in program plan with one of the reference's other cost
    providers (efficient/{meta,mnc,dm,lg}): same argv and line format, and the nnz column is
    the reference program's (every plan gives the same result)."""
    import os
    import subprocess

    from resmcm_ari_amd import _build
    import test_oracle_golden as TG

    name = "bib_tiny"
    z, hg = TG.load(name)
    files = [str(tmp_path / f"{x}.txt") for x in ("graph", "vertex", "edge")]
    hg.write_text(*files)
    mpf = tmp_path / "mp.txt"
    mpf.write_text("".join(str(p) + "\n" for p in z["paths"]))
    out = tmp_path / "out.txt"
    env = dict(os.environ, SMCM_PLAN_COST=prov)
    subprocess.run([os.path.join(_build.LIBDIR, "smcm_resmcm"), "8", "1", *files, str(mpf), str(out)],
                   check=True, stdout=subprocess.DEVNULL, env=env)
    got = [line.split() for line in out.read_text().splitlines()]
    want = TG.ref_lines(name, "resmcm")
    assert [(g[0], g[1]) for g in got] == [(w[0], w[1]) for w in want]


@pytest.mark.parametrize("plan_cost", [0, 1])
def test_products_count_is_the_oracles(ctx, cfg1, plan_cost):
    """The executed product count (last_stats()['products'], the benchmark's numerator) equals
    the oracle's count for the executed plan: sum over its SpGEMMs of sum_{(u,k) in A} |row_k(B)|,
    on every fixture meta-path and on config 1; with plan_cost 0 the plan is also the oracle DP's."""
    import test_oracle_golden as TG
    from resmcm_ari_amd.hin import MetaPath

    cases = []
    for name in ["edge", "bib_tiny", "bib_tiny2", "loc_tiny"]:
        z, hg = TG.load(name)
        cases.append((hg, smcm.Graph(ctx, hg), [MetaPath(str(p)) for p in z["paths"]]))
    hg1, G1 = cfg1
    cases.append((hg1, G1, [configs.metapath_from_letters(p, "bib")
                            for p in ["A-P-V-P-A", "A-P-T-P-A", "P-T-P-A-P-V", "T-P-A-P-T"]]))
    ctx.set_option("plan_cost", plan_cost)
    try:
        checked = 0
        for hg, G, paths in cases:
            for mp in paths:
                if mp.pathLen < 2:
                    continue
                r, _ = G.run(mp)
                st = ctx.last_stats()
                r.free()
                pairs = np.asarray(st["pairs"], np.int32).reshape(-1, 2)
                mats = O.chain_matrices(hg, mp)
                if plan_cost == 0:
                    c, _ = O.estimate(mats, O.coldims(hg, mp))
                    np.testing.assert_array_equal(pairs, O.dp(c)[3], err_msg=mp.toString())
                assert st["products"] == O.plan_products(mats, pairs), mp.toString()
                checked += 1
        assert checked > 20
    finally:
        ctx.set_option("plan_cost", 0)


def test_res_program_matches_reference_lines(tmp_path):
    """The estimate-only drop-in (smcm_res, accuracy/res.cpp:612-709) writes the reference
    program's lines exactly: meta-path string and estimate with 6 significant digits (:697),
    against the lines the reference's own res program wrote at one thread on the same inputs."""
    import os
    import subprocess

    from resmcm_ari_amd import _build
    import test_oracle_golden as TG

    for name in TG.GRAPHS:
        z, hg = TG.load(name)
        files = [str(tmp_path / f"{name}_{x}.txt") for x in ("graph", "vertex", "edge")]
        hg.write_text(*files)
        mpf = tmp_path / f"{name}_mp.txt"
        mpf.write_text("".join(str(p) + "\n" for p in z["paths"]))
        out = tmp_path / f"{name}_res.txt"
        subprocess.run([os.path.join(_build.LIBDIR, "smcm_res"), "8", "1", *files, str(mpf), str(out)],
                       check=True, stdout=subprocess.DEVNULL)
        got = [line.split() for line in out.read_text().splitlines()]
        assert got == TG.ref_lines(name, "res"), name


def test_estimate_is_the_oracles(ctx, tiny_graphs, cfg1):
    """smcm_estimate (accuracy mode, c_[0][L-1] of accuracy/res.cpp:494-523) equals the oracle's
    one-thread row-order sum bit for bit under exact_cost_sum, within C_RTOL otherwise; also on
    source subsets."""
    hg1, G1 = cfg1
    cases = [(tiny_graphs["bib"], "bib", ["A-P-V-P-A", "P-T-P-A-P-V", "T-P-A"]),
             (tiny_graphs["loc"], "loc", ["U-L-C-L-Y-L-U-L-S-L-U", "L-U-L"]),
             ((hg1, G1), "bib", ["A-P-V-P-A", "A-P-T-P-A", "V-P-T-P-V"])]
    rng = np.random.default_rng(11)
    for exact in (1, 0):
        ctx.set_option("exact_cost_sum", exact)
        try:
            for (hg, G), schema, paths in cases:
                for p in paths:
                    mp = configs.metapath_from_letters(p, schema)
                    t0 = mp.vertexTypes[0]
                    for src in (None, np.sort(rng.choice(hg.bins[t0], size=max(1, len(hg.bins[t0]) // 3),
                                                         replace=False))):
                        local = None if src is None else hg.local_ids()[src]
                        mats = O.chain_matrices(hg, mp, local)
                        c, _ = O.estimate(mats, O.coldims(hg, mp))
                        got = G.estimate(mp, src_rows=src)
                        if exact:
                            assert got == c[0, -1], (p, got, c[0, -1])
                        else:
                            np.testing.assert_allclose(got, c[0, -1], rtol=C_RTOL, atol=0, err_msg=p)
        finally:
            ctx.set_option("exact_cost_sum", 0)
