"""The oracle against fixtures written by the reference's own programs (tests/golden/).

Pins: typed-matrix extraction + Boolean structure of every result row, every r_ij bit-exact,
every c_ij bit-exact (reference run with one thread), the DP split table and chain order, the
reference program's nnz column and the estimator program's output line.
"""
import os

import numpy as np
import pytest

import oracle as O
from resmcm_ari_amd import configs
from resmcm_ari_amd.hin import MetaPath

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

import importlib.util

_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)

GRAPHS = ["bib_tiny", "bib_tiny2", "loc_tiny", "edge", "cfg1"]


def load(name):
    z = np.load(os.path.join(HERE, f"golden_{name}.npz"))
    hg = make_golden.GRAPHS[name][0]()
    return z, hg


def ref_lines(name, kind):
    with open(os.path.join(HERE, f"golden_{name}_{kind}.txt")) as f:
        return [line.split() for line in f if line.strip()]


@pytest.mark.parametrize("name", GRAPHS)
def test_estimator_and_plan_match_reference(name):
    z, hg = load(name)
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        if mp.pathLen < 2:
            continue
        mats = O.chain_matrices(hg, mp)
        c, R = O.estimate(mats, O.coldims(hg, mp))
        np.testing.assert_array_equal(c, z[f"p{k}_c"], err_msg=f"{p}: c table")
        _, _, s, pairs = O.dp(c)
        L = mp.pathLen
        ref_s = z[f"p{k}_s"]
        for i in range(L):
            for j in range(i + 1, L):
                assert s[i, j] == ref_s[i, j], f"{p}: split ({i},{j})"
        np.testing.assert_array_equal(pairs, z[f"p{k}_pairs"], err_msg=f"{p}: chain order")
        for i in range(L):
            for j in range(i, L):
                got = np.ascontiguousarray(R[i][:, j - i])
                if f"p{k}_r_{i}_{j}" in z:
                    np.testing.assert_array_equal(got, z[f"p{k}_r_{i}_{j}"], err_msg=f"{p}: r_{i},{j}")
                else:
                    import hashlib
                    sha = np.frombuffer(hashlib.sha256(got.tobytes()).digest(), np.uint8)
                    np.testing.assert_array_equal(sha, z[f"p{k}_rsha_{i}_{j}"], err_msg=f"{p}: r_{i},{j}")


@pytest.mark.parametrize("name", GRAPHS)
def test_structure_matches_reference(name):
    z, hg = load(name)
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        res, _, _, _ = O.run_chain(hg, mp)
        colmap = hg.bins[mp.vertexTypes[-1]]
        lens = z[f"p{k}_lens"]
        if mp.pathLen == 1:
            # the reference returns matrix 0 itself: duplicates kept, adjacency order
            np.testing.assert_array_equal(np.diff(res.ptr), lens, err_msg=f"{p}: widths")
            if f"p{k}_cols" in z:
                got = np.concatenate([np.sort(colmap[res.col[res.ptr[u]:res.ptr[u + 1]]])
                                      for u in range(res.nrows)] or [np.zeros(0, np.int32)])
                np.testing.assert_array_equal(got, z[f"p{k}_cols"], err_msg=f"{p}: columns")
            continue
        np.testing.assert_array_equal(np.diff(res.ptr), lens, err_msg=f"{p}: row widths")
        gcols = colmap[res.col]
        if f"p{k}_cols" in z:
            np.testing.assert_array_equal(gcols, z[f"p{k}_cols"], err_msg=f"{p}: columns")
        else:
            dig = make_golden.structure_digest(lens.astype(np.int64), gcols.astype(np.int64))
            np.testing.assert_array_equal(dig, z[f"p{k}_digest"], err_msg=f"{p}: structure digest")


@pytest.mark.parametrize("name", GRAPHS)
def test_reference_program_lines(name):
    """nnz column of the reference program (efficient/resmcm.cpp:883) and the estimate line of
    accuracy/res.cpp:698 (default 6 significant digits)."""
    z, hg = load(name)
    nnz_lines = ref_lines(name, "resmcm")
    est_lines = ref_lines(name, "res")
    for k, p in enumerate(z["paths"]):
        mp = MetaPath(str(p))
        assert nnz_lines[k][0] == mp.toString() == est_lines[k][0]
        res, c, _, _ = O.run_chain(hg, mp)
        assert int(nnz_lines[k][1]) == res.nnz, p
        est = c[0, mp.pathLen - 1] if mp.pathLen > 1 else 0.0
        assert f"{est:g}" == est_lines[k][1], p
