"""Row densities of the estimator on rows long enough for the hub-row path (more than 32
entries: factors gathered in 2048-entry items, then multiplied in row order), against the oracle,
bit for bit.  Row lengths straddle the light/heavy boundary (32, 33), one and several items
(2048, 2049, 5000), several LDS chunks per row, and rows whose running products reach exactly 0
(a neighbour of density 1) so that the early exit is taken."""
import numpy as np
import pytest

import oracle as O
from resmcm_ari_amd import smcm
from resmcm_ari_amd.hin import HeteroGraph, MetaPath

pytestmark = pytest.mark.gpu

NX, NY, NZ = 40, 6000, 64
# edge types: X->Y 0, Y->X 1, Y->Z 2, Z->Y 3
HUB_LENGTHS = [0, 1, 31, 32, 33, 64, 65, 2047, 2048, 2049, 4096, 5000, 5999]


def build_graph(seed, saturate):
    """X rows of the listed lengths over Y; every Y row has 1-4 Z neighbours, except (when
    `saturate`) Y vertex 0, which is adjacent to every Z vertex (density 1: its factor is 0)."""
    rng = np.random.default_rng(seed)
    n = NX + NY + NZ
    vtype = np.concatenate([np.zeros(NX), np.ones(NY), np.full(NZ, 2)]).astype(np.int32)
    xy = []
    for x in range(NX):
        d = HUB_LENGTHS[x % len(HUB_LENGTHS)]
        ys = np.sort(rng.choice(NY, size=d, replace=False))
        if saturate and x % 3 == 0 and d > 40 and 0 not in ys:
            ys = np.sort(np.concatenate([ys[1:], [0]]))  # the saturated vertex somewhere in the row
        xy.append(ys)
    yz = []
    for y in range(NY):
        if saturate and y == 0:
            yz.append(np.arange(NZ))
        else:
            yz.append(np.sort(rng.choice(NZ, size=int(rng.integers(1, 5)), replace=False)))
    adj = [[] for _ in range(n)]
    for x, ys in enumerate(xy):
        for y in ys:
            adj[x].append((NX + int(y), 0))
            adj[NX + int(y)].append((x, 1))
    for y, zs in enumerate(yz):
        for z in zs:
            adj[NX + y].append((NX + NY + int(z), 2))
            adj[NX + NY + int(z)].append((NX + y, 3))
    ptr = np.zeros(n + 1, np.int64)
    nbr, et = [], []
    for v in range(n):
        lst = sorted(adj[v])
        ptr[v + 1] = ptr[v] + len(lst)
        nbr += [a for a, _ in lst]
        et += [b for _, b in lst]
    return HeteroGraph.from_arrays(vtype, ptr, np.array(nbr, np.int32), np.array(et, np.int32))


@pytest.mark.parametrize("saturate", [False, True], ids=["plain", "early_exit"])
@pytest.mark.parametrize("path", ["0-0-1-2-2", "0-0-1-2-2-3-1-1-0", "0-0-1-1-0-0-1-2-2"])
def test_hub_rows_bit_exact(ctx, path, saturate):
    hg = build_graph(17, saturate)
    G = smcm.Graph(ctx, hg)
    try:
        mp = MetaPath(path)
        _, c_ref, R_ref, _ = O.run_chain(hg, mp)
        for i in range(mp.pathLen):
            Ri = G.estimator_rows(mp, i)
            np.testing.assert_array_equal(Ri, R_ref[i], err_msg=f"{path} r_{i},* differs")
        c, _, _, _ = G.plan(mp)
        np.testing.assert_allclose(c, c_ref, rtol=1e-12, atol=0)
        if saturate:
            # the density-1 vertex sits in long rows: their running products become exactly 0
            mats = O.chain_matrices(hg, mp)
            M0 = mats[0]
            long_rows = [u for u in range(M0.nrows) if M0.ptr[u + 1] - M0.ptr[u] > 32]
            assert any(0 in M0.col[M0.ptr[u]:M0.ptr[u + 1]] for u in long_rows)
    finally:
        G.close()
