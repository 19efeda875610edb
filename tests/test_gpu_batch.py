"""Batched form (smcm_batch_run) on the GPU: every chain's result, nnz and row digests equal the
per-chain path's bit for bit, through the persistent executor (all jobs small), the per-chain
path (all jobs large), the rerun path (an arena too small for the batch) and every executor
shape (workgroup size, sort capacity, tile rows, window width).  Batches repeat meta-paths with
overlapping, disjoint and empty source sets, so jobs serve several chains."""
import numpy as np
import pytest

from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu


def mid_bib():
    return configs.BibSpec(nA=2000, nP=3000, nV=20, nT=200, lam_auth=2.0, terms_per_paper=8,
                           seed=configs.SEED_BASE + 77)


@pytest.fixture(scope="module")
def graphs(ctx):
    out = {}
    for name, spec in (("tiny", configs.tiny_bib()), ("mid", mid_bib())):
        hg = spec.build()
        out[name] = (hg, smcm.Graph(ctx, hg))
    return out


def make_batch(hg, n, seed, lmax=5, smax=300, whole=False):
    rng = np.random.default_rng(seed)
    paths, srcs = [], []
    for _ in range(n):
        L = int(rng.integers(1, lmax + 1))
        mp = configs.random_walk_path(rng, L, "bib")
        t0 = mp.vertexTypes[0]
        d0 = int(hg.dimensions[t0])
        k = int(rng.integers(0, min(d0, smax) + 1))
        rows = np.sort(rng.choice(d0, size=k, replace=False)).astype(np.int32)
        paths.append(mp)
        srcs.append((rows + hg.type_offset[t0]).astype(np.int32))
    return paths, (None if whole else srcs)


def reference(G, paths, srcs):
    out = []
    for i, mp in enumerate(paths):
        r, _ = G.run(mp, None if srcs is None else srcs[i])
        out.append((r.to_host(), r.row_digest(), r.row_ids()))
        r.free()
    return out


def check(ctx, G, paths, srcs, ref, expect_small=None):
    nnz, dig, res, st = G.run_batch(paths, srcs, digests=True, results=True)
    for i, mp in enumerate(paths):
        (ptr, col, val), rdig, rows = ref[i]
        what = f"chain {i} {mp.toString()}"
        assert nnz[i] == len(col), what
        np.testing.assert_array_equal(dig[i], rdig, err_msg=what)
        gp, gc, gv = res[i].to_host()
        np.testing.assert_array_equal(gp, ptr, err_msg=what)
        np.testing.assert_array_equal(gc, col, err_msg=what)
        np.testing.assert_array_equal(gv, val, err_msg=what)
        np.testing.assert_array_equal(res[i].row_ids(), rows, err_msg=what)
        res[i].free()
    if expect_small is not None:
        assert st["small_jobs"] == expect_small, st
    return st


@pytest.fixture()
def batch_opts(ctx):
    keys = ["batch_small_products", "batch_threads", "batch_sort_cap", "batch_tile_rows", "batch_log2w",
            "batch_arena_mb"]
    saved = {k: ctx.get_option(k) for k in keys}
    yield
    for k, v in saved.items():
        ctx.set_option(k, v)


@pytest.mark.parametrize("name,seed", [("tiny", 1), ("tiny", 2), ("mid", 3)])
def test_batch_executor_all_small(ctx, graphs, batch_opts, name, seed):
    hg, G = graphs[name]
    paths, srcs = make_batch(hg, 60, seed)
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", 1 << 40)
    n_multi = len({(tuple(mp.vertexTypes), tuple(mp.edgeTypes)) for mp in paths if mp.pathLen > 1})
    check(ctx, G, paths, srcs, ref, expect_small=n_multi)


def test_batch_whole_rows(ctx, graphs, batch_opts):
    hg, G = graphs["tiny"]
    paths, srcs = make_batch(hg, 30, 5, whole=True)
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", 1 << 40)
    check(ctx, G, paths, srcs, ref)


def test_batch_all_large(ctx, graphs, batch_opts):
    hg, G = graphs["mid"]
    paths, srcs = make_batch(hg, 25, 6)
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", 0)
    check(ctx, G, paths, srcs, ref, expect_small=0)


def test_batch_arena_overflow_reruns(ctx, graphs, batch_opts):
    hg, G = graphs["mid"]
    paths, srcs = make_batch(hg, 40, 7, smax=3000)
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", 1 << 40)
    ctx.set_option("batch_arena_mb", 1)
    st = check(ctx, G, paths, srcs, ref)
    assert st["rerun_jobs"] > 0, st


@pytest.mark.parametrize("threads,cap,tile,log2w", [(256, 512, 1, 8), (256, 2048, 7, 10), (512, 512, 64, 8),
                                                    (256, 2048, 32, 13), (512, 1024, 16, 9)])
def test_batch_executor_shapes(ctx, graphs, batch_opts, threads, cap, tile, log2w):
    hg, G = graphs["mid"]
    paths, srcs = make_batch(hg, 40, 11 + tile, smax=2000)
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", 1 << 40)
    ctx.set_option("batch_threads", threads)
    ctx.set_option("batch_sort_cap", cap)
    ctx.set_option("batch_tile_rows", tile)
    ctx.set_option("batch_log2w", log2w)
    check(ctx, G, paths, srcs, ref)


def test_batch_empty_and_single(ctx, graphs, batch_opts):
    hg, G = graphs["tiny"]
    nnz, dig, res, st = G.run_batch([], [], digests=True, results=True)
    assert len(nnz) == 0 and st["chains"] == 0
    mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
    empty = np.zeros(0, np.int32)
    ref = reference(G, [mp], [empty])
    ctx.set_option("batch_small_products", 1 << 40)
    check(ctx, G, [mp], [empty], ref)


@pytest.mark.parametrize("small", [True, False])
def test_batch_shared_paths(ctx, graphs, batch_opts, small):
    """Many chains over a few meta-paths: overlapping, nested, duplicate and disjoint source sets."""
    hg, G = graphs["mid"]
    rng = np.random.default_rng(21)
    letters = ["A-P-V-P-A", "P-T-P", "A-P-T-P-V", "T-P-A"]
    paths, srcs = [], []
    for q in range(40):
        mp = configs.metapath_from_letters(letters[q % len(letters)], "bib")
        t0 = mp.vertexTypes[0]
        d0 = int(hg.dimensions[t0])
        k = int(rng.integers(0, min(d0, 400) + 1))
        rows = np.sort(rng.choice(d0, size=k, replace=False)).astype(np.int32)
        paths.append(mp)
        srcs.append((rows + hg.type_offset[t0]).astype(np.int32))
    srcs[5] = srcs[1].copy()  # identical sources
    ref = reference(G, paths, srcs)
    ctx.set_option("batch_small_products", (1 << 40) if small else 0)
    st = check(ctx, G, paths, srcs, ref)
    assert st["jobs"] == len(letters), st


def test_batch_autotune(ctx, graphs, batch_opts):
    """The autotuner leaves a valid setting on the context and results stay identical."""
    hg, G = graphs["mid"]
    paths, srcs = make_batch(hg, 30, 41, smax=1000)
    ref = reference(G, paths, srcs)
    t = G.autotune_batch(paths[:20], srcs[:20])
    assert t["trials"] > 0 and t["best_s"] <= t["baseline_s"] + 1e-9, t
    assert ctx.get_option("batch_small_products") == t["small_products"]
    assert ctx.get_option("batch_tile_rows") == t["tile_rows"]
    check(ctx, G, paths, srcs, ref)
