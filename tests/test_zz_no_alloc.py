"""smcm_estimate_noalloc (the allocation-free variant of smcm_estimate, scratch in one static
device array per GPU) through the C ABI: its estimate equals smcm_estimate's bit for bit, with
and without exact_cost_sum, on two calls in a row where the second chain is smaller than the
first (the array keeps the first call's contents); chains whose scratch exceeds the array are
rejected with SMCM_EINVAL and the entry point keeps working afterwards; a meta-path whose typed
matrices are not yet on the device is rejected."""
import ctypes as C

import numpy as np
import pytest

from resmcm_ari_amd import configs, smcm

pytestmark = pytest.mark.gpu


class DeviceDouble:
    """One double of device memory for the entry point's output (test plumbing: the test may
    allocate; the entry point under test does not)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(8)) == 0

    def read(self) -> float:
        out = C.c_double()
        assert self.hip.hipMemcpy(C.byref(out), self.p, C.c_size_t(8), 2) == 0  # device to host
        return out.value

    def write(self, x: float) -> None:
        v = C.c_double(x)
        assert self.hip.hipMemcpy(self.p, C.byref(v), C.c_size_t(8), 1) == 0  # host to device

    def free(self):
        self.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def graphs(ctx):
    hg1 = configs.CFG1.build()
    hgt = configs.tiny_bib().build()
    G1, Gt = smcm.Graph(ctx, hg1), smcm.Graph(ctx, hgt)
    yield G1, Gt
    G1.close()
    Gt.close()


def noalloc(ctx, G, mp, out: DeviceDouble) -> float:
    out.write(-1.0)
    G.estimate_noalloc(mp, out.p.value)
    ctx.sync()
    return out.read()


@pytest.mark.parametrize("exact", [0, 1])
def test_noalloc_equals_estimate(ctx, graphs, exact):
    G1, Gt = graphs
    bib = lambda p: configs.metapath_from_letters(p, "bib")  # noqa: E731
    # larger chains first, then smaller ones: the scratch keeps the previous call's contents
    calls = [(G1, bib("A-P-T-P-A-P-T-P-A")), (G1, bib("A-P-V-P-A")), (Gt, bib("A-P-V-P-A")), (Gt, bib("T-P-A")),
             (Gt, bib("V-P-V")), (G1, bib("P-T-P")), (Gt, bib("P-A"))]
    out = DeviceDouble()
    ctx.set_option("exact_cost_sum", exact)
    try:
        for G, mp in calls:
            want = G.estimate(mp)  # also builds the typed matrices
            got = noalloc(ctx, G, mp, out)
            assert got == want, (mp.toString(), got, want)
            assert noalloc(ctx, G, mp, out) == want  # a repeat on the same contents
    finally:
        ctx.set_option("exact_cost_sum", 0)
        out.free()


def test_noalloc_limits(ctx, graphs):
    G1, _ = graphs
    big = configs.BibSpec(nA=300_000, nP=450_000, nV=2_000, nT=20_000, lam_auth=2.0, terms_per_paper=8,
                          seed=configs.SEED_BASE + 61).build()
    Gb = smcm.Graph(ctx, big)
    out = DeviceDouble()
    try:
        fresh = configs.metapath_from_letters("V-P-A", "bib")
        with pytest.raises(smcm.SmcmError) as e:  # typed matrices not built yet on this graph
            Gb.estimate_noalloc(fresh, out.p.value)
        assert e.value.code == smcm.SMCM_EINVAL and "not on the device" in str(e.value)
        for p, what in (("A-P-T-P-A-P-T-P-A", "MiB of scratch"), ("A-P-T-P-A", "scratch exhausted")):
            mp = configs.metapath_from_letters(p, "bib")
            Gb.estimate(mp)
            with pytest.raises(smcm.SmcmError) as e:
                Gb.estimate_noalloc(mp, out.p.value)
            assert e.value.code == smcm.SMCM_EINVAL and what in str(e.value), str(e.value)
        ctx.sync()
        mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
        assert noalloc(ctx, G1, mp, out) == G1.estimate(mp)  # still working afterwards
    finally:
        out.free()
        Gb.close()
