"""The zero-allocation count (cds_count_cliques_noalloc) against the
allocating one (cds_count_cliques), through the C ABI.

Its scratch is a static device array that keeps the previous call's
contents, so the check runs calls back to back: a larger graph first, then
a smaller one, each at k = 2, 3, 4, and every result must equal
cds_count_cliques bit for bit (T and every clique degree).  A graph whose
scratch does not fit the array is refused with CDS_ERR_RANGE before any
kernel runs.
"""
import numpy as np
import pytest

import cds_amd

pytestmark = pytest.mark.gpu


def test_noalloc_count_equals_count_cliques_on_consecutive_calls():
    with cds_amd.Engine(device=0) as eng:
        graphs = [("rmat15", cds_amd.Graph.rmat(15, 16, seed=31)),
                  ("chung_lu_3000", cds_amd.Graph.chung_lu(3000, 2.4, 8.0, 20, 0.9, seed=5))]
        assert graphs[1][1].n < graphs[0][1].n and graphs[1][1].m < graphs[0][1].m
        for name, g in graphs:
            dg = eng.upload(g)
            for k in (2, 3, 4):
                print(f"[case] noalloc {name} k={k}", flush=True)
                t0, cd0, _ = eng.count_cliques(dg, k)
                t1, cd1, _ = eng.count_cliques_noalloc(dg, k)
                assert t1 == t0, f"{name} k={k}: T {t1} != {t0}"
                assert np.array_equal(cd1, cd0), f"{name} k={k}: clique degrees differ"
                assert int(cd1.sum()) == k * t1
            dg.free()


def test_noalloc_refuses_a_graph_beyond_its_scratch():
    with cds_amd.Engine(device=0) as eng:
        g = cds_amd.Graph.rmat(20, 16, seed=2)  # about 90 MB of scratch
        dg = eng.upload(g)
        print("[case] noalloc rmat20 refused", flush=True)
        with pytest.raises(cds_amd.CdsError) as ei:
            eng.count_cliques_noalloc(dg, 3)
        assert ei.value.status == 5 and "scratch" in str(ei.value)
        dg.free()
