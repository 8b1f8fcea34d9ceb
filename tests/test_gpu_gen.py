"""The batch graph generator on the GPU gives the host generator's graphs
(structure and int32 weights) bit for bit."""
import numpy as np
import pytest

from graphblas_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_device_batch_matches_host(grb):
    specs = W.batch_specs(40, seed=3, n_hi=8192)
    specs[0] = W.GraphSpec(65536, 32, "plus_pair_int32", 12345, 777)   # largest n and ef
    specs[1] = W.GraphSpec(512, 4, "min_plus_int32", 99, 98)           # smallest
    specs[2] = W.GraphSpec(1, 4, "min_plus_int32", 5, 6)               # one vertex, no edges
    specs[3] = W.GraphSpec(3000, 0, "plus_pair_int32", 5, 6)           # no samples
    print(f"[gen] {len(specs)} graphs", flush=True)
    db = W.DeviceBatch(specs)
    try:
        for g, s in enumerate(specs):
            ref = W.batch_graph_host(s)
            got = db.download(g)
            assert np.array_equal(got.p, ref.p), f"graph {g}: row pointers"
            assert np.array_equal(got.j, ref.j), f"graph {g}: columns"
            assert np.array_equal(got.x, ref.x), f"graph {g}: weights"
    finally:
        db.free()
