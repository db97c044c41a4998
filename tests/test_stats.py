"""Reference-shaped statistics JSON (field names and nesting of
include/cetric/statistics.h, document of apps/katric.cpp:278-297)."""
import json

import numpy as np
import pytest

from katric_amd import stats as KS

LOCAL_KEYS = [
    "rank", "io_time", "ghost_rank_gather", "preprocessing_local_phase", "preprocessing_global_phase",
    "primary_load_balancing", "secondary_load_balancing", "local_phase_time", "contraction_time",
    "global_phase_time", "reduce_time", "preprocessing_time", "preprocessing_global_time", "message_statistics",
    "global_phase_threshold", "wedges", "local_wall_time", "local_time", "skipped_nodes", "wedge_checks",
    "intersection_size_local", "intersection_size_global", "nodes_parallel2d", "local_triangles", "type3_triangles",
]


def test_local_statistics_layout():
    st = {"wedge_checks": 7, "intersection_size_local": 11, "intersection_size_global": 0, "local_triangles": 4,
          "type3_triangles": 0}
    tm = {"total_ms": 2.0, "prep_ms": 0.5, "count_ms": 1.0, "intersect_ms": 0.8}
    ls = KS.local_statistics(st, tm, rank=3, wedge_count=12)
    assert list(ls) == LOCAL_KEYS
    assert ls["rank"] == 3 and ls["wedges"] == 12 and ls["local_triangles"] == 4
    assert ls["local_phase_time"] == pytest.approx(0.8e-3) and ls["global_phase_time"] == 0.0
    assert set(ls["preprocessing_local_phase"]) == {"degree_exchange_time", "orientation_time", "sorting_time",
                                                   "message_statistics"}
    assert set(ls["message_statistics"]) == {"sent_messages", "received_messages", "send_volume", "receive_volume"}


def test_document_reduces_ranks():
    a = KS.local_statistics({"local_triangles": 5, "type3_triangles": 2}, {"total_ms": 1.0}, 0)
    b = KS.local_statistics({"local_triangles": 1, "type3_triangles": 3}, {"total_ms": 3.0}, 1)
    doc = KS.document([KS.statistics([a, b])], {"algorithm": "cetric"}, io_time=0.25)
    s = doc["stats"][0]
    assert list(s) == ["local_statistics", "global_wall_time", "triangles", "counted_triangles"]
    assert s["triangles"] == s["counted_triangles"] == 11
    assert s["global_wall_time"] == pytest.approx(3e-3)
    assert list(doc) == ["stats", "io_time", "config"]
    json.dumps(doc)


def test_wedges_closed_form():
    # star K_{1,k}: the centre has k(k-1)/2 wedges, the leaves none
    k = 9
    off = np.concatenate([[0, k], k + np.arange(1, k + 1)]).astype(np.uint64)
    assert KS.wedges(off) == k * (k - 1) // 2


@pytest.mark.gpu
def test_run_shmem_json(ctx):
    import katric_amd as K
    from tests.conftest import golden

    g = K.read_metis(golden("hiv.metis"))
    doc = json.loads(KS.run_shmem_json(ctx, g, iterations=2))
    assert len(doc["stats"]) == 2
    for run in doc["stats"]:
        assert run["triangles"] == run["counted_triangles"] == 45
        assert run["local_statistics"][0]["wedges"] == KS.wedges(g.offsets)
