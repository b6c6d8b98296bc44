"""DirStore (the on-disk layout of the graph outputs) without a GPU."""
import numpy as np
import pytest

from scarf_amd.store import DirStore


def test_dirstore_roundtrip(tmp_path):
    s = DirStore(str(tmp_path / "z"))
    a = np.arange(12, dtype=np.uint64).reshape(6, 2)
    s["RNA/normed__I__hvgs/reduction__pca__5__I/ann__l2__50__50__48__4466/knn__3/indices"] = a
    s["latest_graph"] = "RNA/normed__I__hvgs/x/graph__1.0__1.5"
    s["RNA/projections/q/_ref_n_cells"] = 7
    s["RNA/x/recall"] = np.float64(99.5)
    t = DirStore(str(tmp_path / "z"))  # a second session
    got = t["RNA/normed__I__hvgs/reduction__pca__5__I/ann__l2__50__50__48__4466/knn__3/indices"]
    assert got.dtype == np.uint64 and np.array_equal(got, a)
    assert t["latest_graph"] == "RNA/normed__I__hvgs/x/graph__1.0__1.5"
    assert t["RNA/projections/q/_ref_n_cells"] == 7
    assert float(t["RNA/x/recall"]) == 99.5
    assert "RNA/x/recall" in t and "RNA/x/none" not in t and t.get("nope") is None
    assert sorted(t) == sorted(s) and len(t) == 4
    s["latest_graph"] = np.array([1, 2])  # replacing changes the kind
    assert np.array_equal(t["latest_graph"], [1, 2])
    del t["RNA/x/recall"]
    assert "RNA/x/recall" not in s
    with pytest.raises(KeyError):
        s["../escape"] = 1
    with pytest.raises(TypeError):
        s["obj"] = object()
