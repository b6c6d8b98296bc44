"""CPU-only: the C ABI library loads and exports what include/scarf_hip.h declares, and
the host-side logic (block schedule, FP64 merge, feature maps, CSR validation)."""
import ctypes
import os
import re

import numpy as np
import pytest

from scarf_amd import _lib
from scarf_amd import pca as gpca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "scarf_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libscarfhip.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), f"{s} declared in scarf_hip.h but not exported"
    # the ctypes table covers exactly the declared entry points
    assert sorted(_lib.exported_symbols()) == syms


def test_pure_queries_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libscarfhip.so not built")
    L = _lib.lib()
    assert L.sh_version() >= 10000
    assert L.sh_knn_pack_k(_lib.SH_BF16, 50) == 64
    assert L.sh_knn_pack_k(_lib.SH_F32, 50) == 56
    assert L.sh_knn_pack_k(_lib.SH_BF16, 100) == 112
    assert L.sh_gram_workspace(50000, 2000, 1) > L.sh_gram_workspace(50000, 2000, 0) > 0
    assert L.sh_norm_stats_workspace(100000, 250, 100) > L.sh_norm_stats_workspace(100000, 0, 100) > 0
    # argument validation happens before any device work
    rc = L.sh_stats_merge(None, None, None, 0, 10, None, None, None, None)
    assert rc == _lib.SH_EINVAL
    assert b"bad sizes" in L.sh_last_error()
    rc = L.sh_knn_screen(None, 10, None, 100, 50, _lib.SH_BF16, None, None, 500, 1, None, None,
                         None, None, 0, None)
    assert rc == _lib.SH_EINVAL


def test_partial_fit_schedule_matches_scarf_order():
    from oracle import ipca as oipca
    sizes = [3, 10, 10, 2, 7, 1]
    blocks, s = [], 0
    for z in sizes:
        blocks.append(np.arange(s, s + z))
        s += z
    calls, reservoir, carry = gpca.scarf_partial_fit_calls(blocks, 4)
    ref = oipca.scarf_block_schedule(sizes, 4)
    got = [sorted(c.tolist()) for c in calls] + \
        [sorted(np.concatenate([reservoir] + ([carry] if carry is not None else [])).tolist())]
    want = [sorted(np.concatenate([blocks[b] for b in call]).tolist()) for call in ref]
    assert got == want


def test_host_merge_matches_oracle_ipca():
    """IPCAState.merge (product host FP64 code) vs the oracle's IPCAGram."""
    from oracle import ipca as oipca
    rng = np.random.default_rng(5)
    X = rng.standard_normal((500, 25)) @ rng.standard_normal((25, 25))
    blocks = [X[i:i + 120] for i in range(0, 500, 120)]
    st = gpca.IPCAState(7)
    ref = oipca.IPCAGram(7)
    for b in blocks:
        st.merge(b.shape[0], b.T @ b, b.sum(0))
        ref.partial_fit(b)
    np.testing.assert_allclose(st.components_, ref.components_, atol=1e-9)
    np.testing.assert_allclose(st.singular_values_, ref.singular_values_, rtol=1e-10)
    np.testing.assert_allclose(st.var_, ref.var_, rtol=1e-10)
    np.testing.assert_allclose(st.explained_variance_ratio_, ref.explained_variance_ratio_,
                               rtol=1e-9)


def test_merge_rejects_too_many_components():
    st = gpca.IPCAState(10)
    with pytest.raises(ValueError):
        st.merge(5, np.eye(20), np.zeros(20))


def test_align_col_map():
    from scarf_amd.mapping import align_col_map
    pos, missing = align_col_map(["a", "b", "c", "d"], ["c", "x", "a"])
    np.testing.assert_array_equal(pos, [2, -1, 0])
    np.testing.assert_array_equal(missing, [False, True, False, True])
    with pytest.raises(ValueError):
        align_col_map(["a"], ["z"])


def test_csr_validation_host():
    from scarf_amd.device import validate_csr_host
    v = np.array([1, 2, 3], np.int32)
    with pytest.raises(ValueError):
        validate_csr_host(v, np.array([0, 1, 5], np.int32), np.array([0, 2, 3]), 4)  # col range
    with pytest.raises(ValueError):
        validate_csr_host(v, np.array([0, 1, 2], np.int32), np.array([0, 2, 1, 3]), 4)  # monotone
    with pytest.raises(ValueError):
        validate_csr_host(v, np.array([0, 1, 2], np.int32), np.array([0, 2]), 4)  # span
    validate_csr_host(v, np.array([0, 1, 2], np.int32), np.array([0, 2, 3]), 4)


def test_csr_validation_rejects_repeated_columns():
    """The kernels take each (row, column) at most once; a repeated column would race on
    one accumulator and be normalised twice instead of summed first."""
    from scarf_amd.device import validate_csr_host
    v = np.array([1, 2, 3, 4], np.int32)
    rp = np.array([0, 2, 4])
    with pytest.raises(ValueError, match="repeated"):
        validate_csr_host(v, np.array([1, 1, 0, 2], np.int32), rp, 4)  # sorted duplicate
    with pytest.raises(ValueError, match="repeated"):
        validate_csr_host(v, np.array([0, 1, 3, 3], np.int32), rp, 4)
    with pytest.raises(ValueError, match="repeated"):
        validate_csr_host(np.ones(3, np.int32), np.array([2, 0, 2], np.int32),
                          np.array([0, 3]), 4)  # unsorted duplicate
    # unsorted rows without repeats, and a column equal across a row boundary, pass
    validate_csr_host(v, np.array([3, 1, 0, 2], np.int32), rp, 4)
    validate_csr_host(v, np.array([0, 3, 3, 1], np.int32), rp, 4)
    validate_csr_host(v, np.array([0, 3, 1, 2], np.int32), np.array([0, 2, 2, 4]), 4)


def test_slice_plan():
    from scarf_amd.knn import _fix_slices, default_slices
    for nr in (64, 100, 5000, 500_000, 1_000_000):
        for nq in (1, 100, 1000, 50_000, 1_000_000):
            s = _fix_slices(nr, default_slices(nq, nr))
            sl = -(-(-(-nr // s)) // 128) * 128
            assert -(-nr // sl) == s


def test_load_unified_graph_host():
    """Reference edges plus the first use_k mapping edges of each target cell at a fixed
    weight (mapping_datastore.py:368-412), against the reference's loop construction."""
    from scipy.sparse import csr_matrix
    from scarf_amd.datastore import load_unified_graph
    rng = np.random.default_rng(7)
    ref_n, k, nt, use_k = 40, 5, 13, 3
    edges = np.stack([np.repeat(np.arange(ref_n), k), rng.integers(0, ref_n, ref_n * k)], 1)
    weights = rng.random(ref_n * k)
    pidx = rng.integers(0, ref_n, (nt, k)).astype(np.uint64)
    store = {"latest_graph": "RNA/g", "RNA/g/edges": edges, "RNA/g/weights": weights,
             "RNA/projections/t/indices": pidx, "RNA/projections/t/_ref_n_cells": ref_n}
    n_cells, g = load_unified_graph(store, ["t"], use_k, 0.1)
    ne, nw = [], []
    for n, i in enumerate(pidx[:, :use_k]):
        for j in i:
            ne.append([ref_n + n, j])
            nw.append(0.1)
    me = np.vstack([edges, ne]).astype(int)
    mw = np.hstack([weights, nw])
    ref = csr_matrix((mw, (me[:, 0], me[:, 1])), shape=(ref_n + nt, ref_n + nt))
    assert n_cells == [ref_n, nt]
    assert (g != ref).nnz == 0 and g.shape == ref.shape
