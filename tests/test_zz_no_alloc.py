"""GPU: the allocation-free mapping-score entry point (sh_mapping_score_static) against the
workspace one (sh_mapping_score), through the C ABI, two calls in a row with the second
on a smaller input, plus its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, nq, k, n_ref, n_groups):
    import torch
    idx = torch.from_numpy(rng.integers(0, n_ref, (nq, k)).astype(np.int64)).cuda()
    dist = torch.from_numpy(rng.uniform(0, 40, (nq, k)).astype(np.float32)).cuda()
    groups = torch.from_numpy(rng.integers(0, n_groups, nq).astype(np.int32)).cuda()
    return idx, dist, groups


def _both(idx, dist, groups, n_ref, n_groups, weighted, fixed_weight):
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    nq, k = idx.shape
    a = torch.empty((n_groups, n_ref), dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    nbytes = int(_lib.lib().sh_mapping_score_workspace(n_ref, n_groups))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    _lib.call("sh_mapping_score", ptr(idx), ptr(dist), nq, k, ptr(groups), n_groups, n_ref,
              int(weighted), float(fixed_weight), 1000.0, 1, ptr(a), ptr(ws), nbytes,
              stream_handle())
    _lib.call("sh_mapping_score_static", ptr(idx), ptr(dist), nq, k, ptr(groups), n_groups,
              n_ref, int(weighted), float(fixed_weight), 1000.0, 1, ptr(b), stream_handle())
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_static_scratch_matches_workspace_entry_point(cuda):
    rng = np.random.default_rng(71)
    # first call: 200k references x 3 groups; second: a smaller table on fewer references
    for nq, k, n_ref, ng, weighted in ((60_000, 11, 200_000, 3, True),
                                       (700, 5, 1_000, 2, False)):
        print("static scratch", nq, k, n_ref, ng, weighted, flush=True)
        idx, dist, groups = _inputs(rng, nq, k, n_ref, ng)
        a, b = _both(idx, dist, groups, n_ref, ng, weighted, 0.1)
        np.testing.assert_array_equal(a, b)
        assert np.isfinite(a).all() and (a > 0).any()


def test_static_scratch_limit(cuda):
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle
    import torch
    cap = int(_lib.lib().sh_mapping_score_static_max(1))
    assert cap == 1 << 23
    assert int(_lib.lib().sh_mapping_score_static_max(4)) == cap // 4
    idx = torch.zeros((4, 3), dtype=torch.int64, device="cuda")
    dist = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    out = torch.empty(8, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="static scratch"):
        _lib.call("sh_mapping_score_static", ptr(idx), ptr(dist), 4, 3, None, 1, cap + 1, 1,
                  0.1, 1000.0, 1, ptr(out), stream_handle())
