"""Seeded screen thresholds: the candidate pools, and so the search results, are the same
with and without seeds (seed rows of each 256-query block screened first)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _clustered(n, D, seed, dup_rows=0):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((12, D)) * 3
    E = (c[rng.integers(0, 12, n)] + rng.standard_normal((n, D))).astype(np.float32)
    if dup_rows:
        E[rng.choice(n, dup_rows, replace=False)] = E[1]
    return E


def _screen(Q, R, refs, pool, S, seeds):
    import torch
    from scarf_amd import _lib
    from scarf_amd.device import ptr, stream_handle, workspace
    from scarf_amd.knn import pack
    nq, D = Q.shape
    qp, _, _ = pack(Q, refs.dtype, 1)
    cand = torch.empty((nq, S, pool), dtype=torch.int64, device=Q.device)
    cnt = torch.empty((nq, S), dtype=torch.int32, device=Q.device)
    ws, nbytes = workspace("screen").get(_lib.lib().sh_knn_screen_workspace(nq, pool, S))
    _lib.call("sh_knn_screen", ptr(qp), nq, ptr(refs.data), R.shape[0], D,
              _lib.DTYPES[refs.dtype], None, None, pool, S, ptr(seeds), ptr(cand), ptr(cnt),
              ptr(ws), nbytes, stream_handle())
    torch.cuda.synchronize()
    return cand, cnt


def _assert_same_pools(a, b):
    import torch
    ca, na = a
    cb, nb = b
    assert torch.equal(na, nb)
    pool = ca.shape[2]
    m = torch.arange(pool, device=ca.device).view(1, 1, -1) < na.unsqueeze(-1).long()
    assert torch.equal(ca[m], cb[m])


@pytest.mark.parametrize("screen,pool,D", [("bf16", 64, 50), ("bf16", 128, 50), ("fp32", 64, 50),
                                           ("fp8", 64, 50), ("fp8", 128, 100)])
def test_seeded_screen_same_pools(cuda, screen, pool, D):
    import torch
    from scarf_amd.knn import locality_order, pack_refs, permute_rows
    n = 9000  # 35 full blocks of 256 queries and a tail of 40
    R = _clustered(n, D, 31, dup_rows=80)
    Rt = torch.from_numpy(R).cuda()
    refs = pack_refs(Rt, screen)
    perm = locality_order(Rt)
    Q = permute_rows(Rt, perm)
    seeds = perm.clone()
    seeds[5] = -1          # slots without a usable seed are ignored
    seeds[300] = n + 7
    for S in (1, 3):
        say("seeded screen", screen, "pool", pool, "slices", S)
        plain = _screen(Q, Rt, refs, pool, S, None)
        seeded = _screen(Q, Rt, refs, pool, S, seeds)
        _assert_same_pools(plain, seeded)
        # a block whose seeds are fewer than the pool keeps an open threshold
        few = torch.full_like(seeds, -1)
        few[:pool // 2] = perm[:pool // 2]
        _assert_same_pools(plain, _screen(Q, Rt, refs, pool, S, few))


@pytest.mark.parametrize("screen", ["bf16", "fp32"])
def test_seeded_search_same_results(cuda, screen):
    import torch
    from scarf_amd.knn import pack_refs, search
    R = _clustered(20000, 50, 32, dup_rows=40)
    Rt = torch.from_numpy(R).cuda()
    refs = pack_refs(Rt, screen)
    say("seeded search", screen)
    a = search(Rt, Rt, refs, 21, pool=64, self_offset=0, n_slices=1, order=True, seed=False)
    b = search(Rt, Rt, refs, 21, pool=64, self_offset=0, n_slices=1, order=True, seed=True)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
