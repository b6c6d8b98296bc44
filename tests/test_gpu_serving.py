"""GPU: the continuous-batching mapping server."""
import time

import numpy as np
import pytest

from oracle import knn as oknn
from oracle import mapping as omap
from scarf_amd.synth import CSR, generate_csr, make_gene_model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def server_setup(cuda):
    from scarf_amd.datastore import make_graph
    model = make_gene_model(200, 5, 0.1, 41)
    ref = generate_csr(6000, model, 41)
    ann = make_graph(ref, None, dims=20, k=11, batch_size=1000, screen="bf16",
                     return_ann_object=True)
    src_ids = [f"g{i}" for i in range(200)]
    rng = np.random.default_rng(0)
    qcols = rng.permutation(200)[:185]  # query schema: shuffled, 15 genes missing
    q_ids = [f"g{i}" for i in qcols]
    pool = generate_csr(9000, model, 42, batch_shift_sigma=0.2).to_dense()[:, qcols]
    return ann, src_ids, q_ids, pool


def _req(pool, a, b):
    return CSR.from_dense(pool[a:b].astype(np.uint32))


def _server(ann, src_ids, q_ids, **kw):
    from scarf_amd.serving import MappingServer
    return MappingServer(ann.loadings, ann.mu, ann.sigma, ann.annIdx, src_ids, q_ids,
                         norm=ann.data.norm, k=11, **kw)


def test_serving_matches_offline_mapping(server_setup):
    from scarf_amd.datastore import get_mapping_score, run_mapping
    ann, src_ids, q_ids, pool = server_setup
    req = _req(pool, 0, 1500)
    say("serving vs run_mapping, 1500 cells, chunk 400")
    with _server(ann, src_ids, q_ids, chunk_rows=400, n_streams=3) as srv:
        r = srv.submit(req).result(timeout=60)
    assert r.error is None
    st = run_mapping(ann, req, q_ids, src_ids, "x", save_k=11)
    np.testing.assert_array_equal(r.indices, st["RNA/projections/x/indices"])
    np.testing.assert_array_equal(r.distances.astype(np.float64), st["RNA/projections/x/distances"])
    (_, ms), = list(get_mapping_score(st, "x"))
    np.testing.assert_array_equal(r.scores, ms)
    (_, msr), = omap.mapping_scores(r.indices.astype(np.int64), r.distances.astype(np.float64),
                                    ann.nCells)
    np.testing.assert_allclose(r.scores, msr, rtol=1e-10, atol=1e-12)


def test_alone_equals_in_mix(server_setup):
    ann, src_ids, q_ids, pool = server_setup
    probe = _req(pool, 2000, 2600)
    with _server(ann, src_ids, q_ids, chunk_rows=256, n_streams=4) as srv:
        alone = srv.submit(probe).result(timeout=60)
    say("alone done; now in a mix of 12 requests")
    sizes = [1, 3000, 17, 600, 250, 1, 900, 40, 2200, 5, 333, 71]
    with _server(ann, src_ids, q_ids, chunk_rows=256, n_streams=4) as srv:
        futs = []
        s = 0
        for i, n in enumerate(sizes):
            futs.append(srv.submit(_req(pool, s % 8000, s % 8000 + n)))
            s += n
            if i == 5:
                fp = srv.submit(probe)  # arrives while others are on the GPU
        mixed = fp.result(timeout=120)
        others = [f.result(timeout=120) for f in futs]
    assert all(o.error is None for o in others)
    np.testing.assert_array_equal(alone.indices, mixed.indices)
    np.testing.assert_array_equal(alone.distances, mixed.distances)
    np.testing.assert_array_equal(alone.scores, mixed.scores)
    # every request got exactly its own rows answered
    for o, n in zip(others, sizes):
        assert o.indices.shape == (n, 11)


def test_admission_rejects_oversized_and_recovers(server_setup):
    ann, src_ids, q_ids, pool = server_setup
    big = _req(pool, 0, 4000)
    small = _req(pool, 0, 50)
    probe = _server(ann, src_ids, q_ids)
    budget = probe.footprint(small) * 3
    with _server(ann, src_ids, q_ids, mem_budget=budget, chunk_rows=64) as srv:
        rb = srv.submit(big).result(timeout=30)
        assert rb.error is not None and "budget" in rb.error
        # three small ones fit at once; ten are admitted as memory frees up
        res = [f.result(timeout=60) for f in [srv.submit(small) for _ in range(10)]]
    assert all(r.error is None for r in res)
    for r in res[1:]:
        np.testing.assert_array_equal(r.indices, res[0].indices)


def test_empty_and_single_cell(server_setup):
    ann, src_ids, q_ids, pool = server_setup
    with _server(ann, src_ids, q_ids) as srv:
        e = srv.submit(CSR(np.zeros(0, np.uint32), np.zeros(0, np.int32), np.zeros(1, np.int64),
                           len(q_ids))).result(timeout=30)
        one = srv.submit(_req(pool, 7, 8)).result(timeout=30)
    assert e.indices.shape == (0, 11) and e.error is None
    assert one.indices.shape == (1, 11)
    with pytest.raises(ValueError):
        _server(ann, src_ids, q_ids).submit(CSR(np.zeros(0, np.uint32), np.zeros(0, np.int32),
                                                np.zeros(1, np.int64), 3))


def test_cancel_frees_slot_and_memory(server_setup):
    """A cancelled queued request never runs; a cancelled admitted request stops issuing
    chunks and gives back its memory and slot, and the requests around it still get the
    same answers as alone."""
    ann, src_ids, q_ids, pool = server_setup
    small = _req(pool, 100, 160)
    big = _req(pool, 0, 8000)
    with _server(ann, src_ids, q_ids) as srv:
        alone = srv.submit(small).result(timeout=30)
    probe = _server(ann, src_ids, q_ids)
    # room for the big request only: the next one queues behind it
    budget = probe.footprint(big) + probe.footprint(small) // 2
    with _server(ann, src_ids, q_ids, mem_budget=budget, chunk_rows=16) as srv:
        fb = srv.submit(big)
        fq = srv.submit(small)  # queued: does not fit beside the big one
        assert srv.cancel(fq)
        # cancel the big one once its first chunks are on the GPU (500 chunks in all)
        deadline = time.time() + 30
        while not any(r.next_row > 0 for r in list(srv._active)) and time.time() < deadline:
            time.sleep(0.0005)
        assert srv.cancel(fb)
        after = srv.submit(small).result(timeout=60)
        assert fb.cancelled() and fq.cancelled()
        # everything drained: no memory or slot left behind by the cancelled requests
        deadline = time.time() + 30
        while (srv._active or srv._inflight) and time.time() < deadline:
            time.sleep(0.01)
        assert srv._used == 0 and not srv._active
    assert after.error is None
    np.testing.assert_array_equal(after.indices, alone.indices)
    np.testing.assert_array_equal(after.distances, alone.distances)
    np.testing.assert_array_equal(after.scores, alone.scores)
