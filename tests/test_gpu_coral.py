"""GPU CORAL (mapping_utils.py:15-95) against the oracle (scipy's fractional powers)."""
import numpy as np
import pytest

from conftest import say
from oracle import coral as oc
from oracle import mapping as omap
from oracle import norm as onorm

pytestmark = pytest.mark.gpu


def spd(d, n, rank, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, rank)) @ rng.normal(size=(rank, d)) * 2 + rng.normal(size=(n, d))
    s = (x - x.mean(0)) / x.std(0)
    return oc.cov_diaged(s)


@pytest.mark.parametrize("d,rank", [(64, 8), (300, 20), (1000, 40)])
def test_spd_sqrt_vs_scipy(cuda, d, rank):
    import torch
    from scipy.linalg import fractional_matrix_power as fmp
    from scarf_amd.coral import spd_sqrt
    C = spd(d, 3 * d, rank, seed=d)
    say("spd_sqrt d", d)
    h, mh, it, res = spd_sqrt(torch.from_numpy(C).cuda())
    h, mh = h.cpu().numpy(), mh.cpu().numpy()
    want_h, want_mh = np.real(fmp(C, 0.5)), np.real(fmp(C, -0.5))
    say("iterations", it, "residual", res)
    assert np.linalg.norm(h - want_h) / np.linalg.norm(want_h) < 1e-11
    assert np.linalg.norm(mh - want_mh) / np.linalg.norm(want_mh) < 1e-11
    assert np.abs(h @ mh - np.eye(d)).max() < 1e-10


@pytest.mark.parametrize("scaling", [True, False])
def test_coral_mapping_embedding(cuda, scaling):
    """project_target(run_coral=True) against the dense restatement: query and reference
    standardised by their own stats, C_s^{-1/2} C_t^{1/2}, then the loadings with no z-scoring
    (scarf maps CORAL-corrected data with feature scaling off, whatever the reference was
    built with).  FP32 projection: 1e-4 of the embedding's scale."""
    from scarf_amd.datastore import make_graph, project_target, run_mapping
    from scarf_amd.synth import CSR, generate_csr, make_gene_model
    model = make_gene_model(240, 4, 0.12, 41)
    csr = generate_csr(3000, model, 41)
    feat = np.arange(0, 240, 2)
    src_ids = [f"g{i}" for i in feat]
    say("make_graph reference for CORAL")
    ann = make_graph(csr, feat, dims=12, k=7, batch_size=1000, screen="fp32",
                     feat_scaling=scaling, return_ann_object=True)
    tgt = generate_csr(800, model, 42, batch_shift_sigma=0.4)
    tdense = tgt.to_dense()
    rng = np.random.default_rng(5)
    keep = np.setdiff1d(np.arange(240), feat[rng.choice(120, 6, replace=False)])
    cols = np.intersect1d(keep, feat)
    rng.shuffle(cols)
    tcounts = tdense[:, cols].astype(np.uint32)
    tids = [f"g{i}" for i in cols]
    say("project_target run_coral")
    E = project_target(ann, CSR.from_dense(tcounts), tids, src_ids, run_coral=True)
    E = E.cpu().numpy().astype(np.float64)
    t_idx = omap.order_features(src_ids, tids)
    xa = omap.align_normed(tcounts, t_idx, sf=1000.0)
    xr = onorm.normed(csr.to_dense(), sf=1000.0, feat_idx=feat)
    want = oc.mapped_embedding(xa, xr, ann.loadings)
    err = np.abs(E - want).max() / np.abs(want).max()
    say("CORAL embedding max error / scale", err)
    assert err < 1e-4
    plain = project_target(ann, CSR.from_dense(tcounts), tids, src_ids).cpu().numpy()
    assert np.abs(plain - want).max() / np.abs(want).max() > 1e-2  # CORAL changed it
    st = run_mapping(ann, CSR.from_dense(tcounts), tids, src_ids, "c", save_k=5,
                     run_coral=True)
    assert st["RNA/projections/c/indices"].shape == (800, 5)
