"""Restatement of scarf's CORAL alignment for mapping (oracle only).

* ``mapping_utils.py:15-19``   _cov_diaged: ``cov(x, rowvar=0)`` (ddof 1), +inf -> 0, + I
* ``mapping_utils.py:22-38``   _correlation_alignment:
                               ``s @ (fmp(C_s, -0.5) @ fmp(C_t, 0.5))`` with scipy's
                               ``fractional_matrix_power`` (the reference's own call)
* ``mapping_utils.py:41-95``   coral: both inputs standardised by their own
                               ``clean_array``'d mean and std (ddof 0)
* ``mapping_datastore.py:160-208``  run_mapping(run_coral=True): the corrected query
                               (``coral(target_data, ann_obj.data)``) reduced with the
                               reference's ``transform_z`` and loadings

Pinned in ``tests/test_oracle.py``: the covariance against its textbook sum, the
re-colouring identity A^T C_s A = C_t that defines CORAL's A (a transposed product or a
sign slip in an exponent breaks it), A = I when source and target are the same matrix,
and the eigendecomposition form of the fractional powers.
"""
from __future__ import annotations

import numpy as np

from .norm import clean_array


def cov_diaged(x: np.ndarray) -> np.ndarray:
    a = np.cov(x, rowvar=False)
    a[a == np.inf] = 0
    return a + np.eye(a.shape[0])


def alignment(s: np.ndarray, t: np.ndarray) -> np.ndarray:
    """A = C_s^{-1/2} C_t^{1/2} (``mapping_utils.py:22-38``)."""
    from scipy.linalg import fractional_matrix_power as fmp

    a = np.dot(fmp(cov_diaged(s), -0.5), fmp(cov_diaged(t), 0.5))
    return np.real_if_close(a)


def coral(source: np.ndarray, target: np.ndarray):
    """(corrected source, A) for dense normalised matrices (``mapping_utils.py:41-95``)."""
    sm = clean_array(source.mean(axis=0))
    sd = clean_array(source.std(axis=0), 1)
    tm = clean_array(target.mean(axis=0))
    td = clean_array(target.std(axis=0), 1)
    s = (source - sm) / sd
    t = (target - tm) / td
    A = alignment(s, t)
    return s @ A, A


def mapped_embedding(query_aligned: np.ndarray, reference: np.ndarray, loadings) -> np.ndarray:
    """The query's embedding under run_mapping(run_coral=True): scarf turns feature
    scaling off when CORAL runs (``mapping_datastore.py:141-142``), so make_graph hands
    back an AnnStream with disable_scaling=True (``graph_datastore.py:886``) whose reducer
    is ``data_coral @ loadings`` (``ann.py:126-127``)."""
    data, _ = coral(query_aligned, reference)
    return data @ loadings
