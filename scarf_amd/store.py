"""A directory store for the graph outputs: scarf's zarr hierarchy without zarr.

scarf keeps every step's output in its zarr store under a path named by the step's
parameters (``graph_datastore.py:738-745``), with ``latest_*`` pointers
(``graph_datastore.py:1003-1008``) and the ANN index saved as the file ``ann_idx`` inside
the ANN group (``graph_datastore.py:939-956``); reruns reuse what exists
(``graph_datastore.py:836-881``, ``977-978``).  ``DirStore`` is a mapping with the same keys
(``"RNA/normed__I__hvgs/reduction__pca__30__I/ann__l2__.../knn__11/indices"``) kept as files
under a root directory:

* numpy arrays -> ``<key>.npy`` (``numpy.save``, no pickling);
* strings, numbers, booleans -> ``<key>.json``;
* a :class:`~scarf_amd.knn.GpuIndex` -> ``<key>.index.npz`` (``GpuIndex.save_index``),
  loaded back onto the GPU on first access.

``make_graph``, ``run_mapping`` and ``load_graph`` take it wherever they take a dict, so a
second session finds the normalisation statistics, loadings, index, kNN table, graph and
seed partitions of the first and recomputes nothing.
"""
from __future__ import annotations

import json
import os
from collections.abc import MutableMapping

import numpy as np

__all__ = ["DirStore"]

_SUFFIXES = (".npy", ".json", ".index.npz")


class DirStore(MutableMapping):
    def __init__(self, root: str):
        self.root = os.path.abspath(root)
        os.makedirs(self.root, exist_ok=True)
        self._loaded = {}

    def _base(self, key: str) -> str:
        key = str(key).strip("/")
        if not key or any(part in ("", ".", "..") for part in key.split("/")):
            raise KeyError(f"invalid store key {key!r}")
        return os.path.join(self.root, *key.split("/"))

    def _find(self, key: str):
        base = self._base(key)
        for s in _SUFFIXES:
            if os.path.isfile(base + s):
                return base + s, s
        return None, None

    def __contains__(self, key) -> bool:
        try:
            return self._find(key)[0] is not None
        except KeyError:
            return False

    def __getitem__(self, key):
        path, suffix = self._find(key)
        if path is None:
            raise KeyError(key)
        if key in self._loaded:
            return self._loaded[key]
        if suffix == ".npy":
            return np.load(path, allow_pickle=False)
        if suffix == ".json":
            with open(path) as f:
                return json.load(f)
        from .knn import GpuIndex
        idx = GpuIndex()
        idx.load_index(path)
        self._loaded[key] = idx
        return idx

    def __setitem__(self, key, value):
        from .knn import GpuIndex
        base = self._base(key)
        if isinstance(value, GpuIndex):
            if self._loaded.get(key) is value:  # unchanged since it was read
                return
            self._remove(key)
            os.makedirs(os.path.dirname(base), exist_ok=True)
            value.save_index(base + ".index.npz")
            self._loaded[key] = value
            return
        self._remove(key)
        os.makedirs(os.path.dirname(base), exist_ok=True)
        if isinstance(value, (str, bool, int, float)) or value is None:
            with open(base + ".json", "w") as f:
                json.dump(value, f)
        elif isinstance(value, (np.ndarray, np.generic, list, tuple)):
            a = np.asarray(value)
            if a.dtype == object:
                raise TypeError(f"{key}: object arrays are not stored")
            np.save(base + ".npy", a, allow_pickle=False)
        else:
            raise TypeError(f"{key}: cannot store a {type(value).__name__}")

    def _remove(self, key):
        base = self._base(key)
        for s in _SUFFIXES:
            if os.path.isfile(base + s):
                os.remove(base + s)
        self._loaded.pop(key, None)

    def __delitem__(self, key):
        if key not in self:
            raise KeyError(key)
        self._remove(key)

    def __iter__(self):
        for d, _, files in os.walk(self.root):
            rel = os.path.relpath(d, self.root)
            for fn in sorted(files):
                for s in _SUFFIXES[::-1]:  # longest first: .index.npz before .npy
                    if fn.endswith(s):
                        name = fn[: -len(s)]
                        yield name if rel == "." else "/".join(rel.split(os.sep) + [name])
                        break

    def __len__(self) -> int:
        return sum(1 for _ in self)
