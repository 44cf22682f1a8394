"""Gaussian density filter on element centres (oracle).

Restates ``filters/spacial.py``: kernel ``w = exp(-0.5 (d/s)^2)`` with
``s = r/3`` and support ``3 s`` (``spacial.py:37-43``); neighbours from
``cKDTree.query_ball_point`` restricted to design elements, columns in the
compressed design numbering, rows normalised to sum 1 with zero rows -> 1
(``spacial.py:51-103``); forward ``out[design] = W rho[design]`` with
non-design entries passed through (``spacial.py:135-157``); gradient
``out[design] = W^T v[design]``, zero elsewhere (``spacial.py:159-166``).
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, diags
from scipy.spatial import cKDTree


def gaussian_support(radius: float):
    """Return (sigma, support) exactly as ``spacial.py:38-39`` computes them."""
    s = radius / 3.0
    return s, 3.0 * s


def build_filter_matrix(centers: np.ndarray, radius: float,
                        design_mask: np.ndarray | None):
    """Row-normalised CSR filter matrix over design elements (``spacial.py:51-103``)."""
    c = centers
    n_all = c.shape[1]
    if design_mask is None:
        design_mask = np.ones(n_all, dtype=bool)
    s, support = gaussian_support(radius)
    design_ids = np.nonzero(design_mask)[0]
    tree = cKDTree(c.T)
    rows, cols, data = [], [], []
    for row_idx, i in enumerate(design_ids):
        js_all = np.asarray(tree.query_ball_point(c[:, i], support), dtype=int)
        js = js_all[design_mask[js_all]]
        if js.size == 0:
            continue
        dij = np.linalg.norm(c[:, js] - c[:, [i]], axis=0)
        w = np.exp(-0.5 * (dij / s) ** 2)
        rows.append(np.full(js.size, row_idx, dtype=int))
        cols.append(np.searchsorted(design_ids, js))
        data.append(w)
    if not data:
        raise RuntimeError("No neighbor relations found")
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    data = np.concatenate(data)
    n = len(design_ids)
    W = coo_matrix((data, (rows, cols)), shape=(n, n)).tocsr()
    row_sum = np.array(W.sum(axis=1)).ravel()
    row_sum[row_sum == 0.0] = 1.0
    W = diags(1.0 / row_sum) @ W
    W.sort_indices()
    return W.tocsr()


class SpacialFilter:
    """Oracle counterpart of ``SpacialFilter`` (``spacial.py:106-166``).

    ``W`` depends only on the radius and the design mask, so it is built once
    per radius here instead of on every ``forward`` call; the operator is the
    same matrix the reference rebuilds each time.
    """

    def __init__(self, centers, radius, design_mask):
        self.centers = centers
        self.design_mask = design_mask
        self.radius = None
        self.update_radius(radius)

    def update_radius(self, radius):
        if radius != self.radius:
            self.radius = radius
            self.W = build_filter_matrix(self.centers, radius, self.design_mask)
            self.WT = self.W.T.tocsr()

    def forward(self, rho):
        out = np.copy(rho)
        out[self.design_mask] = self.W @ rho[self.design_mask]
        return out

    def gradient(self, v):
        out = np.zeros(self.centers.shape[1])
        out[self.design_mask] = self.WT @ v[self.design_mask]
        return out
