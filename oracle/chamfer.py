"""NumPy restatement (float64) of the training loss on particle frames.

TEST INFRASTRUCTURE ONLY: imported by tests/, as the checker of
empm_amd.loss, never as the product.

The reference computes, per frame (phys_sim/engine/trainer_warp_mpm.py:727-766):
- pytorch3d.loss.chamfer_distance(pred_visible, target_visible) with its
  defaults (pytorch3d is a dependency of the reference, not vendored in it):
  for every point its squared Euclidean distance to the nearest point of the
  other cloud, averaged over each cloud, the two directions summed;
- (pred_valid - target_valid).norm() / n_valid;
summed over frames, times 1e3.  The nearest neighbour is found by brute
force here, ties resolved to the smallest index.
"""
from __future__ import annotations

import numpy as np


def nn_sqdist(a, b):
    """(dist, idx) of the nearest row of b for every row of a."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = ((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)
    idx = np.argmin(d, axis=1)
    return d[np.arange(a.shape[0]), idx], idx


def chamfer(x, y):
    dx, _ = nn_sqdist(x, y)
    dy, _ = nn_sqdist(y, x)
    return dx.mean() + dy.mean()


def chamfer_grad(x, y):
    """d chamfer / dx for fixed nearest-neighbour assignments."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    _, ix = nn_sqdist(x, y)
    _, iy = nn_sqdist(y, x)
    g = 2.0 * (x - y[ix]) / x.shape[0]
    np.add.at(g, iy, -2.0 * (y - x[iy]) / y.shape[0])
    return g


def tracking(pred, target):
    pred = np.asarray(pred, np.float64)
    return float(np.linalg.norm(pred - np.asarray(target, np.float64)) / pred.shape[0])


def frame_loss(points, target_points, visibilities, motions_valid):
    T = len(points)
    ch = sum(chamfer(points[t][visibilities[t]], target_points[t][visibilities[t]]) for t in range(T))
    tr = sum(tracking(points[t][motions_valid[t]], target_points[t][motions_valid[t]]) for t in range(T))
    return 1e3 * (ch + tr)
