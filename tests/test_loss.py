"""The training loss (empm_amd.loss) on the CPU: the float64 restatement
(oracle/chamfer.py) pinned by closed forms, and the host-side argument
handling of the GPU functions."""
import numpy as np
import pytest
import torch

from empm_amd import loss
from oracle import chamfer as och


def test_chamfer_closed_forms():
    rng = np.random.default_rng(0)
    x = rng.random((50, 3))
    assert och.chamfer(x, x) == 0.0
    # single points: both directions see the same squared distance
    a, b = np.array([[0.1, 0.2, 0.3]]), np.array([[0.4, -0.1, 0.7]])
    assert np.isclose(och.chamfer(a, b), 2 * np.sum((a - b) ** 2))
    # a lattice shifted by less than half its spacing: every point's nearest
    # neighbour is its own image, so the loss is 2 |t|^2
    g = np.stack(np.meshgrid(*[np.arange(4) * 0.1] * 3, indexing="ij"), -1).reshape(-1, 3)
    t = np.array([0.01, -0.02, 0.015])
    assert np.isclose(och.chamfer(g + t, g), 2 * np.sum(t * t))
    # symmetric
    y = rng.random((37, 3))
    assert np.isclose(och.chamfer(x, y), och.chamfer(y, x))


def test_chamfer_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    x, y = rng.random((20, 3)), rng.random((15, 3))
    g = och.chamfer_grad(x, y)
    h = 1e-6
    for i, d in [(0, 0), (3, 1), (7, 2), (19, 0)]:
        e = np.zeros_like(x)
        e[i, d] = h
        fd = (och.chamfer(x + e, y) - och.chamfer(x - e, y)) / (2 * h)
        assert abs(fd - g[i, d]) < 1e-6 * max(1.0, abs(fd))


def test_nearest_neighbour_ties_take_the_smallest_index():
    a = np.array([[0.0, 0.0, 0.0]])
    b = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]])
    d, i = och.nn_sqdist(a, b)
    assert i[0] == 0 and d[0] == 1.0


def test_tracking_and_frame_loss():
    rng = np.random.default_rng(2)
    p, q = rng.random((2, 30, 3)), rng.random((2, 30, 3))
    vis = np.ones((2, 30), bool)
    assert np.isclose(och.tracking(p[0], q[0]), np.sqrt(np.sum((p[0] - q[0]) ** 2)) / 30)
    v = och.frame_loss(p, q, vis, vis)
    assert np.isclose(v, 1e3 * sum(och.chamfer(p[t], q[t]) + och.tracking(p[t], q[t]) for t in range(2)))


def test_gpu_functions_validate_arguments_on_the_host():
    x = torch.zeros((1, 5, 3))
    with pytest.raises(NotImplementedError):
        loss.chamfer_distance(x, x, x_lengths=torch.tensor([5]))
    with pytest.raises(NotImplementedError):
        loss.chamfer_distance(x, x, point_reduction="sum")
    with pytest.raises(ValueError):
        loss.chamfer_distance(torch.zeros((2, 5, 3)), torch.zeros((1, 5, 3)))
    with pytest.raises(RuntimeError):
        loss.nn_sqdist(torch.zeros((5, 3)), torch.zeros((4, 3)))  # CPU tensors: the search runs on the GPU
