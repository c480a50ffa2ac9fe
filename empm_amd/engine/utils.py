"""Coordinate transforms and parameter guesses of phys_sim/engine/utils.py
(:27-120), the constants the trainer's simulator set-up depends on.

The reference picks the transform from the data set named in
configs/experiments.yaml (`data_path`); here it is the module-level DATASET
("ours" or "phystwin"), set by the caller.
"""
from __future__ import annotations

import torch

DATASET = "ours"
CSCALE = 1.0  # utils.py:9 (2.5 for some of the reference's own experiments)
SLOTH_CROP_CENTER = 0.505
SLOTH_CROP_TRANSITION = 0.005


def linear_crop(x, xbar, theta):
    """Moves x[:, 2] to above xbar (utils.py:15-24)."""
    mask = x[:, 2] < (xbar + theta)
    if mask.any():
        xmin = x[mask, 2].min()
        x[mask, 2] = xbar + (x[mask, 2] - xmin) / (xbar + theta - xmin) * theta
    return x


def transform_points(x, name="", dataset=None):
    """Data coordinates -> simulation coordinates (utils.py:27-38)."""
    ds = DATASET if dataset is None else dataset
    if ds == "ours":
        xx = torch.tensor([0.5, 0.5, 0.5], device=x.device, dtype=x.dtype) + x / 5.0 * CSCALE
        if "sloth" in name:
            xx = linear_crop(xx, xbar=SLOTH_CROP_CENTER, theta=SLOTH_CROP_TRANSITION)
        return xx
    if ds == "phystwin":
        return torch.tensor([0.5, 0.5, 0.5], device=x.device, dtype=x.dtype) - x / 5.0
    raise ValueError(f"dataset {ds!r} not defined")


def inv_transform_points(x, name="", dataset=None):
    """Simulation coordinates -> data coordinates (utils.py:41-47)."""
    ds = DATASET if dataset is None else dataset
    if ds == "ours":
        return (-torch.tensor([0.5, 0.5, 0.5], device=x.device, dtype=x.dtype) + x) * 5.0 / CSCALE
    if ds == "phystwin":
        return (torch.tensor([0.5, 0.5, 0.5], device=x.device, dtype=x.dtype) - x) * 5.0
    raise ValueError(f"dataset {ds!r} not defined")


def initial_param_guess(name=""):
    """Initial (normalised: E in MPa) parameters per case (utils.py:50-87)."""
    if "dough" in name:
        return {"type": "plasticine", "E": 0.001, "nu": 0.2, "ys": 1.5e3}
    if "playdoh" in name:
        return {"type": "plasticine", "E": 0.008, "nu": 0.3, "rho": 2e3, "ys": 100.0}
    if "pita" in name:
        return {"type": "plasticine", "E": 0.01, "nu": 0.3, "rho": 2e3, "ys": 100.0}
    if "cloth" in name:
        return {"type": "jelly", "E": 0.03, "nu": 0.3, "rho": 2e3}
    return {"type": "jelly", "E": 0.008, "nu": 0.3, "rho": 2e3}


def get_param(name=""):
    """Inference parameters per case (utils.py:90-120)."""
    if "dough" in name:
        return {"type": "plasticine", "E": 0.001, "nu": 0.2, "ys": 1.5e3}
    if "playdoh" in name:
        return {"type": "plasticine", "E": 0.01, "nu": 0.2, "ys": 10000.0, "rho": 500.0}
    if "pita" in name:
        return {"type": "plasticine", "E": 0.01, "nu": 0.3, "rho": 2e3, "ys": 100.0}
    return {"type": "jelly", "E": 0.05, "nu": 0.3, "rho": 2e3}
