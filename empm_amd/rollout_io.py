"""The training step's output files, in the formats the reference's
evaluation scripts read:

- a rollout trajectory (`inference_MPM.pkl`): a pickled float32 numpy array
  of shape (F, N, 3), one row of particle positions per frame
  (phys_sim/engine/trainer_warp_mpm.py:534-538; read back by the evaluation
  scripts with pickle.load);
- material parameters (`optimal_param.pt`, `best_model_param.pt`): a dict
  of scalars written with torch.save (trainer_warp_mpm.py:680,783,788-795);
- the controller description (`ctrl_param.pkl`): a pickled dict of arrays
  and lists (trainer_warp_mpm.py:368-371,454-456).

The readers here execute nothing from the file: pickles go through an
unpickler that only rebuilds numpy arrays, scalars and containers, and
parameter files load with torch.load(weights_only=True).
"""
from __future__ import annotations

import io
import pickle

import numpy as np
import torch

# the globals a numpy array / scalar pickle refers to, and nothing else
_ALLOWED = {
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"),
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
}


class _ArrayUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ALLOWED:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"refusing to load {module}.{name}: only numpy arrays are allowed")


def _safe_load(path):
    with open(path, "rb") as f:
        return _ArrayUnpickler(io.BytesIO(f.read())).load()


def save_trajectory(vertices, path):
    """Write frames (F, N, 3) (torch or numpy) as the reference's
    inference_MPM.pkl: a pickled float32 numpy array."""
    if isinstance(vertices, torch.Tensor):
        vertices = vertices.detach().cpu().numpy()
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32))
    if v.ndim != 3 or v.shape[-1] != 3:
        raise ValueError(f"a trajectory is (frames, particles, 3); got {v.shape}")
    with open(path, "wb") as f:
        pickle.dump(v, f)


def load_trajectory(path) -> np.ndarray:
    v = _safe_load(path)
    if not isinstance(v, np.ndarray) or v.ndim != 3 or v.shape[-1] != 3:
        raise ValueError(f"{path} does not hold a (frames, particles, 3) array")
    return v


def save_params(params: dict, path):
    """torch.save of a parameter dict (floats or tensors), as the trainer
    writes optimal_param.pt / best_model_param.pt."""
    torch.save({k: (v.detach().cpu() if isinstance(v, torch.Tensor) else v) for k, v in params.items()}, path)


def load_params(path) -> dict:
    return torch.load(path, weights_only=True)


def save_ctrl_param(ctrl: dict, path):
    """ctrl_param.pkl: a pickled dict of numpy arrays, lists and scalars."""
    out = {}
    for k, v in ctrl.items():
        if isinstance(v, torch.Tensor):
            v = v.detach().cpu().numpy()
        out[k] = v
    with open(path, "wb") as f:
        pickle.dump(out, f)


def load_ctrl_param(path) -> dict:
    d = _safe_load(path)
    if not isinstance(d, dict):
        raise ValueError(f"{path} does not hold a dict")
    return d
