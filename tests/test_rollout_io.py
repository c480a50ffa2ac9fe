"""Output files of the training step (empm_amd.rollout_io): the trajectory
pickle is what a plain pickle.load of the reference's evaluation scripts
reads (a float32 (F, N, 3) numpy array), parameter and controller files
round-trip, and the readers refuse pickles that would run code."""
import os
import pickle

import numpy as np
import pytest
import torch

from empm_amd import rollout_io as rio


def test_trajectory_is_a_pickled_float32_array(tmp_path):
    v = torch.randn(4, 100, 3, dtype=torch.float64)
    p = os.path.join(tmp_path, "inference_MPM.pkl")
    rio.save_trajectory(v, p)
    with open(p, "rb") as f:
        raw = pickle.load(f)  # the evaluation scripts' reader (our own file)
    assert isinstance(raw, np.ndarray) and raw.dtype == np.float32 and raw.shape == (4, 100, 3)
    np.testing.assert_array_equal(rio.load_trajectory(p), v.numpy().astype(np.float32))
    with pytest.raises(ValueError):
        rio.save_trajectory(np.zeros((4, 3)), p)


def test_params_and_ctrl_round_trip(tmp_path):
    p = os.path.join(tmp_path, "optimal_param.pt")
    rio.save_params({"E": 0.3, "nu": torch.tensor(0.25)}, p)
    d = rio.load_params(p)
    assert d["E"] == 0.3 and float(d["nu"]) == 0.25
    c = os.path.join(tmp_path, "ctrl_param.pkl")
    ctrl = {"csize": [0.05, 0.04], "handler_pos": np.array([[0.1, 0.2, 0.3]]), "ctrl_mask": np.array([0, 1, 1]),
            "vscale": [1.0, 2.0]}
    rio.save_ctrl_param(ctrl, c)
    back = rio.load_ctrl_param(c)
    assert back["csize"] == [0.05, 0.04]
    np.testing.assert_array_equal(back["handler_pos"], ctrl["handler_pos"])


class _Evil:
    def __reduce__(self):
        return (os.getcwd, ())


def test_readers_refuse_code(tmp_path):
    p = os.path.join(tmp_path, "bad.pkl")
    with open(p, "wb") as f:
        pickle.dump({"x": _Evil()}, f)
    with pytest.raises(pickle.UnpicklingError):
        rio.load_ctrl_param(p)
    with pytest.raises(pickle.UnpicklingError):
        rio.load_trajectory(p)
