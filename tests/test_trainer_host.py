"""Host logic of the trainer mirror (empm_amd.engine) that needs no GPU: the
parameter scaling, the controller velocities of each frame, the coordinate
transforms and the simulator settings the trainer derives from its inputs
(phys_sim/engine/trainer_warp_mpm.py:314-400, phys_sim/engine/utils.py)."""
import numpy as np
import pytest
import torch

from empm_amd.engine import InvPhyTrainerWarpMPM
from empm_amd.engine import utils as eu


def _trainer():
    rng = np.random.default_rng(1)
    N, M, F = 50, 5, 3
    obj = rng.uniform(-0.3, 0.3, (F + 1, N, 3)).astype(np.float32)
    ctrl = rng.uniform(-0.2, 0.2, (F + 1, M, 3)).astype(np.float32)
    cp = {"csize": [(0.03, 0.03, 0.02), (0.02, 0.02, 0.02)], "handler_pos": [(0.4, 0.5, 0.6), (0.6, 0.5, 0.6)],
          "ctrl_index": [-1, 3], "ctrl_mask": np.array([0, 0, 1, 1, 0]), "vscale": [(20.0, 20.0, 20.0), (1.0, 2.0, 3.0)]}
    return InvPhyTrainerWarpMPM(obj, np.ones((F + 1, N), bool), np.ones((F + 1, N), bool), ctrl, cp, device="cpu"), ctrl


def test_denormalize_param_scales_E_only_and_fills_defaults():
    tr, _ = _trainer()
    p = tr.denormalize_param({"E": 0.008, "nu": 0.25, "type": "plasticine"})
    assert p == {"type": "plasticine", "E": 8000.0, "nu": 0.25, "rho": 2e3, "ys": 1e5, "vol": 1e-5}
    assert tr.denormalize_param({}) == {"type": "jelly", "E": 5e4, "nu": 0.3, "rho": 2e3, "ys": 1e5, "vol": 1e-5}


def test_controller_velocity_is_the_scaled_displacement_of_the_handle():
    tr, ctrl = _trainer()
    for f in range(3):
        # part 0 follows the mean of control points 0, 1 and 4
        mean = lambda k: ctrl[k][[0, 1, 4]].astype(np.float64).mean(0)
        np.testing.assert_allclose(tr._ctrl_velocity(f, 0), (mean(f + 1) - mean(f)) * 20.0, rtol=1e-5, atol=1e-6)
        # part 1 follows control point 3, per-axis scales
        d = ctrl[f + 1, 3].astype(np.float64) - ctrl[f, 3]
        np.testing.assert_allclose(tr._ctrl_velocity(f, 1), d * np.array([1.0, 2.0, 3.0]), rtol=1e-5, atol=1e-6)


def test_material_params_keep_the_yield_key_quirk_and_damping():
    tr, _ = _trainer()
    mp = tr._material_params(tr.denormalize_param({"type": "plasticine", "ys": 100.0}))
    # the parameter dicts say "ys", the reference tests "y_s": the gallery's yield stress stays
    assert mp["yield_stress"] == 5000.0 and mp["grid_v_damping_scale"] == 0.98
    mp2 = tr._material_params({"type": "plasticine", "ys": 100.0, "y_s": 1.0})
    assert mp2["yield_stress"] == 100.0
    # the gallery itself is not written
    from empm_amd.diff_simulator.warp_mpm.warp_mpm_wrapper import material_gallery
    assert "grid_v_damping_scale" not in material_gallery["plasticine"]


@pytest.mark.parametrize("dataset", ["ours", "phystwin"])
def test_transforms_are_inverse_and_match_the_closed_form(dataset):
    x = torch.tensor([[0.0, 0.0, 0.0], [0.5, -0.25, 1.0]], dtype=torch.float64)
    y = eu.transform_points(x, dataset=dataset)
    sign = 1.0 if dataset == "ours" else -1.0
    torch.testing.assert_close(y, 0.5 + sign * x / 5.0)
    torch.testing.assert_close(eu.inv_transform_points(y, dataset=dataset), x)
    with pytest.raises(ValueError):
        eu.transform_points(x, dataset="other")


def test_sloth_crop_lifts_the_bottom_layer_above_the_bar():
    x = torch.tensor([[0.0, 0.0, -0.2], [0.0, 0.0, 0.0], [0.0, 0.0, 0.5]], dtype=torch.float64)
    y = eu.transform_points(x.clone(), name="sloth_x", dataset="ours")
    # z in simulation coordinates 0.46, 0.5, 0.6: the two below 0.51 are mapped onto [0.505, 0.51)
    assert float(y[0, 2]) == pytest.approx(0.505)
    assert 0.505 < float(y[1, 2]) < 0.51 and float(y[2, 2]) == pytest.approx(0.6)


def test_initial_guesses_follow_the_case_name():
    assert eu.initial_param_guess("double_lift_dough")["type"] == "plasticine"
    assert eu.initial_param_guess("cloth_x") == {"type": "jelly", "E": 0.03, "nu": 0.3, "rho": 2e3}
    assert eu.get_param("playdoh_1")["rho"] == 500.0
    assert eu.get_param("rope")["E"] == 0.05


def test_special_experiments_are_refused():
    tr, _ = _trainer()
    tr.exp_name = "double_tear_pita_test"
    with pytest.raises(NotImplementedError):
        tr.initialize_mpm({"E": 0.01})
