"""Host-side parameter logic of the Python mirror, against the oracle's
restatement and closed forms (CPU only)."""
import numpy as np
import pytest

from empm_amd.diff_simulator.warp_mpm.mpm_data_structure import friction_alpha
from oracle import oracle as orc


@pytest.mark.parametrize("phi", [0.0, 10.0, 25.0, 30.0, 33.7, 45.0])
def test_friction_alpha_matches_oracle_and_closed_form(phi):
    """alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) in float32
    (mpm_solver_diff.py:69-72, mpm_data_structure.py:475-478): the product's
    value equals the oracle's to one float32 ulp, and at 30 degrees
    (sin = 1/2) the closed form sqrt(2/3) / 2.5."""
    a = friction_alpha(phi)
    b = orc.make_params(1, 16, material=2, friction_angle=phi).alpha
    assert abs(a - b) <= float(np.spacing(np.float32(max(abs(b), 1e-30))))
    s = np.sin(phi / 180.0 * np.pi)
    assert abs(a - np.sqrt(2 / 3) * 2 * s / (3 - s)) < 1e-6 * max(a, 1e-6)
    if phi == 30.0:
        assert abs(a - np.sqrt(2 / 3) / 2.5) < 1e-6
