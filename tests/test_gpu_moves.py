"""The move generator as the GPU runs it: gd_move_draws with the device
kernel's code, against the host build of the same source (a few ulps: the two
math libraries differ) and against the distributions the CPU tests pin."""
import numpy as np
import pytest

from geauxdock_amd import abi
from tests.test_moves import SCALE, check_gaussian, check_uniform, lib_draws

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [abi.GD_MOVE_GAUSSIAN, abi.GD_MOVE_REF_UNIFORM])
def test_device_draws_match_host_and_distribution(mode):
    dev = lib_draws(mode, device=1)
    host = lib_draws(mode, device=0)
    if mode == abi.GD_MOVE_GAUSSIAN:
        tol = 8 * np.spacing(np.abs(host)) + 1e-7 * SCALE
        bad = np.abs(dev - host) > tol
        assert not bad.any(), (np.abs(dev - host) / np.maximum(np.abs(host), 1e-30))[bad].max()
        check_gaussian(dev)
    else:
        assert np.array_equal(dev, host)
        check_uniform(dev)
