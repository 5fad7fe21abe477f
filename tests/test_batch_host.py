"""CPU check of the batched front end's argument checks: Batch.gradient refuses observed
arrays whose shape differs from the shot's trace layout (the library copies ncomp x steps x
nrec samples per shot from each array, so a smaller array would be read past its end)."""
import numpy as np
import pytest

from biot_amd.batch import Batch


def _stand_in(shots):
    b = Batch.__new__(Batch)  # no device handle: the checks run before any library call
    b._h = None
    b.shots = shots
    b.n_model = 8
    b.model_shape = (2, 2, 2)
    return b


SHOTS = [dict(ncomp=3, nt=50, nrec=7, size=(4, 4, 4), origin=(0, 0, 0)),
         dict(ncomp=2, nt=40, nrec=5, size=(4, 4, 4), origin=(0, 0, 0))]


@pytest.mark.parametrize("bad", [
    [np.zeros((3, 30, 7)), np.zeros((2, 40, 5))],   # fewer steps than the receivers hold
    [np.zeros((3, 50, 7)), np.zeros((2, 40, 4))],   # fewer receivers
    [np.zeros((3, 50, 7)), np.zeros((3, 40, 5))],   # more components
    [np.zeros(3 * 50 * 7), np.zeros((2, 40, 5))],   # flat
    [np.zeros((3, 50, 7))],                          # one array short
])
def test_gradient_refuses_mis_shaped_observations(bad):
    with pytest.raises(ValueError):
        _stand_in(SHOTS).gradient(30, bad)
