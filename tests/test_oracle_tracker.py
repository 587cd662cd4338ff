"""Pins the tracker restatement (oracle/tracker.py) on hand-computed cases:
the exponent of a force bound, the cap at 2^4 k_min, the 16-step memory of
the smallest exponent, and too_small."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, _load

T = _load("orc_tracker", os.path.join(ROOT, "oracle", "tracker.py"))


def _force_for(kmin, fscale, ratio):
    """A force array whose bound gives k_max = ratio * k_min."""
    kmax = ratio * kmin
    fmax = fscale / (4 * kmax) ** 2 - 1e-3
    return [np.array([fmax, -0.5 * fmax])]


def test_exponent_and_cap():
    tr = T.Tracker(1e-4, 1.0)
    assert tr.bound(_force_for(1e-4, 1.0, 2 ** 2.5)) == 2
    assert tr.bound(_force_for(1e-4, 1.0, 2 ** 6.2)) == 6
    k, e = tr(_force_for(1e-4, 1.0, 2 ** 6.2))
    assert e == 4 and k == 1e-4 * 16


def test_two_arrays_combine_in_quadrature():
    tr = T.Tracker(1.0, 16.0 * 4)
    # |f| = sqrt(3^2 + 4^2) = 5 -> k_max = sqrt(64 / 5.001) / 4
    e = tr.bound([np.array([3.0, -1.0]), np.array([-4.0])])
    assert e == math.floor(math.log2(math.sqrt(64 / 5.001) / 4))


def test_history_keeps_the_smallest_exponent_for_16_steps():
    tr = T.Tracker(1e-4, 1.0)
    small = _force_for(1e-4, 1.0, 2 ** 1.5)
    big = _force_for(1e-4, 1.0, 2 ** 8)
    assert tr(small)[1] == 1
    for i in range(15):
        assert tr(big)[1] == 1
    assert tr(big)[1] == 4          # the small exponent has left the 16-step window


def test_too_small():
    tr = T.Tracker(1e-4, 1.0)
    with pytest.raises(T.TooSmall):
        tr(_force_for(1e-4, 1.0, 0.7))
