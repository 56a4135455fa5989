"""The suspect rule of ybus_powerflow/src/dc/gpu.rs:406-409 as ybus_amd.screening states it, on the
host (no device): singular, or the angles' maximum not finite or strictly above 1e5, or the
loadings' maximum not finite."""
import numpy as np

from ybus_amd import screening


def maxima(values, not_finite):
    rows = np.zeros(len(values), dtype=np.int64)
    return rows, np.asarray(values, dtype=np.float64), np.asarray(not_finite, dtype=bool)


def test_each_clause_of_the_rule_flags_alone():
    clean = dict(singular=[False], angles=maxima([1.0], [False]), loadings=maxima([0.5], [False]))
    assert not screening.suspect_columns(**clean)[0]
    assert screening.suspect_columns([True], clean["angles"], clean["loadings"])[0]
    assert screening.suspect_columns([False], maxima([1.0], [True]), clean["loadings"])[0]
    assert screening.suspect_columns([False], maxima([1.0000001e5], [False]), clean["loadings"])[0]
    assert screening.suspect_columns([False], clean["angles"], maxima([0.5], [True]))[0]


def test_the_angle_bound_is_strict():
    # dc/gpu.rs:408: angle.value > 1e5, so exactly 1e5 passes
    at = screening.suspect_columns([False, False], maxima([1e5, np.nextafter(1e5, np.inf)], [False, False]),
                                   maxima([0.0, 0.0], [False, False]))
    assert at.tolist() == [False, True]
    assert screening.SUSPECT_ANGLE == 1e5


def test_columns_are_flagged_independently():
    singular = [False, True, False, False, False]
    angles = maxima([3.0, 3.0, 2e5, 3.0, 3.0], [False, False, False, True, False])
    loadings = maxima([1.0, 1.0, 1.0, 1.0, 1.0], [False, False, False, False, True])
    assert screening.suspect_columns(singular, angles, loadings).tolist() == [False, True, True, True, True]
