"""N-1 screening over the device: the host side of the reference's
DcGpuStudy::outage_loadings (ybus_powerflow/src/dc/gpu.rs:334-430) above this library's pair
updates, operator product maxima and column maxima.

For each outage (a rank-one update of the base angles: its pair, alpha and inverse weight, and
the operator row its own flow zeroes), the device forms the updated angles, the flows and the
most loaded branch, and only each column's maxima come back. A column is suspect when its update
is singular, its angles are not finite or reach past 1e5 in magnitude, or its loadings are not
finite (dc/gpu.rs:406-409). The reference solves a suspect outage again on the CPU, which then
decides its result (dc/gpu.rs:421-425). That solve belongs to the caller, not to this library:
`resolve(j)` is called for each suspect outage j and its return value is that outage's result.
This library has no CPU path of its own.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

# dc/gpu.rs:408: an angle above this marks the column suspect
SUSPECT_ANGLE = 1e5


@dataclass
class OutageLoading:
    """An outage's most loaded branch (the operator row, None when no product is nonzero) and its
    loading |flow| x weight (dc/gpu.rs:410-413; the reference maps the row to its case branch)."""

    branch: Optional[int]
    loading: float


def suspect_columns(singular, angles, loadings) -> np.ndarray:
    """The suspect rule of dc/gpu.rs:406-409 over a block's columns: singular, or the angles'
    maximum not finite or above SUSPECT_ANGLE, or the loadings' maximum not finite. `angles` and
    `loadings` are GpuBlock.maxima_arrays() and product_maxima_arrays() results."""
    _, angle_value, angle_not_finite = angles
    _, _, loading_not_finite = loadings
    return (np.asarray(singular, dtype=bool) | np.asarray(angle_not_finite, dtype=bool)
            | (np.asarray(angle_value) > SUSPECT_ANGLE) | np.asarray(loading_not_finite, dtype=bool))


def outage_loadings(factor, operator, base, plus, minus, alpha, inverse_weight, zero,
                    resolve: Callable[[int], object]) -> List[object]:
    """Screens outages j = 0 .. len(plus) - 1 in blocks of factor.block_width() columns
    (dc/gpu.rs:386-430): each block's pair updates of `base` (plus and minus -1 for none), the
    operator's product maxima with row zero[j] zeroed in column j (-1 for none), the angles'
    maxima and the singular flags. Returns, per outage, its OutageLoading, or resolve(j) for a
    suspect one. The operator's offset and weights are the caller's to set beforehand
    (dc/gpu.rs:347-384)."""
    plus = np.ascontiguousarray(plus, dtype=np.int64)
    minus = np.ascontiguousarray(minus, dtype=np.int64)
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    inverse_weight = np.ascontiguousarray(inverse_weight, dtype=np.float64)
    zero = np.ascontiguousarray(zero, dtype=np.int64)
    count = len(plus)
    if not (len(minus) == len(alpha) == len(inverse_weight) == len(zero) == count):
        raise ValueError("one plus, minus, alpha, inverse_weight and zero row per outage")
    width = max(1, factor.block_width())
    results: List[object] = []
    for start in range(0, count, width):
        end = min(count, start + width)
        block = factor.pair_updates_arrays(base, plus[start:end], minus[start:end], alpha[start:end],
                                           inverse_weight[start:end])
        try:
            loadings = block.product_maxima_arrays(operator, zero[start:end])
            angles = block.maxima_arrays()
            singular = block.singular()
        finally:
            block.free()
        suspect = suspect_columns(singular, angles, loadings)
        rows, values, _ = loadings
        for j in range(end - start):
            if suspect[j]:
                results.append(resolve(start + j))
            else:
                results.append(OutageLoading(None if rows[j] < 0 else int(rows[j]), float(values[j])))
    return results
