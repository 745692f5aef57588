"""Angular weights (host side, float64)."""
import numpy as np


def compute_angle_weights_1d(angles: np.ndarray) -> np.ndarray:
    """Weights proportional to the angular distance to the neighbours.

    Same contract as odtbrain.util.compute_angle_weights_1d
    (odtbrain/util.py:4-67): the coverage is wrapped at pi after shifting by
    the minimum angle, an angle present N times gets 1/N of the weight, and
    the weights have mean one.

    Steps: wrap (angles - min) into [0, pi); collapse duplicates with their
    multiplicity; for the sorted distinct angles take half the wrapped
    distance between the two neighbours (``(next - prev) mod pi``); divide by
    the multiplicity; scale so that the sum equals the number of angles.
    """
    a = np.asarray(angles, dtype=float).reshape(-1)
    wrapped = (a - np.min(a)) % np.pi
    distinct, where, mult = np.unique(wrapped, return_inverse=True,
                                      return_counts=True)
    # np.unique returns the distinct values sorted
    span = (np.roll(distinct, -1) - np.roll(distinct, 1)) % np.pi
    scale = wrapped.size / np.sum(span)
    per_distinct = span / mult * scale
    return per_distinct[where.reshape(-1)]
