"""Small tensor helpers shared by the coarse-graining algorithms
(reference src/tenax/algorithms/_tensor_utils.py)."""

from __future__ import annotations

import math

from tenax_amd.core import LOG_EPS
from tenax_amd.core.tensor import Tensor


def max_abs_normalize(T: Tensor):
    """(T / (max|T| + LOG_EPS), log(max|T| + LOG_EPS)); _tensor_utils.py:65-80.

    The log norm comes back as a Python float (one device sync per call)."""
    norm = float(T.max_abs())
    return T * (1.0 / (norm + LOG_EPS)), math.log(norm + LOG_EPS)
