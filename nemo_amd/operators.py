"""LGL operators of the product library (host-side, no device needed)."""
from __future__ import annotations

import numpy as np

from ._lib import check, dptr, lib


def operators(N: int):
    """(x, w, D, Vinv) for polynomial degree N from nemo_gpu_operators."""
    M = N + 1
    x, w = np.zeros(M), np.zeros(M)
    D, Vi = np.zeros((M, M)), np.zeros((M, M))
    check(lib().nemo_gpu_operators(N, dptr(x), dptr(w), dptr(D), dptr(Vi)), "nemo_gpu_operators")
    return x, w, D, Vi


def lgl(N: int):
    x, w, _, _ = operators(N)
    return x, w
