"""Loader for the reference's golden kernel vectors.

``tests/golden/kernel_vectors.json`` is the fixture file the reference ships
at docs/data/kernel_vectors.json (written by scripts/gen_kernel_vectors.py of
the FireWebApp tree from the reference kernels on seeded inputs).
"""
from __future__ import annotations

import functools
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


@functools.lru_cache(maxsize=1)
def vectors():
    with open(os.path.join(GOLDEN_DIR, "kernel_vectors.json")) as f:
        return json.load(f)


def cases(section):
    return vectors()[section]["cases"]


def case_ids(section):
    return [c["name"] for c in cases(section)]


def field(c, key, shape=None):
    a = np.asarray(c[key], dtype=np.float64)
    if shape is None:
        shape = (c["nz"], c["ny"], c["nx"])
    return np.ascontiguousarray(a.reshape(shape))


def arr(c, key):
    return np.ascontiguousarray(np.asarray(c[key], dtype=np.float64))


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den
