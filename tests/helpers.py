"""Shared test helpers: golden fixtures, scene builders, comparisons."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from fluidsimulator_amd.params import Params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))  # allow_pickle=False (default)
    d = json.loads(str(z["params"]))
    d["external_forces"] = tuple(d["external_forces"])
    d["planes"] = [tuple(x) for x in d["planes"]]
    p = Params(**d)
    exp = {k: z[k] for k in z.files if k != "params"}
    exp["lambda"] = exp.pop("lam")
    return p, exp


def same_bits(a, b):
    """Bitwise equality of float arrays (distinguishes -0.0 and NaN payloads)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def same_bits_or_nan(a, b):
    """Bitwise equality of float arrays, except that a NaN equals any NaN at
    the same position: IEEE 754 leaves the sign and payload of an arithmetic
    NaN result open, and two kernels that compute the same operations in the
    same order may still pass a different NaN operand through."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint8), b[~nb].view(np.uint8))


def neighbor_lists(counts, flat):
    off = np.concatenate([[0], np.cumsum(counts)])
    return [flat[off[i]:off[i + 1]] for i in range(len(counts))]
