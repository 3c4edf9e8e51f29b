"""Loader for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import json
import os

import numpy as np

from pdos_amd.abi import read_problem_file

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_manifest = None
_arrays = None


def manifest() -> dict:
    global _manifest
    if _manifest is None:
        with open(os.path.join(GOLDEN, "golden.json")) as fh:
            _manifest = json.load(fh)
    return _manifest


def arrays():
    global _arrays
    if _arrays is None:
        with np.load(os.path.join(GOLDEN, "golden.npz"), allow_pickle=False) as z:
            _arrays = {k: z[k] for k in z.files}
    return _arrays


def case_names():
    return sorted(manifest()["cases"].keys())


def problem(tag):
    p = read_problem_file(os.path.join(GOLDEN, manifest()["cases"][tag]["input"]))
    a = arrays()
    if f"{tag}__x0" in a:
        p = p.with_warm(a[f"{tag}__x0"], a[f"{tag}__y0"], a[f"{tag}__s0"])
    return p


def expected(tag, plugin):
    a = arrays()
    e = dict(manifest()["cases"][tag][plugin])
    for k in ("x", "s", "y"):
        e[k] = a[f"{tag}__{plugin}__{k}"]
    return e


def setup_expected(tag):
    a = arrays()
    return a[f"{tag}__D"], a[f"{tag}__E"], float(a[f"{tag}__lambda"][0]), a[f"{tag}__amd_perm"]
