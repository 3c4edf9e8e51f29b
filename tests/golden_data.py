"""Loader for the committed golden fixtures (tests/golden/)."""
from __future__ import annotations

import json
import os

import numpy as np

from pdos_amd.abi import read_problem_file

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_manifest = {}
_arrays = {}

# variant "" : the reference as built (golden.json / .npz, make_golden.py);
# variant "ruiz": built with ruiz_equil_common.h (golden_ruiz.*, make_golden_ruiz.py)


def _suffix(variant):
    return f"_{variant}" if variant else ""


def manifest(variant="") -> dict:
    if variant not in _manifest:
        with open(os.path.join(GOLDEN, f"golden{_suffix(variant)}.json")) as fh:
            _manifest[variant] = json.load(fh)
    return _manifest[variant]


def arrays(variant=""):
    if variant not in _arrays:
        with np.load(os.path.join(GOLDEN, f"golden{_suffix(variant)}.npz"), allow_pickle=False) as z:
            _arrays[variant] = {k: z[k] for k in z.files}
    return _arrays[variant]


def case_names(variant=""):
    return sorted(manifest(variant)["cases"].keys())


def problem(tag, variant=""):
    p = read_problem_file(os.path.join(GOLDEN, manifest(variant)["cases"][tag]["input"]))
    a = arrays()  # warm starts are stored with the main fixtures
    if f"{tag}__x0" in a:
        p = p.with_warm(a[f"{tag}__x0"], a[f"{tag}__y0"], a[f"{tag}__s0"])
    return p


def expected(tag, plugin, variant=""):
    a = arrays(variant)
    e = dict(manifest(variant)["cases"][tag][plugin])
    for k in ("x", "s", "y"):
        e[k] = a[f"{tag}__{plugin}__{k}"]
    return e


def setup_expected(tag, variant=""):
    a = arrays(variant)
    return a[f"{tag}__D"], a[f"{tag}__E"], float(a[f"{tag}__lambda"][0]), a[f"{tag}__amd_perm"]
