"""Writes tests/golden/verbose_logs.json: the reference's VERBOSE output
(pdos.c:79-127, 306-335; direct/private.c:157) on a few golden cases,
captured from fd 1 of the compiled reference (oracle/_ref), as data for the
log-format test.  Runs only where the reference tree is present.

Usage: python tests/golden/make_golden_verbose.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import make_golden as MG  # noqa: E402
import golden_data as G  # noqa: E402

CASES = [("fixture01", "direct"), ("fixture01", "indirect"), ("random_n300_m900", "indirect"),
         ("b_zero_n60_m150", "indirect"), ("portfolio_n101_steps10", "direct")]


def capture(lib, prob) -> str:
    d, k, keep = prob.with_params(VERBOSE=1).to_c()
    libc = C.CDLL(None)
    with tempfile.TemporaryFile() as tmp:
        libc.fflush(None)
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            lib.pdos(C.byref(d), C.byref(k))
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def main():
    libs = {pl: MG.load_ref(pl) for pl in ("indirect", "direct")}
    out = {f"{tag}/{pl}": capture(libs[pl], G.problem(tag)) for tag, pl in CASES}
    with open(os.path.join(HERE, "verbose_logs.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", len(out), "logs")


if __name__ == "__main__":
    main()
