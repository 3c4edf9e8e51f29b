"""Writes the golden fixtures of the Ruiz scaling option from the compiled
reference built with its alternative scaling header (ruiz_equil_common.h
force-included in place of common.h: oracle/Makefile, target `ref`).

For every parity case of make_golden.py, plus one case whose columns make the
header's cone-collapse scan run on into the next column, it stores x, s, y,
status and the iteration count of both plugins, and D, E, lambda and the AMD
permutation from the reference's initWork:

  golden_ruiz.npz   outputs
  golden_ruiz.json  manifest: status, iteration count, last residual row

Runs only in the container that holds the reference tree.

Usage: python tests/golden/make_golden_ruiz.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import make_golden as MG  # noqa: E402
from pdos_amd.abi import Problem, read_problem_file, write_problem_file  # noqa: E402


def overrun_case() -> Problem:
    """f=1, l=2, SOCs of 3 and 4 rows.  Column 0 ends in the second cone and
    column 1 begins in it; column 2 ends in the first cone and column 3 begins
    in it: the header's scan of a column's cone entries is bounded by nnz, so
    those leading entries of columns 1 and 3 take their left neighbour's cone
    maximum (ruiz_equil_common.h:45-63)."""
    Ap = [0, 3, 5, 9, 11]
    Ai = [0, 1, 8, 6, 7, 0, 2, 3, 5, 4, 9]
    Ax = [1.0, -1.0, 2.5, 0.3, -0.4, 1.0, -1.0, 0.7, 1.2, -0.9, 0.2]
    b = [1.0, 0.0, 0.0, 1.0, 0.5, 0.2, 1.0, 0.3, 0.1, 0.2]
    c = [0.5, -1.0, 0.3, 0.2]
    p = Problem(4, 10, Ap, Ai, Ax, b, c, 1, 2, [3, 4])
    return p.with_params(NORMALIZE=1, MAX_ITERS=200, EPS_ABS=1e-4, VERBOSE=0)


def collapse_defined(p: Problem) -> bool:
    """The reference's cone collapse (ruiz_equil_common.h:29-66) walks each
    column's entries in storage order and moves to the next cone whenever the
    current entry is not in the current one, so a row index below the current
    cone sends it past k->q[qsize - 1] (an out-of-bounds read; it crashes
    here on the jumbled case).  It runs only with NORMALIZE 1, and is defined
    when every column's row indices ascend."""
    if not p.params.get("NORMALIZE", 0):
        return True
    return all(np.all(np.diff(p.Ai[p.Ap[j]:p.Ap[j + 1]]) >= 0) for j in range(p.n))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(MG.ROOT, "oracle"), "ref"], check=True)
    libs = {pl: MG.load_ref(pl, "_ruiz") for pl in ("indirect", "direct")}
    warm = {}
    cases = [c for c in MG.golden_cases(warm) if collapse_defined(c[2])]
    path = "inputs/ruiz_overrun.bin"
    write_problem_file(os.path.join(HERE, path), overrun_case())
    cases.append(("ruiz_overrun", path, read_problem_file(os.path.join(HERE, path))))
    arrays = {}
    manifest = {"cases": {}}
    for tag, path, prob in cases:
        entry = {"input": path, "n": prob.n, "m": prob.m, "nnz": prob.nnz}
        for pl in ("indirect", "direct"):
            out = MG.run_ref(libs[pl], prob)
            for key in ("x", "s", "y"):
                arrays[f"{tag}__{pl}__{key}"] = out[key]
            entry[pl] = {"status": out["status"], "iters": out["iters"], "last_row": out["last_row"]}
        D, E, lam, P = MG.ref_setup(libs["direct"], prob, True)
        arrays[f"{tag}__D"] = D
        arrays[f"{tag}__E"] = E
        arrays[f"{tag}__lambda"] = np.array([lam])
        arrays[f"{tag}__amd_perm"] = P
        manifest["cases"][tag] = entry
        print(tag, {pl: entry[pl]["iters"] for pl in ("indirect", "direct")}, entry["indirect"]["status"],
              flush=True)
    np.savez_compressed(os.path.join(HERE, "golden_ruiz.npz"), **arrays)
    with open(os.path.join(HERE, "golden_ruiz.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    main()
