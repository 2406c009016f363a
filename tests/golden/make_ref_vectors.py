"""Writes ref_vectors.npz: results of the compiled reference engine
(oracle/_ref/libkriging_ref.so, table-mode covariance as shipped) on the small
problems of tests/problems.py.  Run in a container where the reference tree
and oracle/_ref exist:  python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import problems  # noqa: E402
from runners import run_ref  # noqa: E402


def main():
    out = {}
    for spec in problems.parity_cases():
        r = run_ref(spec)
        key = spec["name"]
        out[f"{key}/estimate"] = r["estimate"]
        out[f"{key}/variance"] = r["variance"]
        out[f"{key}/nnear"] = r["nnear"]
        w = int(r["nnear"].max()) if len(r["nnear"]) else 0
        inear = r["inear"][:, :max(w, 1)]
        out[f"{key}/inear"] = inear.astype(np.uint16) if inear.max() < 65535 else inear
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", len(out) // 4, "problems")


if __name__ == "__main__":
    main()
