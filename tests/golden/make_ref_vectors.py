"""Writes ref_vectors.npz: results of the compiled reference engine
(oracle/_ref/libkriging_ref.so, table-mode covariance as shipped) on the small
problems of tests/problems.py (neighbour lists as its weight store reports
them, estimates, variances, and the solver counters: failures, blocks solved
by Cholesky, blocks solved by the symmetric-indefinite fallback -- the
reference's own counters split the last two into factorisations and cache
reuses, which depend on its thread schedule), and ck_vectors.npz: its multi-variable
(cokriging) results on problems.ck_cases() in table and analytic mode
(oracle/_ref/libkriging_ref_analytic.so).  Run in a container where the
reference tree and oracle/_ref exist:  python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import problems  # noqa: E402
from runners import run_ref, run_ref_ck  # noqa: E402


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
        st = r["stats"]
        out[f"{key}/stats"] = np.array([st["fail"], st["chol_fact"] + st["chol_reuse"],
                                        st["ssytrf_fact"] + st["ssytrf_reuse"]], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", len(out) // 5, "problems")
    ck = {}
    for spec in problems.ck_cases():
        for mode in ("table", "analytic"):
            r = run_ref_ck(spec, mode)
            ck[f"{spec['name']}/{mode}/estimate"] = r["estimate"]
            ck[f"{spec['name']}/{mode}/variance"] = r["variance"]
            st = r["stats"]
            ck[f"{spec['name']}/{mode}/stats"] = np.array([st["fail"], st["chol_fact"] + st["chol_reuse"],
                                                           st["ssytrf_fact"] + st["ssytrf_reuse"]], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "ck_vectors.npz"), **ck)
    print("wrote", len(ck) // 3, "cokriging results")


if __name__ == "__main__":
    main()
