"""Writes ref_vectors.npz: results of the compiled reference engine
(oracle/_ref/libkriging_ref.so, table-mode covariance as shipped) on the small
problems of tests/problems.py (neighbour lists as its weight store reports
them, estimates, variances, and the solver counters: failures, blocks solved
by Cholesky, blocks solved by the symmetric-indefinite fallback -- the
reference's own counters split the last two into factorisations and cache
reuses, which depend on its thread schedule), and ck_vectors.npz: its multi-variable
(cokriging) results on problems.ck_cases() in table and analytic mode
(oracle/_ref/libkriging_ref_analytic.so); for some problems also the stored
kriging weights (krige_get_weight_data); drift_vectors.npz: external
drift and weight-correction problems (problems.drift_cases()); and
sgsim_vectors.npz: sequential simulation with given paths and draws
(problems.sgsim_cases()); grad_vectors.npz: increment constraints
(problems.grad_cases()).  Run in a container where the reference tree and
oracle/_ref exist:  python tests/golden/make_ref_vectors.py [drift|grad|sgsim]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import problems  # noqa: E402
from runners import run_ref, run_ref_ck, run_ref_sgsim  # noqa: E402


# problems whose stored kriging weights are kept (every fifth block with two
# or more neighbours; a single-neighbour block is not solved by the
# reference and its weight slot is left over from another block)
WEIGHT_CASES = ("ok2d", "lattice2d_s1_cv0", "random3d_octant", "simple_kriging", "rangescale_localnugget",
                "maxdist", "ladder", "near_duplicates", "big80_2d", "loo2d", "bounds")


def main():
    out = {}
    for spec in problems.parity_cases():
        key = spec["name"]
        r = run_ref(spec, want_weights=key in WEIGHT_CASES)
        if key in WEIGHT_CASES:
            nn = r["nnear"]
            sel = np.arange(0, len(nn), 5)
            sel = sel[nn[sel] >= 2].astype(np.int32)
            w = r["weight"][sel, :int(nn.max())]
            out[f"{key}/wblocks"] = sel
            out[f"{key}/weight"] = np.where(np.arange(w.shape[1])[None, :] < nn[sel, None], w, 0.0)
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
    print("wrote", len(problems.parity_cases()), "problems")
    ck = {}
    for spec in problems.ck_cases():
        for mode in ("table", "analytic"):
            r = run_ref_ck(spec, mode, want_weights=mode == "table")
            if mode == "table":
                # the weight store: neighbour counts and lists per group, and
                # the weights of every third block with two or more neighbours
                nn = r["nnear"]
                sel = np.arange(0, len(nn), 3)
                sel = sel[nn[sel].sum(axis=1) >= 2].astype(np.int32)
                wmax = max(int(nn.max()), 1)
                ck[f"{spec['name']}/nnear"] = nn.astype(np.int16)
                ck[f"{spec['name']}/inear"] = r["inear"][:, :, :wmax].astype(np.uint16)
                ck[f"{spec['name']}/wblocks"] = sel
                ck[f"{spec['name']}/weight"] = r["weight"][sel][:, :, :, :wmax]
            ck[f"{spec['name']}/{mode}/estimate"] = r["estimate"]
            ck[f"{spec['name']}/{mode}/variance"] = r["variance"]
            st = r["stats"]
            ck[f"{spec['name']}/{mode}/stats"] = np.array([st["fail"], st["chol_fact"] + st["chol_reuse"],
                                                           st["ssytrf_fact"] + st["ssytrf_reuse"]], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "ck_vectors.npz"), **ck)
    print("wrote", len(problems.ck_cases()), "cokriging problems")
    write_drift()
    write_grad()
    write_sgsim()


def write_drift():
    """drift_vectors.npz: external drift (KED, UK on blocks, without
    unbiasedness, cokriging) and weight-correction problems, table and
    analytic mode, with the stored weights in table mode"""
    dv = {}
    for spec in problems.drift_cases():
        for mode in ("table", "analytic"):
            r = run_ref_ck(spec, mode, want_weights=mode == "table")
            key = spec["name"]
            if mode == "table":
                nn = r["nnear"]
                sel = np.arange(0, len(nn), 3)
                sel = sel[nn[sel].sum(axis=1) >= 2].astype(np.int32)
                wmax = max(int(nn.max()), 1)
                dv[f"{key}/wblocks"] = sel
                dv[f"{key}/weight"] = r["weight"][sel][:, :, :, :wmax]
            dv[f"{key}/{mode}/estimate"] = r["estimate"]
            dv[f"{key}/{mode}/variance"] = r["variance"]
            st = r["stats"]
            dv[f"{key}/{mode}/stats"] = np.array([st["fail"], st["chol_fact"] + st["chol_reuse"],
                                                  st["ssytrf_fact"] + st["ssytrf_reuse"]], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "drift_vectors.npz"), **dv)
    print("wrote", len(problems.drift_cases()), "drift problems")


def write_grad():
    """grad_vectors.npz: increment (gradient-pair) constraints
    (problems.grad_cases()), estimates and variances in table and analytic
    mode"""
    gv = {}
    for spec in problems.grad_cases():
        for mode in ("table", "analytic"):
            r = run_ref_ck(spec, mode)
            gv[f"{spec['name']}/{mode}/estimate"] = r["estimate"]
            gv[f"{spec['name']}/{mode}/variance"] = r["variance"]
    np.savez_compressed(os.path.join(HERE, "grad_vectors.npz"), **gv)
    print("wrote", len(problems.grad_cases()), "gradient problems")


def write_sgsim():
    """sgsim_vectors.npz: sequential simulation with given paths and draws
    (tests/problems.py sgsim_cases), realisations and variances in table and
    analytic mode"""
    sv = {}
    for spec in problems.sgsim_cases():
        for mode in ("table", "analytic"):
            e, v = run_ref_sgsim(spec, mode)
            sv[f"{spec['name']}/{mode}/sim"] = e
            sv[f"{spec['name']}/{mode}/variance"] = v
    np.savez_compressed(os.path.join(HERE, "sgsim_vectors.npz"), **sv)
    print("wrote", len(problems.sgsim_cases()), "simulation problems")


if __name__ == "__main__":
    # an argument names one writer (drift, grad, sgsim); none runs them all
    if len(sys.argv) > 1:
        globals()["write_" + sys.argv[1]]()
    else:
        main()
