"""Writes tests/golden/c5_ols.npz: the oracle's FP64 fit (its restatement of
the reference's feols_stream: explicit S, block PCG, V'M_D V, beta) of
`y ~ x0 + ... + x299 | worker_id + firm_id + year` on each of the 32 uneven
300-covariate models of `ols_models(32, lo=5_000, hi=50_000, n_cov=300,
seed=6)` -- the slopes, the residual sum of squares and the rows kept per
model.  The oracle takes minutes for them, too long for a GPU test; the
inputs are regenerated from the seed.

Usage: python tests/golden/make_c5_ols_golden.py
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

SPEC = dict(n_models=32, lo=5_000, hi=50_000, n_cov=300, seed=6)
FE = ["worker_id", "firm_id", "year"]


def main():
    from hdfe_stream_amd.simulate import ols_models
    from oracle import oracle as O
    O.build()
    ds = ols_models(**SPEC)
    betas, rss, kept, rows = [], [], [], []
    for i, d in enumerate(ds):
        data = {"y": d["y"], "worker_id": d["worker_id"], "firm_id": d["firm_id"],
                "year": d["year"]}
        xs = []
        for j in range(d["X"].shape[1]):
            data[f"x{j}"] = d["X"][:, j].astype(np.float64)
            xs.append(f"x{j}")
        ref = O.fit(data, "y", xs, FE, tol=1e-12)
        betas.append(ref["beta"])
        rss.append(float(ref["resid"] @ ref["resid"]))
        kept.append(len(ref["resid"]))
        rows.append(len(d["y"]))
        print(f"model {i}: {rows[-1]} rows, rss {rss[-1]:.6f}", flush=True)
    np.savez_compressed(Path(__file__).parent / "c5_ols.npz", beta=np.array(betas),
                        rss=np.array(rss), kept=np.array(kept), rows=np.array(rows),
                        spec=np.array([SPEC["n_models"], SPEC["lo"], SPEC["hi"],
                                       SPEC["n_cov"], SPEC["seed"]]))


if __name__ == "__main__":
    main()
