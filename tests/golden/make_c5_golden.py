"""Writes tests/golden/c5_poisson.npz: the oracle's Poisson IRLS (its FP64
restatement of the reference's fepois, glm.py:743-838) on each of the 32
uneven 300-covariate models of `poisson_models(32, lo=5_000, hi=50_000,
n_cov=300, seed=5)` -- the slopes, deviance and rows kept per model.  The
oracle takes about seven minutes for the 32 models on eight cores, too long
for a GPU test, so its answers are kept here; the inputs are regenerated from
the seed.

Usage: python tests/golden/make_c5_golden.py
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

SPEC = dict(n_models=32, lo=5_000, hi=50_000, n_cov=300, seed=5)
FE = ["worker_id", "firm_id", "year"]


def main():
    from hdfe_stream_amd.simulate import poisson_models
    from oracle import oracle as O
    O.build()
    ds = poisson_models(**SPEC)
    betas, dev, kept, rows = [], [], [], []
    for i, d in enumerate(ds):
        data = {"y": d["y"], "worker_id": d["worker_id"], "firm_id": d["firm_id"],
                "year": d["year"]}
        xs = []
        for j in range(d["X"].shape[1]):
            data[f"x{j}"] = d["X"][:, j].astype(np.float64)
            xs.append(f"x{j}")
        ref = O.fepois(data, "y", xs, FE, iwls_tol=1e-10, iwls_maxiter=50)
        assert ref["converged"], i
        betas.append(ref["beta"])
        dev.append(ref["deviance"])
        kept.append(int(ref["keep"].sum()))
        rows.append(len(d["y"]))
        print(f"model {i}: {rows[-1]} rows, deviance {dev[-1]:.6f}", flush=True)
    np.savez_compressed(Path(__file__).parent / "c5_poisson.npz", beta=np.array(betas),
                        deviance=np.array(dev), kept=np.array(kept), rows=np.array(rows),
                        spec=np.array([SPEC["n_models"], SPEC["lo"], SPEC["hi"],
                                       SPEC["n_cov"], SPEC["seed"]]))


if __name__ == "__main__":
    main()
