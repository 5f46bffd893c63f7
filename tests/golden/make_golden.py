"""Write the golden fixtures under tests/golden/.

rademacher.npz: a table of the counter-based Rademacher draws (rows 1000-1063,
draws 5-36, seed 20260921), generated by the oracle after its generator is
pinned to the published splitmix64 outputs (tests/test_oracle.py); the device
generator is compared against this table.
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))


def main():
    from oracle import oracle as O
    O.build()
    ordinal, n, draws, seed, first = 1000, 64, 32, 20260921, 5
    np.savez(HERE / "rademacher.npz", table=O.rademacher(ordinal, n, draws, seed, first),
             ordinal=ordinal, seed=seed, first_draw=first)


if __name__ == "__main__":
    main()
