"""Write the batched-tube workload fixture (nx, beta, s of every tube) from
utrip_amd.problems.tube_set with its documented seed.

    python tests/golden/gen_tubes.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from utrip_amd.problems import C3_SEED, tube_set  # noqa: E402

if __name__ == "__main__":
    nx, beta, s = tube_set(4096, C3_SEED)
    np.savez(HERE / "tubes_4096.npz", seed=np.int64(C3_SEED), nx=nx, beta=beta, s=s)
    print("nx", nx.min(), nx.max(), "cells", int(nx.sum()), "beta", beta.min(), beta.max())
