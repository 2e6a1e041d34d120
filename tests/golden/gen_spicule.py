"""Write tests/golden/spicule_dac.npz from the DAC files the UTRIP spicule
run shipped (cans_beta_vary/cans1d/md_spicule/hdmlw/*.dac): the static
arrays x, bx, bxm, sc, scm, gx, gxm, rr, rrm, and both records (t = 0 and
the last, step 100000) of t, ro, pr, vx, vy, by.  These files are the
reference's own output data; they are read with utrip_amd.dac.

Run:  python tests/golden/gen_spicule.py [reference-root]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from utrip_amd import dac  # noqa: E402

STATIC = ("x", "bx", "bxm", "sc", "scm", "gx", "gxm", "rr", "rrm")
SERIES = ("ro", "pr", "vx", "vy", "by")


def main(ref_root="/root/reference/cans_beta_vary"):
    d = Path(ref_root) / "cans1d" / "md_spicule" / "hdmlw"
    out = {}
    for k in STATIC:
        snaps, _ = dac.read(d / f"{k}.dac")
        out[k] = np.asarray(snaps[0], dtype=np.float64).reshape(-1)
    ts, _ = dac.read(d / "t.dac")
    out["t"] = np.array([float(np.asarray(v).reshape(-1)[0]) for v in ts])
    for k in SERIES:
        snaps, _ = dac.read(d / f"{k}.dac")
        out[k] = np.stack([np.asarray(v, dtype=np.float64).reshape(-1) for v in snaps])
    dst = Path(__file__).resolve().parent / "spicule_dac.npz"
    np.savez_compressed(dst, **out)
    print(dst, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(*sys.argv[1:])
