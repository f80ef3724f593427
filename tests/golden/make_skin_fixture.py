"""Builds the Skin fixture: the first rows of the reference's own data file
test/Skin_NonSkin.txt (B, G, R, class label; integers) and the FP64 oracle's
exact k-NN self-join of them at d=3 (B, G, R) and d=4 (label included, as
the reference's run_sample.sh reads it), ties broken by (squared distance,
row).  Run once in a checkout beside the reference:

    python tests/golden/make_skin_fixture.py /path/to/reference/test/Skin_NonSkin.txt
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

N, K = 1200, 32


def main(src):
    with open(src, "rb") as f:
        lines = [f.readline() for _ in range(N)]
    with open(os.path.join(HERE, "skin_1200.txt"), "wb") as f:
        f.writelines(lines)
    X = np.array([[float(v) for v in ln.split()] for ln in lines], dtype=np.float32)
    import oracle
    out = {}
    for d in (3, 4):
        idx, dist, _ = oracle.bruteforce(X[:, :d], X[:, :d], K)
        out[f"idx_d{d}"] = idx.astype(np.int16)
        out[f"dist_d{d}"] = dist
    np.savez_compressed(os.path.join(HERE, "skin_1200_k32.npz"), **out)
    h = hashlib.sha256(open(os.path.join(HERE, "skin_1200.txt"), "rb").read()).hexdigest()
    print("skin_1200.txt sha256", h)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/Skin_NonSkin.txt")
