"""Builds the Skin fixtures from the reference's own data file
test/Skin_NonSkin.txt (B, G, R, class label; integers 0..255 and 1..2):
  - skin_1200.txt: its first 1,200 lines as they are, and skin_1200_k32.npz,
    the FP64 oracle's exact k-NN self-join of them at d=3 (B, G, R) and d=4
    (label included, as the reference's run_sample.sh reads it), ties
    broken by (squared distance, row);
  - skin_rows_u8.npz: every row of the file as uint8 [245057, 4] (the data
    only; tests write the text the reference's loader reads from it and
    compute the expected answers with the oracle).
Run once in a checkout beside the reference:

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
        every = [ln for ln in f if ln.strip()]
    allrows = np.array([[int(v) for v in ln.split()] for ln in every], dtype=np.int64)
    assert allrows.min() >= 0 and allrows.max() <= 255
    np.savez_compressed(os.path.join(HERE, "skin_rows_u8.npz"), rows=allrows.astype(np.uint8))
    print("skin_rows_u8.npz rows", allrows.shape, "sha256",
          hashlib.sha256(allrows.astype(np.uint8).tobytes()).hexdigest())
    lines = every[:N]
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
