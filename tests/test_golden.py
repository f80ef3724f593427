"""The Skin fixture (first 1,200 rows of the reference's own data file,
test/Skin_NonSkin.txt) and its golden k-NN, pinned three ways on the CPU:
the oracle reproduces the golden, an independent numpy binary64 brute force
with the (squared distance, row) order agrees on sampled queries, and the
command-line loader parses the file row by row (d=3 drops the label column
that the reference's token-wise fscanf loader would shift into the rows)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "knn_sweet_amd", "knnjoin")


def skin_rows():
    with open(os.path.join(GOLD, "skin_1200.txt")) as f:
        return np.array([[float(v) for v in ln.split()] for ln in f if ln.strip()], dtype=np.float32)


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "skin_1200_k32.npz"))


def test_fixture_shape_and_duplicates():
    X = skin_rows()
    assert X.shape == (1200, 4)
    assert np.all(X[:, :3] == np.round(X[:, :3])) and X[:, :3].min() >= 0 and X[:, :3].max() <= 255
    # many exact duplicates: the tie-break decides these lists
    _, counts = np.unique(X[:, :3], axis=0, return_counts=True)
    assert counts.max() > 32


@pytest.mark.parametrize("d", [3, 4])
def test_oracle_reproduces_golden(gold, d):
    X = skin_rows()[:, :d]
    idx, dist, _ = oracle.bruteforce(X, X, 32)
    assert np.array_equal(idx, gold[f"idx_d{d}"].astype(np.int32))
    assert np.array_equal(dist, gold[f"dist_d{d}"])


@pytest.mark.parametrize("d", [3, 4])
def test_numpy_bruteforce_agrees(gold, d):
    X = skin_rows()[:, :d].astype(np.float64)
    for q in range(0, 1200, 37):
        sq = ((X - X[q]) ** 2).sum(axis=1)
        order = np.lexsort((np.arange(len(X)), sq))[:32]
        assert np.array_equal(order, gold[f"idx_d{d}"][q].astype(np.int64)), f"query {q}"
        np.testing.assert_array_equal(np.sqrt(sq[order]).astype(np.float32), gold[f"dist_d{d}"][q])


needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="knnjoin not built")


def _dump(args, n):
    out = subprocess.run([CLI, *map(str, args), "--dump-input", str(n)], capture_output=True, text=True, check=True)
    rows = {"query": [], "source": []}
    for ln in out.stdout.splitlines():
        if ":" in ln and ln.split()[0] in rows:
            rows[ln.split()[0]].append([float(v) for v in ln.split(":", 1)[1].split()])
    return rows


@needs_cli
@pytest.mark.parametrize("d", [3, 4])
def test_cli_row_aware_loader(d):
    path = os.path.join(GOLD, "skin_1200.txt")
    rows = _dump([1200, 1200, d, 8, 8, 5, path, path], 6)
    assert np.array_equal(np.array(rows["query"], np.float32), skin_rows()[:6, :d])


@needs_cli
def test_cli_random_matches_reference_generator():
    """srand(2015); query coordinates then source coordinates, each
    rand() % 10 + (float)rand() / RAND_MAX (libc's generator)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(2015)
    want = []
    for _ in range(7 * 3 + 5 * 3):
        a = libc.rand() % 10
        b = np.float32(libc.rand()) / np.float32(2147483647)
        want.append(np.float32(a) + b)
    rows = _dump([7, 5, 3, 2, 2, 3, "random", "random"], 10)
    got = np.concatenate([np.ravel(rows["query"]), np.ravel(rows["source"])]).astype(np.float32)
    assert np.array_equal(got, np.array(want, np.float32))


@needs_cli
def test_cli_rejects_short_rows(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1 2 3\n4 5\n")
    r = subprocess.run([CLI, "2", "2", "3", "1", "1", "1", str(p), str(p), "--dump-input", "1"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "bad.txt:2" in r.stderr
