"""Writes the committed golden fixtures for the CPU oracle tests.

Sources of the expected values (none of them is the oracle itself):
  * CSC: a brute-force stable sort by column (Python's sorted() is stable) of
    the COO input, the semantics of make_compact_csc
    (reference include/sparse_fft_internal.cuh:200-222).
  * 2-D spectra: numpy.fft.rfft2 (pocketfft, FP64) on the dense 0/1 matrix,
    the quantity FFTW's r2c_2d computes in the reference's CPU path
    (reference src/cpu_reference.cpp:28-45).

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def stable_csc(rows, cols, r, c):
    order = sorted(range(len(c)), key=lambda i: int(c[i]))
    sc = sorted(set(int(x) for x in c))
    counts = [0] * len(sc)
    slot = {col: i for i, col in enumerate(sc)}
    for x in c:
        counts[slot[int(x)]] += 1
    cp = [0]
    for n in counts:
        cp.append(cp[-1] + n)
    return (np.array(sc, np.int32), np.array(cp, np.int32),
            np.array([int(r[i]) for i in order], np.int32))


def cases():
    rng = np.random.default_rng(2024)
    out = []
    shapes = [(1, 1, 1), (4, 6, 7), (5, 5, 10), (9, 16, 40), (16, 9, 40), (30, 31, 150),
              (64, 48, 600), (7, 100, 90), (100, 7, 90), (23, 29, 0)]
    for k, (R, C, nnz) in enumerate(shapes):
        nnz = min(nnz, R * C)
        idx = rng.choice(R * C, nnz, replace=False)
        r = (idx // C).astype(np.int32)
        c = (idx % C).astype(np.int32)
        if k % 2 == 0:  # canonical (row, col) order like read_mtx; odd cases stay shuffled
            o = np.lexsort((c, r))
            r, c = r[o], c[o]
        out.append((f"rand{k}", R, C, r, c))
    return out


def main():
    payload = {}
    for name, R, C, r, c in cases():
        dense = np.zeros((R, C))
        dense[r, c] = 1.0
        X = np.fft.rfft2(dense)
        sc, cp, ri = stable_csc(R, C, r, c)
        payload[f"{name}__shape"] = np.array([R, C], np.int32)
        payload[f"{name}__row"] = r
        payload[f"{name}__col"] = c
        payload[f"{name}__sparse_cols"] = sc
        payload[f"{name}__col_ptr"] = cp
        payload[f"{name}__row_idx"] = ri
        payload[f"{name}__X"] = X
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **payload)
    print("wrote", os.path.join(HERE, "small_cases.npz"), len(payload) // 7, "cases")


if __name__ == "__main__":
    main()
