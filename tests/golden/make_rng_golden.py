"""Write tests/golden/golden_rng.npz from the reference's own RAN1 and Gauss
(adv_forcing.f90), compiled in place by `make -C oracle/ref rng` into
oracle/_ref/rng_ref.

For each seed the fixture holds the first HEAD values of RAN1 and of Gauss
(each from a fresh generator) and the SHA-256 of the first N values of each
(float64, little-endian), so the whole length is pinned without storing it.

Usage:  make -C oracle/ref rng && python tests/golden/make_rng_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "rng_ref")
SEEDS = (-11, -3, 1, 12345)
N, HEAD = 100_000, 2048


def sequences(seed: int, n: int):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "rng.bin")
        subprocess.run([EXE, str(seed), str(n), out], check=True)
        a = np.fromfile(out, dtype="<f8")
    return a[:n], a[n:]


def main():
    rec = {"seeds": np.array(SEEDS, np.int64), "n": np.int64(N)}
    for s in SEEDS:
        r, g = sequences(s, N)
        tag = f"m{-s}" if s < 0 else f"p{s}"
        rec[f"ran1_{tag}"] = r[:HEAD]
        rec[f"gauss_{tag}"] = g[:HEAD]
        rec[f"ran1_sha_{tag}"] = np.array(hashlib.sha256(r.tobytes()).hexdigest())
        rec[f"gauss_sha_{tag}"] = np.array(hashlib.sha256(g.tobytes()).hexdigest())
    np.savez_compressed(os.path.join(HERE, "golden_rng.npz"), **rec)


if __name__ == "__main__":
    main()
