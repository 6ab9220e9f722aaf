"""Write tests/golden/ewald_pos_f32.npy from the reference's Ewald test input.

The reference's own force-accuracy test (Tests/Ewald) holds 65,536 particle
positions in the periodic unit box as raw float64 triples
(reference_ewald_pos.double3).  The near-field tests use them as a realistic
input; they are stored here as float32 box coordinates (the path computes in
float32 on cell-relative positions) with the source file's SHA-256.

Run from the repository root:  python tests/golden/make_ewald_fixture.py
"""
import hashlib
import os

import numpy as np

SRC = "/root/reference/Tests/Ewald/reference_ewald_pos.double3"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ewald_pos_f32.npy")


def main():
    raw = open(SRC, "rb").read()
    pos = np.frombuffer(raw, dtype="<f8").reshape(-1, 3)
    assert pos.shape == (65536, 3) and np.all(pos >= -0.5) and np.all(pos < 0.5)
    np.save(OUT, pos.astype(np.float32))
    print(f"{OUT}: {pos.shape[0]} particles; source sha256 {hashlib.sha256(raw).hexdigest()}")


if __name__ == "__main__":
    main()
