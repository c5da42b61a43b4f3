"""Writes the MAT v5 format-variant fixtures of tests/test_mat_loader.py with
scipy.io.savemat: uncompressed and zlib-compressed dense arrays of several
numeric classes, a sparse double matrix, a logical array, long and short
(small-element) variable names.  The graph and price fixtures beside them
(Lattice_*, Triangular_*, CHH_*, FF25EUnew) are data files of the reference
(test_instances/, DataSets/) copied unchanged."""
import os

import numpy as np
import scipy.io as sio
import scipy.sparse as sp

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mat")


def main():
    rng = np.random.default_rng(7)
    variables = {
        "A": rng.integers(0, 2, (9, 7)).astype(np.uint8),
        "data": rng.standard_normal((13, 5)),
        "i16": rng.integers(-300, 300, (4, 6)).astype(np.int16),
        "f32": rng.standard_normal((3, 8)).astype(np.float32),
        "flags": rng.random((5, 5)) < 0.4,
        "S": sp.random(11, 6, density=0.3, random_state=3, format="csc"),
        "a_rather_long_variable_name": np.arange(12.0).reshape(3, 4),
        "e": np.zeros((0, 3)),
    }
    os.makedirs(HERE, exist_ok=True)
    sio.savemat(os.path.join(HERE, "variants_plain.mat"), variables, do_compression=False)
    sio.savemat(os.path.join(HERE, "variants_zlib.mat"), variables, do_compression=True)


if __name__ == "__main__":
    main()
