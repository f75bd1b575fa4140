#!/usr/bin/env python3
"""Convert the reference's shipped HDF5 result files into small committed .npz fixtures.

The reference ships FP64 outputs of its own CPU path (no h5py is installed, so this uses the
image's `h5dump -b` to read raw little-endian doubles).  The fixtures are *data*: eigenvalue
lists per iteration, entropy, specific heat and the temperature grid.

  examples/entropyKondo/resultsZero.h5       free chain, spinhalf(0,0) impurity, Lambda=2, K=1024
  examples/entropyKondo/resultsJ0.250000.h5  Kondo J=0.25, S=1/2, Lambda=2, K=1024
  examples/entropySIAM/resultSIAMU0.200000.h5  SIAM U=0.2, eps=-0.08, Gamma=0.01, Lambda=2.5, K=512
  examples/entropySIAM/resultSIAMZero.h5       same driver with impurity (0,0)
  independFiles/zAverage/hoppingElementsL3.h5  z = 1, Lambda = 3 hopping elements of the flat band
                                               (calcTn, hoppingElement.cpp:16-105)

Run once in a container that has /root/reference mounted:
    python tests/golden/make_fixtures.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    "free_chain_L2_K1024": "examples/entropyKondo/resultsZero.h5",
    "kondo_J025_L2_K1024": "examples/entropyKondo/resultsJ0.250000.h5",
    "siam_U02_L25_K512": "examples/entropySIAM/resultSIAMU0.200000.h5",
    "siam_zero_L25_K512": "examples/entropySIAM/resultSIAMZero.h5",
    "zdisc_flat_z1_L3": "independFiles/zAverage/hoppingElementsL3.h5",
}


def datasets(path):
    out = subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout
    return re.findall(r'DATASET "([^"]+)"', out)


def read(path, name):
    with tempfile.TemporaryDirectory() as td:
        binf = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-d", "/" + name, "-b", "LE", "-o", binf, path],
                       check=True, capture_output=True)
        return np.fromfile(binf, dtype="<f8")


def main():
    for key, rel in FILES.items():
        path = os.path.join(REF, rel)
        arrays = {}
        for name in datasets(path):
            arrays[name] = read(path, name)
        out = os.path.join(HERE, key + ".npz")
        np.savez_compressed(out, **arrays)
        print(key, sorted(arrays)[:5], "...", len(arrays), "datasets ->", out)


if __name__ == "__main__":
    sys.exit(main())
