#!/usr/bin/env python3
"""Copy the input data files the reference ships with its 2-D FWI examples
into tests/golden/reference_inputs/ (data only: the run configuration, the
prior bounds and a saved ADVI posterior), so the tests can read them where
the reference checkout is absent.

    examples/fwi2d/input/config.ini                      run parameters of the example
    examples/fwi2d/input/prior.txt                       (lower, upper) per depth row
    examples/fwi2d_vpr/input/Linear_diagonal_parameter.npy
                                                         [u, exp(diag)] of a mean-field
                                                         Linear flow, 25 000 cells

    python tests/golden/copy_reference_inputs.py [REFERENCE_ROOT]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "config.ini": "examples/fwi2d/input/config.ini",
    "prior.txt": "examples/fwi2d/input/prior.txt",
    "Linear_diagonal_parameter.npy": "examples/fwi2d_vpr/input/Linear_diagonal_parameter.npy",
}


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(HERE, "reference_inputs")
    os.makedirs(out, exist_ok=True)
    for name, rel in FILES.items():
        shutil.copyfile(os.path.join(root, rel), os.path.join(out, name))
        print(f"{rel} -> tests/golden/reference_inputs/{name}")


if __name__ == "__main__":
    main()
