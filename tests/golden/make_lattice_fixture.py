"""Copy the expected lattice positions of the reference's lattice test
(tests/test_make_lattice.py: EXPECTED_FCC_POSITIONS, EXPECTED_BCC_POSITIONS,
EXPECTED_HEXAGONAL_POSITIONS) into tests/golden/lattice_expected.npz.
The arrays are read as literals from the test file's syntax tree (nothing of
the reference is imported or executed).

    python tests/golden/make_lattice_fixture.py /root/reference
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    src = open(os.path.join(ref_root, "tests", "test_make_lattice.py")).read()
    tree = ast.parse(src)
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name.startswith("EXPECTED_") and isinstance(node.value, ast.Call):
                out[name] = np.array(ast.literal_eval(node.value.args[0]), dtype=np.float64)
    np.savez(os.path.join(HERE, "lattice_expected.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
