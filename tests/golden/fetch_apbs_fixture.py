"""Copies the APBS 3.4.1 single-atom fixture held by the reference
(test_kappa_function/single_atom-PE0.dx, single_atom.pqr; inputs in
test_kappa_function/apbs_single_atom.in: pdie 4, sdie 78.4, no salt, mdh
boundary, spl0 charges) into tests/golden/.  Data files only; run where the
reference is mounted."""
import os
import shutil

SRC = "/root/reference/test_kappa_function"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    shutil.copy(os.path.join(SRC, "single_atom-PE0.dx"), os.path.join(HERE, "apbs_single_atom_pot.dx"))
    shutil.copy(os.path.join(SRC, "single_atom.pqr"), os.path.join(HERE, "apbs_single_atom.pqr"))
