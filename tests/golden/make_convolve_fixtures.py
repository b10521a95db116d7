"""Copies the reference's Gaussian-broadening sample data into fixtures.

The reference's convolve/ directory holds the stick lists spec.line and
spec.line.adia (energy in eV, intensity) and the envelopes its Fortran
program computed from them, output and output.adia (convolve/main.f90:
2002 grid points x = -0.1 + i (0.5/2001), i = 0..2001, printed as x + 2.7807
with f15.8; fwhm 5e-3 eV; the gaussian of convolve/sub.f90 with
sigma = fwhm / (2 sqrt(2 ln 2)); the constants -0.1, 0.4 and 2.7807 are
default-real literals, i.e. single precision).  These are data, not source:
the fixtures are the files unchanged.  (output.dia does not belong to
spec.line.dia and is not used.)

Run in a checkout that has the reference next to it:
    python tests/golden/make_convolve_fixtures.py /path/to/reference
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = ["spec.line", "output", "spec.line.adia", "output.adia"]


def main(ref):
    dst = os.path.join(HERE, "convolve")
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(ref, "convolve", f), os.path.join(dst, f))
        print("copied", f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
