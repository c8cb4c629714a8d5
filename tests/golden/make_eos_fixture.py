"""
Writes ``mom6_eos_wright.npz``: the Wright (1997) full-range density check points of the reference's own test data
(tests/data/mom6_eos_reference.npz, values of compiled MOM6 WRIGHT_FULL, used by tests/test_eos_nonlinear.py:20-60):
the 2048 (T, S, p) points, their densities ``w_rho`` and the 15 coefficients as the Fortran compiler folds them.

Run in a container that holds the reference at /root/reference:  python tests/golden/make_eos_fixture.py
"""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tests/data/mom6_eos_reference.npz")
WRIGHT = ("a0", "a1", "a2", "b0", "b1", "b2", "b3", "b4", "b5", "c0", "c1", "c2", "c3", "c4", "c5")


def main() -> None:
    z = np.load(SRC)                                   # allow_pickle=False: plain numeric and text arrays only
    names = [str(n) for n in z["coef_names"]]
    values = z["coef_values"]
    coef = np.array([values[names.index(n)] for n in WRIGHT])
    np.savez(Path(__file__).with_name("mom6_eos_wright.npz"), T=z["T"], S=z["S"], p=z["p"], w_rho=z["w_rho"],
             coef_names=np.array(WRIGHT), coef_values=coef)


if __name__ == "__main__":
    main()
