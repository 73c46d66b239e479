"""Writes the golden fixtures of tests/golden/ (run from the repository root):

    python tests/golden/make_golden.py

reference_setup_<case>.npz
    The products of the reference's own constructor and initLawsonRK
    (src/MultimodePropagation.cpp:4-199, src/RungeKutta.cpp), computed by the
    reference source files compiled as they lie (oracle/_ref/libref_setup.so, built by
    `make -C oracle ref` where the reference tree is present): tableau, dimensionless
    scalars, beta~, Q~, unique pairs, psi0, L, E1, E_i. These are reference-produced
    vectors; they let the oracle and the host plan be checked bit for bit on machines
    without the reference.
oracle_two_mode_gaussian_zL.npz
    The two-mode Gaussian workload (M=2, NT=4096, 5000 RK4 steps, Raman on,
    reference pair decode) at z = L from the oracle, for device checks that need no
    oracle run. Regression data of the oracle, not a reference output (the reference's
    stepping code needs FFTW, absent here).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

from mm_propagation_amd import configs  # noqa: E402

import oracle_binding as ob  # noqa: E402


def setup_cases():
    def negative_gamma_five_orders():
        sp = configs.two_mode_gaussian(nt=128)
        sp.nonlinearity_const = -sp.nonlinearity_const
        sp.dispersion_coefficients = [[0.1, -0.2, 0.02, 1e-4, -2e-6], [-2e3, 1e-3, -0.03, 3e-5, 1e-7]]
        return sp

    return {
        "two_mode_gaussian": (lambda: configs.two_mode_gaussian(nt=256), True),
        "two_mode_rk6_raman_off": (lambda: configs.two_mode_gaussian(nt=64, method_order=6), False),
        "two_mode_rk8": (lambda: configs.two_mode_gaussian(nt=64, method_order=8), True),
        "grin6_k_le_l": (lambda: configs.grin_fiber(M=6, nt=64), True),
        "grin4_symmetric": (lambda: configs.grin_fiber(M=4, nt=128, symmetric=True), True),
        "negative_gamma_five_orders": (negative_gamma_five_orders, True),
    }


def write_reference_setup():
    if not ob.ref_available():
        print("oracle/_ref/libref_setup.so missing: run make -C oracle ref (needs the reference tree)")
        return
    for name, (make, raman) in setup_cases().items():
        r = ob.ref_setup(make(), raman=raman)
        path = os.path.join(HERE, f"reference_setup_{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in r.items()})
        print("wrote", path)


def write_oracle_two_mode():
    sp = configs.two_mode_gaussian()
    o = ob.Oracle(sp, n_threads=2)
    o.compute_lawson_rk(sp.n_steps)
    fields = o.get_result()
    psi, z = o.get_state()
    o.close()
    path = os.path.join(HERE, "oracle_two_mode_gaussian_zL.npz")
    np.savez_compressed(path, fields=fields, z=np.float64(z), steps=np.int64(sp.n_steps))
    print("wrote", path)


if __name__ == "__main__":
    write_reference_setup()
    write_oracle_two_mode()
