"""Writes tests/golden/coeffs_fixture.npz: lens and critical-curve polynomial
coefficients evaluated at seeded parameter points from the sympy expressions that
the reference PRINTS as cell outputs of notebooks/ComplexPolynomialCoefficients.ipynb
(cells 10 and 13: lens polynomials N=2, N=3; cells 19 and 22: critical-curve
polynomials N=2, N=3).  Those printed outputs are data held by the reference; this
script parses them with sympy and evaluates them at 50 significant digits with
mpmath, then stores inputs and rounded outputs.  Only the resulting numbers are
committed.  Run in a container where the reference tree is mounted:

    python tests/golden/make_coeff_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys

import mpmath
import numpy as np
import sympy as sp

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
NB = os.path.join(REF, "notebooks", "ComplexPolynomialCoefficients.ipynb")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "coeffs_fixture.npz")


def cell_exprs(nb, idx):
    text = "".join("".join(o.get("text", "")) for o in nb["cells"][idx]["outputs"])
    parts = re.split(r"p_(\d+)\s*=", text)[1:]
    exprs = {}
    for k in range(0, len(parts), 2):
        exprs[int(parts[k])] = parts[k + 1].strip()
    return [exprs[i] for i in sorted(exprs)]


def to_sympy(s):
    s = s.replace("conjugate(w)", "wb").replace("conjugate(r[3])", "r3b")
    s = s.replace("e[1]", "e1").replace("e[2]", "e2").replace("r[3]", "r3")
    s = s.replace("exp(-I*\\phi)", "x")
    syms = {k: sp.Symbol(k) for k in ["w", "wb", "a", "e1", "e2", "r3", "r3b", "x"]}
    return sp.sympify(s, locals=syms), syms


def main():
    mpmath.mp.dps = 50
    nb = json.load(open(NB))
    rng = np.random.default_rng(20240611)
    npts = 64
    w = rng.uniform(-2, 2, npts) + 1j * rng.uniform(-2, 2, npts)
    a = rng.uniform(0.15, 1.5, npts)
    e1 = rng.uniform(0.05, 0.9, npts)
    e2 = (1 - e1) * rng.uniform(0.05, 0.95, npts)
    r3 = rng.uniform(0.3, 2.0, npts) * np.exp(1j * rng.uniform(0, 2 * np.pi, npts))
    phi = rng.uniform(-np.pi, np.pi, npts)
    out = dict(w=w, a=a, e1=e1, e2=e2, r3=r3, phi=phi)
    for name, idx in [("lens2", 10), ("lens3", 13), ("crit2", 19), ("crit3", 22)]:
        exprs = cell_exprs(nb, idx)
        funcs = []
        for s in exprs:
            e, syms = to_sympy(s)
            args = [syms[k] for k in ["w", "wb", "a", "e1", "e2", "r3", "r3b", "x"]]
            funcs.append(sp.lambdify(args, e, modules="mpmath"))
        vals = np.zeros((npts, len(funcs)), complex)
        for i in range(npts):
            W = mpmath.mpc(w[i].real, w[i].imag)
            R3 = mpmath.mpc(r3[i].real, r3[i].imag)
            X = mpmath.exp(mpmath.mpc(0, -phi[i]))
            args = [W, mpmath.conj(W), mpmath.mpf(a[i]), mpmath.mpf(e1[i]), mpmath.mpf(e2[i]),
                    R3, mpmath.conj(R3), X]
            for k, f in enumerate(funcs):
                vals[i, k] = complex(f(*args))
        out[name] = vals
        print(name, vals.shape)
    np.savez(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
