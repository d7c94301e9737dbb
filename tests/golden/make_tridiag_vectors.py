"""Write golden vectors of the reference's own tridiag_periodic.f90.

Runs the reference routine compiled from its source by oracle/Makefile
(oracle/_ref/libref_tridiag.so, amdflang, no FMA) on seeded periodic systems
with the coefficients solxi/solxj/solxk build (diagonal 1, off-diagonals
-G/(1+2G), rhs scaled by 1/(1+2G)) and on general random diagonally dominant
systems, and stores inputs and outputs in tridiag_periodic_ref.npz.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

import oracle  # noqa: E402


def run_ref(fn, am, ac, ap, r, n):
    m1 = n + 1  # the callers pass arrays of length n1 = n1m + 1
    a = [np.zeros(m1) for _ in range(4)]
    for dst, src in zip(a, (am, ac, ap, r)):
        dst[:n] = src
    fn(*(x.ctypes.data for x in a), C.byref(C.c_int(1)), C.byref(C.c_int(n)), C.byref(C.c_int(m1)))
    return a[3][:n].copy()


def main():
    oracle.build(with_ref=True)
    fn = oracle.load_ref_tridiag()
    if fn is None:
        raise SystemExit("reference tridiag_periodic not built (needs the reference sources and amdflang)")
    rng = np.random.default_rng(2024)
    cases = []
    for n in (8, 12, 48, 64, 96, 256):
        for G in (1e-3, 0.05, 0.7, 30.0):
            ackl_b = 1.0 / (1.0 + 2.0 * G)
            am = np.full(n, -G * ackl_b); ap = am.copy(); ac = np.ones(n)
            r = rng.normal(size=n) * ackl_b
            cases.append((am, ac, ap, r))
        am = rng.uniform(-1, 1, n); ap = rng.uniform(-1, 1, n); ac = 3.0 + rng.uniform(0, 1, n)
        cases.append((am, ac, ap, rng.normal(size=n)))
    out = {}
    for i, (am, ac, ap, r) in enumerate(cases):
        n = len(r)
        out[f"am{i}"], out[f"ac{i}"], out[f"ap{i}"], out[f"r{i}"] = am, ac, ap, r
        out[f"x{i}"] = run_ref(fn, am, ac, ap, r, n)
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(HERE / "tridiag_periodic_ref.npz", **out)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
