"""Write the nuclear-gradient parity fixtures from the CPU oracle.

Runs in the build container only (it needs the oracle library; about 15 s).
For a bent, off-axis water (tests/test_oracle.py WATER_BENT) in STO-3G,
6-31G and 6-31G* (Cartesian d on O), it stores the oracle's converged RHF
density P and Fock matrix F (conver 1e-10) and the oracle's gradient dE/dX
for exactly that P and F (orc_rhf_gradient: the derivative integrals of
cython/grad.pyx contracted as mmd/forces.py:8-97 does, nuclear repulsion
included).  tests/test_oracle.py pins the oracle's gradient against the
reference's H2 forces (tests/test001.py:15-18) and central differences of
the energy.  A GPU test feeds the same P and F to mmd_rhf_gradient.

Usage:  python tests/golden/make_gradient_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from test_oracle import WATER_BENT  # noqa: E402

BASES = ("sto-3g", "6-31g", "6-31gs")


def main():
    out = {"geometry": WATER_BENT}
    rec = {"geometry": WATER_BENT, "conver": 1e-10, "bases": {}}
    for basis in BASES:
        om = oracle.OracleMolecule(WATER_BENT, os.path.join(ROOT, "mcmurchie-davidson_amd", "mmd", "basis",
                                                            basis + ".gbs"))
        r = oracle.rhf(om, conver=1e-10)
        t0 = time.perf_counter()
        g = om.rhf_gradient(r["P"], r["F"])
        key = basis.replace("-", "")
        out["P_" + key] = np.real(r["P"])
        out["F_" + key] = np.real(r["F"])
        out["grad_" + key] = g
        out["energy_" + key] = np.array(r["energy"].real)
        rec["bases"][basis] = {"nbasis": om.nbasis, "energy": float(r["energy"].real),
                               "oracle_gradient_s": time.perf_counter() - t0}
        print(basis, om.nbasis, rec["bases"][basis], flush=True)
    np.savez_compressed(os.path.join(HERE, "gradient_water.npz"), **out)
    with open(os.path.join(HERE, "gradient_provenance.json"), "w") as f:
        json.dump(rec, f, indent=1)


if __name__ == "__main__":
    main()
