"""Write the regression fixtures in this folder.

The reference ships no golden vectors for this path (its tests read external
configuration files that are not in its repository) and cannot be built
here, so these fixtures are outputs of the oracle (oracle/qcd_oracle.c),
written once after the oracle was pinned by the closed-form, brute-force and
covariance checks of tests/test_oracle.py.  They anchor later versions of the
oracle and of the GPU path to the same numbers: a change that alters a
result must change these files on purpose.

usage: python tests/golden/make_golden.py   (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

L = (4, 4, 2, 2)            # T, X, Y, Z: small enough to commit in full
KAPPA, MU, EPS = 0.125, 0.02, 0.2
SEED_U, SEED_PSI = 9101, 9102


def main():
    geo = O.Geometry(L, theta=(np.pi, 0.0, 0.0, 0.0))
    u = O.gauge_generate(L, EPS, SEED_U)
    psi = O.spinor_gaussian(L, SEED_PSI)
    m = 1 / (2 * KAPPA) - 4
    d = O.dslash(geo, psi, u, m, MU)
    q = O.qtm(geo, psi, u, m, MU)
    g = O.gauss3d(geo, psi, u, 4.0, 1)
    ape = O.ape3d(geo, u, 0.5)
    stout = O.stout(geo, u, 0.1, 4)
    b = np.zeros_like(psi)
    b[0, 1, 2] = 1.0
    rc, x, _, _ = O.invert(geo, b, u, KAPPA, MU, n=8, op=0, keep_q1=False, max_res=1e-12)
    assert rc == 0
    np.savez(os.path.join(HERE, "wilson_4x4x2x2.npz"), L=np.array(L), kappa=KAPPA, mu=MU, eps=EPS,
             seed_u=SEED_U, seed_psi=SEED_PSI, u=u, psi=psi, dpsi=d, qpsi=q, gauss_t1=g, ape=ape,
             stout4=stout, b=b, x=x, plaquette=O.plaquette(geo, u))
    print("wrote", os.path.join(HERE, "wilson_4x4x2x2.npz"))


if __name__ == "__main__":
    main()
