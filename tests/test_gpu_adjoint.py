"""SIMP compliance adjoint: the device gradient against central finite differences
(h = 1e-6, 16 seeded elements) and against the oracle's FP64 CPU gradient; the
torch.autograd.Function wrapper (test harness only) is checked in a child process so that
torch's HIP runtime is loaded before libmgx.so there (see compressiontests_amd/_lib.py)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from compressiontests_amd.adjoint import compliance_and_gradient
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_gradient_matches_oracle_and_fd(gpu):
    n = 8
    rho = 0.1 + 0.9 * O.rng_uniform(20261203 + 5, n ** 3)
    C, g, u, st = compliance_and_gradient((n, n, n), rho, tolerance=1e-13)
    assert st.HasConverged
    ref = O.simp_gradient(n, n, n, rho, u)
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(g <= 0)             # adding material never increases compliance
    rng = np.random.default_rng(20261203)
    h = 1e-6
    # central differences at h = 1e-6 carry a round-off floor of ~C eps / h (C ~ 40 here, so
    # ~4e-9 absolute); the agreement is measured relative to the largest gradient entry
    gscale = np.abs(g).max()
    for e in rng.choice(n ** 3, 16, replace=False):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        Cp = compliance_and_gradient((n, n, n), rp, tolerance=1e-13)[0]
        Cm = compliance_and_gradient((n, n, n), rm, tolerance=1e-13)[0]
        fd = (Cp - Cm) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * gscale, (e, fd, g[e])


@pytest.mark.slow
def test_gradient_96_simp_structure(gpu):
    """The 96^3 SIMP structure (884,736 elements, rho ~ U[0.1, 1], E = Emin + rho^3 (E0 - Emin),
    E0 = 1, Emin = 1e-9): compliance from the FP64-refined MG-PCG (FP32 V-cycle inside FP64
    iterative refinement) and the device gradient of every element against the oracle's FP64
    CPU gradient on the same displacements.  Central differences (FP64 MG-PCG solves to 1e-13):
    at h = 1e-6 on 16 seeded elements the quotient carries a round-off floor of tens of
    eps C / h (a few 1e-5 of the largest gradient entry), so it is held to that floor; at
    h = 1e-4 (4 elements; truncation ~ (h / rho)^2 of the entry) to 1e-6 of the largest
    gradient entry."""
    n = 96
    rho = 0.1 + 0.9 * O.rng_uniform(20261203 + 5, n ** 3)
    C, g, u, st = compliance_and_gradient((n, n, n), rho, levels=5, tolerance=1e-11, precision="mixed")
    assert st.HasConverged
    ref = O.simp_gradient(n, n, n, rho, u)
    # u_e^T Ke0 u_e is a 576-term quadratic form with cancellation; different summation orders
    # agree to ~1e-12 of the largest entry here
    assert np.abs(g - ref).max() <= 1e-11 * np.abs(ref).max()
    assert np.all(g <= 0)
    gscale = np.abs(g).max()
    rng = np.random.default_rng(20261203 + 5)
    elems = rng.choice(n ** 3, 16, replace=False)
    fdkw = dict(levels=5, tolerance=1e-13)

    def fd(e, h):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        return (compliance_and_gradient((n, n, n), rp, **fdkw)[0] - compliance_and_gradient((n, n, n), rm, **fdkw)[0]) / (2 * h)

    # measured: the +h and -h compliances differ from their exact values by ~30 eps C (solve and
    # summation rounding), so the quotient at h = 1e-6 is held to 100 eps C / h
    floor = 100 * np.finfo(float).eps * abs(C) / 1e-6
    bad = []
    for h, es, tol in ((1e-6, elems, max(1e-6 * gscale, floor)), (1e-4, elems[:4], 1e-6 * gscale)):
        for e in es:
            q = fd(e, h)
            print(f"h={h:g} element {e}: adjoint {g[e]:.9e} fd {q:.9e} |diff|/max|g| {abs(q - g[e]) / gscale:.2e}"
                  f" (bound {tol / gscale:.1e})")
            if abs(q - g[e]) > tol:
                bad.append((h, int(e), q, g[e]))
    assert not bad, bad


def test_torch_autograd_wrapper(gpu):
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_adjoint_check.py")], capture_output=True, text=True,
                       timeout=280)
    sys.stdout.write(r.stdout[-2000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0
    assert "gradcheck ok" in r.stdout
