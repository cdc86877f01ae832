"""SIMP compliance adjoint: the device gradient against central finite differences
(h = 1e-6, 16 seeded elements) and against the oracle's FP64 CPU gradient; the
torch.autograd.Function wrapper (test harness only) is checked in a child process so that
torch's HIP runtime is loaded before libmgx.so there (see compressiontests_amd/_lib.py)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from compressiontests_amd.adjoint import compliance_and_gradient, simp_moduli
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
    CPU gradient on the same displacements.

    Central differences on 16 seeded elements, from FP64 MG-PCG solves to 1e-13 at rho_e +- h:
    * h = 1e-4, every element to 1e-6 of its own gradient entry.  The quotient
      (C+ - C-) / 2h is evaluated without the cancellation of C+ - C-: with K+ u+ = K- u- = f and
      K symmetric, C+ - C- = f.u+ - f.u- = u-.K- u+ - u+.K+ u- = -(E+ - E-) u+_e . Ke0 u-_e, an
      identity of the exact solutions, so its relative rounding is that of the two solves, and
      the only other error is the O(h^2) truncation (below 4e-7 for rho >= 0.1).  The element
      form u+_e . Ke0 u-_e is the oracle's quadratic form polarised.
    * the same quotient as the plain difference of the two compliances agrees with it up to the
      difference's round-off floor (100 eps C / h);
    * h = 1e-6, the plain difference against the adjoint to that floor."""
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
    eps = np.finfo(float).eps

    def quad(u_a, u_b):
        """u_a,e . Ke0 u_b,e of every element (the oracle's form at penal 1, E0 - Emin = 1)."""
        ones = np.ones(n ** 3)
        qs = O.simp_gradient(n, n, n, ones, u_a + u_b, penal=1.0, Emin=0.0, E0=1.0)
        qd = O.simp_gradient(n, n, n, ones, u_a - u_b, penal=1.0, Emin=0.0, E0=1.0)
        return -(qs - qd) / 4

    def fd(e, h):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        Cp, _, up, sp = compliance_and_gradient((n, n, n), rp, **fdkw)
        Cm, _, um, sm = compliance_and_gradient((n, n, n), rm, **fdkw)
        assert sp.HasConverged and sm.HasConverged
        plain = (Cp - Cm) / (2 * h)
        dE = simp_moduli(rp[e]) - simp_moduli(rm[e])
        stable = -dE * quad(up, um)[e] / (2 * h)
        return plain, stable

    bad = []
    floor4 = 100 * eps * abs(C) / 1e-4
    for e in elems:
        plain, stable = fd(e, 1e-4)
        rel = abs(stable - g[e]) / abs(g[e])
        print(f"h=1e-4 element {e} rho {rho[e]:.3f}: adjoint {g[e]:.12e} fd {stable:.12e} relative {rel:.2e}; "
              f"plain difference {plain:.9e} (|diff| {abs(plain - stable):.1e}, floor {floor4:.1e})", flush=True)
        if rel > 1e-6 or abs(plain - stable) > floor4:
            bad.append((1e-4, int(e), stable, plain, g[e]))
    floor6 = max(1e-6 * gscale, 100 * eps * abs(C) / 1e-6)
    for e in elems[:4]:
        plain, _ = fd(e, 1e-6)
        print(f"h=1e-6 element {e}: adjoint {g[e]:.9e} fd {plain:.9e} |diff| {abs(plain - g[e]):.1e} "
              f"(bound {floor6:.1e})", flush=True)
        if abs(plain - g[e]) > floor6:
            bad.append((1e-6, int(e), plain, g[e]))
    assert not bad, bad


def test_torch_autograd_wrapper(gpu):
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_adjoint_check.py")], capture_output=True, text=True,
                       timeout=280)
    sys.stdout.write(r.stdout[-2000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0
    assert "gradcheck ok" in r.stdout
