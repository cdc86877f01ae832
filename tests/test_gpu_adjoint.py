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


def test_torch_autograd_wrapper(gpu):
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_adjoint_check.py")], capture_output=True, text=True,
                       timeout=280)
    sys.stdout.write(r.stdout[-2000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0
    assert "gradcheck ok" in r.stdout
