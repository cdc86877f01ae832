"""SIMP compliance and its adjoint sensitivity on the device.

Modified SIMP: E_e = Emin + rho_e^p (E0 - Emin), K(rho) = sum_e E_e Ke0 (unit-modulus top3d
element, FemCantilever3D.cs:32-120, scaled per element as ElementDensities does,
FemCantileverBase.cs:99-100).  Compliance C = f^T u with K u = f.  Because K is symmetric the
adjoint solve is the forward solve (lambda = -u), so

    dC/drho_e = -u^T (dK/drho_e) u = -p rho_e^(p-1) (E0 - Emin) u_e^T Ke0 u_e,

computed per element by the `mgx_simp_gradient` kernel from the FP64 MG-PCG solution.
"""
import numpy as np

from .device import DeviceArray
from .model import FemCantilever3D
from .solvers import HipMgPcgSolver
from . import kernels


def simp_moduli(rho, penal=3.0, Emin=1e-9, E0=1.0):
    rho = np.asarray(rho, np.float64)
    return Emin + rho ** penal * (E0 - Emin)


def compliance_and_gradient(dims, rho, penal=3.0, Emin=1e-9, E0=1.0, nu=0.3, levels=3, tolerance=1e-12,
                            smoother="gs", nu_smooth=1, load=None, precision="fp64"):
    """Returns (C, dC/drho, u, stats) for the x = 0 clamped cantilever with the tip load (or
    `load`, a free-dof vector).  precision "mixed": the solve is the FP64-refined FP32 MG-PCG."""
    nx, ny, nz = dims
    model = FemCantilever3D(dims, ElementDensities=simp_moduli(rho, penal, Emin, E0), PoissonRatio=nu, rhs="tip")
    f = model.rhs() if load is None else np.ascontiguousarray(load, np.float64)
    s = HipMgPcgSolver(model, levels=levels, smoother=smoother, nu=nu_smooth, tolerance=tolerance,
                       max_iterations=2000, precision=precision)
    s.Initialize(f)
    u, st = s.Solve()
    C = float(f @ u)
    g = kernels.simp_gradient(nx, ny, nz, DeviceArray.from_numpy(rho), DeviceArray.from_numpy(u), penal, Emin, E0, nu)
    return C, g.numpy(), u, st
