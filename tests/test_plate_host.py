"""The library's Kirchhoff plate generators on the host (mgx_plate_free_dofs,
mgx_plate_system_host, mgx_plate_transfer_host) against the oracle's restatement of FemPlate.cs,
bit for bit, and the FemPlate model's sizes and coarsening.  Host code only (no kernel is
launched)."""
import numpy as np
import pytest

from compressiontests_amd import device as D
from compressiontests_amd.model import FemPlate
from oracle import oracle as O

ARGS = (3.0, 2.0, 0.1, 52416.0, 0.3, -10.0)


def same_csr(a, b):
    return (a.nrows == b.nrows and a.ncols == b.ncols and np.array_equal(a.rowptr, b.rowptr)
            and np.array_equal(a.col, b.col) and a.val.tobytes() == b.val.tobytes())


@pytest.mark.parametrize("nx,ny", [(8, 6), (16, 8), (12, 4), (4, 2), (5, 3), (64, 32)])
def test_plate_generators_bit_exact(nx, ny):
    """K_ff, b, the free dofs and both transfers equal the oracle's bit for bit (odd counts too:
    then no node lies on x = Lx/2 or y = Ly/2 and the supports are the three edges)."""
    assert np.array_equal(D.plate_free_dofs(nx, ny, 3.0, 2.0), O.plate_free_dofs(nx, ny, 3.0, 2.0))
    K, b = D.plate_system_host(nx, ny, *ARGS)
    Ko, bo = O.plate(nx, ny, *ARGS)
    assert same_csr(K, Ko) and b.tobytes() == bo.tobytes()
    if nx > 1 and ny > 1:
        R, P = D.plate_transfer_host(nx, ny, 3.0, 2.0)
        Ro, Po = O.plate_transfer(nx, ny, 3.0, 2.0)
        assert same_csr(R, Ro) and same_csr(P, Po)


def test_plate_model_sizes_and_coarsening():
    m = FemPlate((16, 8), (3.0, 2.0, 0.1), 52416.0, 0.3, -10.0)
    nx1, ny1 = 17, 9
    assert m.NumDofsAll == 3 * nx1 * ny1
    # FemPlate.cs:118-122 for even element counts
    assert m.NumDofsFree == m.NumDofsAll - 2 * ny1 - 3 * nx1 + 6
    assert np.array_equal(m.FindFreeDofs(), O.plate_free_dofs(16, 8, 3.0, 2.0))
    c = m.CreateCoarserModel()
    assert c.NumElements == (8, 4) and c.Lengths == m.Lengths and c.Thickness == m.Thickness
    assert (c.ElasticityModulus, c.PoissonRatio, c.DistributedLoad) == (52416.0, 0.3, -10.0)
    assert FemPlate((1, 3), (1.0, 1.0, 1.0)).CreateCoarserModel().NumElements == (1, 1)


def test_plate_generator_rejects_invalid_input():
    from compressiontests_amd import _lib as L
    with pytest.raises(L.MgxError):
        D.plate_system_host(0, 4, *ARGS)
    with pytest.raises(L.MgxError):
        D.plate_system_host(4, 4, 3.0, 2.0, 0.1, 52416.0, 0.5, -10.0)
    with pytest.raises(ValueError):
        FemPlate((4, 4, 4), (1.0, 1.0, 1.0))
