"""The library's plane-stress cantilever generators on the host (mgx_cantilever2d_system_host,
mgx_cantilever2d_transfer_host) against the oracle's restatement of FemCantilever2D.cs and
FemCantileverBase.cs, bit for bit, and the FemCantilever2D model's sizes, free dofs, manufactured
displacements and coarsening.  Host code only (no kernel is launched)."""
import numpy as np
import pytest

from compressiontests_amd import _lib as L
from compressiontests_amd import device as D
from compressiontests_amd.model import FemCantilever2D
from oracle import oracle as O


def same_csr(a, b):
    return (a.nrows == b.nrows and a.ncols == b.ncols and np.array_equal(a.rowptr, b.rowptr)
            and np.array_equal(a.col, b.col) and a.val.tobytes() == b.val.tobytes())


@pytest.mark.parametrize("nx,ny,hetero", [(8, 4, False), (16, 2, True), (5, 3, True), (1, 1, False),
                                          (256, 16, False), (64, 8, True)])
def test_cant2d_generators_bit_exact(nx, ny, hetero):
    dens = 10 ** (3 * O.rng_uniform(77, nx * ny) - 1.5) if hetero else None
    K = D.cantilever2d_system_host(nx, ny, 1.0, 0.3, dens)
    assert same_csr(K, O.cantilever2d(nx, ny, 1.0, 0.3, dens))
    if nx > 1 and ny > 1:
        R, P = D.cantilever2d_transfer_host(nx, ny)
        Ro, Po = O.cantilever2d_transfer(nx, ny)
        assert same_csr(R, Ro) and same_csr(P, Po)


def test_cant2d_model():
    m = FemCantilever2D((256, 16), (20.0, 1.0, 1.0))
    assert m.NumDofsAll == 2 * 257 * 17 and m.NumDofsFree == 2 * 256 * 17
    assert np.array_equal(m.CalcFreeDisplacements(), O.manufactured2d(256, 16, (20.0, 1.0, 1.0)))
    f = m.FindFreeDofs()
    assert len(f) == m.NumDofsFree and np.all(np.diff(f) > 0) and not np.any((f // 2) % 257 == 0)
    c = m.CreateCoarserModel()
    assert c.NumElements == (128, 8) and c.Lengths == m.Lengths and c.CantileverDepth == 1.0
    assert c.ElementDensities is None
    with pytest.raises(ValueError):
        FemCantilever2D((4, 4, 4))
    with pytest.raises(ValueError):
        FemCantilever2D((4, 4), (1.0, 1.0))


def test_cant2d_generator_rejects_invalid_input():
    with pytest.raises(L.MgxError):
        D.cantilever2d_system_host(0, 4)
    with pytest.raises(ValueError):
        D.cantilever2d_system_host(4, 4, dens=np.ones(3))
