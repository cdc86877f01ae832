"""Device-side structured generators against the oracle: K_ff bit-exact (same element-order
sums), reference R/P bit-exact (same arithmetic), Galerkin coarse operators within 1e-12."""
import numpy as np
import pytest

from compressiontests_amd import device as D
from compressiontests_amd.model import FemCantilever3D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def same(A, B):
    return (np.array_equal(A.rowptr, B.rowptr) and np.array_equal(A.col, B.col)
            and np.array_equal(A.val, B.val))


@pytest.mark.parametrize("dims", [(2, 2, 2), (5, 3, 4), (16, 16, 16)])
def test_cantilever_bit_exact(gpu, dims):
    nx, ny, nz = dims
    dens = 10 ** (4 * O.rng_uniform(1, nx * ny * nz) - 2)
    for d in (None, dens):
        got = D.cantilever_matrix(nx, ny, nz, dens=d).to_host()
        ref = O.cantilever(nx, ny, nz, dens=d)
        assert same(got, ref)


@pytest.mark.parametrize("n", [4, 8, 16])
def test_transfer_reference_bit_exact(gpu, n):
    R, P = D.transfer_reference(n, n, n)
    Ro, Po = O.rp_reference(n, n, n)
    assert same(R.to_host(), Ro)
    assert same(P.to_host(), Po)


def test_transfer_trilinear(gpu):
    n = 8
    R, P = D.transfer_trilinear(n, n, n)
    Po = O.p_trilinear(n, n, n)
    assert same(P.to_host(), Po)
    assert same(R.to_host(), O.transpose(Po))


@pytest.mark.parametrize("n,levels", [(8, 3), (16, 4)])
def test_galerkin_hierarchy(gpu, n, levels):
    dens = 10 ** (4 * O.rng_uniform(2, n ** 3) - 2)
    Ks = [K.to_host() for K in D.galerkin_hierarchy(n, n, n, levels, dens=dens)]
    A = O.cantilever(n, n, n, dens=dens)
    assert same(Ks[0], A)
    m = n
    for l in range(1, levels):
        P = O.p_trilinear(m, m, m)
        A = O.galerkin(A, P)
        m //= 2
        assert np.array_equal(Ks[l].rowptr, A.rowptr) and np.array_equal(Ks[l].col, A.col)
        assert np.abs(Ks[l].val - A.val).max() <= 1e-12 * np.abs(A.val).max()


def test_model_rhs_and_free_dofs(gpu):
    m = FemCantilever3D((4, 3, 2))
    A, U, b = m.CreateAndSolveLinearSystem()
    Ao = O.cantilever(4, 3, 2)
    assert same(A, Ao)
    assert np.allclose(U, O.manufactured(4, 3, 2), rtol=1e-15, atol=0)
    assert np.allclose(b, O.spmv(Ao, U), rtol=1e-12, atol=1e-12 * np.abs(b).max())
    free = m.FindFreeDofs()
    assert len(free) == m.NumDofsFree == A.nrows
    assert np.all((free // 3) % 5 != 0)


def test_device_generators_reproduce_frozen_16_fixtures(gpu):
    """The device assembly of the 16^3 K_ff and the device reference R / P against the frozen
    fixtures (tests/golden/cantilever16.json): the same bytes as the oracle wrote there."""
    import hashlib
    import json
    import os

    def sha(a, dtype):
        return hashlib.sha256(np.ascontiguousarray(a, dtype).astype(np.dtype(dtype).newbyteorder("<")).tobytes()).hexdigest()

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cantilever16.json")))
    K = D.cantilever_matrix(16, 16, 16).to_host()
    R, P = D.transfer_reference(16, 16, 16)
    for key, M in (("K_ff", K), ("R", R.to_host()), ("P", P.to_host())):
        d = fx[key]
        assert (M.nrows, M.ncols, len(M.col)) == (d["nrows"], d["ncols"], d["nnz"]), key
        assert sha(M.rowptr, np.int32) == d["rowptr_sha256"], key
        assert sha(M.col, np.int32) == d["col_sha256"], key
        assert sha(M.val, np.float64) == d["val_sha256"], key


def test_hierarchy_save_load_solves_identically(gpu, tmp_path):
    """A device hierarchy saved to the on-disk format and loaded back gives the same matrices (bit
    for bit, grids included) and a solver built on it returns the same solution bits."""
    from compressiontests_amd.solvers import HipMgPcgSolver
    n = 16
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    s = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-10)
    path = str(tmp_path / "h16.mgxh")
    D.save_hierarchy(path, s.hierarchy)
    assert D.hierarchy_info(path)[1:3] == (3, True)
    hier = D.load_hierarchy(path)
    assert len(hier) == 3
    for (A, R, P), (A2, R2, P2) in zip(s.hierarchy, hier):
        assert same(A.to_host(), A2.to_host())
        assert D.matrix_grid(A) == D.matrix_grid(A2) != (0, 0, 0)
        if R is not None:
            assert same(R.to_host(), R2.to_host()) and same(P.to_host(), P2.to_host())
    s.Initialize(f)
    x, st = s.Solve()
    s2 = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-10, hierarchy=hier)
    s2.Initialize(f)
    x2, st2 = s2.Solve()
    assert st.HasConverged and st2.NumIterationsRequired == st.NumIterationsRequired
    assert x2.tobytes() == x.tobytes()
