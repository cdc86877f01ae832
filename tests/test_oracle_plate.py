"""Pins the oracle's restatement of the reference's Kirchhoff plate workload (FemPlate.cs,
PlateElementRectangle4Nodes3Dofs.cs, PlateElementVectorAssembler.cs, IGeometricMultigridModel.cs
on two axes) and the Gauss-Seidel node groups, against what the mathematics fixes:

* the element matrix against an independent derivation: the Adini-Clough-Melosh rectangle from its
  12-term polynomial w = sum c_k x^p y^q, nodal dofs (w, dw/dy, dw/dx) at the corners
  (-a,-b), (a,-b), (a,b), (-a,b), K = A^-T (int B^T D B) A^-1 with B = (w_xx, w_yy, 2 w_xy),
  D = E t^3 / (12 (1 - nu^2)) [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu)/2]], by Gauss quadrature
  (exact for these degrees); the other rotation conventions are shown to miss it;
* the element load vector against the consistent load q int N dA of the same shape functions;
* the free dofs against FemPlate's NumDofsFree expression and the five support lines;
* K_ff and b against a dense element-by-element scatter (bit for bit: the same sums in element
  order), and the transfers against a dense restatement of the window loops followed by the
  free-dof submatrix, with their row-sum properties;
* the node groups (plate: 2 dofs at supported nodes, 3 elsewhere; cantilever: the triples), and
  the group sweep equal to the triple sweep on the cantilever.
"""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

MON = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3), (3, 1), (1, 3)]


def _d(x, y, dx, dy):
    out = []
    for p, q in MON:
        if p < dx or q < dy:
            out.append(0.0)
            continue
        c = math.perm(p, dx) * math.perm(q, dy)
        out.append(c * x ** (p - dx) * y ** (q - dy))
    return np.array(out)


def acm_A(a, b, conv=(1, 1, False)):
    sx, sy, swap = conv
    rows = []
    for x, y in [(-a, -b), (a, -b), (a, b), (-a, b)]:
        rows.append(_d(x, y, 0, 0))
        rows.append(sx * (_d(x, y, 0, 1) if not swap else _d(x, y, 1, 0)))
        rows.append(sy * (_d(x, y, 1, 0) if not swap else _d(x, y, 0, 1)))
    return np.array(rows)


def acm_stiffness(a, b, E, nu, t, conv=(1, 1, False)):
    D = E * t ** 3 / (12 * (1 - nu ** 2))
    Dm = D * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    g, w = np.polynomial.legendre.leggauss(5)
    Kc = np.zeros((12, 12))
    for i in range(5):
        for j in range(5):
            x, y = a * g[i], b * g[j]
            B = np.array([_d(x, y, 2, 0), _d(x, y, 0, 2), 2 * _d(x, y, 1, 1)])
            Kc += w[i] * w[j] * a * b * (B.T @ Dm @ B)
    Ai = np.linalg.inv(acm_A(a, b, conv))
    return Ai.T @ Kc @ Ai


@pytest.mark.parametrize("a,b,E,nu,t", [(0.3, 0.2, 7.0, 0.27, 0.15), (1.0, 1.0, 1.0, 0.3, 1.0),
                                        (0.03125, 0.03125, 52416.0, 0.3, 0.1), (0.2, 0.7, 3.0, 0.0, 2.0)])
def test_plate_ke_matches_acm_derivation(a, b, E, nu, t):
    Ko = O.plate_ke(a, b, E, nu, t)
    assert np.array_equal(Ko, Ko.T)
    K = acm_stiffness(a, b, E, nu, t)
    assert np.abs(K - Ko).max() <= 1e-12 * np.abs(Ko).max()
    # every other sign / order of the two rotations gives a different matrix
    for conv in itertools.product([1, -1], [1, -1], [False, True]):
        if conv != (1, 1, False):
            assert np.abs(acm_stiffness(a, b, E, nu, t, conv) - Ko).max() > 1e-3 * np.abs(Ko).max()


def test_plate_ke_rigid_modes():
    a, b = 0.4, 0.25
    K = O.plate_ke(a, b, 5.0, 0.3, 0.2)
    for c0, c1, c2 in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        u = []
        for x, y in [(-a, -b), (a, -b), (a, b), (-a, b)]:
            u += [c0 + c1 * x + c2 * y, c2, c1]                  # (w, dw/dy, dw/dx)
        assert np.abs(K @ np.array(u)).max() <= 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) <= 1e-10 * ev.max()) == 3 and ev.min() > -1e-10 * ev.max()


def test_plate_load_is_consistent():
    a, b, q = 0.3, 0.2, -10.0
    Ai = np.linalg.inv(acm_A(a, b))
    g, w = np.polynomial.legendre.leggauss(5)
    m = sum(w[i] * w[j] * a * b * _d(a * g[i], b * g[j], 0, 0) for i in range(5) for j in range(5))
    assert np.allclose(O.plate_load(a, b, q), q * (m @ Ai), rtol=1e-13, atol=1e-15)


def support_nodes(nx, ny):
    """Nodes on x = 0, x = Lx/2, y = 0, y = Ly/2, y = Ly (FemPlate.cs:181-196), even nx, ny."""
    ix, iy = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    return ((ix == 0) | (ix == nx // 2) | (iy == 0) | (iy == ny // 2) | (iy == ny)).reshape(-1)


@pytest.mark.parametrize("nx,ny", [(48, 32), (8, 6), (4, 2), (12, 16)])
def test_plate_free_dofs(nx, ny):
    dofs = O.plate_free_dofs(nx, ny, 3.0, 2.0)
    nodes = (nx + 1) * (ny + 1)
    # FemPlate.cs:126-130: all dofs - 2 rows of constant x - 3 rows of constant y + 6 crossings
    assert len(dofs) == 3 * nodes - 2 * (ny + 1) - 3 * (nx + 1) + 6
    fixed = support_nodes(nx, ny)
    ref = [3 * n + d for n in range(nodes) for d in range(3) if not (fixed[n] and d == 0)]
    assert dofs.tolist() == ref


def dense_plate(nx, ny, Lx, Ly, t, E, nu, q):
    a, b = 0.5 * (Lx / nx), 0.5 * (Ly / ny)
    Ke, fe = O.plate_ke(a, b, E, nu, t), O.plate_load(a, b, q)
    N = 3 * (nx + 1) * (ny + 1)
    K = np.zeros((N, N))
    f = np.zeros(N)
    for ey in range(ny):
        for ex in range(nx):
            nodes = [ex + (nx + 1) * ey, ex + 1 + (nx + 1) * ey, ex + 1 + (nx + 1) * (ey + 1), ex + (nx + 1) * (ey + 1)]
            dofs = [3 * n + d for n in nodes for d in range(3)]
            K[np.ix_(dofs, dofs)] += Ke
            f[dofs] += fe
    free = O.plate_free_dofs(nx, ny, Lx, Ly)
    return K[np.ix_(free, free)], f[free]


@pytest.mark.parametrize("nx,ny", [(8, 6), (12, 8)])
def test_plate_assembly_bit_exact_against_dense_scatter(nx, ny):
    args = (nx, ny, 3.0, 2.0, 0.1, 52416.0, 0.3, -10.0)
    A, b = O.plate(*args)
    Kd, fd = dense_plate(*args)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols))
    assert np.all(np.diff(A.rowptr) > 0)
    for i in range(A.nrows):                                     # sorted columns
        assert np.all(np.diff(A.col[A.rowptr[i]:A.rowptr[i + 1]]) > 0)
    nz = Kd != 0
    assert np.array_equal(S.toarray()[nz], Kd[nz])
    assert S.nnz >= np.count_nonzero(Kd)
    assert np.array_equal(b, fd)
    ev = np.linalg.eigvalsh(Kd)
    assert ev.min() > 0                                         # the supports remove every rigid mode


def dense_transfer(nx, ny, Lx, Ly):
    """IGeometricMultigridModel.cs:76-194 on two axes, dense, then the free-dof submatrices."""
    nf = [nx + 1, ny + 1]
    cn = [n >> 1 if n > 1 else n for n in (nx, ny)]
    nc = [c + 1 for c in cn]
    nid = [nx / cn[0], ny / cn[1]]
    R = np.zeros((3 * nc[0] * nc[1], 3 * nf[0] * nf[1]))
    P = np.zeros((3 * nf[0] * nf[1], 3 * nc[0] * nc[1]))
    sumP = np.zeros(nf[0] * nf[1])
    for a0 in range(nc[0]):
        for a1 in range(nc[1]):
            cur = [a0 * nid[0], a1 * nid[1]]
            s = [max(0, math.ceil(math.nextafter(cur[i] - nid[i], math.inf))) for i in range(2)]
            e = [min(nf[i], math.floor(math.nextafter(cur[i] + nid[i], -math.inf)) + 1) for i in range(2)]
            cid = a0 + nc[0] * a1
            sr = 0.0
            for w0 in range(s[0], e[0]):
                for w1 in range(s[1], e[1]):
                    infl = 1 - math.sqrt(((w0 - cur[0]) / nid[0]) ** 2 + ((w1 - cur[1]) / nid[1]) ** 2)
                    fid = w0 + nf[0] * w1
                    sr += infl
                    sumP[fid] += infl
                    R[3 * cid, 3 * fid] = infl
                    P[3 * fid, 3 * cid] = infl
            R[3 * cid] /= sr
            for d in (1, 2):
                R[3 * cid + d, d::3] = R[3 * cid, 0::3]
    for fid in range(nf[0] * nf[1]):
        P[3 * fid] /= sumP[fid]
        for d in (1, 2):
            P[3 * fid + d, d::3] = P[3 * fid, 0::3]
    ff, cf = O.plate_free_dofs(nx, ny, Lx, Ly), O.plate_free_dofs(cn[0], cn[1], Lx, Ly)
    return R[np.ix_(cf, ff)], P[np.ix_(ff, cf)], R, P


@pytest.mark.parametrize("nx,ny", [(8, 6), (16, 8), (12, 4)])
def test_plate_transfer_against_dense_loops(nx, ny):
    R, P = O.plate_transfer(nx, ny, 3.0, 2.0)
    Rd, Pd, Rall, Pall = dense_transfer(nx, ny, 3.0, 2.0)
    Rs = sp.csr_matrix((R.val, R.col, R.rowptr), shape=(R.nrows, R.ncols)).toarray()
    Ps = sp.csr_matrix((P.val, P.col, P.rowptr), shape=(P.nrows, P.ncols)).toarray()
    assert np.abs(Rs - Rd).max() <= 2 * np.finfo(float).eps and np.abs(Ps - Pd).max() <= 2 * np.finfo(float).eps
    assert np.array_equal(Rs != 0, Rd != 0) and np.array_equal(Ps != 0, Pd != 0)
    # before the submatrix every row of R and of P sums to one, per dof
    assert np.allclose(Rall.sum(1), 1) and np.allclose(Pall.sum(1), 1)
    # distance 2: the coarse node's own fine node weighs 1, its axis neighbours 1/2, diagonals 1 - sqrt(1/2)
    assert np.isclose(Pall[3 * (2 + (nx + 1) * 2), 3 * (1 + (nx // 2 + 1) * 1)], 1.0)


def test_row_groups_and_group_sweep():
    A, _ = O.plate(8, 6, 3.0, 2.0, 0.1, 52416.0, 0.3, -10.0)
    g = O.row_groups(A)
    fixed = support_nodes(8, 6)
    assert np.array_equal(np.diff(g), np.where(fixed, 2, 3))
    nc, col = O.group_coloring(A, g)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols))
    grp = np.repeat(np.arange(len(g) - 1), np.diff(g))
    for i, j in zip(*S.nonzero()):                               # coupled groups differ in colour
        assert grp[i] == grp[j] or col[grp[i]] != col[grp[j]]
    # cantilever: the groups are the node triples and the sweeps agree bit for bit
    C = O.cantilever(4, 4, 4, dens=10 ** (2 * O.rng_uniform(3, 64) - 1))
    gc = O.row_groups(C)
    assert np.array_equal(gc, np.arange(0, C.nrows + 1, 3))
    ncb, colb = O.block_coloring(C)
    ncg, colg = O.group_coloring(C, gc)
    assert ncb == ncg and np.array_equal(colb, colg)
    rng = np.random.default_rng(4)
    b, x0 = rng.normal(size=C.nrows), rng.normal(size=C.nrows)
    for bw in (False, True):
        assert np.array_equal(O.gs_groups(C, gc, b, x0, ncg, colg, backward=bw),
                              O.gs_multicolor(C, ncb, colb, b, x0, backward=bw))


@pytest.mark.parametrize("smoother", [0, 1])
def test_plate_gmg_converges_to_the_direct_solution(smoother):
    """The oracle's GPU-semantics V-cycle on the plate hierarchy (groups for Gauss-Seidel) reaches
    the sparse direct solution."""
    sizes = [(16, 8), (8, 4), (4, 2)]
    lv = []
    for l, m in enumerate(sizes):
        A, rhs = O.plate(*m, 3.0, 2.0, 0.1, 52416.0, 0.3, -10.0)
        if l == 0:
            b = rhs
        if l < len(sizes) - 1:
            R, P = O.plate_transfer(*m, 3.0, 2.0)
            lv.append(dict(A=A, R=R, P=P, gptr=O.row_groups(A)))
        else:
            lv.append(dict(A=A))
    A0 = lv[0]["A"]
    S0 = sp.csr_matrix((A0.val, A0.col, A0.rowptr), shape=(A0.nrows, A0.ncols))
    pre = None
    if smoother == 0:
        # per-level weights 2 / (1.2 lambda(D^-1 A)) (the power method of GeometricMultigridSolver.cs:
        # 80-121): the coarsest-level weight the cantilever tests use is no bound on the plate's
        # finer levels, whose stiffness scales differently with the mesh size
        pre = []
        for l in lv[:-1]:
            d = O.jacobi_preconditioner(l["A"])
            pre.append(d * (2 / (1.2 * O.power_bound(l["A"], d))))
    tol = 1e-9 * np.abs(b).max()
    x, st = O.gmg_solve(lv, b, pre=pre, smoother=smoother, level_iterations=4, max_cycles=20000, tol=tol)
    assert st.converged == 1
    xs = sp.linalg.spsolve(S0.tocsc(), b)
    assert np.abs(x - xs).max() <= 1e-6 * np.abs(xs).max()
