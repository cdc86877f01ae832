"""Pins the oracle's restatement of the reference's plane-stress cantilever (FemCantilever2D.cs with
FemCantileverBase.cs, IGeometricMultigridModel.cs on two axes with two dofs per node) against what
the mathematics fixes:

* the element matrix against the textbook bilinear plane-stress quadrilateral (unit square,
  thickness 1, 2 x 2 Gauss quadrature, exact for it) with the corners (0,0), (1,0), (1,1), (0,1) and
  dofs (ux, uy); the other corner orders are shown to miss it; its three rigid-body modes;
* K_ff against a dense element-by-element scatter with random densities (bit for bit: the same sums
  in element order), the stencil's nnz, symmetry and positive definiteness;
* the manufactured displacements against the Euler-Bernoulli cantilever's textbook tip values
  (w(L) = -q L^4 / (8 E I), slope -q L^3 / (6 E I)) and the neutral axis;
* b = K_ff U_f in the DOK insertion order against a brute-force dictionary replay of the element
  loop (bit for bit) and scipy's product;
* the transfers against a dense restatement of the window loops with two dofs per node and the
  clamped x = 0 nodes removed, with their row-sum properties;
* the Gauss-Seidel row groups (the node pairs).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def q4_textbook(E, nu, order):
    D = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    xy = np.array(order, float)
    g = np.array([-1, 1]) / math.sqrt(3)
    K = np.zeros((8, 8))
    for xi in g:
        for eta in g:
            # bilinear shapes on [0,1]^2 from the corners' own coordinates
            s, t = (xi + 1) / 2, (eta + 1) / 2
            dN = np.array([[(2 * X - 1) * (t if Y else 1 - t), (2 * Y - 1) * (s if X else 1 - s)] for X, Y in xy])
            B = np.zeros((3, 8))
            B[0, 0::2] = dN[:, 0]
            B[1, 1::2] = dN[:, 1]
            B[2, 0::2] = dN[:, 1]
            B[2, 1::2] = dN[:, 0]
            K += 0.25 * (B.T @ D @ B)
    return K


CCW = [(0, 0), (1, 0), (1, 1), (0, 1)]


@pytest.mark.parametrize("E,nu", [(1.0, 0.3), (7.5, 0.0), (210.0, 0.45)])
def test_ke_q4_matches_textbook_quadrature(E, nu):
    Ke = O.ke_q4(E, nu)
    Kt = q4_textbook(E, nu, CCW)
    assert np.abs(Ke - Kt).max() <= 1e-13 * np.abs(Kt).max()
    for order in ([(0, 0), (0, 1), (1, 1), (1, 0)], [(1, 0), (1, 1), (0, 1), (0, 0)], [(0, 0), (1, 0), (0, 1), (1, 1)]):
        assert np.abs(Ke - q4_textbook(E, nu, order)).max() > 1e-3 * np.abs(Kt).max()
    assert np.array_equal(Ke, Ke.T)


def test_ke_q4_rigid_modes():
    Ke = O.ke_q4(1.0, 0.3)
    tx = np.tile([1.0, 0.0], 4)
    ty = np.tile([0.0, 1.0], 4)
    rot = np.array([v for X, Y in CCW for v in (-Y, X)], float)
    for m in (tx, ty, rot):
        assert np.abs(Ke @ m).max() <= 1e-14
    ev = np.linalg.eigvalsh(Ke)
    assert np.sum(ev < 1e-12) == 3 and ev[3] > 1e-3


def dense_cant2d(nx, ny, dens):
    Ke = O.ke_q4(1.0, 0.3)
    N = 2 * (nx + 1) * (ny + 1)
    K = np.zeros((N, N))
    for ey in range(ny):
        for ex in range(nx):
            nodes = [ex + (nx + 1) * ey, ex + 1 + (nx + 1) * ey, ex + 1 + (nx + 1) * (ey + 1), ex + (nx + 1) * (ey + 1)]
            dofs = [2 * n + d for n in nodes for d in range(2)]
            K[np.ix_(dofs, dofs)] += dens[ex + nx * ey] * Ke
    free = [2 * n + d for n in range((nx + 1) * (ny + 1)) if n % (nx + 1) != 0 for d in range(2)]
    return K[np.ix_(free, free)]


@pytest.mark.parametrize("nx,ny", [(8, 4), (5, 3), (16, 2), (1, 1)])
def test_cant2d_assembly_bit_exact_against_dense_scatter(nx, ny):
    dens = 10 ** (2 * O.rng_uniform(31, nx * ny) - 1)
    A = O.cantilever2d(nx, ny, dens=dens)
    Kd = dense_cant2d(nx, ny, dens)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols)).toarray()
    assert A.nrows == 2 * nx * (ny + 1)
    assert np.array_equal(S, Kd)
    cx = [min(i + 1, nx) - max(i - 1, 1) + 1 for i in range(1, nx + 1)]
    cy = [min(i + 1, ny) - max(i - 1, 0) + 1 for i in range(ny + 1)]
    assert len(A.col) == 4 * sum(cx) * sum(cy)
    for i in range(A.nrows):
        assert np.all(np.diff(A.col[A.rowptr[i]:A.rowptr[i + 1]]) > 0)
    assert np.array_equal(S, S.T) and np.linalg.eigvalsh(S).min() > 0


def test_manufactured_matches_the_beam_formulas():
    nx, ny, L, H, Dz, E, q = 40, 4, 20.0, 1.0, 1.0, 1.0, 1.0
    U = O.manufactured2d(nx, ny, (L, H, Dz), E, q).reshape(ny + 1, nx, 2)
    I = Dz * H ** 3 / 12
    assert U[0, -1, 1] == pytest.approx(-q * L ** 4 / (8 * E * I), rel=1e-13)        # tip deflection
    assert np.all(U[:, :, 1] == U[0, :, 1])                                             # w independent of z
    assert np.all(U[ny // 2, :, 0] == 0)                                                # neutral axis
    # u = -rot z with rot(L) = -q L^3 / (6 E I); top fibre z = H / 2
    assert U[ny, -1, 0] == pytest.approx(q * L ** 3 / (6 * E * I) * H / 2, rel=1e-13)
    assert np.allclose(U[0, :, 0], -U[ny, :, 0], rtol=0, atol=1e-12 * np.abs(U[ny, :, 0]).max())


def test_dok_order_product_bit_exact():
    nx, ny = 6, 3
    A = O.cantilever2d(nx, ny)
    U = O.manufactured2d(nx, ny)
    y = O.multiply_dok_order2d(nx, ny, A, U)
    S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols))
    # replay of the dictionaries: a row's columns in the order the element loop first touched them
    order = [dict() for _ in range(A.nrows)]
    free = {}
    for n in range((nx + 1) * (ny + 1)):
        if n % (nx + 1):
            for d in range(2):
                free[2 * n + d] = len(free)
    for ey in range(ny):
        for ex in range(nx):
            nodes = [ex + (nx + 1) * ey, ex + 1 + (nx + 1) * ey, ex + 1 + (nx + 1) * (ey + 1), ex + (nx + 1) * (ey + 1)]
            dofs = [2 * n + d for n in nodes for d in range(2)]
            for r in dofs:
                for c in dofs:
                    if r in free and c in free:
                        order[free[r]].setdefault(free[c], None)
    yb = np.zeros(A.nrows)
    for r in range(A.nrows):
        acc = 0.0
        for c in order[r]:
            acc += S[r, c] * U[c]
        yb[r] = acc
    assert y.tobytes() == yb.tobytes()
    assert np.abs(y - S @ U).max() <= 1e-12 * np.abs(y).max()


def dense_transfer2(nx, ny):
    nf = [nx + 1, ny + 1]
    cn = [n >> 1 if n > 1 else n for n in (nx, ny)]
    nc = [c + 1 for c in cn]
    nid = [nx / cn[0], ny / cn[1]]
    R = np.zeros((2 * nc[0] * nc[1], 2 * nf[0] * nf[1]))
    P = np.zeros((2 * nf[0] * nf[1], 2 * nc[0] * nc[1]))
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
                    R[2 * cid, 2 * fid] = infl
                    P[2 * fid, 2 * cid] = infl
            R[2 * cid] /= sr
            R[2 * cid + 1, 1::2] = R[2 * cid, 0::2]
    for fid in range(nf[0] * nf[1]):
        P[2 * fid] /= sumP[fid]
        P[2 * fid + 1, 1::2] = P[2 * fid, 0::2]
    ff = [2 * n + d for n in range(nf[0] * nf[1]) if n % nf[0] for d in range(2)]
    cf = [2 * n + d for n in range(nc[0] * nc[1]) if n % nc[0] for d in range(2)]
    return R[np.ix_(cf, ff)], P[np.ix_(ff, cf)], R, P


@pytest.mark.parametrize("nx,ny", [(8, 4), (16, 2), (12, 6), (5, 3)])
def test_cant2d_transfer_against_dense_loops(nx, ny):
    R, P = O.cantilever2d_transfer(nx, ny)
    Rd, Pd, Rall, Pall = dense_transfer2(nx, ny)
    Rs = sp.csr_matrix((R.val, R.col, R.rowptr), shape=(R.nrows, R.ncols)).toarray()
    Ps = sp.csr_matrix((P.val, P.col, P.rowptr), shape=(P.nrows, P.ncols)).toarray()
    assert np.abs(Rs - Rd).max() <= 2 * np.finfo(float).eps and np.abs(Ps - Pd).max() <= 2 * np.finfo(float).eps
    assert np.array_equal(Rs != 0, Rd != 0) and np.array_equal(Ps != 0, Pd != 0)
    assert np.allclose(Rall.sum(1), 1) and np.allclose(Pall.sum(1), 1)
    # the x-component never feeds the y-component
    assert np.all(Ps[0::2, 1::2] == 0) and np.all(Rs[0::2, 1::2] == 0)


def test_cant2d_row_groups_are_node_pairs():
    A = O.cantilever2d(8, 4)
    g = O.row_groups(A)
    assert np.array_equal(g, np.arange(0, A.nrows + 1, 2))
    nc, col = O.group_coloring(A, g)
    assert nc == 4                                   # greedy colouring of the 9-point node stencil


def test_cant2d_gmg_reaches_the_direct_solution():
    """The oracle's V-cycle on the 2D cantilever hierarchy (Gauss-Seidel over the node pairs) reaches
    scipy's direct solution of K_ff U = b for the manufactured b."""
    sizes = [(32, 4), (16, 2), (8, 1)]
    lv = []
    for l, m in enumerate(sizes):
        A = O.cantilever2d(*m)
        if l < len(sizes) - 1:
            R, P = O.cantilever2d_transfer(*m)
            lv.append(dict(A=A, R=R, P=P, gptr=O.row_groups(A)))
        else:
            lv.append(dict(A=A))
    A0 = lv[0]["A"]
    b = O.multiply_dok_order2d(32, 4, A0, O.manufactured2d(32, 4))
    x, st = O.gmg_solve(lv, b, smoother=1, level_iterations=4, max_cycles=20000, tol=1e-10 * np.abs(b).max())
    assert st.converged == 1
    S0 = sp.csr_matrix((A0.val, A0.col, A0.rowptr), shape=(A0.nrows, A0.ncols))
    xs = sp.linalg.spsolve(S0.tocsc(), b)
    assert np.abs(x - xs).max() <= 1e-6 * np.abs(xs).max()
