"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module, and only as the checker or
the timed CPU baseline.  The product package (compressiontests_amd) never imports it.

Each wrapper names the C function it calls; the C sources cite the reference
file:line they restate.
"""
import ctypes as C
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

Csr = namedtuple("Csr", "rowptr col val nrows ncols")

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("residual_norm", C.c_double), ("rel_residual", C.c_double)]


class OrcLevel(C.Structure):
    _fields_ = [("n", C.c_int32),
                ("Arp", C.c_void_p), ("Aci", C.c_void_p), ("Av", C.c_void_p),
                ("Rrp", C.c_void_p), ("Rci", C.c_void_p), ("Rv", C.c_void_p),
                ("Prp", C.c_void_p), ("Pci", C.c_void_p), ("Pv", C.c_void_p),
                ("gs_order", C.c_void_p), ("gptr", C.c_void_p), ("ngroups", C.c_int32)]


class OrcDuviBlock(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("nnz", C.c_int64), ("nblocks", C.c_int64), ("nuniq", C.c_int64),
                ("vi_bytes", C.c_int32),
                ("row_blocks", C.POINTER(C.c_int32)), ("row_dist", C.POINTER(C.c_int32)),
                ("block_col", C.POINTER(C.c_int32)), ("dist", C.POINTER(C.c_uint8)),
                ("vi", C.c_void_p), ("values", C.POINTER(C.c_double))]


class OrcDuviUnits(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32),
                ("nunits", C.c_int64), ("nbytes", C.c_int64), ("nvals", C.c_int64), ("nuniq", C.c_int64),
                ("unit_off", C.POINTER(C.c_int64)), ("bytes", C.POINTER(C.c_uint8)),
                ("vi", C.POINTER(C.c_int32)), ("uniq", C.POINTER(C.c_double))]


_LIBS = {}


def lib(omp=False):
    """liboracle.so, or with omp=True its OpenMP build liboracle_omp.so (parallel row loops,
    bit-identical results; the all-core CPU baseline)."""
    name = "liboracle_omp.so" if omp else "liboracle.so"
    L = _LIBS.get(name)
    if L is None:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle/{name} is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_cantilever_nnz.restype = C.c_int64
        L.orc_galerkin_nnz.restype = C.c_int64
        L.orc_dedup.restype = C.c_int64
        L.orc_eigen_upper_bound.restype = C.c_double
        L.orc_dot.restype = C.c_double
        L.orc_power_bound.restype = C.c_double
        L.orc_splitmix64_next.restype = C.c_uint64
        L.orc_plate_nnz.restype = C.c_int64
        L.orc_plate_nfree.restype = C.c_int32
        L.orc_cant2d_nnz.restype = C.c_int64
        L.orc_cant2d_nfree.restype = C.c_int32
        L.orc_row_groups.restype = C.c_int32
        _LIBS[name] = L
    return L


def num_threads(omp=False):
    """Threads the chosen build runs its parallel loops on (orc_num_threads)."""
    return lib(omp).orc_num_threads()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ RNG
def rng_uniform(seed, n):
    out = np.empty(n, np.float64)
    lib().orc_rng_uniform(C.c_uint64(seed), C.c_int64(n), _p(out))
    return out


def rng_normal(seed, n):
    out = np.empty(n, np.float64)
    lib().orc_rng_normal(C.c_uint64(seed), C.c_int64(n), _p(out))
    return out


def splitmix64(seed, count):
    s = C.c_uint64(seed)
    return [lib().orc_splitmix64_next(C.byref(s)) for _ in range(count)]


# ------------------------------------------------------------------ FE
def ke_hex8(E=1.0, nu=0.3):
    Ke = np.empty((24, 24), np.float64)
    lib().orc_ke_hex8(C.c_double(E), C.c_double(nu), _p(Ke))
    return Ke


def n_free(nx, ny, nz):
    return 3 * nx * (ny + 1) * (nz + 1)


def cantilever(nx, ny, nz, E=1.0, nu=0.3, dens=None):
    nnz = lib().orc_cantilever_nnz(nx, ny, nz)
    n = n_free(nx, ny, nz)
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    d = None if dens is None else np.ascontiguousarray(dens, np.float64)
    lib().orc_cantilever_assemble(nx, ny, nz, C.c_double(E), C.c_double(nu),
                                  _p(d) if d is not None else None, _p(rp), _p(ci), _p(v))
    return Csr(rp, ci, v, n, n)


def assemble_elements(nx, ny, nz, Kel):
    nnz = lib().orc_cantilever_nnz(nx, ny, nz)
    n = n_free(nx, ny, nz)
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    Kel = np.ascontiguousarray(Kel, np.float64)
    lib().orc_assemble_elements(nx, ny, nz, _p(Kel), _p(rp), _p(ci), _p(v))
    return Csr(rp, ci, v, n, n)


def manufactured(nx, ny, nz, L=(1.0, 1.0, 1.0), E=1.0, q=1.0):
    out = np.empty(n_free(nx, ny, nz), np.float64)
    lib().orc_cantilever_manufactured(nx, ny, nz, C.c_double(L[0]), C.c_double(L[1]), C.c_double(L[2]),
                                      C.c_double(E), C.c_double(q), _p(out))
    return out


def tip_load(nx, ny, nz, total=-1.0):
    out = np.empty(n_free(nx, ny, nz), np.float64)
    lib().orc_cantilever_tip_load(nx, ny, nz, C.c_double(total), _p(out))
    return out


def multiply_dok_order(nx, ny, nz, A, x):
    y = np.empty(A.nrows, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_csr_multiply_dok_order(nx, ny, nz, _p(A.rowptr), _p(A.col), _p(A.val), _p(x), _p(y))
    return y


def rp_reference(nx, ny, nz):
    nr, npp = C.c_int64(), C.c_int64()
    lib().orc_rp_reference_nnz(nx, ny, nz, C.byref(nr), C.byref(npp))
    nf = n_free(nx, ny, nz)
    cx, cy, cz = [(k >> 1) if k > 1 else k for k in (nx, ny, nz)]
    nc = n_free(cx, cy, cz)
    Rrp = np.empty(nc + 1, np.int32); Rci = np.empty(nr.value, np.int32); Rv = np.empty(nr.value, np.float64)
    Prp = np.empty(nf + 1, np.int32); Pci = np.empty(npp.value, np.int32); Pv = np.empty(npp.value, np.float64)
    lib().orc_rp_reference(nx, ny, nz, _p(Rrp), _p(Rci), _p(Rv), _p(Prp), _p(Pci), _p(Pv))
    return Csr(Rrp, Rci, Rv, nc, nf), Csr(Prp, Pci, Pv, nf, nc)


# ------------------------------------------------------------------ Kirchhoff plate
def plate_ke(a, b, E=1.0, nu=0.3, t=1.0):
    """Element stiffness of the 4-node plate rectangle of half-sides a, b (w, rx, ry per node)."""
    Ke = np.empty((12, 12), np.float64)
    lib().orc_plate_ke(C.c_double(a), C.c_double(b), C.c_double(E), C.c_double(nu), C.c_double(t), _p(Ke))
    return Ke


def plate_load(a, b, q):
    fe = np.empty(12, np.float64)
    lib().orc_plate_load(C.c_double(a), C.c_double(b), C.c_double(q), _p(fe))
    return fe


def plate_free_dofs(nx, ny, Lx, Ly):
    n = lib().orc_plate_nfree(nx, ny, C.c_double(Lx), C.c_double(Ly))
    d = np.empty(n, np.int32)
    lib().orc_plate_free_dofs(nx, ny, C.c_double(Lx), C.c_double(Ly), _p(d))
    return d


def plate(nx, ny, Lx=3.0, Ly=2.0, t=0.1, E=52416.0, nu=0.3, q=-10.0):
    """(K_ff, b) of FemPlate(nx x ny elements, Lx x Ly x t) with a uniform load q."""
    n = lib().orc_plate_nfree(nx, ny, C.c_double(Lx), C.c_double(Ly))
    nnz = lib().orc_plate_nnz(nx, ny, C.c_double(Lx), C.c_double(Ly))
    rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    b = np.empty(n, np.float64)
    lib().orc_plate_assemble(nx, ny, C.c_double(Lx), C.c_double(Ly), C.c_double(t), C.c_double(E), C.c_double(nu),
                             C.c_double(q), _p(rp), _p(ci), _p(v), _p(b))
    return Csr(rp, ci, v, n, n), b


def plate_transfer(nx, ny, Lx=3.0, Ly=2.0):
    """(R, P) between the nx x ny plate and its coarser model."""
    nr, npp = C.c_int64(), C.c_int64()
    lib().orc_plate_transfer_nnz(nx, ny, C.c_double(Lx), C.c_double(Ly), C.byref(nr), C.byref(npp))
    cx, cy = [(k >> 1) if k > 1 else k for k in (nx, ny)]
    nf = lib().orc_plate_nfree(nx, ny, C.c_double(Lx), C.c_double(Ly))
    nc = lib().orc_plate_nfree(cx, cy, C.c_double(Lx), C.c_double(Ly))
    Rrp = np.empty(nc + 1, np.int32); Rci = np.empty(nr.value, np.int32); Rv = np.empty(nr.value, np.float64)
    Prp = np.empty(nf + 1, np.int32); Pci = np.empty(npp.value, np.int32); Pv = np.empty(npp.value, np.float64)
    lib().orc_plate_transfer(nx, ny, C.c_double(Lx), C.c_double(Ly), _p(Rrp), _p(Rci), _p(Rv), _p(Prp), _p(Pci),
                             _p(Pv))
    return Csr(Rrp, Rci, Rv, nc, nf), Csr(Prp, Pci, Pv, nf, nc)


# ------------------------------------------------------------------ plane-stress cantilever
def ke_q4(E=1.0, nu=0.3):
    """Element stiffness of FemCantilever2D (8 x 8, dofs (ux, uy) of the corners (0,0), (1,0),
    (1,1), (0,1))."""
    Ke = np.empty((8, 8), np.float64)
    lib().orc_ke_q4(C.c_double(E), C.c_double(nu), _p(Ke))
    return Ke


def cantilever2d(nx, ny, E=1.0, nu=0.3, dens=None):
    """K_ff of FemCantilever2D with nx x ny elements (element densities dens, e = ex + nx ey)."""
    n = lib().orc_cant2d_nfree(nx, ny)
    nnz = lib().orc_cant2d_nnz(nx, ny)
    rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    d = None if dens is None else np.ascontiguousarray(dens, np.float64)
    lib().orc_cant2d_assemble(nx, ny, C.c_double(E), C.c_double(nu), None if d is None else _p(d), _p(rp), _p(ci),
                              _p(v))
    return Csr(rp, ci, v, n, n)


def manufactured2d(nx, ny, L=(20.0, 1.0, 1.0), E=1.0, q=1.0):
    """U_f of FemCantilever2D (lengths (Lx, Ly, depth))."""
    U = np.empty(lib().orc_cant2d_nfree(nx, ny), np.float64)
    lib().orc_cant2d_manufactured(nx, ny, C.c_double(L[0]), C.c_double(L[1]), C.c_double(L[2]), C.c_double(E),
                                  C.c_double(q), _p(U))
    return U


def multiply_dok_order2d(nx, ny, A, x):
    y = np.empty(A.nrows, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_cant2d_multiply_dok_order(nx, ny, _p(A.rowptr), _p(A.col), _p(A.val), _p(x), _p(y))
    return y


def cantilever2d_transfer(nx, ny):
    """(R, P) between the nx x ny cantilever and its coarser model."""
    nr, npp = C.c_int64(), C.c_int64()
    lib().orc_cant2d_transfer_nnz(nx, ny, C.byref(nr), C.byref(npp))
    cx, cy = [(k >> 1) if k > 1 else k for k in (nx, ny)]
    nf, nc = lib().orc_cant2d_nfree(nx, ny), lib().orc_cant2d_nfree(cx, cy)
    Rrp = np.empty(nc + 1, np.int32); Rci = np.empty(nr.value, np.int32); Rv = np.empty(nr.value, np.float64)
    Prp = np.empty(nf + 1, np.int32); Pci = np.empty(npp.value, np.int32); Pv = np.empty(npp.value, np.float64)
    lib().orc_cant2d_transfer(nx, ny, _p(Rrp), _p(Rci), _p(Rv), _p(Prp), _p(Pci), _p(Pv))
    return Csr(Rrp, Rci, Rv, nc, nf), Csr(Prp, Pci, Pv, nf, nc)


def row_groups(A):
    """gptr of the Gauss-Seidel node groups (runs of <= 3 rows with identical column lists)."""
    g = np.empty(A.nrows + 1, np.int32)
    ng = lib().orc_row_groups(A.nrows, _p(A.rowptr), _p(A.col), _p(g))
    return g[:ng + 1].copy()


def group_coloring(A, gptr):
    col = np.empty(len(gptr) - 1, np.int32)
    nc = lib().orc_group_coloring(len(gptr) - 1, _p(gptr), _p(A.rowptr), _p(A.col), _p(col))
    return nc, col


def gs_groups(A, gptr, b, x, ncol=0, color=None, order=None, backward=False):
    x = np.array(x, np.float64)
    o = None if order is None else np.ascontiguousarray(order, np.int32)
    lib().orc_gs_groups(len(gptr) - 1, _p(gptr), _p(A.rowptr), _p(A.col), _p(A.val), None if o is None else _p(o),
                        ncol, None if color is None else _p(color), _p(np.ascontiguousarray(b, np.float64)), _p(x),
                        int(backward))
    return x


def p_trilinear(nx, ny, nz):
    npp = C.c_int64()
    lib().orc_p_trilinear_nnz(nx, ny, nz, C.byref(npp))
    nf = n_free(nx, ny, nz)
    nc = n_free(nx // 2, ny // 2, nz // 2)
    Prp = np.empty(nf + 1, np.int32); Pci = np.empty(npp.value, np.int32); Pv = np.empty(npp.value, np.float64)
    lib().orc_p_trilinear(nx, ny, nz, _p(Prp), _p(Pci), _p(Pv))
    return Csr(Prp, Pci, Pv, nf, nc)


def transpose(A):
    nnz = A.rowptr[A.nrows]
    rp = np.empty(A.ncols + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    lib().orc_csr_transpose(A.nrows, A.ncols, _p(A.rowptr), _p(A.col), _p(A.val), _p(rp), _p(ci), _p(v))
    return Csr(rp, ci, v, A.ncols, A.nrows)


def galerkin(A, P):
    nf, nc = P.nrows, P.ncols
    nnz = lib().orc_galerkin_nnz(nf, nc, _p(A.rowptr), _p(A.col), _p(P.rowptr), _p(P.col))
    rp = np.empty(nc + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    lib().orc_galerkin(nf, nc, _p(A.rowptr), _p(A.col), _p(A.val), _p(P.rowptr), _p(P.col), _p(P.val),
                       _p(rp), _p(ci), _p(v))
    return Csr(rp, ci, v, nc, nc)


# ------------------------------------------------------------------ kernels
def spmv(A, x, y=None, clear=True):
    x = np.ascontiguousarray(x, np.float64)
    if y is None:
        y = np.zeros(A.nrows, np.float64)
    lib().orc_csr_spmv(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), _p(x), _p(y), int(clear))
    return y


def residual(A, b, x):
    r = np.empty(A.nrows, np.float64)
    lib().orc_csr_residual(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val),
                           _p(np.ascontiguousarray(b, np.float64)), _p(np.ascontiguousarray(x, np.float64)), _p(r))
    return r


def residual_check(A, b, x, tol):
    r = np.empty(A.nrows, np.float64)
    flag = lib().orc_csr_residual_check(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val),
                                        _p(np.ascontiguousarray(b, np.float64)),
                                        _p(np.ascontiguousarray(x, np.float64)), _p(r), C.c_double(tol))
    return r, flag


def jacobi_initial(pre, b):
    x = np.empty(len(b), np.float64)
    lib().orc_jacobi_initial(len(b), _p(np.ascontiguousarray(pre)), _p(np.ascontiguousarray(b)), _p(x))
    return x


def jacobi(A, pre, b, x):
    y = np.empty(A.nrows, np.float64)
    lib().orc_csr_jacobi(A.nrows, _p(np.ascontiguousarray(pre)), _p(A.rowptr), _p(A.col), _p(A.val),
                         _p(np.ascontiguousarray(b)), _p(np.ascontiguousarray(x)), _p(y))
    return y


def jacobi_preconditioner(A):
    d = np.empty(A.nrows, np.float64)
    lib().orc_jacobi_preconditioner(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), _p(d))
    return d


def eigen_upper_bound(A):
    return lib().orc_eigen_upper_bound(A.nrows, _p(A.rowptr), _p(A.val))


def power_bound(A, invd):
    invd = np.ascontiguousarray(invd, np.float64)
    return lib().orc_power_bound(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), _p(invd))


def gs_forward(A, b, x):
    x = np.array(x, np.float64)
    lib().orc_gs_forward(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), _p(np.ascontiguousarray(b)), _p(x))
    return x


def block_coloring(A, bs=3):
    col = np.empty(A.nrows // bs, np.int32)
    nc = lib().orc_block_coloring(A.nrows // bs, bs, _p(A.rowptr), _p(A.col), _p(col))
    return nc, col


def gs_ordered(A, order, b, x, backward=False, bs=3):
    """Gauss-Seidel sweep over node blocks in the given node order (backward: exact reverse)."""
    x = np.array(x, np.float64)
    o = np.ascontiguousarray(order, np.int32)
    lib().orc_gs_ordered(A.nrows, bs, _p(A.rowptr), _p(A.col), _p(A.val), _p(o), _p(np.ascontiguousarray(b)), _p(x),
                         int(backward))
    return x


def gs_multicolor(A, ncol, color, b, x, backward=False, bs=3):
    x = np.array(x, np.float64)
    lib().orc_gs_multicolor(A.nrows, bs, _p(A.rowptr), _p(A.col), _p(A.val), ncol, _p(color),
                            _p(np.ascontiguousarray(b)), _p(x), int(backward))
    return x


def cholesky_solve_csr(A, b):
    n = A.nrows
    D = np.empty((n, n), np.float64)
    lib().orc_csr_to_dense(n, _p(A.rowptr), _p(A.col), _p(A.val), _p(D))
    if lib().orc_dense_cholesky(n, _p(D)) != 0:
        raise ValueError("matrix not SPD")
    x = np.empty(n, np.float64)
    lib().orc_dense_cholesky_solve(n, _p(D), _p(np.ascontiguousarray(b, np.float64)), _p(x))
    return x


# ------------------------------------------------------------------ DuVi
def dedup(v, tol=1e-10):
    v = np.ascontiguousarray(v, np.float64)
    idx = np.empty(len(v), np.int32)
    uniq = np.empty(max(len(v), 1), np.float64)
    nu = lib().orc_dedup(len(v), _p(v), C.c_double(tol), _p(idx), _p(uniq))
    return idx, uniq[:nu].copy()


def duvi_block_encode(A, tol=1e-10, mode=0):
    m = OrcDuviBlock()
    lib().orc_duvi_block_encode(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), C.c_double(tol), mode, C.byref(m))
    nnz, nb, nu = m.nnz, m.nblocks, m.nuniq
    out = dict(
        nrows=m.nrows, nnz=nnz, nblocks=nb, nuniq=nu, vi_bytes=m.vi_bytes,
        row_blocks=np.ctypeslib.as_array(m.row_blocks, (m.nrows + 1,)).copy(),
        row_dist=np.ctypeslib.as_array(m.row_dist, (m.nrows + 1,)).copy(),
        block_col=np.ctypeslib.as_array(m.block_col, (max(nb, 1),))[:nb].copy(),
        dist=np.ctypeslib.as_array(m.dist, (max(nnz, 1),))[:nnz].copy(),
        values=np.ctypeslib.as_array(m.values, (max(nu, 1),))[:nu].copy(),
    )
    if m.vi_bytes:
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[m.vi_bytes]
        buf = (C.c_uint8 * (m.vi_bytes * max(nnz, 1))).from_address(m.vi)
        out["vi"] = np.frombuffer(buf, dtype=dt, count=nnz).copy()
    else:
        out["vi"] = None
    lib().orc_duvi_block_free(C.byref(m))
    return out


def _block_struct(d):
    """Rebuild an OrcDuviBlock that points at numpy arrays held in d (kept alive by d)."""
    m = OrcDuviBlock()
    m.nrows = d["nrows"]; m.nnz = d["nnz"]; m.nblocks = d["nblocks"]; m.nuniq = d["nuniq"]
    m.vi_bytes = d["vi_bytes"]
    keep = []
    for name, dt, ctype in (("row_blocks", np.int32, C.c_int32), ("row_dist", np.int32, C.c_int32),
                            ("block_col", np.int32, C.c_int32), ("dist", np.uint8, C.c_uint8),
                            ("values", np.float64, C.c_double)):
        a = np.ascontiguousarray(d[name], dt)
        keep.append(a)
        setattr(m, name, a.ctypes.data_as(C.POINTER(ctype)))
    if d["vi"] is not None:
        a = np.ascontiguousarray(d["vi"])
        keep.append(a)
        m.vi = a.ctypes.data
    else:
        m.vi = None
    return m, keep


def duvi_block_spmv(d, x, y=None, clear=True):
    m, keep = _block_struct(d)
    x = np.ascontiguousarray(x, np.float64)
    if y is None:
        y = np.zeros(d["nrows"], np.float64)
    lib().orc_duvi_block_spmv(C.byref(m), _p(x), _p(y), int(clear))
    return y


def duvi_block_to_csr(d, ncols):
    """The CSR matrix a block-layout encoding stores (values after the tolerance merge)."""
    m, keep = _block_struct(d)
    rp = np.empty(d["nrows"] + 1, np.int32)
    ci = np.empty(d["nnz"], np.int32)
    v = np.empty(d["nnz"], np.float64)
    lib().orc_duvi_block_to_csr(C.byref(m), _p(rp), _p(ci), _p(v))
    return Csr(rp, ci, v, d["nrows"], ncols)


def _units_to_dict(m):
    out = dict(
        nrows=m.nrows, ncols=m.ncols, nunits=m.nunits,
        unit_off=np.ctypeslib.as_array(m.unit_off, (m.nunits + 1,)).copy(),
        bytes=np.ctypeslib.as_array(m.bytes, (max(m.nbytes, 1),))[:m.nbytes].copy(),
        vi=np.ctypeslib.as_array(m.vi, (max(m.nvals, 1),))[:m.nvals].copy(),
        uniq=np.ctypeslib.as_array(m.uniq, (max(m.nuniq, 1),))[:m.nuniq].copy(),
    )
    lib().orc_duvi_units_free(C.byref(m))
    return out


def duvi_units_from_csr(A, tol=1e-10):
    m = OrcDuviUnits()
    rc = lib().orc_duvi_units_from_csr(A.nrows, A.ncols, _p(A.rowptr), _p(A.col), _p(A.val), C.c_double(tol), C.byref(m))
    if rc != 0:
        raise NotImplementedError("column index beyond the varint range")
    return _units_to_dict(m)


def duvi_units_from_dense(D, tol=1e-10):
    D = np.ascontiguousarray(D, np.float64)
    m = OrcDuviUnits()
    rc = lib().orc_duvi_units_from_dense(D.shape[0], D.shape[1], _p(D), C.c_double(tol), C.byref(m))
    if rc != 0:
        raise NotImplementedError("column index beyond the varint range")
    return _units_to_dict(m)


def _units_struct(d):
    m = OrcDuviUnits()
    m.nrows = d["nrows"]; m.ncols = d["ncols"]; m.nunits = d["nunits"]
    m.nbytes = len(d["bytes"]); m.nvals = len(d["vi"]); m.nuniq = len(d["uniq"])
    keep = [np.ascontiguousarray(d["unit_off"], np.int64), np.ascontiguousarray(d["bytes"], np.uint8),
            np.ascontiguousarray(d["vi"], np.int32), np.ascontiguousarray(d["uniq"], np.float64)]
    m.unit_off = keep[0].ctypes.data_as(C.POINTER(C.c_int64))
    m.bytes = keep[1].ctypes.data_as(C.POINTER(C.c_uint8))
    m.vi = keep[2].ctypes.data_as(C.POINTER(C.c_int32))
    m.uniq = keep[3].ctypes.data_as(C.POINTER(C.c_double))
    return m, keep


def duvi_units_spmv(d, x, y=None):
    m, keep = _units_struct(d)
    if y is None:
        y = np.zeros(d["nrows"], np.float64)
    lib().orc_duvi_units_spmv(C.byref(m), _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def duvi_units_gs(d, b, x):
    m, keep = _units_struct(d)
    x = np.array(x, np.float64)
    lib().orc_duvi_units_gs(C.byref(m), _p(np.ascontiguousarray(b, np.float64)), _p(x))
    return x


# ------------------------------------------------------------------ solvers
def _levels(levels):
    """levels: list of dicts with A (and R, P except on the coarsest)."""
    arr = (OrcLevel * len(levels))()
    keep = []
    for i, lv in enumerate(levels):
        A = lv["A"]
        arr[i].n = A.nrows
        arr[i].Arp, arr[i].Aci, arr[i].Av = _p(A.rowptr), _p(A.col), _p(A.val)
        keep.append(A)
        if lv.get("R") is not None:
            R, P = lv["R"], lv["P"]
            arr[i].Rrp, arr[i].Rci, arr[i].Rv = _p(R.rowptr), _p(R.col), _p(R.val)
            arr[i].Prp, arr[i].Pci, arr[i].Pv = _p(P.rowptr), _p(P.col), _p(P.val)
            keep += [R, P]
        if lv.get("gptr") is not None:
            g = np.ascontiguousarray(lv["gptr"], np.int32)
            assert g[0] == 0 and g[-1] == A.nrows
            arr[i].gptr, arr[i].ngroups = _p(g), g.size - 1
            keep.append(g)
        if lv.get("gs_order") is not None:
            o = np.ascontiguousarray(lv["gs_order"], np.int32)
            assert o.size == (A.nrows // 3 if lv.get("gptr") is None else len(lv["gptr"]) - 1)
            arr[i].gs_order = _p(o)
            keep.append(o)
    return arr, keep


def gmg_solve(levels, b, pre=None, smoother=0, level_iterations=4, max_cycles=10000, tol=1e-6, x0=None, omp=False):
    arr, keep = _levels(levels)
    L = len(levels)
    pre_arr = (C.c_void_p * L)()
    pre_keep = []
    if pre is not None:
        for i, p in enumerate(pre):
            p = np.ascontiguousarray(p, np.float64)
            pre_keep.append(p)
            pre_arr[i] = p.ctypes.data
    x = np.zeros(levels[0]["A"].nrows) if x0 is None else np.array(x0, np.float64)
    st = OrcStats()
    rc = lib(omp).orc_gmg_solve(L, arr, pre_arr, smoother, L - 1, level_iterations, max_cycles, C.c_double(tol),
                             _p(np.ascontiguousarray(b, np.float64)), _p(x), int(x0 is not None), C.byref(st))
    if rc != 0:
        raise RuntimeError("orc_gmg_solve failed: %d" % rc)
    return x, st


def pcg_jacobi(A, b, max_it=10000, tol=1e-5, x0=None, omp=False):
    x = np.zeros(A.nrows) if x0 is None else np.array(x0, np.float64)
    st = OrcStats()
    lib(omp).orc_pcg_jacobi(A.nrows, _p(A.rowptr), _p(A.col), _p(A.val), _p(np.ascontiguousarray(b, np.float64)),
                         _p(x), int(x0 is not None), max_it, C.c_double(tol), C.byref(st))
    return x, st


def mgpcg(levels, b, smoother=0, omega=0.6, nu=2, max_it=1000, tol=1e-8, stop=1, omp=False):
    arr, keep = _levels(levels)
    x = np.zeros(levels[0]["A"].nrows)
    st = OrcStats()
    rc = lib(omp).orc_mgpcg(len(levels), arr, smoother, C.c_double(omega), nu, max_it, C.c_double(tol), stop,
                            _p(np.ascontiguousarray(b, np.float64)), _p(x), C.byref(st))
    if rc != 0:
        raise RuntimeError("orc_mgpcg failed: %d" % rc)
    return x, st


def mgpcg_timed(levels, b, iterations, smoother=0, omega=0.6, nu=2, omp=False):
    """orc_mgpcg_timed with the stop test disabled: runs exactly `iterations` PCG iterations and
    returns (x, stats, seconds of each iteration)."""
    arr, keep = _levels(levels)
    x = np.zeros(levels[0]["A"].nrows)
    st = OrcStats()
    t = np.zeros(iterations)
    rc = lib(omp).orc_mgpcg_timed(len(levels), arr, smoother, C.c_double(omega), nu, iterations, C.c_double(0.0), 1,
                                  _p(np.ascontiguousarray(b, np.float64)), _p(x), C.byref(st), _p(t))
    if rc != 0:
        raise RuntimeError("orc_mgpcg_timed failed: %d" % rc)
    return x, st, t


def vcycle_apply(levels, r, smoother=0, omega=0.6, nu=2):
    arr, keep = _levels(levels)
    z = np.zeros(levels[0]["A"].nrows)
    rc = lib().orc_vcycle_apply(len(levels), arr, smoother, C.c_double(omega), nu,
                                _p(np.ascontiguousarray(r, np.float64)), _p(z))
    if rc != 0:
        raise RuntimeError("orc_vcycle_apply failed: %d" % rc)
    return z


def simp_gradient(nx, ny, nz, rho, u, penal=3.0, Emin=1e-9, E0=1.0, nu=0.3):
    Ke0 = ke_hex8(1.0, nu)
    g = np.empty(nx * ny * nz, np.float64)
    lib().orc_simp_gradient(nx, ny, nz, _p(Ke0), _p(np.ascontiguousarray(rho, np.float64)), C.c_double(penal),
                            C.c_double(Emin), C.c_double(E0), _p(np.ascontiguousarray(u, np.float64)), _p(g))
    return g
