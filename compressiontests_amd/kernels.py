"""Device-level kernels of libmgx, one per OpenCL kernel of the reference
(src/MGroup.OCL/CsrGeometricMultigrid.cl, CsrCG.cl, DuViGeometricMultigrid.cl), same
argument meaning.  Arguments are DeviceCsr / DeviceArray objects."""
import ctypes as C

import numpy as np

from . import _lib as L
from .device import DeviceArray


def matrix_vector_product(A, x, y, clear=True, stream=None):
    """y = A x (clear) or y += A x.  CsrGeometricMultigrid.cl:49-62."""
    name = "mgx_csr_spmv_f32" if x.dtype == np.float32 else "mgx_csr_spmv"
    L.call(name, A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), x.vp, y.vp, int(clear), stream)


def spmm(A, X, Y, k, clear=True, stream=None):
    """Y = A X (clear) or Y += A X for k right-hand sides, X and Y row-major (rows x k) blocks."""
    name = "mgx_csr_spmm_f32" if X.dtype == np.float32 else "mgx_csr_spmm"
    L.call(name, A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), k, X.vp, Y.vp, int(clear),
           stream)


class NodeBlockMatrix:
    """Node-block (3x3 block CSR) copy of a device CSR whose rows come in node triples sharing
    whole-node columns: the layout of the batching form's SpMM (FP32 blocks if fp32)."""

    def __init__(self, A, fp32=False):
        h = C.c_void_p()
        L.call("mgx_nodeblock_from_csr", A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val),
               int(fp32), C.byref(h))
        self.handle, self.fp32, self.nrows = h.value, fp32, A.nrows

    def spmm(self, X, Y, k, clear=True, stream=None):
        """Y = A X (clear) or Y += A X; X, Y row-major (rows x k), float32 for an FP32 copy."""
        assert X.dtype == (np.float32 if self.fp32 else np.float64) and Y.dtype == X.dtype
        assert X.n == self.nrows * k and Y.n == self.nrows * k
        L.call("mgx_nodeblock_spmm", C.c_void_p(self.handle), k, X.vp, Y.vp, int(clear), stream)

    def free(self):
        if self.handle is not None:
            L.lib().mgx_nodeblock_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class NodeTransfer:
    """Node-transfer copy of a device CSR prolongation or restriction that acts on the three dofs
    of a node alike (one weight per output/input node pair): the layout of the batching form's
    R and P (FP32 weights if fp32)."""

    def __init__(self, M, fp32=False):
        h = C.c_void_p()
        L.call("mgx_nodetransfer_from_csr", M.nrows, M.ncols, C.c_void_p(M.rowptr), C.c_void_p(M.col),
               C.c_void_p(M.val), int(fp32), C.byref(h))
        self.handle, self.fp32, self.nrows, self.ncols = h.value, fp32, M.nrows, M.ncols

    def apply(self, X, Y, k, add=False, stream=None):
        """Y = M X or (add) Y += M X; X (ncols x k), Y (nrows x k) row-major, float32 for an FP32 copy."""
        assert X.dtype == (np.float32 if self.fp32 else np.float64) and Y.dtype == X.dtype
        assert X.n == self.ncols * k and Y.n == self.nrows * k
        L.call("mgx_nodetransfer_apply", C.c_void_p(self.handle), k, X.vp, Y.vp, int(add), stream)

    def free(self):
        if self.handle is not None:
            L.lib().mgx_nodetransfer_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dense_block_apply(M, X, Y, k, stream=None):
    """Y = M X for the dense FP64 coarse inverse M (nc x nc DeviceArray, row-major) and nc x k
    row-major blocks X, Y (float32 or float64): the batching form's coarse solve."""
    nc = int(round(np.sqrt(M.n)))
    assert M.dtype == np.float64 and nc * nc == M.n and X.dtype == Y.dtype and X.n == nc * k and Y.n == nc * k
    L.call("mgx_dense_block_apply", nc, M.vp, k, X.vp, Y.vp, int(X.dtype == np.float32), stream)


def residual(A, b, x, r, stream=None):
    """r = b - A x.  CsrGeometricMultigrid.cl:96-110."""
    L.call("mgx_csr_residual", A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), b.vp, x.vp, r.vp,
           stream)


def residual_with_check(A, b, x, r, tolerance, stream=None):
    """r = b - A x and the reference flag (1 converged, bit 1 failed).  CsrGeometricMultigrid.cl:121-141."""
    flag = DeviceArray.from_numpy(np.array([1], np.int32))
    L.call("mgx_csr_residual_check", A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), b.vp, x.vp,
           r.vp, C.c_double(tolerance), flag.vp, stream)
    return int(flag.numpy()[0])


def jacobi_initial_iteration(pre, b, x, stream=None):
    """x = pre .* b.  CsrGeometricMultigrid.cl:153-164."""
    L.call("mgx_jacobi_initial", pre.n, pre.vp, b.vp, x.vp, stream)


def jacobi_iteration(pre, A, b, x, y, stream=None):
    """y = x + pre .* (b - A x).  CsrGeometricMultigrid.cl:176-194."""
    L.call("mgx_csr_jacobi", A.nrows, pre.vp, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), b.vp, x.vp,
           y.vp, stream)


def gauss_seidel_color(nodes, A, b, x, backward=False, stream=None):
    """One colour of the multicolour Gauss-Seidel sweep (replaces the racy
    gauss_seidel_iteration, CsrGeometricMultigrid.cl:207-225)."""
    L.call("mgx_csr_gs_color", nodes.n, nodes.vp, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), b.vp,
           x.vp, int(backward), stream)


def dot(a, b, stream=None):
    """a . b.  CsrCG.cl:43-85 (dot_partial + dot_finalize)."""
    res = DeviceArray(1)
    L.call("mgx_dot", a.n, a.vp, b.vp, res.vp, stream)
    return float(res.numpy()[0])


def time_csr_kernel(kind, A, x, y, b=None, pre=None, warmup=5, repeats=100):
    """Average device ms per launch (HIP events) of kind 0 SpMV / 1 Jacobi / 2 residual / 3 FP32 SpMV /
    4 SpMV with the x.y partials and their reduction (the PCG step q = A p, p.q)."""
    out = C.c_double()
    L.call("mgx_time_csr_kernel", kind, A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), x.vp,
           b.vp if b is not None else None, pre.vp if pre is not None else None, y.vp, warmup, repeats, C.byref(out))
    return out.value


def time_spmm(A, X, Y, k, warmup=5, repeats=100, node=False, dot=False, nodeblock=False):
    """Average device ms per launch of Y = A X with k right-hand sides (FP32 if X is float32;
    node: the node-triple CSR kernel; nodeblock: the node-block kernel on a copy built from A;
    dot: the block-PCG form with the per-column X.Y partials)."""
    out = C.c_double()
    flags = int(X.dtype == np.float32) | (2 if node else 0) | (4 if dot else 0) | (8 if nodeblock else 0)
    L.call("mgx_time_spmm", flags, k, A.nrows, C.c_void_p(A.rowptr), C.c_void_p(A.col), C.c_void_p(A.val), X.vp, Y.vp,
           warmup, repeats, C.byref(out))
    return out.value


def time_duvi_spmv(D, x, y, form=0, warmup=5, repeats=100):
    """Average ms per DuVi SpMV launch; form 0 = the kernel the library chooses, 1 = the wave form."""
    out = C.c_double()
    L.call("mgx_time_duvi_spmv", C.byref(D.m), x.vp, y.vp, form, warmup, repeats, C.byref(out))
    return out.value


def simp_gradient(nx, ny, nz, rho, u, penal=3.0, Emin=1e-9, E0=1.0, nu=0.3):
    """dC/drho_e = -p rho_e^(p-1) (E0 - Emin) u_e^T Ke0 u_e on the device."""
    g = DeviceArray(nx * ny * nz)
    L.call("mgx_simp_gradient", nx, ny, nz, C.c_double(nu), rho.vp, C.c_double(penal), C.c_double(Emin),
           C.c_double(E0), u.vp, g.vp, None)
    return g
