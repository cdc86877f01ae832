"""Device buffers and device-resident matrices for the Python side of the C ABI."""
import ctypes as C
import os
from collections import namedtuple

import numpy as np

from . import _lib as L

CsrMatrix = namedtuple("CsrMatrix", "rowptr col val nrows ncols")
"""Host CSR (int32 row offsets / columns, float64 values), the layout of
DokRowMajor.BuildCsrArrays(sorted) (DokRowMajor.cs:286-324)."""


class DeviceArray:
    """A device allocation owned by Python (mgx_device_malloc / mgx_device_free)."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        L.call("mgx_device_malloc", C.byref(p), C.c_uint64(self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            L.call("mgx_memcpy_htod", C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), C.c_uint64(a.nbytes))
        return d

    @classmethod
    def zeros(cls, n, dtype=np.float64):
        d = cls(n, dtype)
        L.call("mgx_memset", C.c_void_p(d.ptr), 0, C.c_uint64(d.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            L.call("mgx_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_uint64(self.nbytes))
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        L.call("mgx_memcpy_htod", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_uint64(a.nbytes))

    @property
    def vp(self):
        return C.c_void_p(self.ptr)

    def free(self):
        if getattr(self, "ptr", None):
            L.lib().mgx_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCsr:
    """CSR arrays on the device, either uploaded by Python (arrays owned here) or owned by
    the library (an mgx_matrix handle from the structured generators)."""

    def __init__(self, handle=None, arrays=None, nrows=0, ncols=0, nnz=0):
        self.handle = handle
        self._arrays = arrays
        if handle is not None:
            v = L.CsrView()
            L.call("mgx_matrix_view", C.c_void_p(handle), C.byref(v))
            self.nrows, self.ncols, self.nnz = v.nrows, v.ncols, v.nnz
            self.rowptr, self.col, self.val = v.rowptr, v.col, v.val
        else:
            self.nrows, self.ncols, self.nnz = nrows, ncols, nnz
            self.rowptr, self.col, self.val = (a.ptr for a in arrays)

    @classmethod
    def from_host(cls, A):
        """Library-owned copy of a host CSR (usable by mgx_solver_set_level_device)."""
        h = C.c_void_p()
        rp = np.ascontiguousarray(A.rowptr, np.int32)
        ci = np.ascontiguousarray(A.col, np.int32)
        v = np.ascontiguousarray(A.val, np.float64)
        L.call("mgx_matrix_from_host", A.nrows, A.ncols, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
               v.ctypes.data_as(C.c_void_p), C.byref(h))
        return cls(handle=h.value)

    @classmethod
    def arrays_from_host(cls, A, val_dtype=np.float64):
        """Python-owned device arrays (for the device-level kernels)."""
        arrs = (DeviceArray.from_numpy(A.rowptr, np.int32), DeviceArray.from_numpy(A.col, np.int32),
                DeviceArray.from_numpy(A.val, val_dtype))
        return cls(arrays=arrs, nrows=A.nrows, ncols=A.ncols, nnz=len(A.col))

    def to_host(self):
        rp = np.empty(self.nrows + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        v = np.empty(self.nnz, np.float64)
        if self.handle is not None:
            L.call("mgx_matrix_download", C.c_void_p(self.handle), rp.ctypes.data_as(C.c_void_p),
                   ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p))
        else:
            rp, ci, v = (a.numpy() for a in self._arrays)
        return CsrMatrix(rp, ci, v, self.nrows, self.ncols)

    def free(self):
        if self.handle is not None:
            L.lib().mgx_matrix_destroy(C.c_void_p(self.handle))
            self.handle = None
        self._arrays = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ structured generators
def cantilever_matrix(nx, ny, nz, E=1.0, nu=0.3, dens=None):
    """K_ff of the clamped cantilever, assembled on the device (bit-identical to the
    element-order DOK sums of FemCantileverBase.cs:87-104)."""
    h = C.c_void_p()
    d = None if dens is None else np.ascontiguousarray(dens, np.float64)
    L.call("mgx_matrix_cantilever", nx, ny, nz, C.c_double(E), C.c_double(nu),
           d.ctypes.data_as(C.c_void_p) if d is not None else None, C.byref(h))
    return DeviceCsr(handle=h.value)


def transfer_reference(nx, ny, nz):
    """(R, P) of IGeometricMultigridModel.CreateRestrictionAndInterpolationMatrix."""
    r, p = C.c_void_p(), C.c_void_p()
    L.call("mgx_matrix_transfer_reference", nx, ny, nz, C.byref(r), C.byref(p))
    return DeviceCsr(handle=r.value), DeviceCsr(handle=p.value)


def transfer_trilinear(nx, ny, nz):
    """(R = P^T, P) with exact trilinear weights."""
    r, p = C.c_void_p(), C.c_void_p()
    L.call("mgx_matrix_transfer_trilinear", nx, ny, nz, C.byref(r), C.byref(p))
    return DeviceCsr(handle=r.value), DeviceCsr(handle=p.value)


def galerkin_hierarchy(nx, ny, nz, nlevels, E=1.0, nu=0.3, dens=None):
    """[K_0, P_0^T K_0 P_0, ...] assembled element by element on the device."""
    hs = (C.c_void_p * nlevels)()
    d = None if dens is None else np.ascontiguousarray(dens, np.float64)
    L.call("mgx_matrix_galerkin_hierarchy", nx, ny, nz, C.c_double(E), C.c_double(nu),
           d.ctypes.data_as(C.c_void_p) if d is not None else None, nlevels, hs)
    return [DeviceCsr(handle=hs[i]) for i in range(nlevels)]


def rng_uniform(seed, n):
    out = np.empty(n, np.float64)
    L.call("mgx_rng_uniform", C.c_uint64(seed), C.c_int64(n), out.ctypes.data_as(C.c_void_p))
    return out


def rng_normal(seed, n):
    out = np.empty(n, np.float64)
    L.call("mgx_rng_normal", C.c_uint64(seed), C.c_int64(n), out.ctypes.data_as(C.c_void_p))
    return out


# ------------------------------------------------------------------ on-disk hierarchy
def _csr_host(M):
    rp = np.ascontiguousarray(M.rowptr, np.int32)
    ci = np.ascontiguousarray(M.col, np.int32)
    v = np.ascontiguousarray(M.val, np.float64)
    h = L.CsrHost(M.nrows, M.ncols, len(ci), rp.ctypes.data, ci.ctypes.data, v.ctypes.data)
    return h, (rp, ci, v)


def write_hierarchy(path, levels, grid=None, duvi_mode=-1, duvi_tol=1e-10):
    """Host levels (finest first: dicts with A, and R, P above the coarsest; CSR with rowptr, col,
    val, nrows, ncols) to a hierarchy file (mgx_hierarchy_write).  grid: per level (nx, ny, nz)."""
    n = len(levels)
    keep = []
    A = (L.CsrHost * n)()
    R = (L.CsrHost * max(n - 1, 1))()
    P = (L.CsrHost * max(n - 1, 1))()
    for l, d in enumerate(levels):
        A[l], k = _csr_host(d["A"])
        keep.append(k)
        if l + 1 < n:
            R[l], k1 = _csr_host(d["R"])
            P[l], k2 = _csr_host(d["P"])
            keep += [k1, k2]
    g = None if grid is None else np.ascontiguousarray(grid, np.int32).reshape(n, 3)
    L.call("mgx_hierarchy_write", os.fsencode(path), n, A, R, P, None if g is None else g.ctypes.data_as(C.c_void_p),
           duvi_mode, C.c_double(duvi_tol))


def hierarchy_info(path):
    """(version, nlevels, has_grid, has_duvi, payload bytes) after verifying every checksum."""
    s = L.HierarchySummary()
    L.call("mgx_hierarchy_info", os.fsencode(path), C.byref(s))
    return s.version, s.nlevels, bool(s.has_grid), bool(s.has_duvi), s.payload_bytes


def _owned_to_csr(m):
    try:
        return CsrMatrix(np.ctypeslib.as_array(m.rowptr, (m.nrows + 1,)).copy(),
                         np.ctypeslib.as_array(m.col, (max(m.nnz, 1),))[:m.nnz].copy(),
                         np.ctypeslib.as_array(m.val, (max(m.nnz, 1),))[:m.nnz].copy(), m.nrows, m.ncols)
    finally:
        L.lib().mgx_csr_owned_free(C.byref(m))


def read_hierarchy_csr(path, level, which):
    """CSR of A (which 0), R (1) or P (2) of a level."""
    m = L.CsrOwned()
    L.call("mgx_hierarchy_read_csr", os.fsencode(path), level, which, C.byref(m))
    return _owned_to_csr(m)


# ------------------------------------------------------------------ Kirchhoff plate (FemPlate.cs)
def plate_free_dofs(nx, ny, Lx, Ly):
    """FindFreeDofs (FemPlate.cs:72-101): all-dof index 3 node + d of every free dof."""
    n = C.c_int32()
    L.call("mgx_plate_free_dofs", nx, ny, C.c_double(Lx), C.c_double(Ly), C.byref(n), None)
    d = np.empty(n.value, np.int32)
    L.call("mgx_plate_free_dofs", nx, ny, C.c_double(Lx), C.c_double(Ly), C.byref(n), d.ctypes.data_as(C.c_void_p))
    return d


def plate_system_host(nx, ny, Lx, Ly, t, E, nu, q):
    """(K_ff as host CSR, load vector) of the plate, without a device."""
    m = L.CsrOwned()
    nfree = len(plate_free_dofs(nx, ny, Lx, Ly))
    b = np.empty(nfree, np.float64)
    L.call("mgx_plate_system_host", nx, ny, C.c_double(Lx), C.c_double(Ly), C.c_double(t), C.c_double(E),
           C.c_double(nu), C.c_double(q), C.byref(m), b.ctypes.data_as(C.c_void_p))
    return _owned_to_csr(m), b


def plate_transfer_host(nx, ny, Lx, Ly):
    r, p = L.CsrOwned(), L.CsrOwned()
    L.call("mgx_plate_transfer_host", nx, ny, C.c_double(Lx), C.c_double(Ly), C.byref(r), C.byref(p))
    return _owned_to_csr(r), _owned_to_csr(p)


def plate_matrix(nx, ny, Lx, Ly, t, E, nu, q):
    """(K_ff on the device, load vector on the host)."""
    h = C.c_void_p()
    nfree = len(plate_free_dofs(nx, ny, Lx, Ly))
    b = np.empty(nfree, np.float64)
    L.call("mgx_matrix_plate", nx, ny, C.c_double(Lx), C.c_double(Ly), C.c_double(t), C.c_double(E), C.c_double(nu),
           C.c_double(q), C.byref(h), b.ctypes.data_as(C.c_void_p))
    return DeviceCsr(handle=h.value), b


def transfer_plate(nx, ny, Lx, Ly):
    """(R, P) between the plate and its coarser model, on the device."""
    r, p = C.c_void_p(), C.c_void_p()
    L.call("mgx_matrix_transfer_plate", nx, ny, C.c_double(Lx), C.c_double(Ly), C.byref(r), C.byref(p))
    return DeviceCsr(handle=r.value), DeviceCsr(handle=p.value)


# ------------------------------------------------------------------ plane-stress cantilever (FemCantilever2D.cs)
def _dens_ptr(dens, n):
    if dens is None:
        return None, None
    d = np.ascontiguousarray(dens, np.float64)
    if d.size != n:
        raise ValueError(f"{n} element densities expected, got {d.size}")
    return d, d.ctypes.data_as(C.c_void_p)


def cantilever2d_system_host(nx, ny, E=1.0, nu=0.3, dens=None):
    """K_ff of the 2D cantilever as host CSR, without a device."""
    m = L.CsrOwned()
    keep, dp = _dens_ptr(dens, nx * ny)
    L.call("mgx_cantilever2d_system_host", nx, ny, C.c_double(E), C.c_double(nu), dp, C.byref(m))
    return _owned_to_csr(m)


def cantilever2d_transfer_host(nx, ny):
    r, p = L.CsrOwned(), L.CsrOwned()
    L.call("mgx_cantilever2d_transfer_host", nx, ny, C.byref(r), C.byref(p))
    return _owned_to_csr(r), _owned_to_csr(p)


def cantilever2d_matrix(nx, ny, E=1.0, nu=0.3, dens=None):
    """K_ff of the 2D cantilever on the device."""
    h = C.c_void_p()
    keep, dp = _dens_ptr(dens, nx * ny)
    L.call("mgx_matrix_cantilever2d", nx, ny, C.c_double(E), C.c_double(nu), dp, C.byref(h))
    return DeviceCsr(handle=h.value)


def transfer_cantilever2d(nx, ny):
    """(R, P) between the 2D cantilever and its coarser model, on the device."""
    r, p = C.c_void_p(), C.c_void_p()
    L.call("mgx_matrix_transfer_cantilever2d", nx, ny, C.byref(r), C.byref(p))
    return DeviceCsr(handle=r.value), DeviceCsr(handle=p.value)


def read_hierarchy_duvi(path, level, which):
    """DuVi block layout of A / R / P of a level (the dict duvi_encode returns)."""
    h = L.DuviHost()
    L.call("mgx_hierarchy_read_duvi", os.fsencode(path), level, which, C.byref(h))
    return _duvi_host_to_dict(h)


def read_hierarchy_grid(path, level):
    g = np.zeros(3, np.int32)
    L.call("mgx_hierarchy_read_grid", os.fsencode(path), level, g.ctypes.data_as(C.c_void_p))
    return tuple(int(v) for v in g)


def save_hierarchy(path, hier, duvi_mode=-1, duvi_tol=1e-10):
    """Library-owned device levels [(A, R, P), ..., (A, None, None)] to a file (mgx_hierarchy_save)."""
    n = len(hier)
    A = (C.c_void_p * n)(*[h[0].handle for h in hier])
    R = (C.c_void_p * max(n - 1, 1))(*[h[1].handle for h in hier[:-1]])
    P = (C.c_void_p * max(n - 1, 1))(*[h[2].handle for h in hier[:-1]])
    L.call("mgx_hierarchy_save", os.fsencode(path), n, A, R, P, duvi_mode, C.c_double(duvi_tol))


def load_hierarchy(path, max_levels=16):
    """Device levels [(A, R, P), ..., (A, None, None)] from a file (mgx_hierarchy_load)."""
    A = (C.c_void_p * max_levels)()
    R = (C.c_void_p * max_levels)()
    P = (C.c_void_p * max_levels)()
    n = C.c_int32()
    L.call("mgx_hierarchy_load", os.fsencode(path), max_levels, C.byref(n), A, R, P)
    out = []
    for l in range(n.value):
        last = l == n.value - 1
        out.append((DeviceCsr(handle=A[l]), None if last else DeviceCsr(handle=R[l]),
                    None if last else DeviceCsr(handle=P[l])))
    return out


def matrix_grid(M):
    g = np.zeros(3, np.int32)
    L.call("mgx_matrix_grid", C.c_void_p(M.handle), g.ctypes.data_as(C.c_void_p))
    return tuple(int(v) for v in g)


def duvi_encode(A, tol=1e-10, mode=0):
    """Host DuVi encoder of the library (OpenCLDuViGeometricMultigridSolver.cs:753-851)."""
    h = L.DuviHost()
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    v = np.ascontiguousarray(A.val, np.float64)
    L.call("mgx_duvi_encode", A.nrows, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
           v.ctypes.data_as(C.c_void_p), C.c_double(tol), mode, C.byref(h))
    return _duvi_host_to_dict(h)


def _duvi_host_to_dict(h):
    """Copies a library-allocated mgx_duvi_host into numpy arrays and frees it."""
    nnz, nb, nu = h.nnz, h.nblocks, h.nuniq
    out = dict(nrows=h.nrows, nnz=nnz, nblocks=nb, nuniq=nu, vi_bytes=h.vi_bytes,
               row_blocks=np.ctypeslib.as_array(h.row_blocks, (h.nrows + 1,)).copy(),
               row_dist=np.ctypeslib.as_array(h.row_dist, (h.nrows + 1,)).copy(),
               block_col=np.ctypeslib.as_array(h.block_col, (max(nb, 1),))[:nb].copy(),
               dist=np.ctypeslib.as_array(h.dist, (max(nnz, 1),))[:nnz].copy(),
               values=np.ctypeslib.as_array(h.values, (max(nu, 1),))[:nu].copy())
    if h.vi_bytes:
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[h.vi_bytes]
        buf = (C.c_uint8 * (h.vi_bytes * max(nnz, 1))).from_address(h.vi)
        out["vi"] = np.frombuffer(buf, dtype=dt, count=nnz).copy()
    else:
        out["vi"] = None
    L.lib().mgx_duvi_host_free(C.byref(h))
    return out


class DeviceDuvi:
    """A DuVi block-layout matrix uploaded to the device (mgx_duvi)."""

    def __init__(self, d):
        self.keep = [DeviceArray.from_numpy(d["row_blocks"], np.int32), DeviceArray.from_numpy(d["row_dist"], np.int32),
                     DeviceArray.from_numpy(d["block_col"], np.int32), DeviceArray.from_numpy(d["dist"], np.uint8),
                     DeviceArray.from_numpy(d["values"], np.float64)]
        vi = None
        if d["vi"] is not None:
            vi = DeviceArray.from_numpy(d["vi"])
            self.keep.append(vi)
        m = L.Duvi()
        m.nrows, m.vi_bytes, m.nnz, m.nblocks, m.nuniq = d["nrows"], d["vi_bytes"], d["nnz"], d["nblocks"], d["nuniq"]
        m.row_blocks, m.row_dist, m.block_col, m.dist, m.values = (a.ptr for a in self.keep[:5])
        m.vi = vi.ptr if vi is not None else None
        self.m = m

    def spmv(self, x, y, clear=True):
        """y = A x (y += A x).  DuViGeometricMultigrid.cl:122-159."""
        L.call("mgx_duvi_spmv", C.byref(self.m), x.vp, y.vp, int(clear), None)

    def residual(self, b, x, r):
        """r = b - A x.  DuViGeometricMultigrid.cl:220-254."""
        L.call("mgx_duvi_residual", C.byref(self.m), b.vp, x.vp, r.vp, None)

    def residual_with_check(self, b, x, r, tolerance):
        """r = b - A x and the reference flag (1 converged, bit 1 failed).  DuViGeometricMultigrid.cl:275-317."""
        flag = DeviceArray.from_numpy(np.array([1], np.int32))
        L.call("mgx_duvi_residual_check", C.byref(self.m), b.vp, x.vp, r.vp, C.c_double(tolerance), flag.vp, None)
        return int(flag.numpy()[0])

    def jacobi(self, pre, b, x, y):
        """y = x + pre .* (b - A x).  DuViGeometricMultigrid.cl:362-402."""
        L.call("mgx_duvi_jacobi", C.byref(self.m), pre.vp, b.vp, x.vp, y.vp, None)

    def gauss_seidel_color(self, nodes, b, x, backward=False):
        """One colour of the multicolour Gauss-Seidel sweep (replaces gauss_seidel_iteration,
        DuViGeometricMultigrid.cl:422-462)."""
        L.call("mgx_duvi_gs_color", C.byref(self.m), nodes.n, nodes.vp, b.vp, x.vp, int(backward), None)

    def spmv_f32(self, x, y, clear=True):
        """FP32 y = A x with an FP32 copy of the value table (made on first use)."""
        if getattr(self, "values32", None) is None:
            self.values32 = DeviceArray.from_numpy(self.keep[4].numpy().astype(np.float32))
        L.call("mgx_duvi_spmv_f32", C.byref(self.m), self.values32.vp, x.vp, y.vp, int(clear), None)
