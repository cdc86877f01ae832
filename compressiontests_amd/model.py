"""The FE models of the reference (IGeometricMultigridModel with FemCantileverBase +
FemCantilever3D, and FemPlate), producing their matrices with the library's generators.

Cantilever mesh conventions (the reference takes them from the unvendored UniformCartesianMesh3D):
node id = ix + (nx+1)(iy + (ny+1) iz); element e = ex + nx (ey + ny ez); element corners in
VTK hexahedron order; 3 dofs (ux, uy, uz) per node; nodes with ix == 0 clamped
(FemCantileverBase.cs:137-158).

2D cantilever (FemCantilever2D): node id = ix + (nx+1) iy, element corners counter-clockwise from
(ex, ey), 2 dofs (ux, uy) per node, nodes with ix == 0 clamped.

Plate conventions (UniformCartesianMesh2D): node id = ix + (nx+1) iy; 3 dofs (w, dw/dy, dw/dx)
per node, node-major; w fixed on the lines x = 0, x = Lx/2, y = 0, y = Ly/2, y = Ly
(FemPlate.cs:126-188).
"""
import numpy as np

from .device import (DeviceArray, cantilever2d_matrix, cantilever_matrix, plate_free_dofs, plate_matrix, rng_normal,
                     transfer_cantilever2d, transfer_plate, transfer_reference)
from . import kernels


class FemCantilever3D:
    """FemCantilever3D (FemCantilever3D.cs:27-131) with the base class behaviour of
    FemCantileverBase (FemCantileverBase.cs:11-193).

    rhs: "manufactured" -> b = K_ff U_f with U_f from Euler-Bernoulli beam formulas
         (FemCantileverBase.cs:53-82,160-182), the reference's own right-hand side;
         "tip" -> total force `TipLoad` along z split equally over the x = Lx face nodes.
    """

    NumDofsPerNode = 3

    def __init__(self, numElementsPerAxis, lengthPerAxis=(1.0, 1.0, 1.0), ElementDensities=None,
                 ElasticityModulus=1.0, PoissonRatio=0.3, DistributedLoad=1.0, rhs="manufactured", TipLoad=-1.0):
        self.NumElements = tuple(int(n) for n in numElementsPerAxis)
        if len(self.NumElements) != 3:
            raise ValueError("FemCantilever3D needs 3 element counts")
        self.Lengths = tuple(float(v) for v in lengthPerAxis)
        self.ElementDensities = None if ElementDensities is None else np.ascontiguousarray(ElementDensities, np.float64)
        self.ElasticityModulus = ElasticityModulus
        self.PoissonRatio = PoissonRatio
        self.DistributedLoad = DistributedLoad
        self.rhs_kind = rhs
        self.TipLoad = TipLoad
        nx, ny, nz = self.NumElements
        self.NumNodes = (nx + 1, ny + 1, nz + 1)
        self.NumDofsAll = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        self.NumDofsFree = self.NumDofsAll - self.NumDofsAll // (nx + 1)
        L1, L2 = self.Lengths[1], self.Lengths[2]
        self.MomentOfInertia = L1 * L2 / 12 * L2 * L2            # FemCantilever3D.cs:28
        self._K = None

    # ---------------------------------------------------------------- matrices
    def device_matrix(self):
        """K_ff on the device (library-owned), element moduli = ElementDensities x E."""
        if self._K is None:
            nx, ny, nz = self.NumElements
            self._K = cantilever_matrix(nx, ny, nz, self.ElasticityModulus, self.PoissonRatio, self.ElementDensities)
        return self._K

    def CalcFreeDisplacements(self):
        """U_f of FemCantileverBase.cs:160-182 with FemCantilever3D.cs:122-128."""
        nx, ny, nz = self.NumElements
        Lx, Ly, Lz = self.Lengths
        ix = np.arange(1, nx + 1, dtype=np.float64)
        iz = np.arange(0, nz + 1, dtype=np.float64)
        x = ix * (Lx / nx)
        z = iz * (Lz / nz) - Lz / 2
        q, E, I, L = self.DistributedLoad, self.ElasticityModulus, self.MomentOfInertia, Lx
        w = -q / (24 * E * I) * (np.power(x, 4) - 4 * L * np.power(x, 3) + 6 * L * L * x * x)
        rot = -q / (6 * E * I) * (np.power(x, 3) - 3 * L * x * x + 3 * L * L * x)
        u = -rot[None, None, :] * z[:, None, None]                 # [iz, iy, ix]
        U = np.zeros((nz + 1, ny + 1, nx, 3))
        U[..., 0] = np.broadcast_to(u, (nz + 1, ny + 1, nx))
        U[..., 2] = np.broadcast_to(w[None, None, :], (nz + 1, ny + 1, nx))
        return U.reshape(-1)

    def tip_load(self):
        nx, ny, nz = self.NumElements
        f = np.zeros((nz + 1, ny + 1, nx, 3))
        f[:, :, nx - 1, 2] = self.TipLoad / ((ny + 1) * (nz + 1))
        return f.reshape(-1)

    def load_cases(self, nrhs, seed):
        """nrhs load cases, (nrhs, NumDofsFree): iid N(0,1) force components on every node of the
        top face z = Lz (all three directions; x = 0 nodes are clamped), zero elsewhere, each case
        scaled to unit 2-norm.  The normals come from one stream of `seed`, case after case, node
        order within a case."""
        nx, ny, nz = self.NumElements
        g = rng_normal(seed, nrhs * (ny + 1) * nx * 3).reshape(nrhs, ny + 1, nx, 3)
        f = np.zeros((nrhs, nz + 1, ny + 1, nx, 3))
        f[:, nz] = g
        f = f.reshape(nrhs, -1)
        return f / np.linalg.norm(f, axis=1, keepdims=True)

    def rhs(self):
        if self.rhs_kind == "tip":
            return self.tip_load()
        K = self.device_matrix()
        U = DeviceArray.from_numpy(self.CalcFreeDisplacements())
        b = DeviceArray(self.NumDofsFree)
        kernels.matrix_vector_product(K, U, b)
        return b.numpy()

    def CreateLinearSystem(self):
        """(A, b): K_ff as host CSR and the right-hand side (IGeometricMultigridModel.cs:39)."""
        return self.device_matrix().to_host(), self.rhs()

    def CreateAndSolveLinearSystem(self):
        A, b = self.CreateLinearSystem()
        return A, self.CalcFreeDisplacements(), b

    def FindFreeDofs(self):
        nx1, ny1, nz1 = self.NumNodes
        n = np.arange(nx1 * ny1 * nz1)
        free_nodes = n[(n % nx1) != 0]
        return (3 * free_nodes[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)

    def CreateCoarserModel(self):
        """FemCantileverBase.cs:186-191: halve every element count > 1, same lengths.  Like the
        reference, the coarser model has default densities and the unit-cube Ke."""
        c = tuple(n >> 1 if n > 1 else n for n in self.NumElements)
        return FemCantilever3D(c, self.Lengths, None, self.ElasticityModulus, self.PoissonRatio, self.DistributedLoad,
                               self.rhs_kind, self.TipLoad)


class FemCantilever2D:
    """FemCantilever2D (FemCantilever2D.cs:6-64) with the base class behaviour of FemCantileverBase
    (FemCantileverBase.cs:11-193): the plane-stress cantilever of 4-node elements, lengthPerAxis =
    (length, height, depth), clamped at x = 0, with the reference's manufactured right-hand side
    b = K_ff U_f (U_f from the Euler-Bernoulli beam, FemCantileverBase.cs:64-77, at z = y - height/2,
    FemCantilever2D.cs:51-57)."""

    NumDofsPerNode = 2

    def __init__(self, numElementsPerAxis, lengthPerAxis=(20.0, 1.0, 1.0), ElementDensities=None,
                 ElasticityModulus=1.0, PoissonRatio=0.3, DistributedLoad=1.0):
        self.NumElements = tuple(int(n) for n in numElementsPerAxis)
        if len(self.NumElements) != 2:
            raise ValueError("FemCantilever2D needs 2 element counts")
        if len(lengthPerAxis) != 3:
            raise ValueError("FemCantilever2D needs (length, height, depth)")
        self.Lengths = tuple(float(v) for v in lengthPerAxis[:2])
        self.CantileverDepth = float(lengthPerAxis[2])
        self.ElementDensities = None if ElementDensities is None else np.ascontiguousarray(ElementDensities, np.float64)
        self.ElasticityModulus = ElasticityModulus
        self.PoissonRatio = PoissonRatio
        self.DistributedLoad = DistributedLoad
        nx, ny = self.NumElements
        self.NumNodes = (nx + 1, ny + 1)
        self.NumDofsAll = 2 * (nx + 1) * (ny + 1)
        self.NumDofsFree = self.NumDofsAll - self.NumDofsAll // (nx + 1)
        H, D = self.Lengths[1], self.CantileverDepth
        self.MomentOfInertia = H * D / 12 * H * H                 # FemCantilever2D.cs:24
        self._K = None

    def device_matrix(self):
        """K_ff on the device (library-owned), element moduli = ElementDensities x E."""
        if self._K is None:
            nx, ny = self.NumElements
            self._K = cantilever2d_matrix(nx, ny, self.ElasticityModulus, self.PoissonRatio, self.ElementDensities)
        return self._K

    def CalcFreeDisplacements(self):
        """U_f of FemCantileverBase.cs:160-182 with FemCantilever2D.cs:51-57."""
        nx, ny = self.NumElements
        Lx, Ly = self.Lengths
        x = np.arange(1, nx + 1, dtype=np.float64) * (Lx / nx)
        z = np.arange(0, ny + 1, dtype=np.float64) * (Ly / ny) - Ly / 2
        q, E, I, L = self.DistributedLoad, self.ElasticityModulus, self.MomentOfInertia, Lx
        w = -q / (24 * E * I) * (np.power(x, 4) - 4 * L * np.power(x, 3) + 6 * L * L * x * x)
        rot = -q / (6 * E * I) * (np.power(x, 3) - 3 * L * x * x + 3 * L * L * x)
        U = np.empty((ny + 1, nx, 2))
        U[..., 0] = -rot[None, :] * z[:, None]
        U[..., 1] = w[None, :]
        return U.reshape(-1)

    def rhs(self):
        K = self.device_matrix()
        U = DeviceArray.from_numpy(self.CalcFreeDisplacements())
        b = DeviceArray(self.NumDofsFree)
        kernels.matrix_vector_product(K, U, b)
        return b.numpy()

    def CreateLinearSystem(self):
        """(A, b): K_ff as host CSR and the manufactured right-hand side."""
        return self.device_matrix().to_host(), self.rhs()

    def CreateAndSolveLinearSystem(self):
        A, b = self.CreateLinearSystem()
        return A, self.CalcFreeDisplacements(), b

    def FindFreeDofs(self):
        nx1, ny1 = self.NumNodes
        n = np.arange(nx1 * ny1)
        free_nodes = n[(n % nx1) != 0]
        return (2 * free_nodes[:, None] + np.arange(2)[None, :]).reshape(-1).astype(np.int32)

    def CreateCoarserModel(self):
        """FemCantileverBase.cs:186-191: halve every element count > 1, same lengths, default
        densities."""
        c = tuple(n >> 1 if n > 1 else n for n in self.NumElements)
        return FemCantilever2D(c, self.Lengths + (self.CantileverDepth,), None, self.ElasticityModulus,
                               self.PoissonRatio, self.DistributedLoad)


class FemPlate:
    """FemPlate (FemPlate.cs:20-189): a Kirchhoff plate of ACM rectangles (12-dof element of
    PlateElementMatrixProvider) under the uniform load DistributedLoad, assembled node-major over
    the free dofs.  lengthPerAxis = (Lx, Ly, thickness).

    NumDofsFree is the count of the free dofs the supports leave; FemPlate.cs:118-122 states it
    as NumDofsAll - 2 (ny+1) - 3 (nx+1) + 6, the same number whenever both element counts are
    even (the lines x = Lx/2 and y = Ly/2 then hold nodes)."""

    NumDofsPerNode = 3

    def __init__(self, numElementsPerAxis, lengthPerAxis, E=1.0, v=0.3, q=0.0):
        self.NumElements = tuple(int(n) for n in numElementsPerAxis)
        if len(self.NumElements) != 2:
            raise ValueError("FemPlate needs 2 element counts")
        if len(lengthPerAxis) != 3:
            raise ValueError("FemPlate needs (length x, length y, thickness)")
        self.Lengths = tuple(float(v_) for v_ in lengthPerAxis[:2])
        self.Thickness = float(lengthPerAxis[2])
        self.ElasticityModulus = float(E)
        self.PoissonRatio = float(v)
        self.DistributedLoad = float(q)
        nx, ny = self.NumElements
        self.NumNodes = (nx + 1, ny + 1)
        self.NumDofsAll = 3 * (nx + 1) * (ny + 1)
        self._free = plate_free_dofs(nx, ny, *self.Lengths)
        self.NumDofsFree = len(self._free)
        self._K = None
        self._b = None

    def _system(self):
        if self._K is None:
            nx, ny = self.NumElements
            self._K, self._b = plate_matrix(nx, ny, *self.Lengths, self.Thickness, self.ElasticityModulus,
                                            self.PoissonRatio, self.DistributedLoad)
        return self._K, self._b

    def device_matrix(self):
        """K_ff on the device (library-owned)."""
        return self._system()[0]

    def rhs(self):
        """The load vector of PlateElementVectorAssembler (FemPlate.cs:65-67) over the free dofs."""
        return self._system()[1].copy()

    def CreateLinearSystem(self):
        """(A, b): K_ff as host CSR and the load vector (FemPlate.cs:40-71)."""
        K, b = self._system()
        return K.to_host(), b.copy()

    def FindFreeDofs(self):
        """FemPlate.cs:72-101: the all-dof index 3 node + d of every free dof, ascending."""
        return self._free.copy()

    def CreateCoarserModel(self):
        """FemPlate.cs:103-108: halve every element count > 1, same lengths and material."""
        c = tuple(n >> 1 if n > 1 else n for n in self.NumElements)
        return FemPlate(c, self.Lengths + (self.Thickness,), self.ElasticityModulus, self.PoissonRatio,
                        self.DistributedLoad)


def CreateCoarserModelAndSmoothenerMatrices(model):
    """IGeometricMultigridModel.cs:60-65: (coarser model, R, P) with R, P on the device."""
    coarser = model.CreateCoarserModel()
    if isinstance(model, FemPlate):
        R, P = transfer_plate(*model.NumElements, *model.Lengths)
    elif isinstance(model, FemCantilever2D):
        R, P = transfer_cantilever2d(*model.NumElements)
    else:
        R, P = transfer_reference(*model.NumElements)
    return coarser, R, P
