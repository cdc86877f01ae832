"""The workload configurations at their full sizes, through the C ABI, anchored on the oracle:

* 128^3 precision form (6,390,144 dofs, 509,597,550 nonzeros, moduli log-normal sigma 1.5
  clipped to a 1e3 contrast, tip load, 5-level Galerkin V-cycle with multicolour Gauss-Seidel,
  FP32 inner MG-PCG inside FP64 iterative refinement): the FP64 true residual, an FP64 GPU
  solve to 1e-12, and the oracle's FP64 MG-PCG (OpenMP build, bit-identical to the serial one)
  on the same hierarchy.
* 64^3 kernels (811,200 dofs, 63,695,790 nonzeros, moduli log-uniform in [1e-2, 1e2]): the
  DuVi encodings (raw-value fallback and u32 indices; the homogeneous matrix's u8 table)
  bit-exact against the oracle encoder (OpenCLDuViGeometricMultigridSolver.cs:753-851), the DuVi
  SpMV against the oracle's DuVi SpMV (DuViGeometricMultigrid.cl:57-99), and the CSR damped
  Jacobi sweep and residual against the oracle (CsrGeometricMultigrid.cl:96-110,176-194).
"""
import numpy as np
import pytest

from compressiontests_amd import kernels as K
from compressiontests_amd.device import DeviceArray, DeviceCsr, DeviceDuvi, duvi_encode
from compressiontests_amd.model import FemCantilever3D
from compressiontests_amd.solvers import HipMgPcgSolver
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED_LOGUNIFORM = 20261203 + 2
SEED_LOGNORMAL = 20261203 + 3


def log(msg):
    print(msg, flush=True)


def device_energy_err(Kd, x, xref):
    """||x - xref||_K / ||xref||_K with the library's FP64 SpMV on the device matrix."""
    e, Ke = DeviceArray.from_numpy(x - xref), DeviceArray(len(x))
    K.matrix_vector_product(Kd, e, Ke)
    r, Kr = DeviceArray.from_numpy(xref), DeviceArray(len(x))
    K.matrix_vector_product(Kd, r, Kr)
    return float(np.sqrt(max((x - xref) @ Ke.numpy(), 0.0) / (xref @ Kr.numpy())))


def oracle_levels(hier):
    lv = []
    for A, R, P in hier:
        d = {"A": _csr(A)}
        if R is not None:
            d["R"], d["P"] = _csr(R), _csr(P)
        lv.append(d)
    return lv


def _csr(M):
    h = M.to_host()
    return O.Csr(h.rowptr, h.col, h.val, h.nrows, h.ncols)


@pytest.mark.slow
def test_config3_precision_form_128(gpu):
    n = 128
    g = np.clip(1.5 * O.rng_normal(SEED_LOGNORMAL, n ** 3), -0.5 * np.log(1e3), 0.5 * np.log(1e3))
    dens = np.exp(g)
    assert abs(dens.max() / dens.min() - 1e3) < 1e-9 * 1e3
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    f = m.rhs()
    s = HipMgPcgSolver(m, levels=5, smoother="gs", nu=1, tolerance=1e-8, precision="mixed")
    s.Initialize(f)
    x, st = s.Solve()
    log(f"128^3 precision form: converged {st.HasConverged}, {st.NumIterationsRequired} inner iterations, "
        f"{st.OuterIterations} refinement steps, true residual {st.RelativeResidual:.3e}")
    assert st.HasConverged and st.OuterIterations >= 1
    assert st.RelativeResidual <= 1.01e-8
    Kd = m.device_matrix()
    assert Kd.nrows == 6_390_144 and Kd.nnz == 509_597_550
    Kx = DeviceArray(Kd.nrows)
    K.matrix_vector_product(Kd, DeviceArray.from_numpy(x), Kx)
    rel = np.linalg.norm(f - Kx.numpy()) / np.linalg.norm(f)
    assert rel <= 1.01e-8, rel
    t = HipMgPcgSolver(m, levels=5, smoother="gs", nu=1, tolerance=1e-12, hierarchy=s.hierarchy)
    t.Initialize(f)
    xt, stt = t.Solve()
    assert stt.HasConverged
    e_gpu = device_energy_err(Kd, x, xt)
    log(f"energy-norm error against the FP64 1e-12 GPU solve: {e_gpu:.3e}")
    assert e_gpu <= 1e-8
    # the oracle: FP64 MG-PCG on the same Galerkin hierarchy, multicolour Gauss-Seidel in its own
    # colour order (so iteration counts agree within 10 %, not exactly)
    lv = oracle_levels(s.hierarchy)
    xo8, so8 = O.mgpcg(lv, f, smoother=1, omega=0.0, nu=1, tol=1e-8, omp=True)
    p8 = HipMgPcgSolver(m, levels=5, smoother="gs", nu=1, tolerance=1e-8, hierarchy=s.hierarchy)
    p8.Initialize(f)
    x8, st8 = p8.Solve()
    log(f"to 1e-8: oracle {so8.iterations} iterations, FP64 GPU {st8.NumIterationsRequired}, precision form "
        f"{st.NumIterationsRequired} inner iterations in {st.OuterIterations} refinement steps")
    assert so8.converged == 1 and st8.HasConverged
    assert abs(st8.NumIterationsRequired - so8.iterations) <= max(1, round(0.1 * so8.iterations))
    # each refinement step restarts the Krylov space: a few iterations more than one FP64 PCG run
    assert st.NumIterationsRequired <= 1.25 * so8.iterations + 2
    xo, so = O.mgpcg(lv, f, smoother=1, omega=0.0, nu=1, tol=1e-11, omp=True)
    assert so.converged == 1
    e_orc = device_energy_err(Kd, x, xo)
    e_fp64 = device_energy_err(Kd, xt, xo)
    log(f"energy-norm error against the oracle's 1e-11 solve: precision form {e_orc:.3e}, FP64 GPU {e_fp64:.3e}")
    assert e_orc <= 1e-8
    assert e_fp64 <= 1e-9


@pytest.fixture(scope="module")
def k64(gpu):
    n = 64
    dens = 10 ** (4 * O.rng_uniform(SEED_LOGUNIFORM, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    assert A.nrows == 811_200 and A.val.size == 63_695_790
    x = O.rng_normal(20261206, A.nrows)
    return A, x, np.abs(A.val).max() * np.abs(x).max()


def same_encoding(d, o):
    for k in ("row_blocks", "row_dist", "block_col", "dist"):
        if not np.array_equal(d[k], o[k]):
            return k
    if d["vi_bytes"] != o["vi_bytes"] or d["nblocks"] != o["nblocks"] or d["values"].tobytes() != o["values"].tobytes():
        return "vi_bytes/nblocks/values"
    if (d["vi"] is None) != (o["vi"] is None) or (d["vi"] is not None and d["vi"].tobytes() != o["vi"].tobytes()):
        return "vi"
    return None


@pytest.mark.slow
@pytest.mark.parametrize("mode,vib", [(0, 0), (4, 4)])
def test_config2_duvi_encode_and_spmv_64(k64, mode, vib):
    """Heterogeneous K has ~1e7 distinct values after the 1e-10 merge: mode 0 falls back to raw
    values (the intent of the reference's >65,536 path), mode 4 stores u32 indices."""
    A, x, scale = k64
    d = duvi_encode(A, 1e-10, mode)
    o = O.duvi_block_encode(A, 1e-10, mode)
    assert d["vi_bytes"] == vib
    assert same_encoding(d, o) is None, same_encoding(d, o)
    log(f"64^3 DuVi mode {mode}: {d['nblocks']} blocks, {d['nuniq']} values, bit-exact")
    Dd = DeviceDuvi(d)
    y = DeviceArray(A.nrows)
    Dd.spmv(DeviceArray.from_numpy(x), y)
    ref = O.duvi_block_spmv(o, x)
    err = np.abs(y.numpy() - ref).max() / scale
    log(f"64^3 DuVi SpMV mode {mode}: max error {err:.2e} of max|a| max|x|")
    assert err <= 1e-12
    # the fused residual and Jacobi sweep on the same encoding, against the oracle on the matrix
    # the encoding stores (DuViGeometricMultigrid.cl:220-254, :362-402)
    M = O.duvi_block_to_csr(o, A.ncols)
    b = O.rng_normal(20261208, A.nrows)
    pre = 0.6 * O.jacobi_preconditioner(M)
    db, dx = DeviceArray.from_numpy(b), DeviceArray.from_numpy(x)
    r = DeviceArray(A.nrows)
    Dd.residual(db, dx, r)
    assert np.abs(r.numpy() - O.residual(M, b, x)).max() <= 1e-12 * (scale + np.abs(b).max())
    Dd.jacobi(DeviceArray.from_numpy(pre), db, dx, y)
    err = np.abs(y.numpy() - O.jacobi(M, pre, b, x)) / (np.abs(x).max() + np.abs(pre) * (np.abs(b).max() + scale))
    assert err.max() <= 1e-12


@pytest.mark.slow
def test_config2_duvi_homogeneous_u8_64(gpu):
    """The reference's own case: the homogeneous stiffness matrix has a handful of distinct
    values, so mode 0 stores u8 indices (OpenCLDuViGeometricMultigridSolver.cs:798-812)."""
    n = 64
    A = O.cantilever(n, n, n)
    d = duvi_encode(A, 1e-10, 0)
    o = O.duvi_block_encode(A, 1e-10, 0)
    assert d["vi_bytes"] == 1 and d["nuniq"] <= 256
    assert same_encoding(d, o) is None, same_encoding(d, o)
    x = O.rng_normal(20261207, A.nrows)
    y = DeviceArray(A.nrows)
    DeviceDuvi(d).spmv(DeviceArray.from_numpy(x), y)
    ref = O.duvi_block_spmv(o, x)
    assert np.abs(y.numpy() - ref).max() <= 1e-12 * np.abs(A.val).max() * np.abs(x).max()


@pytest.mark.slow
def test_config2_jacobi_and_residual_64(k64):
    A, x, scale = k64
    dA = DeviceCsr.arrays_from_host(A)
    b = O.rng_normal(20261208, A.nrows)
    pre = 0.6 * O.jacobi_preconditioner(A)
    db, dx, dpre = DeviceArray.from_numpy(b), DeviceArray.from_numpy(x), DeviceArray.from_numpy(pre)
    y = DeviceArray(A.nrows)
    K.jacobi_iteration(dpre, dA, db, dx, y)
    ref = O.jacobi(A, pre, b, x)
    # per row: y_i = x_i + pre_i (b_i - (A x)_i); the row sum carries the rounding
    err = np.abs(y.numpy() - ref) / (np.abs(x).max() + np.abs(pre) * (np.abs(b).max() + scale))
    log(f"64^3 Jacobi sweep: max scaled error {err.max():.2e}")
    assert err.max() <= 1e-12
    r = DeviceArray(A.nrows)
    K.residual(dA, db, dx, r)
    rref = O.residual(A, b, x)
    assert np.abs(r.numpy() - rref).max() <= 1e-12 * (scale + np.abs(b).max())


@pytest.mark.slow
def test_headline_path_32_against_oracle(gpu):
    """The headline path at 32^3 against the oracle: batching + precision form (8 load cases in one
    block, node-block SpMM and Gauss-Seidel, colour pairs, the zero-start sweep with its
    later-colour residual, node transfers), moduli log-normal clipped to a 1e3 contrast.  Every
    column: inner iterations within 10 % of the oracle's FP64 MG-PCG run in the solver's own
    Gauss-Seidel node order, and at a 1e-9 stop the energy-norm error against the oracle's 1e-12
    solution within 1e-8."""
    n, nrhs = 32, 8
    g = np.clip(1.5 * O.rng_normal(SEED_LOGNORMAL, n ** 3), -0.5 * np.log(1e3), 0.5 * np.log(1e3))
    dens = np.exp(g)
    m = FemCantilever3D((n, n, n), ElementDensities=dens, rhs="tip")
    F = m.load_cases(nrhs, 20261203 + 4)
    s = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-8, nrhs=nrhs, block_width=8,
                       precision="mixed", gs_pairs=True)
    s.Initialize(F)
    assert s.BlockWidth[0] == 8
    assert [s.LevelLayout(l) for l in range(2)] == ["node-block"] * 2
    assert [s.TransferLayout(l) for l in range(2)] == ["node-transfer"] * 2
    assert s.LevelColors(0) == 8 and s.GsPairs(0) == 4
    X, st = s.Solve()
    lv = oracle_levels(s.hierarchy)
    lvo = [dict(d, gs_order=s.GsOrder(l)) if "R" in d else d for l, d in enumerate(lv)]
    t = HipMgPcgSolver(m, levels=3, smoother="gs", nu=1, tolerance=1e-9, nrhs=nrhs, block_width=8,
                       precision="mixed", gs_pairs=True, hierarchy=s.hierarchy)
    t.Initialize(F)
    Xt, stt = t.Solve()
    Kd = m.device_matrix()
    for j in range(nrhs):
        xo8, so8 = O.mgpcg(lvo, F[j], smoother=1, omega=0.0, nu=1, tol=1e-8)
        xo, so = O.mgpcg(lv, F[j], smoother=1, omega=0.0, nu=1, tol=1e-12)
        e = device_energy_err(Kd, Xt[j], xo)
        log(f"32^3 column {j}: iterations {st[j].NumIterationsRequired} (oracle {so8.iterations}), energy-norm "
            f"error at 1e-9 {e:.2e}")
        assert st[j].HasConverged and st[j].RelativeResidual <= 1.01e-8
        assert abs(st[j].NumIterationsRequired - so8.iterations) <= max(1, round(0.1 * so8.iterations)), j
        assert stt[j].HasConverged and e <= 1e-8, j
