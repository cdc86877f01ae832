"""Device-level kernels of libmgx against the oracle, through the C ABI.
FP64 results agree within 1e-12 relative (summation order differs: lanes of a wavefront
reduce a row in a tree); index work and copies are bit-exact."""
import json
import os

import numpy as np
import pytest

from compressiontests_amd import kernels as K
from compressiontests_amd.device import DeviceArray, DeviceCsr, DeviceDuvi, duvi_encode
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def close(a, b, rtol=RTOL):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


@pytest.fixture(scope="module")
def k16(gpu):
    n = 16
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    return A, DeviceCsr.arrays_from_host(A)


def test_opencl_diagonal_kat(gpu):
    """GMCantileverOpenCLTests.cs:16-52: y_i == v_i^2 exactly."""
    idx = np.arange(17, dtype=np.int32)
    vals = np.arange(16, 0, -1).astype(np.float64)
    A = DeviceCsr.arrays_from_host(O.Csr(idx, idx[:16].copy(), vals, 16, 16))
    x = DeviceArray.from_numpy(vals)
    y = DeviceArray(16)
    K.matrix_vector_product(A, x, y)
    assert np.array_equal(y.numpy(), vals * vals)


def test_spmv_clear_and_accumulate(k16):
    A, dA = k16
    rng = np.random.default_rng(0)
    x = rng.normal(size=A.nrows)
    y0 = rng.normal(size=A.nrows)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y0)
    K.matrix_vector_product(dA, dx, dy, clear=True)
    assert close(dy.numpy(), O.spmv(A, x))
    dy.copy_from(y0)
    K.matrix_vector_product(dA, dx, dy, clear=False)
    assert close(dy.numpy(), O.spmv(A, x, y0.copy(), clear=False))


def test_residual_jacobi(k16):
    A, dA = k16
    rng = np.random.default_rng(1)
    x, b = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    pre = 0.6 * O.jacobi_preconditioner(A)
    dx, db, dpre, dy = (DeviceArray.from_numpy(v) for v in (x, b, pre, np.zeros(A.nrows)))
    K.residual(dA, db, dx, dy)
    assert close(dy.numpy(), O.residual(A, b, x))
    K.jacobi_iteration(dpre, dA, db, dx, dy)
    assert close(dy.numpy(), O.jacobi(A, pre, b, x))
    K.jacobi_initial_iteration(dpre, db, dy)
    assert np.array_equal(dy.numpy(), O.jacobi_initial(pre, b))


def test_residual_check_flags(k16):
    A, dA = k16
    x = np.zeros(A.nrows)
    dx = DeviceArray.from_numpy(x)
    r = DeviceArray(A.nrows)
    for b, tol in ((np.full(A.nrows, 1e-9), 1e-8), (np.full(A.nrows, 1e-3), 1e-8)):
        _, ref = O.residual_check(A, b, x, tol)
        assert K.residual_with_check(dA, DeviceArray.from_numpy(b), dx, r, tol) == ref
    b = np.full(A.nrows, 1e-3); b[17] = np.nan
    assert K.residual_with_check(dA, DeviceArray.from_numpy(b), dx, r, 1e-8) & 2


def test_dot(k16):
    A, _ = k16
    rng = np.random.default_rng(2)
    for n in (1, 255, 256, 257, A.nrows, 3_000_001):
        a, b = rng.normal(size=n), rng.normal(size=n)
        got = K.dot(DeviceArray.from_numpy(a), DeviceArray.from_numpy(b))
        assert abs(got - a @ b) <= 1e-12 * np.abs(a * b).sum()


def test_edge_shapes(gpu):
    """Empty rows, a single row, a row longer than a wavefront, nrows = 0."""
    rows = [[], [0, 3, 7], list(range(0, 300)), [], [299]]
    rp, ci = [0], []
    for r in rows:
        ci += r
        rp.append(len(ci))
    val = np.random.default_rng(3).normal(size=len(ci))
    A = O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, len(rows), 300)
    dA = DeviceCsr.arrays_from_host(A)
    x = np.random.default_rng(4).normal(size=300)
    y = DeviceArray(len(rows))
    K.matrix_vector_product(dA, DeviceArray.from_numpy(x), y)
    assert close(y.numpy(), O.spmv(A, x))
    E = DeviceCsr.arrays_from_host(O.Csr(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), 0, 0))
    K.matrix_vector_product(E, DeviceArray(1), DeviceArray(1))


@pytest.mark.parametrize("f32", [False, True])
def test_stream_kernel_long_rows_and_chunks(gpu, f32):
    """Rows long enough that one 64-row tile (32 rows in FP32) spans several LDS chunks, one
    row longer than a chunk, empty rows and a ragged last tile: SpMV, y += Ax, residual and
    Jacobi against the FP64 CPU product."""
    rng = np.random.default_rng(11)
    ncols = 7000
    lens = [400] * 70 + [0, 6500, 3, 0] + [90] * 9
    rp, ci = [0], []
    for L in lens:
        ci += sorted(rng.choice(ncols, size=L, replace=False).tolist())
        rp.append(len(ci))
    val = rng.normal(size=len(ci))
    A = O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, len(lens), ncols)
    dt = np.float32 if f32 else np.float64
    dA = DeviceCsr.arrays_from_host(A, val_dtype=dt)
    x = rng.normal(size=ncols)
    Ah = O.Csr(A.rowptr, A.col, val.astype(dt).astype(np.float64), A.nrows, ncols)
    xh = x.astype(dt).astype(np.float64)
    ref = O.spmv(Ah, xh)
    tol = 1e-5 if f32 else 1e-12
    scale = np.abs(Ah.val).max() * np.abs(xh).max() * np.sqrt(6500)
    dx = DeviceArray.from_numpy(x.astype(dt))
    y = DeviceArray(A.nrows, dt)
    K.matrix_vector_product(dA, dx, y)
    assert np.abs(y.numpy() - ref).max() <= tol * scale
    K.matrix_vector_product(dA, dx, y, clear=False)
    assert np.abs(y.numpy() - 2 * ref).max() <= 2 * tol * scale
    if not f32:
        # square part for the residual / Jacobi epilogues: first nrows columns of x
        n = A.nrows
        sq = O.Csr(A.rowptr, np.minimum(A.col, n - 1).astype(np.int32), val, n, n)
        dS = DeviceCsr.arrays_from_host(sq)
        xs, b = rng.normal(size=n), rng.normal(size=n)
        pre = np.full(n, 0.3)
        r = DeviceArray(n)
        K.residual(dS, DeviceArray.from_numpy(b), DeviceArray.from_numpy(xs), r)
        assert close(r.numpy(), O.residual(sq, b, xs), 1e-11)
        yj = DeviceArray(n)
        K.jacobi_iteration(DeviceArray.from_numpy(pre), dS, DeviceArray.from_numpy(b), DeviceArray.from_numpy(xs), yj)
        assert close(yj.numpy(), O.jacobi(sq, pre, b, xs), 1e-11)


def test_spmv_f32(k16):
    A, _ = k16
    dA32 = DeviceCsr.arrays_from_host(A, val_dtype=np.float32)
    x = np.random.default_rng(5).normal(size=A.nrows)
    dx = DeviceArray.from_numpy(x.astype(np.float32))
    dy = DeviceArray(A.nrows, np.float32)
    K.matrix_vector_product(dA32, dx, dy)
    A32 = O.Csr(A.rowptr, A.col, A.val.astype(np.float32).astype(np.float64), A.nrows, A.ncols)
    ref = O.spmv(A32, x.astype(np.float32).astype(np.float64))
    scale = np.abs(A32.val).max() * np.abs(x).max() * 81
    assert np.abs(dy.numpy() - ref).max() <= 1e-5 * scale


def test_gs_color_matches_oracle(gpu):
    n = 8
    A = O.cantilever(n, n, n, dens=10 ** (O.rng_uniform(3, n ** 3) - 0.5))
    dA = DeviceCsr.arrays_from_host(A)
    ncol, color = O.block_coloring(A)
    rng = np.random.default_rng(6)
    b, x = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    for backward in (False, True):
        dx, db = DeviceArray.from_numpy(x), DeviceArray.from_numpy(b)
        order = range(ncol - 1, -1, -1) if backward else range(ncol)
        for c in order:
            nodes = DeviceArray.from_numpy(np.nonzero(color == c)[0].astype(np.int32))
            K.gauss_seidel_color(nodes, dA, db, dx, backward)
        ref = O.gs_multicolor(A, ncol, color, b, x, backward)
        assert close(dx.numpy(), ref, 1e-11)


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_duvi_spmv_matches_oracle(gpu, mode):
    n = 8
    dens = None if mode == 0 else 10 ** (4 * O.rng_uniform(9, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    d = duvi_encode(A, 1e-10, mode)
    Dd = DeviceDuvi(d)
    x = np.random.default_rng(7).normal(size=A.nrows)
    y = DeviceArray(A.nrows)
    Dd.spmv(DeviceArray.from_numpy(x), y)
    ref = O.duvi_block_spmv(O.duvi_block_encode(A, 1e-10, mode), x)
    assert close(y.numpy(), ref)


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_duvi_ragged_rows(gpu, mode):
    """Rows with > 256 consecutive columns (blocks split at 255 further entries), gaps of 255/256
    and more, empty rows, a tile longer than the LDS capacity (one-thread-per-row fallback), a
    row decoded over several rounds, and accumulation (y += A x), against the oracle's DuVi SpMV."""
    rng = np.random.default_rng(21)
    ncols = 70000
    rows = [list(range(0, 300)), [], [5, 261, 262, 518, 60000], list(range(1000, 1600, 2))]
    rows += [sorted(rng.choice(ncols, size=40, replace=False).tolist()) for _ in range(40)]
    rows += [list(range(2000, 9000))] + [[7, 8]] * 5
    # a tile of four rows holding one 300-entry row (several 96-slot decode rounds) and three empty rows
    rows += [[], []] + [list(range(3000, 3600, 2))] + [[], [], []]
    rp, ci = [0], []
    for r in rows:
        ci += r
        rp.append(len(ci))
    val = np.round(rng.normal(size=len(ci)), 1) if mode == 0 else rng.normal(size=len(ci))
    A = O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, len(rows), ncols)
    d = duvi_encode(A, 1e-10, mode)
    Dd = DeviceDuvi(d)
    x = rng.normal(size=ncols)
    ref = O.duvi_block_spmv(O.duvi_block_encode(A, 1e-10, mode), x)
    y = DeviceArray(A.nrows)
    Dd.spmv(DeviceArray.from_numpy(x), y)
    assert close(y.numpy(), ref)
    Dd.spmv(DeviceArray.from_numpy(x), y, clear=False)
    assert close(y.numpy(), 2 * ref)


def test_duvi_kat(gpu):
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "duvi_kat.json")))
    A = O.Csr(np.array(g["rowptr"], np.int32), np.array(g["col"], np.int32), np.array(g["val"], float), 2, 1024)
    Dd = DeviceDuvi(duvi_encode(A))
    x = np.zeros(1024)
    for k, v in g["x_nonzeros"].items():
        x[int(k)] = v
    y = DeviceArray(2)
    Dd.spmv(DeviceArray.from_numpy(x), y)
    assert y.numpy().tolist() == g["y"]


def test_simp_gradient_matches_oracle(gpu):
    n = 6
    rho = 0.1 + 0.9 * O.rng_uniform(11, n ** 3)
    u = np.random.default_rng(8).normal(size=O.n_free(n, n, n))
    g = K.simp_gradient(n, n, n, DeviceArray.from_numpy(rho), DeviceArray.from_numpy(u))
    ref = O.simp_gradient(n, n, n, rho, u)
    assert close(g.numpy(), ref)


@pytest.mark.slow
def test_spmv_64_parity(gpu):
    """Config 2 size: 811,200 rows, 63.7M nnz, log-uniform moduli in [1e-2, 1e2]."""
    n = 64
    dens = 10 ** (4 * O.rng_uniform(20261205, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    dA = DeviceCsr.arrays_from_host(A)
    x = O.rng_normal(20261206, A.nrows)
    y = DeviceArray(A.nrows)
    K.matrix_vector_product(dA, DeviceArray.from_numpy(x), y)
    ref = O.spmv(A, x)
    err = np.abs(y.numpy() - ref) / (np.abs(A.val).max() * np.abs(x).max())
    assert err.max() <= 1e-12


def duvi_case(mode, n=8, seed=9):
    # mode 1: the homogeneous matrix (u8 indices); heterogeneous at 8^3 has < 65,536 distinct
    # values, so mode 0 stores u16 indices
    dens = None if mode == 1 else 10 ** (4 * O.rng_uniform(seed, n ** 3) - 2)
    A = O.cantilever(n, n, n, dens=dens)
    d = duvi_encode(A, 1e-10, 0 if mode == 1 else mode)
    M = O.duvi_block_to_csr(O.duvi_block_encode(A, 1e-10, 0 if mode == 1 else mode), A.ncols)
    return A, d, M


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_duvi_fused_residual_jacobi_check(gpu, mode):
    """The fused epilogues of the DuVi row kernel against the oracle on the matrix the encoding
    stores: r = b - A x (DuViGeometricMultigrid.cl:220-254), y = x + pre (b - A x) (:362-402) and
    the residual_with_check flag (:275-317) on a passing, a failing and a NaN case."""
    A, d, M = duvi_case(mode)
    Dd = DeviceDuvi(d)
    rng = np.random.default_rng(40 + mode)
    x, b = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    pre = 0.6 * O.jacobi_preconditioner(M)
    dx, db, dpre = DeviceArray.from_numpy(x), DeviceArray.from_numpy(b), DeviceArray.from_numpy(pre)
    r, y = DeviceArray(A.nrows), DeviceArray(A.nrows)
    Dd.residual(db, dx, r)
    assert close(r.numpy(), O.residual(M, b, x))
    Dd.jacobi(dpre, db, dx, y)
    assert close(y.numpy(), O.jacobi(M, pre, b, x))
    rr = O.residual(M, b, x)
    assert Dd.residual_with_check(db, dx, r, 10 * np.abs(rr).max()) == 1
    assert Dd.residual_with_check(db, dx, r, 0.5 * np.abs(rr).max()) == 0
    bn = b.copy()
    bn[A.nrows // 2] = np.nan
    assert Dd.residual_with_check(DeviceArray.from_numpy(bn), dx, r, 1e300) == 3


@pytest.mark.parametrize("mode", [0, 1, 4, 8])
@pytest.mark.parametrize("backward", [False, True])
def test_duvi_gs_color_matches_oracle(gpu, mode, backward):
    """A multicolour Gauss-Seidel sweep on the DuVi layout, colour by colour, against the oracle's
    multicolour sweep on the matrix the encoding stores (mode 1: the homogeneous matrix, u8
    indices)."""
    A, d, M = duvi_case(mode)
    Dd = DeviceDuvi(d)
    ncol, color = O.block_coloring(M)
    rng = np.random.default_rng(50 + mode)
    b, x0 = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    dx, db = DeviceArray.from_numpy(x0), DeviceArray.from_numpy(b)
    for c in (range(ncol - 1, -1, -1) if backward else range(ncol)):
        nodes = DeviceArray.from_numpy(np.nonzero(color == c)[0].astype(np.int32))
        Dd.gauss_seidel_color(nodes, db, dx, backward)
    ref = O.gs_multicolor(M, ncol, color, b, x0, backward)
    assert close(dx.numpy(), ref, 1e-11)


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_duvi_spmv_f32(gpu, mode):
    A, d, M = duvi_case(mode)
    Dd = DeviceDuvi(d)
    x = np.random.default_rng(60).normal(size=A.nrows)
    y = DeviceArray(A.nrows, np.float32)
    Dd.spmv_f32(DeviceArray.from_numpy(x.astype(np.float32)), y)
    ref = O.spmv(M, x)
    assert np.abs(y.numpy() - ref).max() <= 1e-5 * np.abs(M.val).max() * np.abs(x).max() * 9


def test_duvi_fused_long_rows(gpu):
    """Residual and Jacobi epilogues on rows whose tile exceeds the LDS staging (the one-lane-per-row
    path) and on ragged rows, against the oracle."""
    rng = np.random.default_rng(61)
    rows = [list(range(2000, 9000)), [], [5, 261, 262, 518], list(range(0, 300))] + [[3, 4]] * 4
    rp, ci = [0], []
    for r_ in rows:
        ci += r_
        rp.append(len(ci))
    A = O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), rng.normal(size=len(ci)), len(rows), 10000)
    d = duvi_encode(A, 1e-10, 8)
    M = O.duvi_block_to_csr(O.duvi_block_encode(A, 1e-10, 8), A.ncols)
    Dd = DeviceDuvi(d)
    x, b = rng.normal(size=A.ncols), rng.normal(size=A.nrows)
    pre = rng.uniform(0.1, 1, size=A.nrows)
    r = DeviceArray(A.nrows)
    Dd.residual(DeviceArray.from_numpy(b), DeviceArray.from_numpy(x), r)
    assert close(r.numpy(), b - O.spmv(M, x))
    # Jacobi needs x with nrows entries at the row indices: square the case by reading x[:nrows]
    y = DeviceArray(A.nrows)
    Dd.jacobi(DeviceArray.from_numpy(pre), DeviceArray.from_numpy(b), DeviceArray.from_numpy(x), y)
    assert close(y.numpy(), x[:A.nrows] + pre * (b - O.spmv(M, x)))


def ragged_csr(rows, ncols, rng, mode):
    rp, ci = [0], []
    for r_ in rows:
        ci += r_
        rp.append(len(ci))
    val = np.round(rng.normal(size=len(ci)), 1) if mode == 0 else rng.normal(size=len(ci))
    return O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, len(rows), ncols)


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_duvi_stream_and_wave_forms(gpu, mode):
    """The two DuVi SpMV kernels -- the tile-staged stream form the library picks for matrices of 16
    or more slots per row, and the one-wavefront-per-row form -- on the same inputs, each against
    the oracle: ragged tiles that fit the stream form's LDS staging (blocks split at 256 slots,
    gaps of 255 / 256 / more, empty rows, a partial last tile), a tile of more slots than the
    staging holds and a tile of more blocks than it holds (both decoded row by row), and the 8^3
    cantilever of several tiles."""
    rng = np.random.default_rng(70 + mode)
    ncols = 250000
    fit = [list(range(0, 300)), [], [5, 261, 262, 518, 60000], list(range(1000, 1600, 2))]
    for _ in range(150):     # clustered columns: a few blocks per row, some gaps past 256
        base = int(rng.integers(0, ncols - 3000))
        fit.append((base + np.sort(rng.choice(3000, size=int(rng.integers(0, 90)), replace=False))).tolist())
    many_slots = [list(range(10000, 16000))] + [[1, 2, 3]] * 63
    many_blocks = [list(range(0, 240000, 300))] + [list(range(50, 90))] * 63
    dens = 10 ** (4 * O.rng_uniform(9, 8 ** 3) - 2)
    cases = [ragged_csr(fit, ncols, rng, mode), ragged_csr(many_slots + fit[:70], ncols, rng, mode),
             ragged_csr(fit[:64] + many_blocks, ncols, rng, mode), O.cantilever(8, 8, 8, dens=dens)]
    for A in cases:
        assert len(A.col) >= 16 * A.nrows            # the library picks the stream form
        d = duvi_encode(A, 1e-10, mode)
        Dd = DeviceDuvi(d)
        x = rng.normal(size=A.ncols)
        ref = O.duvi_block_spmv(O.duvi_block_encode(A, 1e-10, mode), x)
        dx = DeviceArray.from_numpy(x)
        for form in (0, 1):
            y = DeviceArray.from_numpy(np.full(A.nrows, np.nan))
            K.time_duvi_spmv(Dd, dx, y, form=form, warmup=0, repeats=1)
            assert close(y.numpy(), ref), (A.nrows, form)
