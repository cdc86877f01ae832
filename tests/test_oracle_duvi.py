"""Pins the oracle's DuVi encoders/decoders: the reference KAT (golden bytes, y = [11, 0]),
round trips through independent Python decoders on matrices that hit every block / varint /
delta-width boundary, and the .NET Dictionary dedup semantics restated by brute force."""
import json
import math
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "duvi_kat.json")))
LOW2, LOW3, LOW4 = 128, 16384 + 128, (1 << 21) + 16384 + 128


def kat_csr():
    return O.Csr(np.array(GOLD["rowptr"], np.int32), np.array(GOLD["col"], np.int32),
                 np.array(GOLD["val"], np.float64), GOLD["nrows"], GOLD["ncols"])


def kat_x():
    x = np.zeros(GOLD["ncols"])
    for k, v in GOLD["x_nonzeros"].items():
        x[int(k)] = v
    return x


def test_kat_units_from_csr_and_dense():
    A = kat_csr()
    u = O.duvi_units_from_csr(A)
    assert u["bytes"].tolist() == GOLD["units"]["bytes"]
    assert u["unit_off"].tolist() == GOLD["units"]["unit_off"]
    assert u["vi"].tolist() == GOLD["units"]["value_indices"]
    assert u["uniq"].tolist() == GOLD["units"]["unique_values"]
    D = np.zeros((2, 1024))
    for r in range(2):
        for k in range(A.rowptr[r], A.rowptr[r + 1]):
            D[r, A.col[k]] = A.val[k]
    ud = O.duvi_units_from_dense(D)
    assert ud["bytes"].tolist() == GOLD["units"]["bytes"]
    # the reference asserts y == [11, 0] exactly; both encoders (it multiplies mat1 twice)
    for enc in (u, ud):
        assert O.duvi_units_spmv(enc, kat_x()).tolist() == GOLD["y"]


def test_kat_block_layout():
    d = O.duvi_block_encode(kat_csr())
    g = GOLD["block"]
    for k in ("row_blocks", "row_dist", "block_col", "dist", "vi", "values"):
        assert d[k].tolist() == g[k], k
    assert d["vi_bytes"] == g["vi_bytes"]
    assert O.duvi_block_spmv(d, kat_x()).tolist() == GOLD["y"]


# ------------------------------------------------------------------ independent decoders
def decode_units(u):
    """Python restatement of DuViCompressedSparseMatrix.Unit accessors -> (row, col) list."""
    out, row, col = [], -1, 0
    for k in range(len(u["unit_off"]) - 1):
        d = u["bytes"][u["unit_off"][k]:u["unit_off"][k + 1]].tolist()
        if d[0] & 1:
            row += 1
            col = 0
        db, total = d[0] >> 1, d[1]
        b0 = d[2]
        if not b0 & 128:
            nxt, ln = b0, 1
        elif not b0 & 64:
            nxt, ln = LOW2 + (((b0 & 63) << 8) | d[3]), 2
        elif not b0 & 32:
            nxt, ln = LOW3 + (((b0 & 31) << 16) | (d[3] << 8) | d[4]), 3
        else:
            nxt, ln = LOW4 + (((b0 & 31) << 24) | (d[3] << 16) | (d[4] << 8) | d[5]), 4
        col += nxt
        out.append((row, col))
        w = (1, 2, 4)[db]
        base = 2 + ln
        assert len(d) == base + w * total
        for i in range(total):
            col += int.from_bytes(bytes(d[base + w * i: base + w * (i + 1)]), "little")
            out.append((row, col))
    return out


def decode_blocks(d):
    out = []
    for r in range(d["nrows"]):
        p = int(d["row_dist"][r])
        for b in range(d["row_blocks"][r], d["row_blocks"][r + 1]):
            col = int(d["block_col"][b])
            end = p + int(d["dist"][p]) + 1
            while True:
                out.append((r, col, p))
                p += 1
                if p == end:
                    break
                col += int(d["dist"][p]) + 1
        assert p == d["row_dist"][r + 1]
    return out


def boundary_csr():
    """Rows exercising: empty rows, gaps 255/256/65535/65536, > 255 entries with gap 1
    (block split), varint boundaries of the first column, 1/2/4-byte delta switches."""
    rows = [
        [0, 256, 257, 65792, 65793, 131329, 131330],
        [],
        [127], [128], [16511], [16512], [2113663], [2113664],
        list(range(5, 5 + 300)),
        [1, 1 + 256, 1 + 256 + 65536, 1 + 256 + 65536 + 3],
        [],
        [10, 265, 266, 522, 523, 65800, 65801, 65802],
    ]
    ncols = 2113665
    rp, ci = [0], []
    for r in rows:
        ci += r
        rp.append(len(ci))
    rng = np.random.default_rng(11)
    val = rng.choice([1.0, 2.0, -3.5, 4.25], size=len(ci))
    return O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, len(rows), ncols), rows


def test_roundtrip_boundaries():
    A, rows = boundary_csr()
    expect = [(r, c) for r, row in enumerate(rows) for c in row]
    u = O.duvi_units_from_csr(A)
    got = decode_units(u)
    # empty CSR rows carry one (row, 0) placeholder with value 0 (DuViCompressedSparseMatrix.cs:299-317)
    placeholders = [(r, 0) for r, row in enumerate(rows) if not row]
    assert sorted(got) == sorted(expect + placeholders)
    d = O.duvi_block_encode(A)
    blk = decode_blocks(d)
    assert [(r, c) for r, c, _ in blk] == expect
    # the block heads hold counts <= 255
    assert d["dist"].max() <= 255
    x = np.random.default_rng(0).normal(size=A.ncols)
    ref = O.spmv(A, x)
    assert np.allclose(O.duvi_units_spmv(u, x), ref, rtol=1e-13, atol=1e-13)
    assert np.allclose(O.duvi_block_spmv(d, x), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("seed", range(6))
def test_roundtrip_random(seed):
    rng = np.random.default_rng(seed)
    n, m = 60, 70000
    rp, ci = [0], []
    for r in range(n):
        k = rng.integers(0, 40)
        cols = np.unique(rng.choice(m, size=k, replace=False)) if k else np.array([], int)
        if rng.random() < 0.3 and len(cols):
            cols = np.unique(np.concatenate([cols, cols[0] + np.arange(1, 6)]))
            cols = cols[cols < m]
        ci += cols.tolist()
        rp.append(len(ci))
    val = rng.normal(size=len(ci))
    A = O.Csr(np.array(rp, np.int32), np.array(ci, np.int32), val, n, m)
    d = O.duvi_block_encode(A, mode=4)
    blk = decode_blocks(d)
    assert [c for _, c, _ in blk] == ci
    vals = d["values"][d["vi"][[p for _, _, p in blk]]]
    assert np.abs(vals - val).max() < 1e-10
    u = O.duvi_units_from_csr(A)
    got = [rc for rc in decode_units(u)]
    nonempty = [(r, c) for r in range(n) for c in ci[rp[r]:rp[r + 1]]]
    empt = [(r, 0) for r in range(n) if rp[r] == rp[r + 1]]
    assert sorted(got) == sorted(nonempty + empt)


def net_hash(d):
    bits = struct.unpack("<Q", struct.pack("<d", d))[0]
    if ((bits - 1) & 0x7FFFFFFFFFFFFFFF) >= 0x7FF0000000000000:
        bits &= 0x7FF0000000000000
    return ((bits & 0xFFFFFFFF) ^ (bits >> 32)) & 0xFFFFFFFF


def brute_dedup(vals, tol):
    """Dictionary<double,int>(ToleranceComparer) by brute force: same-hash candidates,
    newest first, first with |k - v| < tol (DuViCompressedSparseMatrix.cs:168-174)."""
    keys, hashes, idx = [], [], []
    for v in vals:
        h = net_hash(float(np.rint(v / (2 * tol))))
        hit = None
        for j in range(len(keys) - 1, -1, -1):
            if hashes[j] == h and abs(keys[j] - v) < tol:
                hit = j
                break
        if hit is None:
            keys.append(v); hashes.append(h); hit = len(keys) - 1
        idx.append(hit)
    return idx, keys


def test_dedup_semantics():
    tol = 1e-10
    rng = np.random.default_rng(5)
    base = rng.choice([0.5, -1.25, 3.0, 1e-3], size=200)
    jitter = rng.choice([0.0, 3e-11, -4e-11, 9e-11, 1.5e-10, -2e-10], size=200)
    vals = base + jitter
    # values straddling a rounding bucket edge: within tol but different hashes -> not merged
    edge = (np.rint(7.0) + 0.5) * 2 * tol
    vals = np.concatenate([vals, [edge - 1e-12, edge + 1e-12, 0.0, -0.0]])
    idx, uniq = O.dedup(vals, tol)
    bi, bk = brute_dedup(vals, tol)
    assert idx.tolist() == bi
    assert uniq.tolist() == bk


def test_u16_switch_and_raw_fallback():
    """B-1/B-2 implemented as intended: > 256 values -> u16 indices, > 65536 -> raw values."""
    n = 300
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.zeros(n, np.int32)
    A = O.Csr(rp, ci, np.arange(n, dtype=np.float64) + 1.0, n, 1)
    d = O.duvi_block_encode(A)
    assert d["vi_bytes"] == 2 and d["nuniq"] == n
    assert np.array_equal(d["values"][d["vi"]], A.val)
    n = 70000
    rp = np.arange(n + 1, dtype=np.int32)
    A = O.Csr(rp, np.zeros(n, np.int32), np.arange(n, dtype=np.float64) + 1.0, n, 1)
    d = O.duvi_block_encode(A)
    assert d["vi_bytes"] == 0 and np.array_equal(d["values"], A.val)


def test_units_gs_matches_csr_gs():
    A = O.cantilever(3, 3, 3)
    u = O.duvi_units_from_csr(A, tol=0.0)
    rng = np.random.default_rng(2)
    b, x = rng.normal(size=A.nrows), rng.normal(size=A.nrows)
    assert np.allclose(O.duvi_units_gs(u, b, x), O.gs_forward(A, b, x), rtol=1e-13, atol=1e-13)


def test_cantilever_duvi_sizes():
    """Homogeneous K at 16^3 has few distinct values (u8 indices) and ~3 blocks per row."""
    A = O.cantilever(16, 16, 16)
    d = O.duvi_block_encode(A)
    assert d["vi_bytes"] == 1
    assert 2.5 < d["nblocks"] / A.nrows < 3.5
    x = np.random.default_rng(1).normal(size=A.nrows)
    assert np.allclose(O.duvi_block_spmv(d, x), O.spmv(A, x), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_block_decode_to_csr(mode):
    """The block layout decodes back to the CSR it encodes: exactly (indices and value bits) with
    the exact dedup (tol = 0), and with the 1e-10 merge to the matrix whose product is the DuVi
    SpMV's (the same slot order, so the same bits)."""
    n = 4
    dens = 10 ** (2 * O.rng_uniform(31, n ** 3) - 1)
    A = O.cantilever(n, n, n, dens=dens)
    exact = O.duvi_block_to_csr(O.duvi_block_encode(A, 0.0, mode), A.ncols)
    assert np.array_equal(exact.rowptr, A.rowptr) and np.array_equal(exact.col, A.col)
    assert exact.val.tobytes() == A.val.tobytes()
    d = O.duvi_block_encode(A, 1e-10, mode)
    M = O.duvi_block_to_csr(d, A.ncols)
    assert np.array_equal(M.col, A.col) and np.abs(M.val - A.val).max() < 1e-10
    x = O.rng_normal(32, A.ncols)
    assert O.spmv(M, x).tobytes() == O.duvi_block_spmv(d, x).tobytes()
