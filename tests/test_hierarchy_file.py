"""The on-disk hierarchy format (mgx_hierarchy_write / _info / _read_*, include/mgx.h): bit-exact
round trips of CSR levels, DuVi sections and grids; the byte layout of a small file against the
format's definition; and the rejection of corrupted, truncated, mislabelled and future-version
files.  Host code only (no kernel is launched)."""
import os
import struct

import numpy as np
import pytest

from compressiontests_amd import _lib as L
from compressiontests_amd import device as D
from oracle import oracle as O


def galerkin_levels(n, nlev, dens=None):
    lv, m, A = [], n, O.cantilever(n, n, n, dens=dens)
    grid = []
    for _ in range(nlev - 1):
        P = O.p_trilinear(m, m, m)
        lv.append(dict(A=A, R=O.transpose(P), P=P))
        grid.append((m, m, m))
        A = O.galerkin(A, P)
        m //= 2
    lv.append(dict(A=A))
    grid.append((m, m, m))
    return lv, grid


def same_csr(a, b):
    return (a.nrows == b.nrows and a.ncols == b.ncols and np.array_equal(a.rowptr, b.rowptr)
            and np.array_equal(a.col, b.col) and a.val.tobytes() == b.val.tobytes())


@pytest.fixture(scope="module")
def h8(tmp_path_factory):
    dens = 10 ** (4 * O.rng_uniform(77, 8 ** 3) - 2)
    lv, grid = galerkin_levels(8, 3, dens)
    path = str(tmp_path_factory.mktemp("h") / "h8.mgxh")
    D.write_hierarchy(path, lv, grid=grid, duvi_mode=0)
    return path, lv, grid


def test_round_trip_csr_grid_duvi(h8):
    path, lv, grid = h8
    assert not os.path.exists(path + ".part")
    version, nlevels, has_grid, has_duvi, nbytes = D.hierarchy_info(path)
    assert (version, nlevels, has_grid, has_duvi) == (1, 3, True, True)
    assert nbytes > 0
    for l, d in enumerate(lv):
        for which, key in enumerate("ARP"):
            if key not in d:
                continue
            M = D.read_hierarchy_csr(path, l, which)
            assert same_csr(M, d[key]), (l, key)
            dv = D.read_hierarchy_duvi(path, l, which)
            ref = D.duvi_encode(d[key], 1e-10, 0)
            for k in ("row_blocks", "row_dist", "block_col", "dist", "values"):
                assert dv[k].tobytes() == ref[k].tobytes(), (l, key, k)
            assert (dv["vi"] is None) == (ref["vi"] is None)
            if dv["vi"] is not None:
                assert dv["vi"].tobytes() == ref["vi"].tobytes()
            # and it decodes to the stored matrix up to the 1e-10 value merge
            back = O.duvi_block_to_csr(dv, d[key].ncols)
            assert np.array_equal(back.col, d[key].col) and np.abs(back.val - d[key].val).max() < 1e-10
        assert D.read_hierarchy_grid(path, l) == grid[l]


def test_without_duvi_and_missing_sections(tmp_path):
    lv, grid = galerkin_levels(4, 2)
    path = str(tmp_path / "h4.mgxh")
    D.write_hierarchy(path, lv)
    assert D.hierarchy_info(path)[1:4] == (2, False, False)
    assert same_csr(D.read_hierarchy_csr(path, 1, 0), lv[1]["A"])
    with pytest.raises(L.MgxInvalidArgument):
        D.read_hierarchy_duvi(path, 0, 0)
    with pytest.raises(L.MgxInvalidArgument):
        D.read_hierarchy_csr(path, 1, 1)          # the coarsest level has no R
    with pytest.raises(L.MgxInvalidArgument):
        D.read_hierarchy_grid(path, 0)


def fnv_words(data):
    """The format's checksum, from its definition: four lanes of (h ^ w) * FNV prime over the
    8-byte little-endian words of each 32-byte group, the tail byte by byte into lane 0, then
    the length and the lanes folded the same way, and a final xor-shift by 29."""
    prime, basis, mask = 0x100000001B3, 0xCBF29CE484222325, (1 << 64) - 1
    h = [basis, basis ^ 1, basis ^ 2, basis ^ 3]
    n = len(data)
    i = 0
    while i + 32 <= n:
        for k in range(4):
            w = struct.unpack_from("<Q", data, i + 8 * k)[0]
            h[k] = ((h[k] ^ w) * prime) & mask
        i += 32
    while i < n:
        h[0] = ((h[0] ^ data[i]) * prime) & mask
        i += 1
    r = (n * prime) & mask
    for k in range(4):
        r = ((r ^ h[k]) * prime) & mask
    return r ^ (r >> 29)


def test_byte_layout_of_a_small_file(tmp_path):
    """One level, a 3 x 3 CSR (rows {0,1}, {1}, {0,2}), grid (1, 0, 0): the exact bytes."""
    A = O.Csr(np.array([0, 2, 3, 5], np.int32), np.array([0, 1, 1, 0, 2], np.int32),
              np.array([4.0, -1.0, 3.0, -1.0, 5.0]), 3, 3)
    path = str(tmp_path / "tiny.mgxh")
    D.write_hierarchy(path, [dict(A=A)], grid=[(1, 0, 0)])
    raw = open(path, "rb").read()
    assert raw[:8] == b"MGXHIER\0"
    assert struct.unpack_from("<4I", raw, 8) == (1, 1, 2, 0)          # version, levels, sections, flags
    csr = (struct.pack("<iiq", 3, 3, 5) + struct.pack("<4i", 0, 2, 3, 5) + struct.pack("<5i", 0, 1, 1, 0, 2)
           + b"\0" * 4 + struct.pack("<5d", 4.0, -1.0, 3.0, -1.0, 5.0))
    grid = struct.pack("<3i", 1, 0, 0) + b"\0" * 4
    off = 24
    heads = []
    for kind, payload in ((1, csr), (7, grid)):
        head = struct.pack("<IiQQ", kind, 0, len(payload), fnv_words(payload))
        assert raw[off:off + 24] == head
        assert raw[off + 24:off + 24 + len(payload)] == payload
        heads.append(head)
        off += 24 + len(payload)
    chain = fnv_words(raw[:24])
    for head in heads:
        chain = ((chain ^ fnv_words(head)) * 0x100000001B3) & ((1 << 64) - 1)
    assert raw[off:] == b"MGXHEND\0" + struct.pack("<Q", chain)


def test_corruption_is_detected(h8, tmp_path):
    path, lv, grid = h8
    raw = bytearray(open(path, "rb").read())
    bad = tmp_path / "flip.mgxh"
    flip = bytearray(raw)
    flip[len(raw) // 2] ^= 0x10                    # one bit in the middle of some payload
    bad.write_bytes(bytes(flip))
    with pytest.raises(L.MgxInvalidArgument, match="checksum"):
        D.hierarchy_info(str(bad))
    trunc = tmp_path / "trunc.mgxh"
    trunc.write_bytes(bytes(raw[:-9]))
    with pytest.raises(L.MgxInvalidArgument):
        D.hierarchy_info(str(trunc))
    magic = tmp_path / "magic.mgxh"
    magic.write_bytes(b"NOTAHIER" + bytes(raw[8:]))
    with pytest.raises(L.MgxInvalidArgument, match="not a hierarchy"):
        D.read_hierarchy_csr(str(magic), 0, 0)
    ver = tmp_path / "ver.mgxh"
    ver.write_bytes(bytes(raw[:8]) + struct.pack("<I", 2) + bytes(raw[12:]))
    with pytest.raises(L.MgxNotImplemented):
        D.hierarchy_info(str(ver))
    # a relabelled section (level 0 A called level 1 A) breaks the trailer chain
    lab = bytearray(raw)
    struct.pack_into("<i", lab, 24 + 4, 1)
    relab = tmp_path / "relabel.mgxh"
    relab.write_bytes(bytes(lab))
    with pytest.raises(L.MgxInvalidArgument, match="trailer"):
        D.hierarchy_info(str(relab))
    with pytest.raises(L.MgxInvalidArgument):
        D.hierarchy_info(str(tmp_path / "missing.mgxh"))
