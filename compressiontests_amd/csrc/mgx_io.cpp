// mgx_io.cpp -- the on-disk format of a multigrid hierarchy (host code, no device kernels).
//
// The reference rebuilds every level from the model's DOK matrices at each Initialize and keeps
// its DuVi encodings only in memory (the DuViMat types, OpenCLDuViGeometricMultigridSolver.cs:16-31,
// filled by FromDokRowMajor, :753-851).  This file format stores a hierarchy once: for every level
// the CSR matrix A (and the R / P towards the next coarser level), optionally the DuVi block
// layout of each of them, and the structured grid of the level; loading skips the assembly,
// the Galerkin products and the DuVi encoding.
//
// Layout (little endian; every offset a multiple of 8):
//   header   "MGXHIER\0", u32 version (1), u32 nlevels, u32 nsections, u32 flags (0)
//   section  u32 kind, i32 level, u64 payload bytes, u64 payload checksum, payload, zero pad to 8
//            kind 1/2/3: CSR A / R / P  -- i32 nrows, i32 ncols, i64 nnz, i32 rowptr[nrows+1],
//                                          i32 col[nnz], f64 val[nnz] (each array padded to 8)
//            kind 4/5/6: DuVi A / R / P -- i32 nrows, i32 vi_bytes, i64 nnz, i64 nblocks, i64 nuniq,
//                                          i32 row_blocks[nrows+1], i32 row_dist[nrows+1],
//                                          i32 block_col[nblocks], u8 dist[nnz], vi[nnz * vi_bytes],
//                                          f64 values[nuniq] (each array padded to 8)
//            kind 7: grid -- i32 nx, ny, nz elements (0 0 0: unstructured)
//   trailer  "MGXHEND\0", u64 checksum of the header and of every section's 24-byte head
// Checksum: four interleaved lanes of FNV-1a's multiply-xor step applied to 8-byte words (the
// tail byte by byte), combined with the length; it detects corrupted, truncated, reordered and
// mislabelled data, not deliberate tampering.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mgx_internal.h"

namespace {

constexpr char kMagic[8] = {'M', 'G', 'X', 'H', 'I', 'E', 'R', 0};
constexpr char kEnd[8] = {'M', 'G', 'X', 'H', 'E', 'N', 'D', 0};
constexpr uint32_t kVersion = 1;
enum : uint32_t { SEC_CSR = 1, SEC_DUVI = 4, SEC_GRID = 7 };   // + 0 (A), 1 (R), 2 (P)
constexpr uint64_t kPrime = 0x100000001b3ULL, kBasis = 0xcbf29ce484222325ULL;

struct Fail {
    int code;
    std::string msg;
};

uint64_t checksum64(const uint8_t *p, size_t n)
{
    uint64_t h[4] = {kBasis, kBasis ^ 1, kBasis ^ 2, kBasis ^ 3};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int k = 0; k < 4; ++k) {
            uint64_t w;
            std::memcpy(&w, p + i + 8 * k, 8);
            h[k] = (h[k] ^ w) * kPrime;
        }
    for (; i < n; ++i) h[0] = (h[0] ^ p[i]) * kPrime;
    uint64_t r = (uint64_t)n * kPrime;
    for (int k = 0; k < 4; ++k) r = (r ^ h[k]) * kPrime;
    return r ^ (r >> 29);
}

// growing byte buffer of one section's payload
struct Buf {
    std::vector<uint8_t> b;
    template <typename T> void put(const T &v) { append(&v, sizeof(T)); }
    void append(const void *p, size_t n)
    {
        const size_t o = b.size();
        b.resize(o + n);
        if (n) std::memcpy(b.data() + o, p, n);
    }
    void array(const void *p, size_t n)
    {
        append(p, n);
        b.resize((b.size() + 7) & ~(size_t)7, 0);
    }
};

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode))
    {
        if (!f) throw Fail{MGX_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path};
    }
    ~File() { if (f) std::fclose(f); }
    void write(const void *p, size_t n)
    {
        if (n && std::fwrite(p, 1, n, f) != n) throw Fail{MGX_ERR_DEVICE, "write failed"};
    }
    void read(void *p, size_t n)
    {
        if (n && std::fread(p, 1, n, f) != n) throw Fail{MGX_ERR_INVALID_ARGUMENT, "hierarchy file truncated"};
    }
    void skip(uint64_t n)
    {
        if (std::fseek(f, (long)n, SEEK_CUR) != 0) throw Fail{MGX_ERR_INVALID_ARGUMENT, "hierarchy file truncated"};
    }
};

struct SecHead {
    uint32_t kind;
    int32_t level;
    uint64_t bytes, sum;
};
static_assert(sizeof(SecHead) == 24, "section head");

struct Header {
    char magic[8];
    uint32_t version, nlevels, nsections, flags;
};
static_assert(sizeof(Header) == 24, "header");

Buf csr_payload(const mgx_csr_host &m)
{
    if (m.nrows < 0 || m.ncols < 0 || !m.rowptr || m.nnz != (int64_t)m.rowptr[m.nrows])
        throw Fail{MGX_ERR_INVALID_ARGUMENT, "invalid CSR level"};
    Buf p;
    p.put(m.nrows);
    p.put(m.ncols);
    p.put((int64_t)m.nnz);
    p.array(m.rowptr, sizeof(int32_t) * ((size_t)m.nrows + 1));
    p.array(m.col, sizeof(int32_t) * (size_t)m.nnz);
    p.array(m.val, sizeof(double) * (size_t)m.nnz);
    return p;
}

Buf duvi_payload(const mgx_duvi_host &d)
{
    Buf p;
    p.put(d.nrows);
    p.put(d.vi_bytes);
    p.put(d.nnz);
    p.put(d.nblocks);
    p.put(d.nuniq);
    p.array(d.row_blocks, sizeof(int32_t) * ((size_t)d.nrows + 1));
    p.array(d.row_dist, sizeof(int32_t) * ((size_t)d.nrows + 1));
    p.array(d.block_col, sizeof(int32_t) * (size_t)d.nblocks);
    p.array(d.dist, (size_t)d.nnz);
    if (d.vi_bytes) p.array(d.vi, (size_t)d.vi_bytes * (size_t)d.nnz);
    p.array(d.values, sizeof(double) * (size_t)d.nuniq);
    return p;
}

// reads the payload of (kind, level) after checking every section head on the way and the
// payload's checksum; the header and trailer are checked too
std::vector<uint8_t> read_section(const char *path, uint32_t kind, int32_t level, Header *hdr_out = nullptr,
                                  bool all = false, std::vector<SecHead> *heads = nullptr)
{
    File f(path, "rb");
    Header h;
    f.read(&h, sizeof h);
    if (std::memcmp(h.magic, kMagic, 8) != 0) throw Fail{MGX_ERR_INVALID_ARGUMENT, "not a hierarchy file"};
    if (h.version != kVersion)
        throw Fail{MGX_ERR_NOT_IMPLEMENTED, "hierarchy file version " + std::to_string(h.version) + " (reads 1)"};
    uint64_t chain = checksum64((const uint8_t *)&h, sizeof h);
    std::vector<uint8_t> found;
    bool have = false;
    for (uint32_t s = 0; s < h.nsections; ++s) {
        SecHead sh;
        f.read(&sh, sizeof sh);
        chain = (chain ^ checksum64((const uint8_t *)&sh, sizeof sh)) * kPrime;
        if (sh.bytes % 8 != 0) throw Fail{MGX_ERR_INVALID_ARGUMENT, "corrupt section head"};
        if (heads) heads->push_back(sh);
        const bool want = all || (sh.kind == kind && sh.level == level && !have);
        if (want) {
            std::vector<uint8_t> p(sh.bytes);
            f.read(p.data(), p.size());
            if (checksum64(p.data(), p.size()) != sh.sum)
                throw Fail{MGX_ERR_INVALID_ARGUMENT, "checksum mismatch in section " + std::to_string(s)};
            if (sh.kind == kind && sh.level == level && !have) { found.swap(p); have = true; }
        } else f.skip(sh.bytes);
    }
    char end[8];
    uint64_t tsum;
    f.read(end, 8);
    f.read(&tsum, 8);
    if (std::memcmp(end, kEnd, 8) != 0 || tsum != chain)
        throw Fail{MGX_ERR_INVALID_ARGUMENT, "hierarchy file trailer mismatch"};
    if (hdr_out) *hdr_out = h;
    if (!have && !all)
        throw Fail{MGX_ERR_INVALID_ARGUMENT, "no section " + std::to_string(kind) + " for level " + std::to_string(level)};
    return found;
}

// sequential reader of one payload
struct Reader {
    const uint8_t *p;
    size_t n, o = 0;
    template <typename T> T get()
    {
        T v;
        need(sizeof(T));
        std::memcpy(&v, p + o, sizeof(T));
        o += sizeof(T);
        return v;
    }
    void array(void *dst, size_t bytes)
    {
        need(bytes);
        if (bytes) std::memcpy(dst, p + o, bytes);
        o = (o + bytes + 7) & ~(size_t)7;
    }
    void need(size_t k)
    {
        if (o + k > n) throw Fail{MGX_ERR_INVALID_ARGUMENT, "section payload too short"};
    }
};

template <class F>
int io_guard(F &&f)
{
    try {
        f();
        return MGX_OK;
    } catch (const Fail &e) {
        mgx::set_error(e.msg);
        return e.code;
    } catch (const mgx::Error &e) {
        mgx::set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        mgx::set_error("host out of memory");
        return MGX_ERR_OUT_OF_MEMORY;
    }
}

void check_status(int rc)
{
    if (rc != MGX_OK) throw Fail{rc, mgx_last_error()};
}

}  // namespace

extern "C" {

int mgx_hierarchy_write(const char *path, int32_t nlevels, const mgx_csr_host *A, const mgx_csr_host *R,
                        const mgx_csr_host *P, const int32_t *grid, int32_t duvi_mode, double duvi_tol)
{
    return io_guard([&] {
        if (!path || nlevels < 1 || !A || (nlevels > 1 && (!R || !P)))
            throw Fail{MGX_ERR_INVALID_ARGUMENT, "invalid hierarchy arguments"};
        if (duvi_mode != -1 && duvi_mode != 0 && duvi_mode != 4 && duvi_mode != 8)
            throw Fail{MGX_ERR_INVALID_ARGUMENT, "duvi_mode must be -1 (none), 0, 4 or 8"};
        std::vector<std::pair<SecHead, Buf>> secs;
        auto add = [&](uint32_t kind, int32_t level, Buf &&b) {
            SecHead h{kind, level, (uint64_t)b.b.size(), checksum64(b.b.data(), b.b.size())};
            secs.emplace_back(h, std::move(b));
        };
        for (int32_t l = 0; l < nlevels; ++l) {
            const mgx_csr_host *m[3] = {&A[l], l + 1 < nlevels ? &R[l] : nullptr, l + 1 < nlevels ? &P[l] : nullptr};
            for (int w = 0; w < 3; ++w) {
                if (!m[w]) continue;
                add(SEC_CSR + w, l, csr_payload(*m[w]));
                if (duvi_mode >= 0) {
                    mgx_duvi_host d;
                    check_status(mgx_duvi_encode(m[w]->nrows, m[w]->rowptr, m[w]->col, m[w]->val, duvi_tol, duvi_mode, &d));
                    Buf b = duvi_payload(d);
                    mgx_duvi_host_free(&d);
                    add(SEC_DUVI + w, l, std::move(b));
                }
            }
            if (grid) {
                Buf g;
                g.array(grid + 3 * l, 3 * sizeof(int32_t));
                add(SEC_GRID, l, std::move(g));
            }
        }
        const std::string tmp = std::string(path) + ".part";
        {
            File f(tmp.c_str(), "wb");
            Header h{};
            std::memcpy(h.magic, kMagic, 8);
            h.version = kVersion;
            h.nlevels = (uint32_t)nlevels;
            h.nsections = (uint32_t)secs.size();
            h.flags = 0;
            f.write(&h, sizeof h);
            uint64_t chain = checksum64((const uint8_t *)&h, sizeof h);
            for (auto &s : secs) {
                f.write(&s.first, sizeof s.first);
                chain = (chain ^ checksum64((const uint8_t *)&s.first, sizeof s.first)) * kPrime;
                f.write(s.second.b.data(), s.second.b.size());
            }
            f.write(kEnd, 8);
            f.write(&chain, 8);
            if (std::fflush(f.f) != 0) throw Fail{MGX_ERR_DEVICE, "write failed"};
        }
        if (std::rename(tmp.c_str(), path) != 0) throw Fail{MGX_ERR_DEVICE, std::string("cannot rename to ") + path};
    });
}

int mgx_hierarchy_info(const char *path, mgx_hierarchy_summary *info)
{
    return io_guard([&] {
        if (!path || !info) throw Fail{MGX_ERR_INVALID_ARGUMENT, "null argument"};
        Header h;
        std::vector<SecHead> heads;
        read_section(path, 0, 0, &h, true, &heads);   // every checksum
        std::memset(info, 0, sizeof *info);
        info->version = (int32_t)h.version;
        info->nlevels = (int32_t)h.nlevels;
        for (const auto &s : heads) {
            if (s.kind == SEC_GRID) info->has_grid = 1;
            if (s.kind >= SEC_DUVI && s.kind < SEC_DUVI + 3) info->has_duvi = 1;
            info->payload_bytes += (int64_t)s.bytes;
        }
    });
}

int mgx_hierarchy_read_csr(const char *path, int32_t level, int32_t which, mgx_csr_owned *out)
{
    return io_guard([&] {
        if (!path || !out || which < 0 || which > 2) throw Fail{MGX_ERR_INVALID_ARGUMENT, "invalid arguments"};
        std::memset(out, 0, sizeof *out);
        const std::vector<uint8_t> p = read_section(path, SEC_CSR + which, level);
        Reader r{p.data(), p.size()};
        const int32_t nrows = r.get<int32_t>(), ncols = r.get<int32_t>();
        const int64_t nnz = r.get<int64_t>();
        if (nrows < 0 || ncols < 0 || nnz < 0) throw Fail{MGX_ERR_INVALID_ARGUMENT, "corrupt CSR section"};
        std::unique_ptr<int32_t[]> rp(new int32_t[(size_t)nrows + 1]), ci(new int32_t[nnz > 0 ? nnz : 1]);
        std::unique_ptr<double[]> v(new double[nnz > 0 ? nnz : 1]);
        r.array(rp.get(), sizeof(int32_t) * ((size_t)nrows + 1));
        r.array(ci.get(), sizeof(int32_t) * (size_t)nnz);
        r.array(v.get(), sizeof(double) * (size_t)nnz);
        out->nrows = nrows;
        out->ncols = ncols;
        out->nnz = nnz;
        out->rowptr = rp.release();
        out->col = ci.release();
        out->val = v.release();
    });
}

void mgx_csr_owned_free(mgx_csr_owned *m)
{
    if (!m) return;
    delete[] m->rowptr;
    delete[] m->col;
    delete[] m->val;
    std::memset(m, 0, sizeof *m);
}

int mgx_hierarchy_read_duvi(const char *path, int32_t level, int32_t which, mgx_duvi_host *out)
{
    return io_guard([&] {
        if (!path || !out || which < 0 || which > 2) throw Fail{MGX_ERR_INVALID_ARGUMENT, "invalid arguments"};
        std::memset(out, 0, sizeof *out);
        const std::vector<uint8_t> p = read_section(path, SEC_DUVI + which, level);
        Reader r{p.data(), p.size()};
        mgx_duvi_host d{};
        d.nrows = r.get<int32_t>();
        d.vi_bytes = r.get<int32_t>();
        d.nnz = r.get<int64_t>();
        d.nblocks = r.get<int64_t>();
        d.nuniq = r.get<int64_t>();
        if (d.nrows < 0 || d.nnz < 0 || d.nblocks < 0 || d.nuniq < 0 ||
            !(d.vi_bytes == 0 || d.vi_bytes == 1 || d.vi_bytes == 2 || d.vi_bytes == 4))
            throw Fail{MGX_ERR_INVALID_ARGUMENT, "corrupt DuVi section"};
        // the same allocation as mgx_duvi_encode, so that mgx_duvi_host_free releases it
        d.row_blocks = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)d.nrows + 1));
        d.row_dist = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)d.nrows + 1));
        d.block_col = (int32_t *)std::malloc(sizeof(int32_t) * (d.nblocks > 0 ? d.nblocks : 1));
        d.dist = (uint8_t *)std::malloc(d.nnz > 0 ? d.nnz : 1);
        d.vi = d.vi_bytes ? std::malloc((size_t)d.vi_bytes * (d.nnz > 0 ? d.nnz : 1)) : nullptr;
        d.values = (double *)std::malloc(sizeof(double) * (d.nuniq > 0 ? d.nuniq : 1));
        *out = d;                                   // owned by the caller from here on, even on a throw below
        if (!d.row_blocks || !d.row_dist || !d.block_col || !d.dist || (d.vi_bytes && !d.vi) || !d.values)
            throw std::bad_alloc();
        r.array(d.row_blocks, sizeof(int32_t) * ((size_t)d.nrows + 1));
        r.array(d.row_dist, sizeof(int32_t) * ((size_t)d.nrows + 1));
        r.array(d.block_col, sizeof(int32_t) * (size_t)d.nblocks);
        r.array(d.dist, (size_t)d.nnz);
        if (d.vi_bytes) r.array(d.vi, (size_t)d.vi_bytes * (size_t)d.nnz);
        r.array(d.values, sizeof(double) * (size_t)d.nuniq);
    });
}

int mgx_hierarchy_read_grid(const char *path, int32_t level, int32_t *nxyz)
{
    return io_guard([&] {
        if (!path || !nxyz) throw Fail{MGX_ERR_INVALID_ARGUMENT, "null argument"};
        const std::vector<uint8_t> p = read_section(path, SEC_GRID, level);
        Reader r{p.data(), p.size()};
        r.array(nxyz, 3 * sizeof(int32_t));
    });
}

int mgx_hierarchy_save(const char *path, int32_t nlevels, const mgx_matrix *A, const mgx_matrix *R,
                       const mgx_matrix *P, int32_t duvi_mode, double duvi_tol)
{
    return io_guard([&] {
        if (!path || nlevels < 1 || !A || (nlevels > 1 && (!R || !P)))
            throw Fail{MGX_ERR_INVALID_ARGUMENT, "invalid hierarchy arguments"};
        struct Host {
            std::vector<int32_t> rp, ci;
            std::vector<double> v;
            mgx_csr_host h{};
        };
        auto fetch = [](mgx_matrix m, Host &o) {
            mgx_csr_view view;
            check_status(mgx_matrix_view(m, &view));
            o.rp.resize((size_t)view.nrows + 1);
            o.ci.resize((size_t)view.nnz);
            o.v.resize((size_t)view.nnz);
            check_status(mgx_matrix_download(m, o.rp.data(), o.ci.data(), o.v.data()));
            o.h = mgx_csr_host{view.nrows, view.ncols, view.nnz, o.rp.data(), o.ci.data(), o.v.data()};
        };
        std::vector<Host> a(nlevels), r(nlevels > 1 ? nlevels - 1 : 0), p(nlevels > 1 ? nlevels - 1 : 0);
        std::vector<mgx_csr_host> ha, hr, hp;
        std::vector<int32_t> grid(3 * (size_t)nlevels, 0);
        for (int32_t l = 0; l < nlevels; ++l) {
            fetch(A[l], a[l]);
            ha.push_back(a[l].h);
            check_status(mgx_matrix_grid(A[l], &grid[3 * (size_t)l]));
            if (l + 1 < nlevels) {
                fetch(R[l], r[l]);
                fetch(P[l], p[l]);
                hr.push_back(r[l].h);
                hp.push_back(p[l].h);
            }
        }
        check_status(mgx_hierarchy_write(path, nlevels, ha.data(), hr.data(), hp.data(), grid.data(), duvi_mode,
                                         duvi_tol));
    });
}

int mgx_hierarchy_load(const char *path, int32_t max_levels, int32_t *nlevels, mgx_matrix *A, mgx_matrix *R,
                       mgx_matrix *P)
{
    return io_guard([&] {
        if (!path || !nlevels || !A) throw Fail{MGX_ERR_INVALID_ARGUMENT, "null argument"};
        mgx_hierarchy_summary info;
        check_status(mgx_hierarchy_info(path, &info));
        if (info.nlevels > max_levels) throw Fail{MGX_ERR_INVALID_ARGUMENT, "more levels than max_levels"};
        if (info.nlevels > 1 && (!R || !P)) throw Fail{MGX_ERR_INVALID_ARGUMENT, "R and P arrays needed"};
        std::vector<mgx_matrix> made;
        try {
            auto load = [&](int32_t l, int32_t which) {
                mgx_csr_owned m;
                check_status(mgx_hierarchy_read_csr(path, l, which, &m));
                mgx_matrix out = nullptr;
                const int rc = mgx_matrix_from_host(m.nrows, m.ncols, m.rowptr, m.col, m.val, &out);
                mgx_csr_owned_free(&m);
                check_status(rc);
                made.push_back(out);
                return out;
            };
            for (int32_t l = 0; l < info.nlevels; ++l) {
                A[l] = load(l, 0);
                if (info.has_grid) {
                    int32_t g[3];
                    check_status(mgx_hierarchy_read_grid(path, l, g));
                    check_status(mgx_matrix_set_grid(A[l], g));
                }
                if (l + 1 < info.nlevels) { R[l] = load(l, 1); P[l] = load(l, 2); }
            }
        } catch (...) {
            for (mgx_matrix m : made) mgx_matrix_destroy(m);
            throw;
        }
        *nlevels = info.nlevels;
    });
}

}  // extern "C"
