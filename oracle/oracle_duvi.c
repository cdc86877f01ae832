/*
 * oracle_duvi.c -- CPU restatement of the two DuVi ("Delta-Unit indices + Value-Indexed
 * values") formats.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates:
 *   DuViCompressedSparseMatrix.cs:168-174  ToleranceComparer (|x-y| < tol, hash of Round(x/2tol))
 *     with .NET Dictionary<double,int> lookup semantics: candidates with the same 32-bit hash,
 *     newest insertion first; the first inserted value is the representative.
 *   OpenCLDuViGeometricMultigridSolver.cs:753-851  GPU block layout encoder (FromDokRowMajor)
 *     with defects B-1 (16-bit switch) and B-2 (>65536 values) implemented as intended:
 *     u16 indices above 256 values, "DU + raw values" above 65536; plus a u32-index mode.
 *   DuViGeometricMultigrid.cl:57-99        block decode SpMV
 *   DuViCompressedSparseMatrix.cs:11-152   Unit byte stream (header, count, varint, deltas)
 *   DuViCompressedSparseMatrix.cs:176-400  dense and CSR constructors
 *   DuViCompressedSparseMatrix.cs:421-441  Multiply
 *   GaussSeidelIterationCsrDuVi.cs:22-50   forward Gauss-Seidel over units
 *   Deviation B-5: a unit is split when it would hold more than 255 deltas (the reference
 *   stores the count in one byte without a guard); identical bytes for <= 255.
 *
 * Pinned by (tests/test_oracle_duvi.py): the hand-derived golden bytes of the reference KAT
 * matrix (CompressedSparseMatrixTests.cs:18-27) for both layouts, y = [11, 0]; decode(encode)
 * round trips on random CSR with empty rows, gaps 255/256/65535/65536 and varint boundaries;
 * dedup against a brute-force restatement of the Dictionary scan.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- .NET hashing */
static int32_t net_double_hash(double d)
{
    uint64_t bits; memcpy(&bits, &d, 8);
    if (((bits - 1) & 0x7FFFFFFFFFFFFFFFULL) >= 0x7FF0000000000000ULL) bits &= 0x7FF0000000000000ULL;
    return (int32_t)((uint32_t)bits ^ (uint32_t)(bits >> 32));
}

/* open-addressing table: hash -> chain of entry indices (newest first) */
typedef struct {
    int64_t cap;
    int32_t *keys;    /* hash value per slot */
    int64_t *head;    /* newest entry with that hash, -1 = empty slot */
    int64_t *next;    /* per entry: previous entry with same hash */
    double  *val;     /* per entry: stored key */
    int64_t  count, entcap;
} dmap;

static void dmap_init(dmap *m, int64_t expect)
{
    m->cap = 1024;
    while (m->cap < 2 * expect + 16) m->cap <<= 1;
    m->keys = (int32_t *)malloc(sizeof(int32_t) * m->cap);
    m->head = (int64_t *)malloc(sizeof(int64_t) * m->cap);
    for (int64_t i = 0; i < m->cap; ++i) m->head[i] = -1;
    m->entcap = expect + 16;
    m->next = (int64_t *)malloc(sizeof(int64_t) * m->entcap);
    m->val = (double *)malloc(sizeof(double) * m->entcap);
    m->count = 0;
}

static void dmap_free(dmap *m) { free(m->keys); free(m->head); free(m->next); free(m->val); }

static int64_t dmap_slot(const dmap *m, int32_t h)
{
    uint64_t s = ((uint32_t)h * 0x9E3779B97F4A7C15ULL) & (m->cap - 1);
    while (m->head[s] != -1 && m->keys[s] != h) s = (s + 1) & (m->cap - 1);
    return (int64_t)s;
}

static int32_t dmap_hash(double v, double tol)
{
    return tol > 0 ? net_double_hash(nearbyint(v / (2 * tol))) : net_double_hash(v);
}

static inline int dmap_eq(double a, double b, double tol)
{
    if (tol > 0) return fabs(a - b) < tol;
    return a == b || (isnan(a) && isnan(b));
}

/* lookup-or-insert; returns the index */
static int64_t dmap_get(dmap *m, double v, double tol)
{
    int32_t h = dmap_hash(v, tol);
    int64_t s = dmap_slot(m, h);
    for (int64_t e = m->head[s]; e != -1; e = m->next[e])
        if (dmap_eq(m->val[e], v, tol)) return e;
    if (m->count == m->entcap) {
        m->entcap *= 2;
        m->next = (int64_t *)realloc(m->next, sizeof(int64_t) * m->entcap);
        m->val = (double *)realloc(m->val, sizeof(double) * m->entcap);
    }
    int64_t e = m->count++;
    m->val[e] = v;
    m->next[e] = m->head[s];
    m->keys[s] = h;
    m->head[s] = e;
    if (2 * m->count > m->cap) {   /* grow the slot table, keeping per-hash chains */
        int64_t oc = m->cap; int32_t *ok = m->keys; int64_t *oh = m->head;
        m->cap <<= 1;
        m->keys = (int32_t *)malloc(sizeof(int32_t) * m->cap);
        m->head = (int64_t *)malloc(sizeof(int64_t) * m->cap);
        for (int64_t i = 0; i < m->cap; ++i) m->head[i] = -1;
        for (int64_t i = 0; i < oc; ++i)
            if (oh[i] != -1) { int64_t ns = dmap_slot(m, ok[i]); m->keys[ns] = ok[i]; m->head[ns] = oh[i]; }
        free(ok); free(oh);
    }
    return e;
}

int64_t orc_dedup(int64_t n, const double *v, double tol, int32_t *idx, double *uniq)
{
    dmap m; dmap_init(&m, n < (1 << 20) ? n : (1 << 20));
    for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)dmap_get(&m, v[i], tol);
    int64_t nu = m.count;
    if (uniq) memcpy(uniq, m.val, sizeof(double) * nu);
    dmap_free(&m);
    return nu;
}

/* ---------------------------------------------------------------- GPU block layout */
int orc_duvi_block_encode(int32_t nrows, const int32_t *rp, const int32_t *ci, const double *v,
                          double tol, int mode, orc_duvi_block *out)
{
    memset(out, 0, sizeof *out);
    int64_t nnz = rp[nrows];
    out->nrows = nrows; out->nnz = nnz;
    out->row_blocks = (int32_t *)malloc(sizeof(int32_t) * (nrows + 1));
    out->row_dist = (int32_t *)malloc(sizeof(int32_t) * (nrows + 1));
    out->dist = (uint8_t *)malloc(nnz > 0 ? nnz : 1);
    int64_t bcap = nnz + 1, nb = 0;
    out->block_col = (int32_t *)malloc(sizeof(int32_t) * bcap);
    int64_t valueIndex = 0, valueBeginIndex = 0;
    for (int32_t i = 0;; ++i) {
        out->row_blocks[i] = (int32_t)nb;
        out->row_dist[i] = (int32_t)valueIndex;
        if (i == nrows) break;
        int32_t from = rp[i], to = rp[i + 1];
        for (int32_t j = from; j < to; ++j) {
            int blockCount = 0;
            out->block_col[nb++] = ci[j];
            for (; j + 1 < to && blockCount < 255; ++j) {
                int32_t diff = ci[j + 1] - ci[j] - 1;
                if (diff < 256) { ++blockCount; ++valueIndex; out->dist[valueIndex] = (uint8_t)diff; }
                else break;
            }
            out->dist[valueBeginIndex] = (uint8_t)blockCount;
            ++valueIndex;
            valueBeginIndex = valueIndex;
        }
    }
    out->nblocks = nb;
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
    double *uniq = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    int64_t nu = orc_dedup(nnz, v, tol, idx, uniq);
    int vib;
    if (mode == 8) vib = 0;
    else if (mode == 4) vib = 4;
    else vib = nu <= 256 ? 1 : (nu <= 65536 ? 2 : 0);
    out->vi_bytes = vib;
    if (vib == 0) {
        out->nuniq = nnz;
        out->values = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
        memcpy(out->values, v, sizeof(double) * nnz);
        out->vi = NULL;
        free(uniq);
    } else {
        out->nuniq = nu;
        out->values = uniq;
        out->vi = malloc((size_t)vib * (nnz > 0 ? nnz : 1));
        for (int64_t k = 0; k < nnz; ++k) {
            if (vib == 1) ((uint8_t *)out->vi)[k] = (uint8_t)idx[k];
            else if (vib == 2) ((uint16_t *)out->vi)[k] = (uint16_t)idx[k];
            else ((uint32_t *)out->vi)[k] = (uint32_t)idx[k];
        }
    }
    free(idx);
    return 0;
}

void orc_duvi_block_free(orc_duvi_block *m)
{
    free(m->row_blocks); free(m->row_dist); free(m->block_col); free(m->dist); free(m->vi); free(m->values);
    memset(m, 0, sizeof *m);
}

static inline double block_value(const orc_duvi_block *m, int64_t d)
{
    switch (m->vi_bytes) {
    case 1: return m->values[((const uint8_t *)m->vi)[d]];
    case 2: return m->values[((const uint16_t *)m->vi)[d]];
    case 4: return m->values[((const uint32_t *)m->vi)[d]];
    default: return m->values[d];
    }
}

void orc_duvi_block_spmv(const orc_duvi_block *m, const double *x, double *y, int clear)
{
    for (int32_t r = 0; r < m->nrows; ++r) {
        double acc = clear ? 0.0 : y[r];
        int64_t d = m->row_dist[r];
        for (int64_t b = m->row_blocks[r]; b < m->row_blocks[r + 1]; ++b) {
            int64_t col = m->block_col[b];
            int64_t dend = d + m->dist[d] + 1;
            for (;;) {
                acc += block_value(m, d) * x[col];
                ++d;
                if (d == dend) break;
                col += m->dist[d] + 1;
            }
        }
        y[r] = acc;
    }
}

/* The matrix the block layout stores, as CSR (the decode of DuVi2Mat,
   OpenCLDuViGeometricMultigridSolver.cs:394-415, with the column rule of
   DuViGeometricMultigrid.cl:84-98): rp[nrows+1], ci[nnz], v[nnz] with the decoded (merged)
   values in slot order. */
void orc_duvi_block_to_csr(const orc_duvi_block *m, int32_t *rp, int32_t *ci, double *v)
{
    for (int32_t r = 0; r <= m->nrows; ++r) rp[r] = m->row_dist[r];
    for (int32_t r = 0; r < m->nrows; ++r) {
        int64_t d = m->row_dist[r];
        for (int64_t b = m->row_blocks[r]; b < m->row_blocks[r + 1]; ++b) {
            int64_t col = m->block_col[b];
            int64_t dend = d + m->dist[d] + 1;
            for (;;) {
                ci[d] = (int32_t)col;
                v[d] = block_value(m, d);
                ++d;
                if (d == dend) break;
                col += m->dist[d] + 1;
            }
        }
    }
}

/* ---------------------------------------------------------------- CPU Unit stream */
enum { DB_ONE = 0, DB_TWO = 1, DB_FOUR = 2 };
#define LOW2 (1 << 7)
#define LOW3 ((1 << 14) + LOW2)
#define LOW4 ((1 << 21) + LOW3)
#define LOW5 ((1 << 29) + LOW4)

typedef struct {
    int64_t nunits, ucap, nbytes, bcap, nvals, vcap;
    int64_t *off; uint8_t *bytes; int32_t *vi;
    dmap map;
    double tol;
    int err;
} ubuild;

static void ub_init(ubuild *u, double tol, int64_t expect)
{
    memset(u, 0, sizeof *u);
    u->ucap = 1024; u->off = (int64_t *)malloc(sizeof(int64_t) * (u->ucap + 1)); u->off[0] = 0;
    u->bcap = 4096; u->bytes = (uint8_t *)malloc(u->bcap);
    u->vcap = 1024; u->vi = (int32_t *)malloc(sizeof(int32_t) * u->vcap);
    u->tol = tol;
    dmap_init(&u->map, expect);
}

static void ub_value(ubuild *u, double v)
{
    if (u->nvals == u->vcap) { u->vcap *= 2; u->vi = (int32_t *)realloc(u->vi, sizeof(int32_t) * u->vcap); }
    u->vi[u->nvals++] = (int32_t)dmap_get(&u->map, v, u->tol);
}

static void ub_byte(ubuild *u, uint8_t b)
{
    if (u->nbytes == u->bcap) { u->bcap *= 2; u->bytes = (uint8_t *)realloc(u->bytes, u->bcap); }
    u->bytes[u->nbytes++] = b;
}

/* Unit ctor, DuViCompressedSparseMatrix.cs:13-88 */
static void ub_unit(ubuild *u, int newRow, int db, int32_t next, const int32_t *deltas, int nd)
{
    if (next >= LOW5) { u->err = 1; return; }
    ub_byte(u, (uint8_t)((newRow ? 1 : 0) | (db << 1)));
    ub_byte(u, (uint8_t)nd);
    if (next >= LOW4) {
        next -= LOW4;
        ub_byte(u, (uint8_t)((next >> 24) | 128 | 64 | 32)); ub_byte(u, (uint8_t)(next >> 16));
        ub_byte(u, (uint8_t)(next >> 8)); ub_byte(u, (uint8_t)next);
    } else if (next >= LOW3) {
        next -= LOW3;
        ub_byte(u, (uint8_t)((next >> 16) | 128 | 64)); ub_byte(u, (uint8_t)(next >> 8)); ub_byte(u, (uint8_t)next);
    } else if (next >= LOW2) {
        next -= LOW2;
        ub_byte(u, (uint8_t)((next >> 8) | 128)); ub_byte(u, (uint8_t)next);
    } else ub_byte(u, (uint8_t)next);
    for (int i = 0; i < nd; ++i) {
        int32_t d = deltas[i];
        ub_byte(u, (uint8_t)d);
        if (db >= DB_TWO) ub_byte(u, (uint8_t)(d >> 8));
        if (db == DB_FOUR) { ub_byte(u, (uint8_t)(d >> 16)); ub_byte(u, (uint8_t)(d >> 24)); }
    }
    if (u->nunits == u->ucap) { u->ucap *= 2; u->off = (int64_t *)realloc(u->off, sizeof(int64_t) * (u->ucap + 1)); }
    u->off[++u->nunits] = u->nbytes;
}

/* per-row delta state machine shared by both constructors (:232-272, :350-391) */
typedef struct { int db, newRow; int32_t deltaNext; int32_t *deltas; int nd, cap; } rowstate;

static void rs_push(rowstate *s, int32_t d)
{
    if (s->nd == s->cap) { s->cap = s->cap ? 2 * s->cap : 64; s->deltas = (int32_t *)realloc(s->deltas, sizeof(int32_t) * s->cap); }
    s->deltas[s->nd++] = d;
}

static void rs_delta(ubuild *u, rowstate *s, int32_t delta)
{
    int cls = delta < 256 ? DB_ONE : (delta < 65536 ? DB_TWO : DB_FOUR);
    if (cls != s->db) {
        if (s->nd > 0) {   /* func(): close the unit, the delta opens the next one */
            ub_unit(u, s->newRow, s->db, s->deltaNext, s->deltas, s->nd);
            s->newRow = 0; s->deltaNext = delta; s->nd = 0;
        } else rs_push(s, delta);
        s->db = cls;
    } else if (s->nd == 255) {   /* deviation B-5: split instead of overflowing the count byte */
        ub_unit(u, s->newRow, s->db, s->deltaNext, s->deltas, s->nd);
        s->newRow = 0; s->deltaNext = delta; s->nd = 0;
    } else rs_push(s, delta);
}

static void ub_finish(ubuild *u, int32_t nrows, int32_t ncols, orc_duvi_units *out)
{
    out->nrows = nrows; out->ncols = ncols;
    out->nunits = u->nunits; out->nbytes = u->nbytes; out->nvals = u->nvals;
    out->unit_off = u->off; out->bytes = u->bytes; out->vi = u->vi;
    out->nuniq = u->map.count;
    out->uniq = (double *)malloc(sizeof(double) * (u->map.count ? u->map.count : 1));
    memcpy(out->uniq, u->map.val, sizeof(double) * u->map.count);
    dmap_free(&u->map);
}

int orc_duvi_units_from_csr(int32_t nrows, int32_t ncols, const int32_t *rp, const int32_t *ci,
                            const double *v, double tol, orc_duvi_units *out)
{
    ubuild u; ub_init(&u, tol, rp[nrows]);
    rowstate s = {0};
    for (int32_t y = 0; y < nrows; ++y) {
        int32_t ri = rp[y];
        if (rp[y + 1] - ri == 0) {                       /* :299-317 */
            ub_value(&u, 0.0);
            ub_unit(&u, 1, DB_ONE, 0, NULL, 0);
            continue;
        }
        s.db = DB_ONE; s.newRow = 1; s.nd = 0;
        int32_t xOld = ci[ri];
        s.deltaNext = xOld;
        ub_value(&u, v[ri]);
        for (int32_t k = ri + 1; k < rp[y + 1]; ++k) {
            ub_value(&u, v[k]);
            int32_t x = ci[k];
            rs_delta(&u, &s, x - xOld);
            xOld = x;
        }
        ub_unit(&u, s.newRow, s.db, s.deltaNext, s.deltas, s.nd);
    }
    free(s.deltas);
    ub_finish(&u, nrows, ncols, out);
    return u.err ? -1 : 0;
}

int orc_duvi_units_from_dense(int32_t nrows, int32_t ncols, const double *dense, double tol,
                              orc_duvi_units *out)
{
    ubuild u; ub_init(&u, tol, 1024);
    rowstate s = {0};
    int64_t indexIn = 0;
    for (int32_t y = 0; y < nrows; ++y) {
        s.db = DB_ONE; s.newRow = 1; s.nd = 0; s.deltaNext = -1;
        int32_t xOld = -1, x;
        for (x = 0; x < ncols; ++x) {                     /* :193-213 */
            double v = dense[indexIn++];
            if (fabs(v) > tol) { ub_value(&u, v); xOld = x; s.deltaNext = x; ++x; break; }
        }
        for (; x < ncols; ++x) {
            double v = dense[indexIn++];
            if (fabs(v) > tol) { ub_value(&u, v); rs_delta(&u, &s, x - xOld); xOld = x; }
        }
        if (s.deltaNext != -1) ub_unit(&u, s.newRow, s.db, s.deltaNext, s.deltas, s.nd);
    }
    free(s.deltas);
    ub_finish(&u, nrows, ncols, out);
    return u.err ? -1 : 0;
}

void orc_duvi_units_free(orc_duvi_units *m)
{
    free(m->unit_off); free(m->bytes); free(m->vi); free(m->uniq);
    memset(m, 0, sizeof *m);
}

/* ExtractVariableInteger, :108-119 */
static int32_t varint(const uint8_t *d, int *len)
{
    if ((d[0] & 128) == 0) { *len = 1; return d[0]; }
    if ((d[0] & 64) == 0) { *len = 2; return LOW2 + (((d[0] & 63) << 8) | d[1]); }
    if ((d[0] & 32) == 0) { *len = 3; return LOW3 + (((d[0] & 31) << 16) | (d[1] << 8) | d[2]); }
    *len = 4;
    return LOW4 + (((d[0] & 31) << 24) | (d[1] << 16) | (d[2] << 8) | d[3]);
}

static inline int32_t unit_delta(const uint8_t *p, int db, int i)
{
    if (db == DB_ONE) return p[i];
    if (db == DB_TWO) return p[2 * i] | (p[2 * i + 1] << 8);
    return (int32_t)((uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) | ((uint32_t)p[4 * i + 3] << 24));
}

/* Multiply, :421-441 (y rows without a unit keep their previous value) */
void orc_duvi_units_spmv(const orc_duvi_units *m, const double *x, double *y)
{
    int32_t row = -1, col = 0; int64_t vidx = 0;
    for (int64_t k = 0; k < m->nunits; ++k) {
        const uint8_t *d = m->bytes + m->unit_off[k];
        if (d[0] & 1) { ++row; col = 0; y[row] = 0; }
        int total = d[1], len;
        int32_t delta = varint(d + 2, &len);
        int db = d[0] >> 1;
        col += delta;
        y[row] += m->uniq[m->vi[vidx++]] * x[col];
        for (int i = 0; i < total; ++i) {
            col += unit_delta(d + 2 + len, db, i);
            y[row] += m->uniq[m->vi[vidx++]] * x[col];
        }
    }
}

/* GaussSeidelIterationCsrDuVi.Execute, :22-50 */
void orc_duvi_units_gs(const orc_duvi_units *m, const double *b, double *x)
{
    int32_t row = -1, col = 0; int64_t nz = 0; double divisor = 0;
    for (int64_t k = 0; k < m->nunits; ++k) {
        const uint8_t *d = m->bytes + m->unit_off[k];
        int db = d[0] >> 1;
        if (d[0] & 1) {
            if (row != -1) x[row] /= divisor;
            ++row; col = 0;
            x[row] = b[row];
        }
        int len; int32_t delta = varint(d + 2, &len);
        col += delta;
        int total = d[1], i = 0;
        for (;;) {
            double a = m->uniq[m->vi[nz++]];
            if (row != col) x[row] -= a * x[col];
            else divisor = a;
            if (i == total) break;
            col += unit_delta(d + 2 + len, db, i++);
        }
    }
    if (row >= 0) x[row] /= divisor;
}
