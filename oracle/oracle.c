/*
 * oracle.c -- CPU restatement (FP64, plain C) of libuipc's per-Newton linear
 * solve.  TEST INFRASTRUCTURE ONLY: it is the checker that tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg call; the product
 * library never links or loads it.
 *
 * Reference paths are relative to src/backends/cuda/ of the reference tree.
 *
 * Pinning (what checks each function beyond re-typing it):
 *   or_convert_sym  structure vs the closed-form grid counts (16^3: U = 26,416;
 *                   64^3: U = 1,798,336) and vs numpy's unique/lexsort brute force;
 *                   values vs a dense accumulation (tests/test_oracle.py).
 *   or_spmv_sym     vs a dense U + U^T - D expansion; symmetry x^T A y = y^T A x.
 *   or_dot/or_norm  the reference's KATs (apps/tests/backends/cuda/cuda_tool_linear_system.cu).
 *   or_diag3/12     closed form: (D^-1 r) re-multiplied by D gives r.
 *   or_mas_*        hierarchy: every level's tables equal a brute force from the
 *                   definition (in-bank connected components of the mesh edges
 *                   mapped to the level below, numbered by lowest member), on
 *                   geometric, random-disconnected and partly unpartitioned
 *                   partitions up to the 6-level cap; that check fails on the
 *                   OR_MUTANT_* builds below (planted mistakes in the closure,
 *                   the neighbour compaction and the coarse mask, compiled only
 *                   into test libraries).  Invariants (one ancestor per level,
 *                   bank-aligned levels, level count), scatter vs dense R_L A R_L^T per level,
 *                   inverse C C^-1 = I, apply vs dense sum_L P_L^T C_L^-1 P_L r,
 *                   one-level apply = 16-node block Jacobi, 1-vertex clusters = 3x3 Jacobi.
 *   or_pcg          exact convergence in <= n steps on small SPD systems, the
 *                   true residual ||b - A x|| against the recurrence, and
 *                   x vs a dense LU solve (numpy) on small systems.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BANKSIZE 16
#define SYM_BLOCK_COUNT (BANKSIZE * (BANKSIZE + 1) / 2)
#define MAX_LEVELS 6

static inline int sym_index(int row, int col) /* mas_preconditioner_engine.cu:35-38 */
{
    return BANKSIZE * row - row * (row + 1) / 2 + col;
}
static inline int bank_align(int n) /* mas_preconditioner_engine.cu:41-44 */
{
    return (n + BANKSIZE - 1) / BANKSIZE * BANKSIZE;
}
static inline int popc32(uint32_t x) { return __builtin_popcount(x); }
static inline int ffs32(uint32_t x) { return __builtin_ffs((int)x); }
static inline uint32_t lanemask_lt(int lane) { return (1u << lane) - 1u; }

/* ===================================================================== convert */

int or_key_bits(uint64_t max_key)
{
    int bits = 0;
    while(max_key > 0)
    {
        ++bits;
        max_key >>= 1;
    }
    return bits > 0 ? bits : 1;
}

/* Stable LSD radix sort of (key, value) pairs on the low `bits` bits, 8-bit
 * digits: the contract of cub::DeviceRadixSort::SortPairs(begin_bit = 0,
 * end_bit = bits) that matrix_converter.inl:1610-1616 relies on. */
static void radix_sort_pairs(uint64_t* keys, int32_t* vals, int64_t m, int bits)
{
    uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1));
    int32_t*  v2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    uint64_t* ks = keys;
    int32_t*  vs = vals;
    uint64_t* kd = k2;
    int32_t*  vd = v2;
    for(int shift = 0; shift < bits; shift += 8)
    {
        int      w    = bits - shift < 8 ? bits - shift : 8;
        uint64_t mask = (1ull << w) - 1ull;
        int64_t  cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for(int64_t i = 0; i < m; ++i)
            cnt[((ks[i] >> shift) & mask) + 1]++;
        for(int d = 0; d < 256; ++d)
            cnt[d + 1] += cnt[d];
        for(int64_t i = 0; i < m; ++i)
        {
            int64_t o = cnt[(ks[i] >> shift) & mask]++;
            kd[o]     = ks[i];
            vd[o]     = vs[i];
        }
        uint64_t* tk = ks;
        ks           = kd;
        kd           = tk;
        int32_t* tv  = vs;
        vs           = vd;
        vd           = tv;
    }
    if(ks != keys)
    {
        memcpy(keys, ks, sizeof(uint64_t) * (size_t)m);
        memcpy(vals, vs, sizeof(int32_t) * (size_t)m);
    }
    free(k2);
    free(v2);
}

int64_t or_convert_sym(int32_t        B,
                       int64_t        T,
                       const int32_t* rows,
                       const int32_t* cols,
                       const double*  vals,
                       int32_t*       out_rows,
                       int32_t*       out_cols,
                       double*        out_vals,
                       int32_t*       row_ptr)
{
    if(B < 0 || T < 0)
        return -1;
    const uint64_t key_rows = (uint64_t)(B > 0 ? B : 1);
    const uint64_t key_cols = key_rows;
    /* 1-3: flag upper & valid, compact (key, source) (matrix_converter.inl:129-176) */
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(T > 0 ? T : 1));
    int32_t*  src  = (int32_t*)malloc(sizeof(int32_t) * (size_t)(T > 0 ? T : 1));
    int64_t   m    = 0;
    for(int64_t i = 0; i < T; ++i)
    {
        if(rows[i] >= 0 && rows[i] <= cols[i])
        {
            keys[m] = (uint64_t)rows[i] * key_cols + (uint64_t)cols[i];
            src[m]  = (int32_t)i;
            ++m;
        }
    }
    /* 4: stable sort by key over bit_width(rows * cols - 1) bits (1610-1616) */
    radix_sort_pairs(keys, src, m, or_key_bits(key_rows * key_cols - 1));
    /* 5-8: unpack, run-length encode, sum each run in sorted (= source) order
     * (379-393, 1627-1698; fast_segmental_reduce.h:33-103) */
    int64_t U = 0;
    for(int64_t k = 0; k < m;)
    {
        int64_t e = k + 1;
        while(e < m && keys[e] == keys[k])
            ++e;
        out_rows[U] = (int32_t)(keys[k] / key_cols);
        out_cols[U] = (int32_t)(keys[k] - (uint64_t)out_rows[U] * key_cols);
        double acc[9];
        for(int q = 0; q < 9; ++q)
            acc[q] = vals[9 * (int64_t)src[k] + q];
        for(int64_t j = k + 1; j < e; ++j)
            for(int q = 0; q < 9; ++q)
                acc[q] = acc[q] + vals[9 * (int64_t)src[j] + q];
        for(int q = 0; q < 9; ++q)
            out_vals[9 * U + q] = acc[q];
        ++U;
        k = e;
    }
    /* BSR offsets: exclusive scan of the per-row run lengths (1041-1096) */
    if(row_ptr)
    {
        int64_t u = 0;
        for(int32_t r = 0; r <= B; ++r)
        {
            while(u < U && out_rows[u] < r)
                ++u;
            row_ptr[r] = (int32_t)u;
        }
    }
    free(keys);
    free(src);
    return U;
}

/* ======================================================================== spmv */

double or_spmv_sym(int32_t        B,
                   int64_t        U,
                   const int32_t* rows,
                   const int32_t* cols,
                   const double*  vals,
                   const double*  x,
                   double*        y)
{
    memset(y, 0, sizeof(double) * 3 * (size_t)B);
    for(int64_t t = 0; t < U; ++t)
    {
        int           i = rows[t], j = cols[t];
        const double* A = vals + 9 * t;
        for(int r = 0; r < 3; ++r)
            for(int c = 0; c < 3; ++c)
                y[3 * i + r] += A[r + 3 * c] * x[3 * j + c];
        if(i != j)
            for(int r = 0; r < 3; ++r)
                for(int c = 0; c < 3; ++c)
                    y[3 * j + c] += A[r + 3 * c] * x[3 * i + r];
    }
    double d = 0.0;
    for(int64_t k = 0; k < 3 * (int64_t)B; ++k)
        d += x[k] * y[k];
    return d;
}

double or_dot(const double* x, const double* y, int64_t n)
{
    double s = 0.0;
    for(int64_t i = 0; i < n; ++i)
        s += x[i] * y[i];
    return s;
}

/* ScaledSquareSum reduction (cuda_tool/linear_reduction.h:14-40), left fold */
double or_norm(const double* x, int64_t n)
{
    double scale = 0.0, ssq = 0.0;
    for(int64_t i = 0; i < n; ++i)
    {
        double v = fabs(x[i]);
        if(v == 0.0)
            continue;
        if(isnan(v) || isnan(scale))
        {
            scale = scale + v;
            ssq   = ssq + 1.0;
            continue;
        }
        if(scale < v)
        {
            double ratio = scale / v;
            ssq          = 1.0 + ssq * ratio * ratio;
            scale        = v;
        }
        else if(v < scale)
        {
            double ratio = v / scale;
            ssq          = ssq + ratio * ratio;
        }
        else
            ssq = ssq + 1.0;
    }
    if(scale == 0.0)
        return 0.0;
    return scale * sqrt(ssq);
}

/* ======================================================================== diag */

static int inv3(const double m[9], double out[9])
{
    double c00 = m[4] * m[8] - m[7] * m[5];
    double c01 = m[7] * m[2] - m[1] * m[8];
    double c02 = m[1] * m[5] - m[4] * m[2];
    double det = m[0] * c00 + m[3] * c01 + m[6] * c02;
    /* a singular block gives non-finite entries, as the reference's analytic
     * 3x3 inverse does (cuda_tool/eigen/analytic_inverse.h:76-92, no check on
     * the determinant): the PCG's PSD guard then sees a non-finite z */
    double id = 1.0 / det;
    out[0]    = c00 * id;
    out[1]    = c01 * id;
    out[2]    = c02 * id;
    out[3]    = (m[6] * m[5] - m[3] * m[8]) * id;
    out[4]    = (m[0] * m[8] - m[6] * m[2]) * id;
    out[5]    = (m[3] * m[2] - m[0] * m[5]) * id;
    out[6]    = (m[3] * m[7] - m[6] * m[4]) * id;
    out[7]    = (m[6] * m[1] - m[0] * m[7]) * id;
    out[8]    = (m[0] * m[4] - m[3] * m[1]) * id;
    return det == 0.0 ? -1 : 0;
}

int or_diag3_assemble(int64_t        U,
                      const int32_t* rows,
                      const int32_t* cols,
                      const double*  vals,
                      int32_t        blk_off,
                      int32_t        blk_cnt,
                      double*        diag_inv)
{
    memset(diag_inv, 0, sizeof(double) * 9 * (size_t)blk_cnt);
    for(int64_t t = 0; t < U; ++t)
    {
        int64_t i = (int64_t)rows[t] - blk_off, j = (int64_t)cols[t] - blk_off;
        if(i < 0 || j < 0 || i >= blk_cnt || j >= blk_cnt || i != j)
            continue;
        inv3(vals + 9 * t, diag_inv + 9 * i);
    }
    return 0;
}

void or_diag3_apply(int32_t blk_cnt, const double* diag_inv, const double* r, double* z)
{
    for(int32_t v = 0; v < blk_cnt; ++v)
    {
        const double* D = diag_inv + 9 * (int64_t)v;
        for(int a = 0; a < 3; ++a)
            z[3 * v + a] = D[a] * r[3 * v] + D[a + 3] * r[3 * v + 1] + D[a + 6] * r[3 * v + 2];
    }
}

/* Gauss-Jordan with partial pivoting, n x n column-major, in place */
static int inv_dense(double* A, int n)
{
    double* I = (double*)calloc((size_t)n * n, sizeof(double));
    for(int i = 0; i < n; ++i)
        I[i + n * i] = 1.0;
    for(int c = 0; c < n; ++c)
    {
        int    p  = c;
        double pv = fabs(A[c + n * c]);
        for(int r = c + 1; r < n; ++r)
            if(fabs(A[r + n * c]) > pv)
            {
                pv = fabs(A[r + n * c]);
                p  = r;
            }
        if(pv == 0.0)
        {
            free(I);
            return -1;
        }
        if(p != c)
            for(int k = 0; k < n; ++k)
            {
                double t     = A[c + n * k];
                A[c + n * k] = A[p + n * k];
                A[p + n * k] = t;
                t            = I[c + n * k];
                I[c + n * k] = I[p + n * k];
                I[p + n * k] = t;
            }
        double d = 1.0 / A[c + n * c];
        for(int k = 0; k < n; ++k)
        {
            A[c + n * k] *= d;
            I[c + n * k] *= d;
        }
        for(int r = 0; r < n; ++r)
            if(r != c)
            {
                double f = A[r + n * c];
                if(f == 0.0)
                    continue;
                for(int k = 0; k < n; ++k)
                {
                    A[r + n * k] -= f * A[c + n * k];
                    I[r + n * k] -= f * I[c + n * k];
                }
            }
    }
    memcpy(A, I, sizeof(double) * (size_t)n * n);
    free(I);
    return 0;
}

int or_diag12_assemble(int64_t        U,
                       const int32_t* rows,
                       const int32_t* cols,
                       const double*  vals,
                       int32_t        blk_off,
                       int32_t        bodies,
                       double*        body_inv)
{
    memset(body_inv, 0, sizeof(double) * 144 * (size_t)bodies);
    for(int64_t t = 0; t < U; ++t)
    {
        int64_t i = (int64_t)rows[t] - blk_off, j = (int64_t)cols[t] - blk_off;
        if(i < 0 || j < 0 || i >= 4 * (int64_t)bodies || j >= 4 * (int64_t)bodies)
            continue;
        if(i / 4 != j / 4)
            continue;
        double*       H  = body_inv + 144 * (i / 4);
        int           bi = (int)(i % 4), bj = (int)(j % 4);
        const double* A  = vals + 9 * t;
        for(int r = 0; r < 3; ++r)
            for(int c = 0; c < 3; ++c)
            {
                H[(3 * bi + r) + 12 * (3 * bj + c)] = A[r + 3 * c];
                H[(3 * bj + c) + 12 * (3 * bi + r)] = A[r + 3 * c];
            }
    }
    for(int32_t b = 0; b < bodies; ++b)
        if(inv_dense(body_inv + 144 * (int64_t)b, 12) != 0)
            return -1;
    return 0;
}

void or_diag12_apply(int32_t bodies, const double* body_inv, const double* r, double* z)
{
    for(int32_t b = 0; b < bodies; ++b)
    {
        const double* H = body_inv + 144 * (int64_t)b;
        for(int a = 0; a < 12; ++a)
        {
            double acc = 0.0;
            for(int c = 0; c < 12; ++c)
                acc += H[a + 12 * c] * r[12 * (int64_t)b + c];
            z[12 * (int64_t)b + a] = acc;
        }
    }
}

/* ========================================================================= MAS */

struct or_mas
{
    int32_t   V;          /* m_total_nodes                 */
    int32_t   N;          /* m_total_map_nodes (part map)  */
    int32_t   L;          /* m_level_num                   */
    int32_t   total;      /* m_total_num_clusters (slots)  */
    int32_t   level1_begin;
    int32_t   padded;     /* bank_align(max(N, V))         */
    int32_t*  part_to_real;
    int32_t*  real_to_part;
    int32_t*  going_next;    /* padded * L */
    int32_t   level_size[2 * (MAX_LEVELS + 2)];
    int32_t*  coarse_space;  /* V * L      */
    int32_t*  table;         /* V * MAX_LEVELS (LevelTable) */
    uint32_t* fine_connect;  /* V          */
    int32_t*  prefix_orig;   /* N / 16     */
    int32_t   has_unpart;
    double*   hess;          /* clusters * 136 * 9 */
    double*   inv;           /* clusters * 136 * 9 */
    double*   diag_inv;      /* V * 9: 3x3 fallback for unpartitioned vertices */
};

/* compute_num_levels (mas_preconditioner_engine.cu:1958-1972) */
static int compute_num_levels(int vert_num)
{
    int n_level  = 1;
    int level_sz = bank_align(vert_num);
    while(level_sz > BANKSIZE)
    {
        level_sz /= BANKSIZE;
        n_level++;
        level_sz = bank_align(level_sz);
    }
    n_level++;
    return n_level < MAX_LEVELS ? n_level : MAX_LEVELS;
}

static int cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* transitive closure of a bank's 16 masks: the visit loop of
 * prepare_prefix_sum_L0 / next_level_cluster (mas_preconditioner_engine.cu:131-141,
 * 318-327), reading the neighbours' initial masks */
static uint32_t bank_closure(const uint32_t cache[BANKSIZE], int lane, uint32_t start)
{
#ifdef OR_MUTANT_NO_CLOSURE
    /* planted mistake, compiled only into a test's mutant library: direct
     * neighbours only, no transitive closure */
    (void)cache;
    (void)lane;
    return start;
#endif
    uint32_t conn = start, visited = 1u << lane;
    for(;;)
    {
        uint32_t todo = visited ^ conn;
        if(!todo)
            break;
        int nv = ffs32(todo) - 1;
        visited |= 1u << nv;
        conn |= cache[nv];
    }
    return conn;
}

or_mas* or_mas_create(int32_t        V,
                      const int32_t* part_ids,
                      const int32_t* tets,
                      int64_t        n_tets,
                      const int32_t* tris,
                      int64_t        n_tris,
                      const int32_t* edges,
                      int64_t        n_edges)
{
    if(V < 1)
        return NULL;
    or_mas* m = (or_mas*)calloc(1, sizeof(or_mas));
    m->V      = V;

    /* ---- adjacency (fem_mas_preconditioner.cu:163-203): sorted unique neighbours */
    int64_t* deg = (int64_t*)calloc((size_t)V + 1, sizeof(int64_t));
#define OR_EDGE_PASS(BODY)                                                          \
    do                                                                             \
    {                                                                              \
        for(int64_t t = 0; t < n_tets; ++t)                                        \
            for(int a = 0; a < 4; ++a)                                             \
                for(int b2 = a + 1; b2 < 4; ++b2)                                  \
                {                                                                  \
                    int32_t u = tets[4 * t + a], w = tets[4 * t + b2];             \
                    BODY;                                                          \
                }                                                                  \
        for(int64_t t = 0; t < n_tris; ++t)                                        \
            for(int a = 0; a < 3; ++a)                                             \
                for(int b2 = a + 1; b2 < 3; ++b2)                                  \
                {                                                                  \
                    int32_t u = tris[3 * t + a], w = tris[3 * t + b2];             \
                    BODY;                                                          \
                }                                                                  \
        for(int64_t t = 0; t < n_edges; ++t)                                       \
        {                                                                          \
            int32_t u = edges[2 * t], w = edges[2 * t + 1];                        \
            BODY;                                                                  \
        }                                                                          \
    } while(0)
    OR_EDGE_PASS(if(u != w && u >= 0 && w >= 0 && u < V && w < V) {
        deg[u]++;
        deg[w]++;
    });
    int64_t* start = (int64_t*)calloc((size_t)V + 1, sizeof(int64_t));
    for(int32_t v = 0; v < V; ++v)
        start[v + 1] = start[v] + deg[v];
    uint32_t* raw  = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(start[V] > 0 ? start[V] : 1));
    int64_t*  fill = (int64_t*)calloc((size_t)V, sizeof(int64_t));
    OR_EDGE_PASS(if(u != w && u >= 0 && w >= 0 && u < V && w < V) {
        raw[start[u] + fill[u]++] = (uint32_t)w;
        raw[start[w] + fill[w]++] = (uint32_t)u;
    });
#undef OR_EDGE_PASS
    uint32_t* nbr_start = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)V);
    uint32_t* nbr_num   = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)V);
    uint32_t* nbr_list  = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(start[V] > 0 ? start[V] : 1));
    int64_t   total_nb  = 0;
    for(int32_t v = 0; v < V; ++v)
    {
        uint32_t* s = raw + start[v];
        int64_t   c = fill[v];
        qsort(s, (size_t)c, sizeof(uint32_t), cmp_u32);
        nbr_start[v] = (uint32_t)total_nb;
        int64_t k    = 0;
        for(int64_t q = 0; q < c; ++q)
            if(q == 0 || s[q] != s[q - 1])
                nbr_list[total_nb + k++] = s[q];
        nbr_num[v] = (uint32_t)k;
        total_nb += k;
    }
    free(raw);
    free(fill);
    free(deg);
    free(start);

    /* ---- partition mappings (fem_mas_preconditioner.cu:291-356) */
    int32_t max_part = 0;
    for(int32_t v = 0; v < V; ++v)
        if(part_ids[v] > max_part)
            max_part = part_ids[v];
    int32_t* psize = (int32_t*)calloc((size_t)max_part + 2, sizeof(int32_t));
    for(int32_t v = 0; v < V; ++v)
        if(part_ids[v] >= 0)
        {
            psize[part_ids[v]]++;
            if(psize[part_ids[v]] > BANKSIZE)
            {
                free(psize);
                free(nbr_start);
                free(nbr_num);
                free(nbr_list);
                free(m);
                return NULL;
            }
        }
    int32_t* poff = (int32_t*)calloc((size_t)max_part + 2, sizeof(int32_t));
    int32_t  N    = 0;
    for(int32_t p = 0; p <= max_part; ++p)
    {
        poff[p] = N;
        N += bank_align(psize[p]);
    }
    m->N            = N;
    m->part_to_real = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
    m->real_to_part = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    for(int32_t i = 0; i < N; ++i)
        m->part_to_real[i] = -1;
    int32_t* pfill = (int32_t*)calloc((size_t)max_part + 2, sizeof(int32_t));
    for(int32_t v = 0; v < V; ++v)
    {
        int32_t p = part_ids[v];
        if(p < 0)
        {
            m->real_to_part[v] = -1;
            m->has_unpart      = 1;
            continue;
        }
        int32_t slot          = poff[p] + pfill[p]++;
        m->part_to_real[slot] = v;
        m->real_to_part[v]    = slot;
    }
    free(psize);
    free(poff);
    free(pfill);

    /* ---- init_neighbor (1974-2024) */
    int max_nodes   = N > V ? N : V;
    m->padded       = bank_align(max_nodes);
    m->L            = compute_num_levels(max_nodes);
    const int L     = m->L;
    m->going_next   = (int32_t*)malloc(sizeof(int32_t) * (size_t)m->padded * L);
    m->coarse_space = (int32_t*)malloc(sizeof(int32_t) * (size_t)V * L);
    m->table        = (int32_t*)malloc(sizeof(int32_t) * (size_t)V * MAX_LEVELS);
    m->fine_connect = (uint32_t*)calloc((size_t)max_nodes, sizeof(uint32_t));
    int nbanks      = N / BANKSIZE;
    m->prefix_orig  = (int32_t*)calloc((size_t)(nbanks > 0 ? nbanks : 1), sizeof(int32_t));
    for(int64_t i = 0; i < (int64_t)m->padded * L; ++i)
        m->going_next[i] = -1;
    for(int64_t i = 0; i < (int64_t)V * L; ++i)
        m->coarse_space[i] = -1;
    for(int64_t i = 0; i < (int64_t)V * MAX_LEVELS; ++i)
        m->table[i] = -1;
    memset(m->level_size, 0, sizeof(m->level_size));

    uint32_t* next_mask = (uint32_t*)calloc((size_t)max_nodes, sizeof(uint32_t));
    int32_t*  next_pref = (int32_t*)calloc((size_t)max_nodes / BANKSIZE + 2, sizeof(int32_t));
    int32_t*  next_psum = (int32_t*)calloc((size_t)max_nodes / BANKSIZE + 2, sizeof(int32_t));

    if(N >= 1)
    {
        /* build_connect_mask_L0 (57-106) */
        for(int32_t tid = 0; tid < N; ++tid)
        {
            int32_t idx = m->part_to_real[tid];
            if(idx < 0)
                continue;
            int      bank = tid / BANKSIZE, lane = tid % BANKSIZE;
            uint32_t msk  = 1u << lane;
            uint32_t nk   = 0, st = nbr_start[idx];
            for(uint32_t q = 0; q < nbr_num[idx]; ++q)
            {
                uint32_t nb = nbr_list[st + q];
                int32_t  np = m->real_to_part[nb];
                if(np < 0)
                    continue;
                if(np / BANKSIZE == bank)
                    msk |= 1u << (np % BANKSIZE);
#ifndef OR_MUTANT_DROP_NEIGHBOURS /* planted mistake (test mutant): coarse levels lose every edge */
                else
                    nbr_list[st + nk++] = nb;
#endif
            }
            nbr_num[idx]          = nk;
            m->fine_connect[idx]  = msk;
        }
        /* prepare_prefix_sum_L0 (108-156) */
        for(int b = 0; b < nbanks; ++b)
        {
            uint32_t cache[BANKSIZE];
            for(int l = 0; l < BANKSIZE; ++l)
            {
                int32_t idx = m->part_to_real[b * BANKSIZE + l];
                cache[l]    = idx >= 0 ? m->fine_connect[idx] : 0u;
            }
            int cnt = 0;
            for(int l = 0; l < BANKSIZE; ++l)
            {
                int32_t idx = m->part_to_real[b * BANKSIZE + l];
                if(idx < 0)
                    continue;
                uint32_t c           = bank_closure(cache, l, cache[l]);
                m->fine_connect[idx] = c;
                if(popc32(c & lanemask_lt(l)) == 0)
                    cnt++;
            }
            m->prefix_orig[b] = cnt;
        }
        /* build_level1 (158-212, host 1926-1972) */
        int32_t* psum = (int32_t*)calloc((size_t)nbanks + 1, sizeof(int32_t));
        for(int b = 0; b < nbanks; ++b)
            psum[b + 1] = psum[b] + m->prefix_orig[b];
        int level1_begin = bank_align(bank_align(N) > V ? bank_align(N) : V);
        m->level1_begin  = level1_begin;
        for(int b = 0; b < nbanks; ++b)
        {
            uint32_t elected = 0;
            for(int l = 0; l < BANKSIZE; ++l)
            {
                int32_t idx = m->part_to_real[b * BANKSIZE + l];
                if(idx >= 0 && popc32(m->fine_connect[idx] & lanemask_lt(l)) == 0)
                    elected |= 1u << l;
            }
            for(int l = 0; l < BANKSIZE; ++l)
            {
                int32_t idx = m->part_to_real[b * BANKSIZE + l];
                if(idx < 0)
                    continue;
                int el     = ffs32(m->fine_connect[idx]) - 1;
                int prefix = popc32(elected & lanemask_lt(el)) + psum[b];
                m->coarse_space[idx] = prefix;
                m->going_next[idx]   = prefix + level1_begin;
            }
        }
        m->level_size[2 * 1 + 0] = psum[nbanks];
        m->level_size[2 * 1 + 1] = level1_begin;
        free(psum);

        for(int level = 1; level < L; ++level)
        {
            memset(next_mask, 0, sizeof(uint32_t) * (size_t)max_nodes);
            /* build_connect_mask_Lx (214-294): OR of every fine vertex's
             * same-bank coarse neighbours into its coarse node */
            for(int32_t tid = 0; tid < N; ++tid)
            {
                int32_t idx = m->part_to_real[tid];
                if(idx < 0)
                    continue;
                int32_t cidx = m->coarse_space[(int64_t)(level - 1) * V + idx];
                if(cidx < 0 || cidx >= N)
                    continue;
                uint32_t conn = 0, nk = 0, st = nbr_start[idx];
                for(uint32_t q = 0; q < nbr_num[idx]; ++q)
                {
                    uint32_t nb = nbr_list[st + q];
                    int32_t  cc = m->coarse_space[(int64_t)(level - 1) * V + nb];
                    if(cc < 0 || cc >= N)
                    {
                        nbr_list[st + nk++] = nb;
                        continue;
                    }
                    if(cidx / BANKSIZE == cc / BANKSIZE)
#ifdef OR_MUTANT_COARSE_SELF /* planted mistake (test mutant): the coarse mask names the node itself */
                        conn |= 1u << (cidx % BANKSIZE);
#else
                        conn |= 1u << (cc % BANKSIZE);
#endif
                    else
                        nbr_list[st + nk++] = nb;
                }
                nbr_num[idx] = nk;
                next_mask[cidx] |= conn;
            }
            int32_t Nx    = m->level_size[2 * level + 0];
            int32_t begin = m->level_size[2 * level + 1];
            int     nbx   = (Nx + BANKSIZE - 1) / BANKSIZE;
            /* next_level_cluster (296-338) */
            for(int b = 0; b < nbx; ++b)
            {
                uint32_t cache[BANKSIZE];
                for(int l = 0; l < BANKSIZE; ++l)
                {
                    int i    = b * BANKSIZE + l;
                    cache[l] = i < Nx ? ((1u << l) | next_mask[i]) : 0u;
                }
                int cnt = 0;
                for(int l = 0; l < BANKSIZE; ++l)
                {
                    int i = b * BANKSIZE + l;
                    if(i >= Nx)
                        continue;
                    uint32_t c   = bank_closure(cache, l, cache[l]);
                    next_mask[i] = c;
                    if(popc32(c & lanemask_lt(l)) == 0)
                        cnt++;
                }
                next_pref[b] = cnt;
            }
            /* prefix_sum_Lx (340-388, host 2216-2238) */
            next_psum[0] = 0;
            for(int b = 0; b < nbx; ++b)
                next_psum[b + 1] = next_psum[b] + next_pref[b];
            int next_begin = begin + bank_align(Nx);
            for(int b = 0; b < nbx; ++b)
            {
                uint32_t elected = 0;
                for(int l = 0; l < BANKSIZE; ++l)
                {
                    int i = b * BANKSIZE + l;
                    if(i < Nx && popc32(next_mask[i] & lanemask_lt(l)) == 0)
                        elected |= 1u << l;
                }
                int32_t newv[BANKSIZE];
                for(int l = 0; l < BANKSIZE; ++l)
                {
                    int i = b * BANKSIZE + l;
                    if(i >= Nx)
                        continue;
                    int el  = ffs32(next_mask[i]) - 1;
                    newv[l] = popc32(elected & lanemask_lt(el)) + next_psum[b];
                    m->going_next[begin + i] = newv[l] + next_begin;
                }
                for(int l = 0; l < BANKSIZE; ++l)
                {
                    int i = b * BANKSIZE + l;
                    if(i < Nx)
                        next_mask[i] = (uint32_t)newv[l];
                }
            }
            m->level_size[2 * (level + 1) + 0] = next_psum[nbx];
            m->level_size[2 * (level + 1) + 1] = next_begin;
            /* compute_next_level (390-409) */
            for(int32_t idx = 0; idx < V; ++idx)
            {
                int32_t nx = m->coarse_space[(int64_t)(level - 1) * V + idx];
                if(nx < 0 || nx >= V)
                    m->coarse_space[(int64_t)level * V + idx] = -1;
                else
                    m->coarse_space[(int64_t)level * V + idx] = (int32_t)next_mask[nx];
            }
        }
        m->total = m->level_size[2 * L + 1];
        /* aggregation_kernel (411-466) */
        for(int32_t idx = 0; idx < V; ++idx)
        {
            int32_t first = m->going_next[idx];
            if(first < 0)
                continue;
            int32_t* tab = m->table + (int64_t)MAX_LEVELS * idx;
            tab[0]       = first;
            int32_t cur  = first;
            for(int l = 1; l < L - 1; ++l)
            {
                cur    = m->going_next[cur];
                tab[l] = cur;
            }
        }
    }
    free(next_mask);
    free(next_pref);
    free(next_psum);
    free(nbr_start);
    free(nbr_num);
    free(nbr_list);

    int64_t C   = m->total / BANKSIZE;
    m->hess     = (double*)calloc((size_t)(C > 0 ? C : 1) * SYM_BLOCK_COUNT * 9, sizeof(double));
    m->inv      = (double*)calloc((size_t)(C > 0 ? C : 1) * SYM_BLOCK_COUNT * 9, sizeof(double));
    m->diag_inv = (double*)calloc((size_t)V * 9, sizeof(double));
    return m;
}

void or_mas_destroy(or_mas* m)
{
    if(!m)
        return;
    free(m->part_to_real);
    free(m->real_to_part);
    free(m->going_next);
    free(m->coarse_space);
    free(m->table);
    free(m->fine_connect);
    free(m->prefix_orig);
    free(m->hess);
    free(m->inv);
    free(m->diag_inv);
    free(m);
}

void or_mas_info(const or_mas* m, int32_t info[5])
{
    info[0] = m->V;
    info[1] = m->N;
    info[2] = m->L;
    info[3] = m->total;
    info[4] = m->level1_begin;
}

int64_t or_mas_going_next_size(const or_mas* m)
{
    return (int64_t)m->padded * m->L;
}

void or_mas_tables(const or_mas* m,
                   int32_t*      part_to_real,
                   int32_t*      real_to_part,
                   int32_t*      going_next,
                   int32_t*      level_sizes,
                   int32_t*      coarse_tables,
                   uint32_t*     fine_connect,
                   int32_t*      prefix_orig)
{
    if(part_to_real)
        memcpy(part_to_real, m->part_to_real, sizeof(int32_t) * (size_t)m->N);
    if(real_to_part)
        memcpy(real_to_part, m->real_to_part, sizeof(int32_t) * (size_t)m->V);
    if(going_next)
        memcpy(going_next, m->going_next, sizeof(int32_t) * (size_t)m->padded * m->L);
    if(level_sizes)
        memcpy(level_sizes, m->level_size, sizeof(int32_t) * 2 * (size_t)(m->L + 1));
    if(coarse_tables)
        memcpy(coarse_tables, m->table, sizeof(int32_t) * (size_t)m->V * MAX_LEVELS);
    if(fine_connect)
        memcpy(fine_connect, m->fine_connect, sizeof(uint32_t) * (size_t)m->V);
    if(prefix_orig)
        memcpy(prefix_orig, m->prefix_orig, sizeof(int32_t) * (size_t)(m->N / BANKSIZE));
}

int64_t or_mas_num_clusters(const or_mas* m)
{
    return m->total / BANKSIZE;
}
void or_mas_cluster_hessians(const or_mas* m, double* out)
{
    memcpy(out, m->hess, sizeof(double) * (size_t)or_mas_num_clusters(m) * SYM_BLOCK_COUNT * 9);
}
void or_mas_cluster_inverses(const or_mas* m, double* out)
{
    memcpy(out, m->inv, sizeof(double) * (size_t)or_mas_num_clusters(m) * SYM_BLOCK_COUNT * 9);
}

/* add H (transposed when tr) into the upper block (a, b) of cluster cid */
static void mas_add(or_mas* m, int cid, int a, int b, const double* H, int tr, int sym_twice)
{
    double* M = m->hess + ((int64_t)cid * SYM_BLOCK_COUNT + sym_index(a, b)) * 9;
    for(int r = 0; r < 3; ++r)
        for(int c = 0; c < 3; ++c)
        {
            double v = tr ? H[c + 3 * r] : H[r + 3 * c];
            if(sym_twice)
                v = H[r + 3 * c] + H[c + 3 * r];
            M[r + 3 * c] += v;
        }
}

/* the symmetric Gauss-Jordan sweep of invert_cluster_matrices_sweep_kernel
 * (mas_preconditioner_engine.cu:1077-1175): column-major 48x48 working matrix,
 * zero diagonal entries -> 1, result -A^{-1}; stores the upper blocks, the
 * diagonal block mirrored from its upper triangle */
static void mas_invert_cluster(const double* hess, double* inv, int fp32)
{
    enum { D = BANKSIZE * 3 };
    static __thread double Ad[D * D];
    static __thread float  Af[D * D];
    for(int col = 0; col < D; ++col)
        for(int row = 0; row < D; ++row)
        {
            int    nr = row / 3, nc = col / 3;
            double v;
            if(nc >= nr)
                v = hess[sym_index(nr, nc) * 9 + (row % 3) + 3 * (col % 3)];
            else
                v = hess[sym_index(nc, nr) * 9 + (col % 3) + 3 * (row % 3)];
            if(fp32)
            {
                float f = (float)v;
                if(row == col && f == 0.0f)
                    f = 1.0f;
                Af[row + D * col] = f;
            }
            else
            {
                if(row == col && v == 0.0)
                    v = 1.0;
                Ad[row + D * col] = v;
            }
        }
    for(int p = 0; p < D; ++p)
    {
        if(fp32)
        {
            float sr[D];
            for(int c = 0; c < D; ++c)
                sr[c] = Af[p + D * c];
            float akk = sr[p];
            for(int col = 0; col < D; ++col)
            {
                float* c = Af + D * col;
                float  f = sr[col];
                float  nv;
                if(col == p)
                {
                    float iv = 1.0f / akk;
                    for(int i = 0; i < D; ++i)
                        c[i] *= iv;
                    nv = -iv;
                }
                else
                {
                    float s = f / akk;
                    for(int i = 0; i < D; ++i)
                        c[i] = fmaf(-sr[i], s, c[i]);
                    nv = s;
                }
                c[p] = nv;
            }
        }
        else
        {
            double sr[D];
            for(int c = 0; c < D; ++c)
                sr[c] = Ad[p + D * c];
            double akk = sr[p];
            for(int col = 0; col < D; ++col)
            {
                double* c = Ad + D * col;
                double  f = sr[col];
                double  nv;
                if(col == p)
                {
                    double iv = 1.0 / akk;
                    for(int i = 0; i < D; ++i)
                        c[i] *= iv;
                    nv = -iv;
                }
                else
                {
                    double s = f / akk;
                    for(int i = 0; i < D; ++i)
                        c[i] = fma(-sr[i], s, c[i]);
                    nv = s;
                }
                c[p] = nv;
            }
        }
    }
    memset(inv, 0, sizeof(double) * SYM_BLOCK_COUNT * 9);
    for(int col = 0; col < D; ++col)
    {
        int nc = col / 3, a = col % 3;
        for(int row = 0; row < D; ++row)
        {
            int    nr = row / 3, b = row % 3;
            double v  = fp32 ? -(double)Af[row + D * col] : -Ad[row + D * col];
            if(nr < nc)
                inv[sym_index(nr, nc) * 9 + b + 3 * a] = v;
            else if(nr == nc)
            {
                int si = sym_index(nc, nc);
                if(b <= a)
                    inv[si * 9 + b + 3 * a] = v;
                if(b < a)
                    inv[si * 9 + a + 3 * b] = v;
            }
        }
    }
}

int or_mas_assemble(or_mas*        m,
                    int64_t        U,
                    const int32_t* rows,
                    const int32_t* cols,
                    const double*  vals,
                    int32_t        blk_off,
                    int            inv_fp32)
{
    int64_t C = or_mas_num_clusters(m);
    memset(m->hess, 0, sizeof(double) * (size_t)(C > 0 ? C : 1) * SYM_BLOCK_COUNT * 9);
    const int L = m->L;
    /* pass 1: scatter_hessian_to_clusters_k1 (684-799) */
    for(int64_t t = 0; t < U; ++t)
    {
        int32_t r = rows[t] - blk_off, c = cols[t] - blk_off;
        if(r < 0 || r >= m->V || c < 0 || c >= m->V)
            continue;
        const double* H  = vals + 9 * t;
        int32_t       vc = m->real_to_part[c], vr = m->real_to_part[r];
        if(vc < 0 || vr < 0)
            continue;
        if(vc / BANKSIZE == vr / BANKSIZE)
        {
            if(vc >= vr)
                mas_add(m, vc / BANKSIZE, vr % BANKSIZE, vc % BANKSIZE, H, 0, 0);
            else
                mas_add(m, vc / BANKSIZE, vc % BANKSIZE, vr % BANKSIZE, H, 1, 0);
            continue;
        }
        for(int level = 1; level <= L - 1; ++level)
        {
            if(level == 1)
            {
                vc = m->going_next[c];
                vr = m->going_next[r];
            }
            else
            {
                vc = m->going_next[vc];
                vr = m->going_next[vr];
            }
            if(vc < 0 || vr < 0)
                break;
            if(vc / BANKSIZE == vr / BANKSIZE)
            {
                if(vc >= vr)
                    mas_add(m, vc / BANKSIZE, vr % BANKSIZE, vc % BANKSIZE, H, 0, vc == vr);
                else
                    mas_add(m, vc / BANKSIZE, vc % BANKSIZE, vr % BANKSIZE, H, 1, 0);
            }
        }
    }
    /* pass 2: scatter_hessian_to_clusters_k2 (801-938) */
    int nbanks = m->N / BANKSIZE;
    for(int b = 0; b < nbanks; ++b)
    {
        double fine[BANKSIZE][BANKSIZE][9];
        for(int lr = 0; lr < BANKSIZE; ++lr)
            for(int lc = 0; lc < BANKSIZE; ++lc)
            {
                const double* M;
                int           tr;
                if(lc >= lr)
                {
                    M  = m->hess + ((int64_t)b * SYM_BLOCK_COUNT + sym_index(lr, lc)) * 9;
                    tr = 0;
                }
                else
                {
                    M  = m->hess + ((int64_t)b * SYM_BLOCK_COUNT + sym_index(lc, lr)) * 9;
                    tr = 1;
                }
                int valid = m->part_to_real[b * BANKSIZE + lr] >= 0
                            && m->part_to_real[b * BANKSIZE + lc] >= 0;
                for(int r = 0; r < 3; ++r)
                    for(int c = 0; c < 3; ++c)
                        fine[lr][lc][r + 3 * c] = valid ? (tr ? M[c + 3 * r] : M[r + 3 * c]) : 0.0;
            }
        if(m->prefix_orig[b] == 1)
        {
            double S[9] = {0};
            for(int lr = 0; lr < BANKSIZE; ++lr)
                for(int lc = 0; lc < BANKSIZE; ++lc)
                    for(int q = 0; q < 9; ++q)
                        S[q] += fine[lr][lc][q];
            int32_t rdx = m->part_to_real[b * BANKSIZE];
            if(rdx < 0)
                continue;
            int32_t nx = m->going_next[rdx];
            for(int level = 1; nx >= 0 && level <= L - 1; ++level)
            {
                mas_add(m, nx / BANKSIZE, nx % BANKSIZE, nx % BANKSIZE, S, 0, 0);
                nx = m->going_next[nx];
            }
        }
        else
        {
            for(int lr = 0; lr < BANKSIZE; ++lr)
                for(int lc = 0; lc < BANKSIZE; ++lc)
                {
                    int32_t rdx = m->part_to_real[b * BANKSIZE + lr];
                    int32_t cdx = m->part_to_real[b * BANKSIZE + lc];
                    if(rdx < 0 || cdx < 0)
                        continue;
                    for(int level = 1; level <= L - 1; ++level)
                    {
                        rdx = m->going_next[rdx];
                        cdx = m->going_next[cdx];
                        if(rdx < 0 || cdx < 0)
                            break;
                        if(rdx / BANKSIZE == cdx / BANKSIZE && cdx >= rdx)
                            mas_add(m, cdx / BANKSIZE, rdx % BANKSIZE, cdx % BANKSIZE, fine[lr][lc], 0, 0);
                    }
                }
        }
    }
    /* invert (2548-2589) */
    for(int64_t k = 0; k < C; ++k)
        mas_invert_cluster(m->hess + k * SYM_BLOCK_COUNT * 9, m->inv + k * SYM_BLOCK_COUNT * 9, inv_fp32);
    /* un-partitioned fallback (fem_mas_preconditioner.cu:24-45) */
    if(m->has_unpart)
        for(int64_t t = 0; t < U; ++t)
        {
            int32_t r = rows[t] - blk_off, c = cols[t] - blk_off;
            if(r < 0 || r >= m->V || r != c || m->real_to_part[r] >= 0)
                continue;
            inv3(vals + 9 * t, m->diag_inv + 9 * (int64_t)r);
        }
    return 0;
}

void or_mas_apply(const or_mas* m, const double* r, double* z, int fp32, int active_levels)
{
    const int L   = m->L;
    int       Lc  = (active_levels > 0 && active_levels <= L) ? active_levels : L;
    int64_t   S   = m->total;
    double*   R   = (double*)calloc((size_t)(S > 0 ? S : 1) * 3, sizeof(double));
    double*   Z   = (double*)calloc((size_t)(S > 0 ? S : 1) * 3, sizeof(double));
    /* restriction: fine slots + pushes to every coarser level (1205-1321, 1603-1857) */
    for(int32_t tid = 0; tid < m->N; ++tid)
    {
        int32_t idx = m->part_to_real[tid];
        if(idx < 0)
            continue;
        for(int c = 0; c < 3; ++c)
        {
            double v         = fp32 ? (double)(float)r[3 * idx + c] : r[3 * idx + c];
            R[3 * tid + c]   = v;
        }
        int32_t cur = idx;
        for(int l = 0; l < L - 1; ++l)
        {
            cur = m->going_next[cur];
            for(int c = 0; c < 3; ++c)
                R[3 * (int64_t)cur + c] += R[3 * tid + c];
        }
    }
    if(fp32)
        for(int64_t k = 0; k < 3 * S; ++k)
            R[k] = (double)(float)R[k];
    /* local solves on every cluster (row-dot, 1603-1700) */
    int64_t C = S / BANKSIZE;
    for(int64_t k = 0; k < C; ++k)
    {
        const double* I = m->inv + k * SYM_BLOCK_COUNT * 9;
        for(int lr = 0; lr < BANKSIZE; ++lr)
            for(int a = 0; a < 3; ++a)
            {
                double acc = 0.0;
                for(int cc = 0; cc < BANKSIZE; ++cc)
                {
                    const double* rv = R + 3 * (k * BANKSIZE + cc);
                    for(int b = 0; b < 3; ++b)
                    {
                        double mv = cc >= lr ? I[sym_index(lr, cc) * 9 + a + 3 * b]
                                             : I[sym_index(cc, lr) * 9 + b + 3 * a];
                        if(fp32)
                            mv = (double)(float)mv;
                        acc += mv * rv[b];
                    }
                }
                Z[3 * (k * BANKSIZE + lr) + a] = fp32 ? (double)(float)acc : acc;
            }
    }
    /* prolongation (collect_final_Z, 1895-1951) */
    for(int32_t idx = 0; idx < m->V; ++idx)
    {
        int32_t rdx = m->real_to_part[idx];
        if(rdx < 0)
        {
            const double* D = m->diag_inv + 9 * (int64_t)idx;
            for(int a = 0; a < 3; ++a)
                z[3 * idx + a] = D[a] * r[3 * idx] + D[a + 3] * r[3 * idx + 1] + D[a + 6] * r[3 * idx + 2];
            continue;
        }
        double acc[3];
        for(int c = 0; c < 3; ++c)
            acc[c] = Z[3 * (int64_t)rdx + c];
        const int32_t* tab = m->table + (int64_t)MAX_LEVELS * idx;
        for(int l = 1; l < Lc; ++l)
            for(int c = 0; c < 3; ++c)
                acc[c] += Z[3 * (int64_t)tab[l - 1] + c];
        for(int c = 0; c < 3; ++c)
            z[3 * idx + c] = fp32 ? (double)(float)acc[c] : acc[c];
    }
    free(R);
    free(Z);
}

/* ========================================================================= PCG */

typedef struct csr_full
{
    int32_t  nrows;
    int64_t* ptr;
    int32_t* col;
    double*  val; /* 9 per block, column-major */
} csr_full;

/* the U + U^T expansion of cpu_sym_spmv (spmv.cu:1235-1279) as block CSR */
static void csr_build(csr_full* A, int32_t B, int64_t U, const int32_t* rows, const int32_t* cols, const double* vals)
{
    A->nrows   = B;
    A->ptr     = (int64_t*)calloc((size_t)B + 1, sizeof(int64_t));
    for(int64_t t = 0; t < U; ++t)
    {
        A->ptr[rows[t] + 1]++;
        if(rows[t] != cols[t])
            A->ptr[cols[t] + 1]++;
    }
    for(int32_t i = 0; i < B; ++i)
        A->ptr[i + 1] += A->ptr[i];
    int64_t  nnz  = A->ptr[B];
    int64_t* fill = (int64_t*)calloc((size_t)B, sizeof(int64_t));
    A->col        = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    A->val        = (double*)malloc(sizeof(double) * 9 * (size_t)(nnz > 0 ? nnz : 1));
    for(int64_t t = 0; t < U; ++t)
    {
        int32_t i = rows[t], j = cols[t];
        int64_t o = A->ptr[i] + fill[i]++;
        A->col[o] = j;
        memcpy(A->val + 9 * o, vals + 9 * t, 9 * sizeof(double));
        if(i != j)
        {
            o         = A->ptr[j] + fill[j]++;
            A->col[o] = i;
            for(int r = 0; r < 3; ++r)
                for(int c = 0; c < 3; ++c)
                    A->val[9 * o + r + 3 * c] = vals[9 * t + c + 3 * r];
        }
    }
    free(fill);
}

static void csr_free(csr_full* A)
{
    free(A->ptr);
    free(A->col);
    free(A->val);
}

static double csr_spmv_dot(const csr_full* A, const double* x, double* y, int threads)
{
    double dot = 0.0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : dot)
    for(int32_t i = 0; i < A->nrows; ++i)
    {
        double s0 = 0, s1 = 0, s2 = 0;
        for(int64_t k = A->ptr[i]; k < A->ptr[i + 1]; ++k)
        {
            const double* v  = A->val + 9 * k;
            const double* xj = x + 3 * (int64_t)A->col[k];
            s0 += v[0] * xj[0] + v[3] * xj[1] + v[6] * xj[2];
            s1 += v[1] * xj[0] + v[4] * xj[1] + v[7] * xj[2];
            s2 += v[2] * xj[0] + v[5] * xj[1] + v[8] * xj[2];
        }
        y[3 * i]     = s0;
        y[3 * i + 1] = s1;
        y[3 * i + 2] = s2;
        dot += x[3 * i] * s0 + x[3 * i + 1] * s1 + x[3 * i + 2] * s2;
    }
    return dot;
}

/* GlobalLinearSystem::Impl::apply_preconditioner (global_linear_system.cu:546-590) */
static void apply_pcs(const or_precond* pcs, int32_t n_pcs, int64_t n, const double* r, double* z, int bypass)
{
    memcpy(z, r, sizeof(double) * (size_t)n); /* identity where nothing covers */
    if(bypass)
        return;
    for(int32_t k = 0; k < n_pcs; ++k)
    {
        const or_precond* p  = &pcs[k];
        int64_t           o  = 3 * (int64_t)p->blk_off;
        switch(p->kind)
        {
            case 1: or_diag3_apply(p->blk_cnt, p->data, r + o, z + o); break;
            case 2: or_diag12_apply(p->blk_cnt / 4, p->data, r + o, z + o); break;
            case 3: or_mas_apply(p->mas, r + o, z + o, p->mas_fp32, p->mas_active_levels); break;
            default: break;
        }
    }
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int or_pcg(int32_t           B,
           int64_t           U,
           const int32_t*    rows,
           const int32_t*    cols,
           const double*     vals,
           const or_precond* pcs,
           int32_t           n_pcs,
           const double*     b,
           double*           x,
           int32_t           criterion,
           double            tol,
           double            max_iter_ratio,
           int32_t           check_interval,
           int32_t           psd_fallback,
           int32_t           threads,
           int64_t           max_iter_cap,
           or_pcg_result*    res)
{
    const int64_t n  = 3 * (int64_t)B;
    int           th = threads > 0 ? threads : 1;
    memset(res, 0, sizeof(*res));
    res->rel_res_true = NAN;
    csr_full A;
    csr_build(&A, B, U, rows, cols, vals);
    double* r  = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* z  = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* p  = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* Ap = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double  t0 = now_s();

    memset(x, 0, sizeof(double) * (size_t)n);
    memcpy(r, b, sizeof(double) * (size_t)n);
    int bypass = 0;
    apply_pcs(pcs, n_pcs, n, r, z, 0);
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = or_dot(r, z, n);
    if(psd_fallback && (rz < 0.0 || !isfinite(rz)))
    {
        int fb = rz < 0.0;
        if(!fb && isfinite(or_norm(r, n)))
            fb = 1;
        if(fb)
        {
            bypass = 1;
            memcpy(z, r, sizeof(double) * (size_t)n);
            memcpy(p, z, sizeof(double) * (size_t)n);
            rz = or_dot(r, z, n);
        }
    }
    res->bypassed   = bypass;
    res->rz0        = rz;
    double bnorm    = sqrt(or_dot(b, b, n));
    double abs_rz0  = fabs(rz);
    int64_t max_iter = (int64_t)(max_iter_ratio * (double)n);
    int     ci       = check_interval > 0 ? check_interval : 1;
    int64_t total    = max_iter > 0 ? max_iter - 1 : 0;
    int64_t conv_at  = -1;
    double  rz_tol   = tol * abs_rz0;

    if(abs_rz0 == 0.0 || !isfinite(rz))
    {
        res->iters     = 0;
        res->converged = abs_rz0 == 0.0;
    }
    else
    {
        for(int64_t it = 1; it <= total; ++it)
        {
            if(max_iter_cap > 0 && it > max_iter_cap)
                break;
            double pAp   = csr_spmv_dot(&A, p, Ap, th);
            double alpha = rz / pAp;
#pragma omp parallel for num_threads(th) schedule(static)
            for(int64_t i = 0; i < n; ++i)
            {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            apply_pcs(pcs, n_pcs, n, r, z, bypass);
            double rz_new = 0.0, rr = 0.0;
#pragma omp parallel for num_threads(th) schedule(static) reduction(+ : rz_new, rr)
            for(int64_t i = 0; i < n; ++i)
            {
                rz_new += r[i] * z[i];
                rr += r[i] * r[i];
            }
            int conv = criterion == 0 ? (fabs(rz_new) <= rz_tol) : (sqrt(rr) <= tol * bnorm);
            double beta = rz_new / rz;
            res->rz_final = rz_new;
            res->rel_res  = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
            if(conv)
            {
                conv_at = it;
                break;
            }
            rz = rz_new;
#pragma omp parallel for num_threads(th) schedule(static)
            for(int64_t i = 0; i < n; ++i)
                p[i] = z[i] + beta * p[i];
        }
        if(conv_at > 0)
        {
            int64_t rep = ((conv_at + ci - 1) / ci) * ci;
            res->iters  = rep < total ? rep : total;
            res->iters_exact = conv_at;
            res->converged   = 1;
        }
        else
        {
            res->iters       = max_iter;
            res->iters_exact = -1;
            res->converged   = 0;
        }
    }
    res->seconds = now_s() - t0;
    /* true residual */
    csr_spmv_dot(&A, x, Ap, th);
    double tr = 0.0;
    for(int64_t i = 0; i < n; ++i)
        tr += (b[i] - Ap[i]) * (b[i] - Ap[i]);
    res->rel_res_true = bnorm > 0 ? sqrt(tr) / bnorm : 0.0;
    csr_free(&A);
    free(r);
    free(z);
    free(p);
    free(Ap);
    return 0;
}
