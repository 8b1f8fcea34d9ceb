// gm_gen.c: deterministic synthetic inputs for the GrB_mxm workloads.
//
// Every random draw comes from a counter-based generator (splitmix64 applied
// to seed + counter * golden-ratio increment), so a graph is a pure function
// of its parameters and does not depend on thread count or scheduling.
//
//  * RMAT (Chakrabarti/Zhan/Faloutsos): each of ef*2^scale samples descends
//    `scale` quadrant levels with probabilities a, b, c, d = 1-a-b-c
//    (a = (0,0), b = (0,1), c = (1,0), d = (1,1)).  No vertex permutation is
//    applied, so low-index vertices are the hubs.  The graph is symmetrised,
//    self-loops are dropped and duplicates removed; rows come out sorted.
//  * Erdos-Renyi with exactly `per_row` distinct uniform columns per row,
//    values uniform in [0,1).
//
// Output arrays are caller-allocated (numpy on the Python side); the RMAT
// builder returns the number of entries written.

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static inline uint64_t gm_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t gm_rand64(uint64_t seed, uint64_t ctr)
{
    return gm_mix64(seed + (ctr + 1) * 0x9E3779B97F4A7C15ULL);
}

// uniform double in [0,1) with 53 random bits
static inline double gm_rand_unif(uint64_t seed, uint64_t ctr)
{
    return (double) (gm_rand64(seed, ctr) >> 11) * (1.0 / 9007199254740992.0);
}

uint64_t gm_gen_rand64(uint64_t seed, uint64_t ctr) { return gm_rand64(seed, ctr); }
double gm_gen_unif(uint64_t seed, uint64_t ctr) { return gm_rand_unif(seed, ctr); }

static int cmp_i64(const void *x, const void *y)
{
    int64_t a = *(const int64_t *) x, b = *(const int64_t *) y;
    return (a > b) - (a < b);
}

static void sort_i64(int64_t *x, int64_t n)
{
    if (n < 32)
    {
        for (int64_t k = 1; k < n; k++)
        {
            int64_t v = x[k], t = k - 1;
            while (t >= 0 && x[t] > v) { x[t + 1] = x[t]; t--; }
            x[t + 1] = v;
        }
    }
    else
    {
        qsort(x, (size_t) n, sizeof(int64_t), cmp_i64);
    }
}

// One RMAT sample: returns (u,v) packed as u*n+v is not used; out via pointers.
static inline void rmat_edge(int scale, double a, double ab, double abc,
                             uint64_t seed, uint64_t e, int64_t *u, int64_t *v)
{
    int64_t r = 0, c = 0;
    for (int lev = 0; lev < scale; lev++)
    {
        double x = gm_rand_unif(seed, e * 64 + (uint64_t) lev);
        int bu, bv;
        if (x < a)        { bu = 0; bv = 0; }
        else if (x < ab)  { bu = 0; bv = 1; }
        else if (x < abc) { bu = 1; bv = 0; }
        else              { bu = 1; bv = 1; }
        r = 2 * r + bu;
        c = 2 * c + bv;
    }
    *u = r;
    *v = c;
}

// Symmetrised, deduplicated, loop-free RMAT graph in CSR on n vertices (any
// n >= 1): ef*n samples descend scale = ceil(log2 n) levels; a sample with
// an endpoint >= n is dropped, as are self-loops.  For n = 2^scale nothing
// is dropped but loops.  Ap: n+1 entries; Aj: capacity >= 2*ef*n.  Returns nnz.
int64_t gm_gen_rmat_sym_n(int64_t n, int64_t ef, double a, double b, double c,
                          uint64_t seed, int64_t *Ap, int64_t *Aj)
{
    int scale = 0;
    while (((int64_t) 1 << scale) < n) scale++;
    const int64_t ne = ef * n;
    const double ab = a + b, abc = a + b + c;
    int64_t *deg = (int64_t *) calloc((size_t) n + 1, sizeof(int64_t));
    int64_t *fill = (int64_t *) malloc(sizeof(int64_t) * ((size_t) n + 1));
    if (!deg || !fill) { free(deg); free(fill); return -1; }

    // pass 1: out-degree counts (both directions, self-loops dropped)
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; e++)
    {
        int64_t u, v;
        rmat_edge(scale, a, ab, abc, seed, (uint64_t) e, &u, &v);
        if (u == v || u >= n || v >= n) continue;
        #pragma omp atomic
        deg[u]++;
        #pragma omp atomic
        deg[v]++;
    }
    // exclusive scan into Ap (temporary row pointers with duplicates)
    int64_t s = 0;
    for (int64_t i = 0; i < n; i++) { Ap[i] = s; s += deg[i]; }
    Ap[n] = s;
    memcpy(fill, Ap, sizeof(int64_t) * (size_t) (n + 1));

    // pass 2: scatter neighbours
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; e++)
    {
        int64_t u, v, pu, pv;
        rmat_edge(scale, a, ab, abc, seed, (uint64_t) e, &u, &v);
        if (u == v || u >= n || v >= n) continue;
        #pragma omp atomic capture
        pu = fill[u]++;
        #pragma omp atomic capture
        pv = fill[v]++;
        Aj[pu] = v;
        Aj[pv] = u;
    }

    // pass 3: sort + dedup each row; deg[i] = unique count
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++)
    {
        int64_t *row = Aj + Ap[i];
        int64_t len = Ap[i + 1] - Ap[i];
        sort_i64(row, len);
        int64_t w = 0;
        for (int64_t k = 0; k < len; k++)
        {
            if (w == 0 || row[w - 1] != row[k]) row[w++] = row[k];
        }
        deg[i] = w;
    }

    // pass 4: compact rows in place (rows only move left)
    int64_t pos = 0;
    for (int64_t i = 0; i < n; i++)
    {
        int64_t start = Ap[i];
        int64_t len = deg[i];
        if (pos != start) memmove(Aj + pos, Aj + start, sizeof(int64_t) * (size_t) len);
        Ap[i] = pos;
        pos += len;
    }
    Ap[n] = pos;
    free(deg);
    free(fill);
    return pos;
}

int64_t gm_gen_rmat_sym(int scale, int64_t ef, double a, double b, double c,
                        uint64_t seed, int64_t *Ap, int64_t *Aj)
{
    return gm_gen_rmat_sym_n((int64_t) 1 << scale, ef, a, b, c, seed, Ap, Aj);
}

// Strict lower (lower=1) or upper (lower=0) triangle of a CSR matrix with
// sorted rows.  Lp: n+1; Lj: capacity nnz(A).  Returns nnz of the result.
int64_t gm_gen_triangle(int64_t n, const int64_t *Ap, const int64_t *Aj,
                        int lower, int64_t *Lp, int64_t *Lj)
{
    int64_t *cnt = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n + 1));
    if (!cnt) return -1;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
    {
        int64_t k = 0;
        for (int64_t p = Ap[i]; p < Ap[i + 1]; p++)
        {
            int64_t j = Aj[p];
            if (lower ? (j < i) : (j > i)) k++;
        }
        cnt[i] = k;
    }
    int64_t s = 0;
    for (int64_t i = 0; i < n; i++) { Lp[i] = s; s += cnt[i]; }
    Lp[n] = s;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
    {
        int64_t w = Lp[i];
        for (int64_t p = Ap[i]; p < Ap[i + 1]; p++)
        {
            int64_t j = Aj[p];
            if (lower ? (j < i) : (j > i)) Lj[w++] = j;
        }
    }
    free(cnt);
    return s;
}

// Symmetric integer edge weights in [lo, hi]: w(i,j) = w(j,i) is a hash of
// the unordered pair.
void gm_gen_sym_weights_i32(int64_t n, const int64_t *Ap, const int64_t *Aj,
                            uint64_t seed, int32_t lo, int32_t hi, int32_t *Ax)
{
    const uint64_t range = (uint64_t) (hi - lo + 1);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
    {
        for (int64_t p = Ap[i]; p < Ap[i + 1]; p++)
        {
            int64_t j = Aj[p];
            uint64_t u = (uint64_t) (i < j ? i : j), v = (uint64_t) (i < j ? j : i);
            uint64_t h = gm_rand64(seed, u * 0x100000001B3ULL ^ (v << 1));
            Ax[p] = lo + (int32_t) (h % range);
        }
    }
}

// Erdos-Renyi: exactly per_row distinct uniform columns per row, sorted,
// values uniform [0,1).  Ap: n+1; Aj, Ax: n*per_row.
void gm_gen_er(int64_t nrows, int64_t ncols, int64_t per_row, uint64_t seed,
               int64_t *Ap, int64_t *Aj, double *Ax)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i <= nrows; i++) Ap[i] = i * per_row;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; i++)
    {
        int64_t *row = Aj + i * per_row;
        uint64_t ctr = (uint64_t) i << 20;
        int64_t k = 0;
        while (k < per_row)
        {
            int64_t c = (int64_t) (gm_rand_unif(seed, ctr++) * (double) ncols);
            if (c >= ncols) c = ncols - 1;
            int dup = 0;
            for (int64_t t = 0; t < k; t++) if (row[t] == c) { dup = 1; break; }
            if (!dup) row[k++] = c;
        }
        sort_i64(row, per_row);
        if (Ax)
        {
            for (int64_t t = 0; t < per_row; t++)
            {
                Ax[i * per_row + t] = gm_rand_unif(seed ^ 0x5DEECE66DULL,
                                                   (uint64_t) (i * per_row + t));
            }
        }
    }
}

// Per-row degree statistics helper: max row length.
int64_t gm_gen_max_rowlen(int64_t n, const int64_t *Ap)
{
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
    {
        int64_t d = Ap[i + 1] - Ap[i];
        if (d > m) m = d;
    }
    return m;
}
