// gm_oracle.c: CPU oracle for GrB_mxm's dot3 and saxpy3 methods.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
// algorithms (SuiteSparse:GraphBLAS v9.3.1), used by tests/, by the smoke
// check and by bench.py's cpu_baseline leg as the checker and CPU baseline.
// The product (graphblas_amd/) never links, loads or calls it.
//
// Semiring instances follow the FactoryKernels macro blocks:
//   plus_pair_int64  FactoryKernels/GB_AxB__plus_pair_int64.c:17-72
//   plus_pair_int32  FactoryKernels/GB_AxB__plus_pair_int32.c (int32 wraps)
//   any_pair_bool    Source/mxm/GB_AxB__any_pair_iso.c:15-60 (iso C = true)
//   plus_times_fp64  FactoryKernels/GB_AxB__plus_times_fp64.c:17-67
//   min_plus_int32   FactoryKernels/GB_AxB__min_plus_int32.c:17-60
//                    (identity INT32_MAX, terminal INT32_MIN)
//   any_secondi_int64 -> value is the inner index k, as any_firstj after the
//                    reference's operand swap (Source/binaryop/GB_flip_binop.c
//                    :366-367, FactoryKernels/GB_AxB__any_firstj_int64.c)
//   plus_times_int64, min_plus_int64, max_plus_int32, min_plus_fp64 extend
//                    the same macro pattern.
//
// Pinning: see tests/test_oracle.py (brute force on tiny inputs, scipy
// sparse products, triangle-count closed forms, compensated sums).

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

typedef struct
{
    int64_t nrows, ncols;
    const int64_t *p;       // nrows+1
    const int64_t *j;       // nnz
    const void *x;          // values (or NULL when the semiring ignores them)
    int32_t iso;            // 1: x holds a single value
    int32_t xsize;          // bytes per value (mask value test)
} or_csr;

// GB_MCAST: a valued mask entry is true iff its bytes are not all zero.
static inline int or_mask_value(const or_csr *M, int64_t p)
{
    if (M->x == NULL) return 1;
    const uint8_t *x = (const uint8_t *) M->x + (size_t) (M->iso ? 0 : p) * (size_t) M->xsize;
    for (int t = 0; t < M->xsize; t++) if (x[t]) return 1;
    return 0;
}

static int or_cmp(const void *x, const void *y)
{
    int64_t a = *(const int64_t *) x, b = *(const int64_t *) y;
    return (a > b) - (a < b);
}

static void or_sort_i64(int64_t *x, int64_t n)
{
    if (n < 24)
    {
        for (int64_t k = 1; k < n; k++)
        {
            int64_t v = x[k], t = k - 1;
            while (t >= 0 && x[t] > v) { x[t + 1] = x[t]; t--; }
            x[t + 1] = v;
        }
    }
    else qsort(x, (size_t) n, sizeof(int64_t), or_cmp);
}

#define IMIN(a, b) (((a) < (b)) ? (a) : (b))
#define IMAX(a, b) (((a) > (b)) ? (a) : (b))

// ---- plus_pair_int64: cij++ with overflow ignored; exists iff cij != 0
#define SR plus_pair_int64
#define ZT int64_t
#define XT int8_t
#define YT int8_t
#define PATTERN_A 1
#define PATTERN_B 1
#define MULT(z, a, b, i, k, j) z = 1
#define UPDATE(z, t) z += t
#define IS_ANY 0
#define PLUS_PAIR_WIDE 1
#define TERMINAL(z) 0
#include "gm_oracle_template.c"

// ---- plus_pair_int32: wraps modulo 2^32, existence by flag
#define SR plus_pair_int32
#define ZT int32_t
#define XT int8_t
#define YT int8_t
#define PATTERN_A 1
#define PATTERN_B 1
#define MULT(z, a, b, i, k, j) z = 1
#define UPDATE(z, t) z = (int32_t) ((uint32_t) (z) + (uint32_t) (t))
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) 0
#include "gm_oracle_template.c"

// ---- any_pair_bool: pattern only, iso true
#define SR any_pair_bool
#define ZT uint8_t
#define XT int8_t
#define YT int8_t
#define PATTERN_A 1
#define PATTERN_B 1
#define MULT(z, a, b, i, k, j) z = 1
#define UPDATE(z, t) z = t
#define IS_ANY 1
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) 1
#include "gm_oracle_template.c"

// ---- plus_times_fp64: z = a*b, then z += a*b (no fused multiply-add)
#define SR plus_times_fp64
#define ZT double
#define XT double
#define YT double
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (a) * (b)
#define UPDATE(z, t) z += t
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) 0
#include "gm_oracle_template.c"

// ---- min_plus_int32: z = a+b (int32 wrap), min monoid, terminal INT32_MIN
#define SR min_plus_int32
#define ZT int32_t
#define XT int32_t
#define YT int32_t
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (int32_t) ((uint32_t) (a) + (uint32_t) (b))
#define UPDATE(z, t) z = IMIN(z, t)
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) ((z) == INT32_MIN)
#include "gm_oracle_template.c"

// ---- any_secondi_int64: value = inner index k
#define SR any_secondi_int64
#define ZT int64_t
#define XT int8_t
#define YT int8_t
#define PATTERN_A 1
#define PATTERN_B 1
#define MULT(z, a, b, i, k, j) z = (k)
#define UPDATE(z, t) z = t
#define IS_ANY 1
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) 1
#include "gm_oracle_template.c"

// ---- plus_times_int64
#define SR plus_times_int64
#define ZT int64_t
#define XT int64_t
#define YT int64_t
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (int64_t) ((uint64_t) (a) * (uint64_t) (b))
#define UPDATE(z, t) z = (int64_t) ((uint64_t) (z) + (uint64_t) (t))
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) 0
#include "gm_oracle_template.c"

// ---- min_plus_int64
#define SR min_plus_int64
#define ZT int64_t
#define XT int64_t
#define YT int64_t
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (int64_t) ((uint64_t) (a) + (uint64_t) (b))
#define UPDATE(z, t) z = IMIN(z, t)
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) ((z) == INT64_MIN)
#include "gm_oracle_template.c"

// ---- max_plus_int32
#define SR max_plus_int32
#define ZT int32_t
#define XT int32_t
#define YT int32_t
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (int32_t) ((uint32_t) (a) + (uint32_t) (b))
#define UPDATE(z, t) z = IMAX(z, t)
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) ((z) == INT32_MAX)
#include "gm_oracle_template.c"

// ---- min_plus_fp64 (fmin semantics: a NaN operand is ignored,
// GB_AxB_saxpy3_template.h:393-530 and GB_math.h's GB_FMIN)
static inline double or_fmin(double x, double y)
{
    if (isnan(x)) return y;
    if (isnan(y)) return x;
    return (x < y) ? x : y;
}
#define SR min_plus_fp64
#define ZT double
#define XT double
#define YT double
#define PATTERN_A 0
#define PATTERN_B 0
#define MULT(z, a, b, i, k, j) z = (a) + (b)
#define UPDATE(z, t) z = or_fmin(z, t)
#define IS_ANY 0
#define PLUS_PAIR_WIDE 0
#define TERMINAL(z) ((z) == -INFINITY)
#include "gm_oracle_template.c"

// semiring codes (same numbering as the test-side table in oracle/pyoracle.py)
enum
{
    OR_PLUS_PAIR_INT64 = 0,
    OR_PLUS_PAIR_INT32 = 1,
    OR_ANY_PAIR_BOOL = 2,
    OR_PLUS_TIMES_FP64 = 3,
    OR_MIN_PLUS_INT32 = 4,
    OR_ANY_SECONDI_INT64 = 5,
    OR_PLUS_TIMES_INT64 = 6,
    OR_MIN_PLUS_INT64 = 7,
    OR_MAX_PLUS_INT32 = 8,
    OR_MIN_PLUS_FP64 = 9,
};

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// C<M> = A*B' by the dot3 method; returns nnz(C) (compact, no zombies).
int64_t oracle_dot3(int sr, const or_csr *M, int mask_struct, const or_csr *A,
                    const or_csr *B, int64_t *Cp, int64_t *Ci, void *Cx)
{
    switch (sr)
    {
        case OR_PLUS_PAIR_INT64:  return dot3_plus_pair_int64(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_PLUS_PAIR_INT32:  return dot3_plus_pair_int32(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_ANY_PAIR_BOOL:    return dot3_any_pair_bool(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_PLUS_TIMES_FP64:  return dot3_plus_times_fp64(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_MIN_PLUS_INT32:   return dot3_min_plus_int32(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_ANY_SECONDI_INT64:return dot3_any_secondi_int64(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_PLUS_TIMES_INT64: return dot3_plus_times_int64(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_MIN_PLUS_INT64:   return dot3_min_plus_int64(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_MAX_PLUS_INT32:   return dot3_max_plus_int32(M, mask_struct, A, B, Cp, Ci, Cx);
        case OR_MIN_PLUS_FP64:    return dot3_min_plus_fp64(M, mask_struct, A, B, Cp, Ci, Cx);
        default: return -2;
    }
}

// C<M or !M> = A*B by saxpy3.  phase 0: counts into Cp[0..m-1], returns
// total; the caller scans Cp.  phase 1: fills Ci/Cx using Cp.  M may be NULL.
int64_t oracle_saxpy3(int sr, const or_csr *M, int mask_comp, int mask_struct,
                      const or_csr *A, const or_csr *B, int64_t *Cp, int64_t *Ci,
                      void *Cx, int phase)
{
    switch (sr)
    {
        case OR_PLUS_PAIR_INT64:  return saxpy3_plus_pair_int64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_PLUS_PAIR_INT32:  return saxpy3_plus_pair_int32(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_ANY_PAIR_BOOL:    return saxpy3_any_pair_bool(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_PLUS_TIMES_FP64:  return saxpy3_plus_times_fp64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_MIN_PLUS_INT32:   return saxpy3_min_plus_int32(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_ANY_SECONDI_INT64:return saxpy3_any_secondi_int64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_PLUS_TIMES_INT64: return saxpy3_plus_times_int64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_MIN_PLUS_INT64:   return saxpy3_min_plus_int64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_MAX_PLUS_INT32:   return saxpy3_max_plus_int32(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        case OR_MIN_PLUS_FP64:    return saxpy3_min_plus_fp64(M, mask_comp, mask_struct, A, B, Cp, Ci, Cx, phase);
        default: return -2;
    }
}

// Sum of C's values as int64 / double (the GrB_reduce(+) that closes the
// triangle count, Source/reduce/template/GB_reduce_to_scalar_template.c:40-59).
int64_t oracle_sum_i64(const int64_t *x, int64_t n)
{
    int64_t s = 0;
    #pragma omp parallel for reduction(+:s)
    for (int64_t t = 0; t < n; t++) s += x[t];
    return s;
}

// Neumaier-compensated dot product of two sparse rows, used to state the
// error of both paths for plus_times fp64 (Neumaier 1974).
double oracle_neumaier_sum(const double *x, int64_t n)
{
    double s = 0.0, c = 0.0;
    for (int64_t t = 0; t < n; t++)
    {
        double v = x[t], u = s + v;
        if (fabs(s) >= fabs(v)) c += (s - u) + v; else c += (v - u) + s;
        s = u;
    }
    return s + c;
}

// C = A*B under plus_times fp64 with every entry accumulated in double-double
// arithmetic: each product a*b is split exactly into p + e (e = fma(a, b, -p),
// the error-free product) and both parts are added to the entry's
// (hi, lo) pair with Knuth's two-sum, so the only rounding left is that of
// the final hi + lo.  This is the compensated FP64 reference against which
// the summation error of both the sequential path (the reference's
// ascending-k order, GB_AxB_saxpy3_coarseHash_phase5.c:22-93) and the GPU
// path is stated.  The pattern (Cp, Ci, sorted rows) comes from
// oracle_saxpy3; Cx receives the compensated values in the same positions.
// Returns 0, or -1 when a column of A*B is missing from Ci (caller error).
int64_t oracle_saxpy3_plus_times_fp64_dd(const or_csr *A, const or_csr *B, const int64_t *Cp,
                                         const int64_t *Ci, double *Cx)
{
    const int64_t m = A->nrows, n = B->ncols;
    const double *Ax = (const double *) A->x, *Bx = (const double *) B->x;
    int64_t bad = 0;
    #pragma omp parallel reduction(+:bad)
    {
        int64_t *pos = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n > 0 ? n : 1));
        double *hi = NULL, *lo = NULL;
        int64_t cap = 0;
        if (pos) for (int64_t t = 0; t < n; t++) pos[t] = -1;
        #pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < m; i++)
        {
            if (!pos) { bad++; continue; }
            const int64_t c0 = Cp[i], c1 = Cp[i + 1], len = c1 - c0;
            if (len > cap)
            {
                free(hi); free(lo);
                cap = 2 * len;
                hi = (double *) malloc(sizeof(double) * (size_t) cap);
                lo = (double *) malloc(sizeof(double) * (size_t) cap);
            }
            for (int64_t q = 0; q < len; q++) { pos[Ci[c0 + q]] = q; hi[q] = 0.0; lo[q] = 0.0; }
            for (int64_t pa = A->p[i]; pa < A->p[i + 1]; pa++)
            {
                const int64_t k = A->j[pa];
                const double a = Ax[A->iso ? 0 : pa];
                for (int64_t pb = B->p[k]; pb < B->p[k + 1]; pb++)
                {
                    const int64_t q = pos[B->j[pb]];
                    if (q < 0) { bad++; continue; }
                    const double b = Bx[B->iso ? 0 : pb];
                    const double p = a * b, e = fma(a, b, -p);
                    // (hi, lo) += p: two-sum, then fold the error and e into lo
                    const double s = hi[q] + p, bb = s - hi[q];
                    const double err = (hi[q] - (s - bb)) + (p - bb);
                    hi[q] = s;
                    lo[q] += err + e;
                }
            }
            for (int64_t q = 0; q < len; q++)
            {
                Cx[c0 + q] = hi[q] + lo[q];
                pos[Ci[c0 + q]] = -1;
            }
        }
        free(pos); free(hi); free(lo);
    }
    return bad ? -1 : 0;
}
