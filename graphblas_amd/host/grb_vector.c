// grb_vector.c: sparse vectors and the vector products GrB_vxm / GrB_mxv.
//
// A vector of size n is held as a 1-by-n CSR matrix in HBM (one row pointer
// pair, sorted indices, values), i.e. the reference's n-by-1 by-column
// vector seen by row.  The products then run on GrB_mxm's kernels, as the
// reference does (Source/mxm/GrB_vxm.c:51-67, GrB_mxv.c: vectors passed to
// GB_mxm as n-by-1 matrices):
//
//   GrB_vxm  w'<m'> = accum (w', u'*A)   ->  GrB_mxm (W, M, accum, s, U, A, desc)
//            (desc's GrB_INP1 transposes A; GrB_INP0 does not apply)
//   GrB_mxv  w<m>   = accum (w, A*u)     ->  w' = u'*A': GrB_mxm (W, M, accum, s, U, A, desc')
//            with desc' transposing A unless desc asks for A' (GrB_INP0)
//
// The supported multipliers (TIMES, PLUS, PAIR/ONEB, and the positional
// ones that return the inner index) give the same result with their operands
// swapped, so the swap needs no flipped operator.

#include "grb_internal.h"

struct GB_Vector_opaque
{
    uint64_t magic;
    GrB_Matrix m;       // 1 x n
};

static GrB_Info GB_check_vector(const GrB_Vector v)
{
    if (!v) return GrB_NULL_POINTER;
    if (v->magic == GB_FREED) return GrB_INVALID_OBJECT;
    if (v->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    return GB_check_matrix(v->m);
}

GrB_Info GrB_Vector_new(GrB_Vector *v, GrB_Type type, GrB_Index n)
{
    if (!GB_initialized()) return GrB_PANIC;
    if (!v) return GrB_NULL_POINTER;
    *v = NULL;
    GrB_Vector w = (GrB_Vector) calloc(1, sizeof(struct GB_Vector_opaque));
    if (!w) return GrB_OUT_OF_MEMORY;
    GrB_Info info = GrB_Matrix_new(&w->m, type, 1, n);
    if (info != GrB_SUCCESS)
    {
        free(w);
        return info;
    }
    w->magic = GB_MAGIC;
    *v = w;
    return GrB_SUCCESS;
}

GrB_Info GrB_Vector_free(GrB_Vector *v)
{
    if (!v || !*v) return GrB_SUCCESS;
    GrB_Vector w = *v;
    if (w->magic == GB_MAGIC)
    {
        GrB_Matrix_free(&w->m);
        w->magic = GB_FREED;
        free(w);
    }
    *v = NULL;
    return GrB_SUCCESS;
}

GrB_Info GrB_Vector_dup(GrB_Vector *w, const GrB_Vector u)
{
    if (!w) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_vector(u);
    if (info != GrB_SUCCESS) return info;
    GrB_Vector v = (GrB_Vector) calloc(1, sizeof(struct GB_Vector_opaque));
    if (!v) return GrB_OUT_OF_MEMORY;
    info = GrB_Matrix_dup(&v->m, u->m);
    if (info != GrB_SUCCESS)
    {
        free(v);
        return info;
    }
    v->magic = GB_MAGIC;
    *w = v;
    return GrB_SUCCESS;
}

GrB_Info GrB_Vector_clear(GrB_Vector v)
{
    GrB_Info info = GB_check_vector(v);
    return info != GrB_SUCCESS ? info : GrB_Matrix_clear(v->m);
}

GrB_Info GrB_Vector_size(GrB_Index *n, const GrB_Vector v)
{
    GrB_Info info = GB_check_vector(v);
    return info != GrB_SUCCESS ? info : GrB_Matrix_ncols(n, v->m);
}

GrB_Info GrB_Vector_nvals(GrB_Index *nvals, const GrB_Vector v)
{
    GrB_Info info = GB_check_vector(v);
    return info != GrB_SUCCESS ? info : GrB_Matrix_nvals(nvals, v->m);
}

// build from (I, X): the 1 x n matrix's tuples are (0, I[k], X[k])
#define GB_VBUILD(T, ctype)                                                                     \
GrB_Info GrB_Vector_build_##T(GrB_Vector w, const GrB_Index *I, const ctype *X,                 \
                              GrB_Index nvals, const GrB_BinaryOp dup)                          \
{                                                                                               \
    GrB_Info info = GB_check_vector(w);                                                         \
    if (info != GrB_SUCCESS) return info;                                                       \
    GrB_Index *zero = (GrB_Index *) calloc(nvals > 0 ? nvals : 1, sizeof(GrB_Index));           \
    if (!zero) return GrB_OUT_OF_MEMORY;                                                        \
    info = GrB_Matrix_build_##T(w->m, zero, I, X, nvals, dup);                                   \
    free(zero);                                                                                 \
    return info;                                                                                \
}
GB_VBUILD(BOOL, bool)
GB_VBUILD(INT32, int32_t)
GB_VBUILD(INT64, int64_t)
GB_VBUILD(FP64, double)
#undef GB_VBUILD

#define GB_VEXTRACT(T, ctype)                                                                   \
GrB_Info GrB_Vector_extractTuples_##T(GrB_Index *I, ctype *X, GrB_Index *nvals,                 \
                                      const GrB_Vector v)                                       \
{                                                                                               \
    GrB_Info info = GB_check_vector(v);                                                         \
    if (info != GrB_SUCCESS) return info;                                                       \
    if (!nvals) return GrB_NULL_POINTER;                                                        \
    GrB_Index *rows = (GrB_Index *) malloc(sizeof(GrB_Index) * (*nvals > 0 ? *nvals : 1));      \
    if (!rows) return GrB_OUT_OF_MEMORY;                                                        \
    info = GrB_Matrix_extractTuples_##T(rows, I, X, nvals, v->m);                               \
    free(rows);                                                                                 \
    return info;                                                                                \
}
GB_VEXTRACT(BOOL, bool)
GB_VEXTRACT(INT32, int32_t)
GB_VEXTRACT(INT64, int64_t)
GB_VEXTRACT(FP64, double)
#undef GB_VEXTRACT

#define GB_VREDUCE(T, ctype)                                                                    \
GrB_Info GrB_Vector_reduce_##T(ctype *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,     \
                               const GrB_Vector u, const GrB_Descriptor desc)                   \
{                                                                                               \
    GrB_Info info = GB_check_vector(u);                                                         \
    return info != GrB_SUCCESS ? info : GrB_Matrix_reduce_##T(c, accum, monoid, u->m, desc);    \
}
GB_VREDUCE(BOOL, bool)
GB_VREDUCE(INT32, int32_t)
GB_VREDUCE(INT64, int64_t)
GB_VREDUCE(FP64, double)
#undef GB_VREDUCE

// the descriptor of the row form: mask/replace from desc, the given
// transpose setting for A (the second operand), none for u
static struct GB_Descriptor_opaque GB_row_desc(const GrB_Descriptor desc, bool a_tran)
{
    struct GB_Descriptor_opaque d;
    memset(&d, 0, sizeof(d));
    if (desc) d = *desc;
    d.magic = GB_MAGIC;
    d.in0 = GrB_DEFAULT;
    d.in1 = a_tran ? GrB_TRAN : GrB_DEFAULT;
    d.builtin = false;
    return d;
}

static GrB_Info GB_vector_product(GrB_Vector w, const GrB_Vector mask, const GrB_BinaryOp accum,
                                  const GrB_Semiring semiring, const GrB_Vector u,
                                  const GrB_Matrix A, const GrB_Descriptor desc, bool a_tran)
{
    GrB_Info info;
    if ((info = GB_check_vector(w)) != GrB_SUCCESS) return info;
    if ((info = GB_check_vector(u)) != GrB_SUCCESS) return info;
    if (mask && (info = GB_check_vector(mask)) != GrB_SUCCESS) return info;
    if (desc && desc->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    struct GB_Descriptor_opaque d = GB_row_desc(desc, a_tran);
    info = GrB_mxm(w->m, mask ? mask->m : NULL, accum, semiring, u->m, A, &d);
    return info;
}

GrB_Info GrB_vxm(GrB_Vector w, const GrB_Vector mask, const GrB_BinaryOp accum,
                 const GrB_Semiring semiring, const GrB_Vector u, const GrB_Matrix A,
                 const GrB_Descriptor desc)
{
    if (!GB_initialized()) return GrB_PANIC;
    const bool a_tran = desc && desc->magic == GB_MAGIC && desc->in1 == GrB_TRAN;
    return GB_vector_product(w, mask, accum, semiring, u, A, desc, a_tran);
}

GrB_Info GrB_mxv(GrB_Vector w, const GrB_Vector mask, const GrB_BinaryOp accum,
                 const GrB_Semiring semiring, const GrB_Matrix A, const GrB_Vector u,
                 const GrB_Descriptor desc)
{
    if (!GB_initialized()) return GrB_PANIC;
    // A*u = (u'*A')': A is transposed unless the descriptor already asks for A'
    const bool a_tran = !(desc && desc->magic == GB_MAGIC && desc->in0 == GrB_TRAN);
    return GB_vector_product(w, mask, accum, semiring, u, A, desc, a_tran);
}

GrB_Info GrB_Vector_error(const char **error, const GrB_Vector v)
{
    if (!error) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_vector(v);
    if (info != GrB_SUCCESS) return info;
    return GrB_Matrix_error(error, v->m);
}
