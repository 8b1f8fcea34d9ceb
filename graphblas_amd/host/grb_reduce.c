// grb_reduce.c: GrB_Matrix_reduce_<T>, c = accum (c, reduce (A)) to a scalar.
// Reference: Source/reduce/GB_reduce_to_scalar.c and its template
// (Source/reduce/template/GB_reduce_to_scalar_template.c:40-95).  The values
// are reduced on the device; the monoid's type must be A's type.

#include "grb_internal.h"

static GrB_Info GB_reduce(void *c, int32_t ccode, const GrB_BinaryOp accum,
                          const GrB_Monoid monoid, const GrB_Matrix A)
{
    if (!c || !monoid) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (monoid->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    if (accum && accum->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    const int32_t mt = monoid->op->ztype->code;
    if (mt != A->type->code)
        return GB_error(A, GrB_NOT_IMPLEMENTED, "reduce: monoid type must equal A's type");
    int add;
    switch (monoid->op->opcode)
    {
        case GB_OP_PLUS: add = GM_ADD_PLUS; break;
        case GB_OP_MIN: case GB_OP_LAND: add = GM_ADD_MIN; break;
        case GB_OP_MAX: case GB_OP_LOR: add = GM_ADD_MAX; break;
        case GB_OP_ANY: add = GM_ADD_ANY; break;
        default: return GrB_NOT_IMPLEMENTED;
    }
    if ((info = GB_matrix_csr(A)) != GrB_SUCCESS) return info;   // bitmap / full: its CSR form
    union { int64_t i64; int32_t i32; double f64; uint8_t b; } r;
    memset(&r, 0, sizeof(r));
    int rc = gm_reduce_scalar(&r, add, mt, A->x, A->p ? A->nvals : 0, A->iso ? 1 : 0);
    if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "reduce: %s", gm_last_error());
    // typecast to the scalar's type, then accum on the host
    double vd = 0;
    int64_t vi = 0;
    switch (mt)
    {
        case GM_INT64: vi = r.i64; vd = (double) r.i64; break;
        case GM_INT32: vi = r.i32; vd = (double) r.i32; break;
        case GM_FP64: vi = (int64_t) r.f64; vd = r.f64; break;
        case GM_BOOL: vi = r.b != 0; vd = (double) (r.b != 0); break;
        default: return GrB_NOT_IMPLEMENTED;
    }
    if (accum && accum->opcode != GB_OP_PLUS && accum->opcode != GB_OP_SECOND)
        return GB_error(A, GrB_NOT_IMPLEMENTED, "reduce: only PLUS/SECOND accum supported");
    const bool acc = accum && accum->opcode == GB_OP_PLUS;
    switch (ccode)
    {
        case GM_INT64: *(int64_t *) c = acc ? *(int64_t *) c + vi : vi; break;
        case GM_INT32: *(int32_t *) c = acc ? *(int32_t *) c + (int32_t) vi : (int32_t) vi; break;
        case GM_FP64: *(double *) c = acc ? *(double *) c + vd : vd; break;
        case GM_BOOL: *(bool *) c = acc ? (*(bool *) c || vi) : (vi != 0); break;
        default: return GrB_NOT_IMPLEMENTED;
    }
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_reduce_INT64(int64_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Matrix A, const GrB_Descriptor desc)
{
    (void) desc;
    return GB_reduce(c, GM_INT64, accum, monoid, A);
}
GrB_Info GrB_Matrix_reduce_INT32(int32_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Matrix A, const GrB_Descriptor desc)
{
    (void) desc;
    return GB_reduce(c, GM_INT32, accum, monoid, A);
}
GrB_Info GrB_Matrix_reduce_FP64(double *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Matrix A, const GrB_Descriptor desc)
{
    (void) desc;
    return GB_reduce(c, GM_FP64, accum, monoid, A);
}
GrB_Info GrB_Matrix_reduce_BOOL(bool *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Matrix A, const GrB_Descriptor desc)
{
    (void) desc;
    return GB_reduce(c, GM_BOOL, accum, monoid, A);
}
