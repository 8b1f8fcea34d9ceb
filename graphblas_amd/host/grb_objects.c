// grb_objects.c: built-in types, operators, monoids, semirings, descriptors,
// init/finalize, errors and global options.
//
// The built-in objects are static structs whose addresses are published as
// the reference's global handles (GrB_INT64, GxB_PLUS_PAIR_INT64,
// GrB_DESC_ST1, ...), so client code written against the reference's header
// compiles and links unchanged for the functions this library provides.

#include "grb_internal.h"
#include <stdarg.h>

int GB_burble = 0;
static bool g_init = false;

bool GB_initialized(void) { return g_init; }

GrB_Info GrB_init(GrB_Mode mode)
{
    if (mode != GrB_NONBLOCKING && mode != GrB_BLOCKING) return GrB_INVALID_VALUE;
    if (g_init) return GrB_INVALID_VALUE;   // the reference allows one init
    // the GPU is opened lazily, on the first operation that needs HBM
    g_init = true;
    return GrB_SUCCESS;
}

GrB_Info GrB_finalize(void)
{
    if (g_init) gm_finalize();
    g_init = false;
    return GrB_SUCCESS;
}

GrB_Info GB_from_gm(int rc)
{
    switch (rc)
    {
        case GM_SUCCESS: return GrB_SUCCESS;
        case GM_NO_VALUE: return GrB_NOT_IMPLEMENTED;
        case GM_INVALID_VALUE: return GrB_INVALID_VALUE;
        case GM_NOT_IMPLEMENTED: return GrB_NOT_IMPLEMENTED;
        case GM_OUT_OF_MEMORY: return GrB_OUT_OF_MEMORY;
        default: return GrB_PANIC;
    }
}

GrB_Info GB_error(GrB_Matrix C, GrB_Info info, const char *fmt, ...)
{
    if (C && C->magic == GB_MAGIC)
    {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(C->err, sizeof(C->err), fmt, ap);
        va_end(ap);
    }
    return info;
}

GrB_Info GrB_Matrix_error(const char **error, const GrB_Matrix A)
{
    if (!error) return GrB_NULL_POINTER;
    if (!A || A->magic != GB_MAGIC) { *error = ""; return GrB_UNINITIALIZED_OBJECT; }
    *error = A->err;
    return GrB_SUCCESS;
}

GrB_Info GxB_Global_Option_set_INT32(int field, int32_t value)
{
    if (field == GxB_BURBLE) { GB_burble = value != 0; return GrB_SUCCESS; }
    return GrB_INVALID_VALUE;
}

GrB_Info GxB_Global_Option_get_INT32(int field, int32_t *value)
{
    if (!value) return GrB_NULL_POINTER;
    if (field == GxB_BURBLE) { *value = GB_burble; return GrB_SUCCESS; }
    return GrB_INVALID_VALUE;
}

//------------------------------------------------------------------------------
// types
//------------------------------------------------------------------------------

#define GB_TYPE_DEF(NAME, CODE, CT)                                                  \
    static struct GB_Type_opaque GB_type_##NAME = {GB_MAGIC, CODE, sizeof(CT), #NAME}; \
    GrB_Type GrB_##NAME = &GB_type_##NAME;

GB_TYPE_DEF(BOOL, GM_BOOL, bool)
GB_TYPE_DEF(INT8, GM_INT8, int8_t)
GB_TYPE_DEF(UINT8, GM_UINT8, uint8_t)
GB_TYPE_DEF(INT16, GM_INT16, int16_t)
GB_TYPE_DEF(UINT16, GM_UINT16, uint16_t)
GB_TYPE_DEF(INT32, GM_INT32, int32_t)
GB_TYPE_DEF(UINT32, GM_UINT32, uint32_t)
GB_TYPE_DEF(INT64, GM_INT64, int64_t)
GB_TYPE_DEF(UINT64, GM_UINT64, uint64_t)
GB_TYPE_DEF(FP32, GM_FP32, float)
GB_TYPE_DEF(FP64, GM_FP64, double)

GrB_Type GB_type_from_code(int32_t code)
{
    switch (code)
    {
        case GM_BOOL: return GrB_BOOL;
        case GM_INT8: return GrB_INT8;
        case GM_UINT8: return GrB_UINT8;
        case GM_INT16: return GrB_INT16;
        case GM_UINT16: return GrB_UINT16;
        case GM_INT32: return GrB_INT32;
        case GM_UINT32: return GrB_UINT32;
        case GM_INT64: return GrB_INT64;
        case GM_UINT64: return GrB_UINT64;
        case GM_FP32: return GrB_FP32;
        case GM_FP64: return GrB_FP64;
        default: return NULL;
    }
}

GrB_Info GxB_Type_size(size_t *size, GrB_Type type)
{
    if (!size || !type) return GrB_NULL_POINTER;
    if (type->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    *size = type->size;
    return GrB_SUCCESS;
}

GrB_Info GrB_Type_get_size(size_t *size, GrB_Type type) { return GxB_Type_size(size, type); }

//------------------------------------------------------------------------------
// binary operators
//------------------------------------------------------------------------------

#define GB_OPDEF(HANDLE, OPC, XT, YT, ZT)                                              \
    static struct GB_BinaryOp_opaque GB_op_##HANDLE = {GB_MAGIC, OPC, &GB_type_##XT,    \
                                                       &GB_type_##YT, &GB_type_##ZT,   \
                                                       #HANDLE, true};                 \
    GrB_BinaryOp HANDLE = &GB_op_##HANDLE;

#define GB_OPS_FOR(T)                                      \
    GB_OPDEF(GrB_PLUS_##T, GB_OP_PLUS, T, T, T)            \
    GB_OPDEF(GrB_MINUS_##T, GB_OP_MINUS, T, T, T)          \
    GB_OPDEF(GrB_TIMES_##T, GB_OP_TIMES, T, T, T)          \
    GB_OPDEF(GrB_MIN_##T, GB_OP_MIN, T, T, T)              \
    GB_OPDEF(GrB_MAX_##T, GB_OP_MAX, T, T, T)              \
    GB_OPDEF(GrB_FIRST_##T, GB_OP_FIRST, T, T, T)          \
    GB_OPDEF(GrB_SECOND_##T, GB_OP_SECOND, T, T, T)        \
    GB_OPDEF(GxB_ANY_##T, GB_OP_ANY, T, T, T)              \
    GB_OPDEF(GxB_PAIR_##T, GB_OP_PAIR, T, T, T)            \
    GB_OPDEF(GrB_ONEB_##T, GB_OP_PAIR, T, T, T)

GB_OPS_FOR(INT32)
GB_OPS_FOR(INT64)
GB_OPS_FOR(FP64)
GB_OPDEF(GrB_LOR, GB_OP_LOR, BOOL, BOOL, BOOL)
GB_OPDEF(GrB_LAND, GB_OP_LAND, BOOL, BOOL, BOOL)
GB_OPDEF(GxB_ANY_BOOL, GB_OP_ANY, BOOL, BOOL, BOOL)
GB_OPDEF(GxB_PAIR_BOOL, GB_OP_PAIR, BOOL, BOOL, BOOL)
GB_OPDEF(GrB_ONEB_BOOL, GB_OP_PAIR, BOOL, BOOL, BOOL)
GB_OPDEF(GrB_FIRST_BOOL, GB_OP_FIRST, BOOL, BOOL, BOOL)
GB_OPDEF(GrB_SECOND_BOOL, GB_OP_SECOND, BOOL, BOOL, BOOL)
GB_OPDEF(GxB_FIRSTI_INT64, GB_OP_FIRSTI, INT64, INT64, INT64)
GB_OPDEF(GxB_FIRSTJ_INT64, GB_OP_FIRSTJ, INT64, INT64, INT64)
GB_OPDEF(GxB_SECONDI_INT64, GB_OP_SECONDI, INT64, INT64, INT64)
GB_OPDEF(GxB_SECONDJ_INT64, GB_OP_SECONDJ, INT64, INT64, INT64)
GB_OPDEF(GxB_FIRSTI_INT32, GB_OP_FIRSTI, INT32, INT32, INT32)
GB_OPDEF(GxB_FIRSTJ_INT32, GB_OP_FIRSTJ, INT32, INT32, INT32)
GB_OPDEF(GxB_SECONDI_INT32, GB_OP_SECONDI, INT32, INT32, INT32)
GB_OPDEF(GxB_SECONDJ_INT32, GB_OP_SECONDJ, INT32, INT32, INT32)

// Every operator here is a static built-in: freeing one changes nothing, and
// the handle is left as it is, as the reference leaves a built-in's handle
// (Source/op/GB_Op_free.c frees and clears only allocated operators), so a
// caller that frees the address of a global handle keeps a working global.
GrB_Info GrB_BinaryOp_free(GrB_BinaryOp *op)
{
    (void) op;
    return GrB_SUCCESS;
}

//------------------------------------------------------------------------------
// monoids
//------------------------------------------------------------------------------

#define GB_MONDEF(HANDLE, OP)                                                          \
    static struct GB_Monoid_opaque GB_mon_##HANDLE = {GB_MAGIC, &GB_op_##OP, #HANDLE, true}; \
    GrB_Monoid HANDLE = &GB_mon_##HANDLE;

GB_MONDEF(GrB_PLUS_MONOID_INT32, GrB_PLUS_INT32)
GB_MONDEF(GrB_PLUS_MONOID_INT64, GrB_PLUS_INT64)
GB_MONDEF(GrB_PLUS_MONOID_FP64, GrB_PLUS_FP64)
GB_MONDEF(GrB_MIN_MONOID_INT32, GrB_MIN_INT32)
GB_MONDEF(GrB_MIN_MONOID_INT64, GrB_MIN_INT64)
GB_MONDEF(GrB_MIN_MONOID_FP64, GrB_MIN_FP64)
GB_MONDEF(GrB_MAX_MONOID_INT32, GrB_MAX_INT32)
GB_MONDEF(GrB_MAX_MONOID_INT64, GrB_MAX_INT64)
GB_MONDEF(GrB_MAX_MONOID_FP64, GrB_MAX_FP64)
GB_MONDEF(GxB_ANY_INT32_MONOID, GxB_ANY_INT32)
GB_MONDEF(GxB_ANY_INT64_MONOID, GxB_ANY_INT64)
GB_MONDEF(GxB_ANY_FP64_MONOID, GxB_ANY_FP64)
GB_MONDEF(GxB_ANY_BOOL_MONOID, GxB_ANY_BOOL)
GB_MONDEF(GrB_LOR_MONOID_BOOL, GrB_LOR)
GB_MONDEF(GrB_LAND_MONOID_BOOL, GrB_LAND)

// built-in monoids only, left as they are (as GrB_BinaryOp_free)
GrB_Info GrB_Monoid_free(GrB_Monoid *monoid)
{
    (void) monoid;
    return GrB_SUCCESS;
}

//------------------------------------------------------------------------------
// semirings
//------------------------------------------------------------------------------

#define GB_SRDEF(HANDLE, MON, OP)                                                   \
    static struct GB_Semiring_opaque GB_sr_##HANDLE = {GB_MAGIC, &GB_mon_##MON,      \
                                                       &GB_op_##OP, #HANDLE, true};  \
    GrB_Semiring HANDLE = &GB_sr_##HANDLE;

GB_SRDEF(GxB_PLUS_PAIR_INT64, GrB_PLUS_MONOID_INT64, GxB_PAIR_INT64)
GB_SRDEF(GxB_PLUS_PAIR_INT32, GrB_PLUS_MONOID_INT32, GxB_PAIR_INT32)
GB_SRDEF(GxB_PLUS_ONEB_INT64, GrB_PLUS_MONOID_INT64, GrB_ONEB_INT64)
GB_SRDEF(GxB_ANY_PAIR_BOOL, GxB_ANY_BOOL_MONOID, GxB_PAIR_BOOL)
GB_SRDEF(GxB_LOR_PAIR_BOOL, GrB_LOR_MONOID_BOOL, GxB_PAIR_BOOL)
GB_SRDEF(GrB_PLUS_TIMES_SEMIRING_FP64, GrB_PLUS_MONOID_FP64, GrB_TIMES_FP64)
GB_SRDEF(GrB_PLUS_TIMES_SEMIRING_INT64, GrB_PLUS_MONOID_INT64, GrB_TIMES_INT64)
GB_SRDEF(GrB_MIN_PLUS_SEMIRING_INT32, GrB_MIN_MONOID_INT32, GrB_PLUS_INT32)
GB_SRDEF(GrB_MIN_PLUS_SEMIRING_INT64, GrB_MIN_MONOID_INT64, GrB_PLUS_INT64)
GB_SRDEF(GrB_MIN_PLUS_SEMIRING_FP64, GrB_MIN_MONOID_FP64, GrB_PLUS_FP64)
GB_SRDEF(GrB_MAX_PLUS_SEMIRING_INT32, GrB_MAX_MONOID_INT32, GrB_PLUS_INT32)
GB_SRDEF(GxB_ANY_SECONDI_INT64, GxB_ANY_INT64_MONOID, GxB_SECONDI_INT64)
GB_SRDEF(GxB_ANY_FIRSTJ_INT64, GxB_ANY_INT64_MONOID, GxB_FIRSTJ_INT64)

GrB_Info GrB_Semiring_new(GrB_Semiring *semiring, GrB_Monoid add, GrB_BinaryOp multiply)
{
    if (!semiring || !add || !multiply) return GrB_NULL_POINTER;
    *semiring = NULL;
    if (add->magic != GB_MAGIC || multiply->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    // the monoid's type must be the multiplier's output type
    if (add->op->ztype != multiply->ztype) return GrB_DOMAIN_MISMATCH;
    struct GB_Semiring_opaque *s = (struct GB_Semiring_opaque *) calloc(1, sizeof(*s));
    if (!s) return GrB_OUT_OF_MEMORY;
    s->magic = GB_MAGIC;
    s->add = add;
    s->mult = multiply;
    s->name = "user";
    s->builtin = false;
    *semiring = s;
    return GrB_SUCCESS;
}

GrB_Info GrB_Semiring_free(GrB_Semiring *semiring)
{
    if (!semiring) return GrB_SUCCESS;
    GrB_Semiring s = *semiring;
    // a built-in keeps its handle (the reference frees and clears only
    // allocated semirings, Source/semiring/GrB_Semiring_free.c)
    if (s && s->magic == GB_MAGIC && !s->builtin)
    {
        s->magic = GB_FREED;
        free(s);
        *semiring = NULL;
    }
    return GrB_SUCCESS;
}

// Decode a semiring into the kernel layer's (add, mult, ztype) triple.
// Follows GB_AxB_semiring_builtin (Source/mxm/GB_AxB_semiring_builtin.c:
// 20-123): boolean LOR/LAND are MAX/MIN; PAIR under LOR/LAND/MIN/MAX/TIMES
// is renamed to ANY_PAIR (:108-120); positional FIRSTJ and SECONDI both
// return the inner index k (the reference flips SECONDI to FIRSTJ in its
// operand swap, Source/binaryop/GB_flip_binop.c:366-367).
bool GB_semiring_codes(const GrB_Semiring s, gm_semiring *out)
{
    if (!s || s->magic != GB_MAGIC || !out) return false;
    const GB_opcode a = s->add->op->opcode;
    const GB_opcode m = s->mult->opcode;
    const int32_t zt = s->add->op->ztype->code;
    int add, mult;
    switch (a)
    {
        case GB_OP_PLUS: add = GM_ADD_PLUS; break;
        case GB_OP_MIN: add = GM_ADD_MIN; break;
        case GB_OP_MAX: add = GM_ADD_MAX; break;
        case GB_OP_ANY: add = GM_ADD_ANY; break;
        case GB_OP_LOR: add = GM_ADD_MAX; break;
        case GB_OP_LAND: add = GM_ADD_MIN; break;
        default: return false;
    }
    switch (m)
    {
        case GB_OP_TIMES: mult = GM_MULT_TIMES; break;
        case GB_OP_PLUS: mult = GM_MULT_PLUS; break;
        case GB_OP_PAIR: mult = GM_MULT_PAIR; break;
        case GB_OP_FIRSTJ: case GB_OP_SECONDI: mult = GM_MULT_K; break;
        default: return false;
    }
    if (mult == GM_MULT_PAIR && add != GM_ADD_PLUS) add = GM_ADD_ANY;
    out->add_op = add;
    out->mult_op = mult;
    out->ztype = zt;
    return true;
}

//------------------------------------------------------------------------------
// descriptors
//------------------------------------------------------------------------------

#define GB_DESCDEF(NAME, OUT, MASK, IN0, IN1)                                           \
    static struct GB_Descriptor_opaque GB_desc_##NAME = {GB_MAGIC, OUT, MASK, IN0, IN1, \
                                                         0, 0, true};                   \
    GrB_Descriptor GrB_DESC_##NAME = &GB_desc_##NAME;

#define GB_R GrB_REPLACE
#define GB_S GrB_STRUCTURE
#define GB_C GrB_COMP
#define GB_T GrB_TRAN
GB_DESCDEF(T1, 0, 0, 0, GB_T)
GB_DESCDEF(T0, 0, 0, GB_T, 0)
GB_DESCDEF(T0T1, 0, 0, GB_T, GB_T)
GB_DESCDEF(C, 0, GB_C, 0, 0)
GB_DESCDEF(CT1, 0, GB_C, 0, GB_T)
GB_DESCDEF(CT0, 0, GB_C, GB_T, 0)
GB_DESCDEF(CT0T1, 0, GB_C, GB_T, GB_T)
GB_DESCDEF(S, 0, GB_S, 0, 0)
GB_DESCDEF(ST1, 0, GB_S, 0, GB_T)
GB_DESCDEF(ST0, 0, GB_S, GB_T, 0)
GB_DESCDEF(ST0T1, 0, GB_S, GB_T, GB_T)
GB_DESCDEF(SC, 0, GB_S + GB_C, 0, 0)
GB_DESCDEF(SCT1, 0, GB_S + GB_C, 0, GB_T)
GB_DESCDEF(SCT0, 0, GB_S + GB_C, GB_T, 0)
GB_DESCDEF(SCT0T1, 0, GB_S + GB_C, GB_T, GB_T)
GB_DESCDEF(R, GB_R, 0, 0, 0)
GB_DESCDEF(RT1, GB_R, 0, 0, GB_T)
GB_DESCDEF(RT0, GB_R, 0, GB_T, 0)
GB_DESCDEF(RT0T1, GB_R, 0, GB_T, GB_T)
GB_DESCDEF(RC, GB_R, GB_C, 0, 0)
GB_DESCDEF(RCT1, GB_R, GB_C, 0, GB_T)
GB_DESCDEF(RCT0, GB_R, GB_C, GB_T, 0)
GB_DESCDEF(RCT0T1, GB_R, GB_C, GB_T, GB_T)
GB_DESCDEF(RS, GB_R, GB_S, 0, 0)
GB_DESCDEF(RST1, GB_R, GB_S, 0, GB_T)
GB_DESCDEF(RST0, GB_R, GB_S, GB_T, 0)
GB_DESCDEF(RST0T1, GB_R, GB_S, GB_T, GB_T)
GB_DESCDEF(RSC, GB_R, GB_S + GB_C, 0, 0)
GB_DESCDEF(RSCT1, GB_R, GB_S + GB_C, 0, GB_T)
GB_DESCDEF(RSCT0, GB_R, GB_S + GB_C, GB_T, 0)
GB_DESCDEF(RSCT0T1, GB_R, GB_S + GB_C, GB_T, GB_T)

GrB_Info GrB_Descriptor_new(GrB_Descriptor *descriptor)
{
    if (!descriptor) return GrB_NULL_POINTER;
    struct GB_Descriptor_opaque *d = (struct GB_Descriptor_opaque *) calloc(1, sizeof(*d));
    if (!d) return GrB_OUT_OF_MEMORY;
    d->magic = GB_MAGIC;
    *descriptor = d;
    return GrB_SUCCESS;
}

GrB_Info GrB_Descriptor_set(GrB_Descriptor d, GrB_Desc_Field field, GrB_Desc_Value val)
{
    if (!d) return GrB_NULL_POINTER;
    if (d->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    if (d->builtin) return GrB_INVALID_VALUE;   // predefined descriptors are read-only
    switch (field)
    {
        case GrB_OUTP:
            if (val != GrB_DEFAULT && val != GrB_REPLACE) return GrB_INVALID_VALUE;
            d->out = val;
            return GrB_SUCCESS;
        case GrB_MASK:
            // GrB_COMP and GrB_STRUCTURE accumulate (GrB_Descriptor_set semantics)
            if (val == GrB_DEFAULT) d->mask = 0;
            else if (val == GrB_COMP) d->mask |= GrB_COMP;
            else if (val == GrB_STRUCTURE) d->mask |= GrB_STRUCTURE;
            else if (val == GrB_COMP_STRUCTURE) d->mask = GrB_COMP_STRUCTURE;
            else return GrB_INVALID_VALUE;
            return GrB_SUCCESS;
        case GrB_INP0:
        case GrB_INP1:
            if (val != GrB_DEFAULT && val != GrB_TRAN) return GrB_INVALID_VALUE;
            if (field == GrB_INP0) d->in0 = val; else d->in1 = val;
            return GrB_SUCCESS;
        case GxB_AxB_METHOD:
            if (val != GrB_DEFAULT && val != GxB_AxB_GUSTAVSON && val != GxB_AxB_DOT &&
                val != GxB_AxB_HASH && val != GxB_AxB_SAXPY)
                return GrB_INVALID_VALUE;
            d->axb = val;
            return GrB_SUCCESS;
        case GxB_SORT:
            d->sort = val;
            return GrB_SUCCESS;
        default:
            return GrB_INVALID_VALUE;
    }
}

GrB_Info GrB_Descriptor_free(GrB_Descriptor *descriptor)
{
    if (!descriptor) return GrB_SUCCESS;
    GrB_Descriptor d = *descriptor;
    // a built-in (GrB_DESC_*) keeps its handle, as in the reference
    if (d && d->magic == GB_MAGIC && !d->builtin)
    {
        d->magic = GB_FREED;
        free(d);
        *descriptor = NULL;
    }
    return GrB_SUCCESS;
}
