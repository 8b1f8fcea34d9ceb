//------------------------------------------------------------------------------
// GraphBLAS.h: the GrB_mxm drop-in surface of the MI355X library
//------------------------------------------------------------------------------

// A subset of the GraphBLAS C API (SuiteSparse:GraphBLAS v9.3.1 names, enum
// values and argument meanings), enough to build CSR matrices, call GrB_mxm
// with the built-in semirings the gfx950 kernels implement, reduce the result
// and read it back.  Matrices are stored by row (CSR, int64 indices) in HBM.
//
// Reference declarations mirrored (Include/GraphBLAS.h of the reference):
//   GrB_Info codes                   :358-381
//   GrB_Desc_Field / GrB_Desc_Value  :490-534 and predefined descriptors :549-589
//   GrB_mxm                          :5749-5758
//   GxB_Matrix_pack_CSR / unpack_CSR :9449-9462, 9976
//   GrB_Matrix_build_<T>             :4418-4426
//   GrB_Matrix_reduce_<T>            :8924
//   GrB_init / GrB_finalize          (Source/init/GrB_init.c:37-47)
// Everything else in the reference's API is outside this library.

#ifndef GRAPHBLAS_H
#define GRAPHBLAS_H

#include <stdint.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GxB_IMPLEMENTATION_NAME "SuiteSparse:GraphBLAS GrB_mxm subset for MI355X (gfx950)"
#define GxB_IMPLEMENTATION_MAJOR 9
#define GxB_IMPLEMENTATION_MINOR 3
#define GxB_IMPLEMENTATION_SUB 1

typedef uint64_t GrB_Index;
#define GrB_INDEX_MAX ((GrB_Index) (1ULL << 60) - 1)

typedef enum
{
    GrB_SUCCESS = 0,
    GrB_NO_VALUE = 1,
    GxB_EXHAUSTED = 7089,
    GrB_UNINITIALIZED_OBJECT = -1,
    GrB_NULL_POINTER = -2,
    GrB_INVALID_VALUE = -3,
    GrB_INVALID_INDEX = -4,
    GrB_DOMAIN_MISMATCH = -5,
    GrB_DIMENSION_MISMATCH = -6,
    GrB_OUTPUT_NOT_EMPTY = -7,
    GrB_NOT_IMPLEMENTED = -8,
    GrB_ALREADY_SET = -9,
    GrB_PANIC = -101,
    GrB_OUT_OF_MEMORY = -102,
    GrB_INSUFFICIENT_SPACE = -103,
    GrB_INVALID_OBJECT = -104,
    GrB_INDEX_OUT_OF_BOUNDS = -105,
    GrB_EMPTY_OBJECT = -106
}
GrB_Info;

typedef enum { GrB_NONBLOCKING = 0, GrB_BLOCKING = 1 } GrB_Mode;

GrB_Info GrB_init(GrB_Mode mode);
GrB_Info GrB_finalize(void);

//------------------------------------------------------------------------------
// objects
//------------------------------------------------------------------------------

typedef struct GB_Type_opaque *GrB_Type;
typedef struct GB_BinaryOp_opaque *GrB_BinaryOp;
typedef struct GB_Monoid_opaque *GrB_Monoid;
typedef struct GB_Semiring_opaque *GrB_Semiring;
typedef struct GB_Descriptor_opaque *GrB_Descriptor;
typedef struct GB_Matrix_opaque *GrB_Matrix;
typedef struct GB_Vector_opaque *GrB_Vector;

extern GrB_Type GrB_BOOL, GrB_INT8, GrB_UINT8, GrB_INT16, GrB_UINT16, GrB_INT32,
    GrB_UINT32, GrB_INT64, GrB_UINT64, GrB_FP32, GrB_FP64;

GrB_Info GrB_Type_get_size(size_t *size, GrB_Type type);   // helper
GrB_Info GxB_Type_size(size_t *size, GrB_Type type);

// binary operators (for semirings, monoids, build's dup and accum)
#define GB_OPS_T(X, T)                                                            \
    X(GrB_PLUS_##T) X(GrB_MINUS_##T) X(GrB_TIMES_##T) X(GrB_MIN_##T) X(GrB_MAX_##T) \
    X(GrB_FIRST_##T) X(GrB_SECOND_##T) X(GxB_ANY_##T) X(GxB_PAIR_##T) X(GrB_ONEB_##T)
#define GB_DECLARE_OP(name) extern GrB_BinaryOp name;
GB_OPS_T(GB_DECLARE_OP, INT32)
GB_OPS_T(GB_DECLARE_OP, INT64)
GB_OPS_T(GB_DECLARE_OP, FP64)
#undef GB_DECLARE_OP
extern GrB_BinaryOp GrB_LOR, GrB_LAND, GxB_ANY_BOOL, GxB_PAIR_BOOL, GrB_ONEB_BOOL,
    GrB_FIRST_BOOL, GrB_SECOND_BOOL;
// positional multipliers (value is an index; operand values are ignored)
extern GrB_BinaryOp GxB_FIRSTI_INT64, GxB_FIRSTJ_INT64, GxB_SECONDI_INT64, GxB_SECONDJ_INT64,
    GxB_FIRSTI_INT32, GxB_FIRSTJ_INT32, GxB_SECONDI_INT32, GxB_SECONDJ_INT32;

GrB_Info GrB_BinaryOp_free(GrB_BinaryOp *op);

// monoids
extern GrB_Monoid GrB_PLUS_MONOID_INT32, GrB_PLUS_MONOID_INT64, GrB_PLUS_MONOID_FP64,
    GrB_MIN_MONOID_INT32, GrB_MIN_MONOID_INT64, GrB_MIN_MONOID_FP64,
    GrB_MAX_MONOID_INT32, GrB_MAX_MONOID_INT64, GrB_MAX_MONOID_FP64,
    GxB_ANY_INT32_MONOID, GxB_ANY_INT64_MONOID, GxB_ANY_FP64_MONOID, GxB_ANY_BOOL_MONOID,
    GrB_LOR_MONOID_BOOL, GrB_LAND_MONOID_BOOL;

GrB_Info GrB_Monoid_free(GrB_Monoid *monoid);

// semirings: the built-ins the gfx950 kernels implement
extern GrB_Semiring GxB_PLUS_PAIR_INT64, GxB_PLUS_PAIR_INT32, GxB_ANY_PAIR_BOOL,
    GrB_PLUS_TIMES_SEMIRING_FP64, GrB_PLUS_TIMES_SEMIRING_INT64,
    GrB_MIN_PLUS_SEMIRING_INT32, GrB_MIN_PLUS_SEMIRING_INT64, GrB_MIN_PLUS_SEMIRING_FP64,
    GrB_MAX_PLUS_SEMIRING_INT32, GxB_ANY_SECONDI_INT64, GxB_ANY_FIRSTJ_INT64,
    GxB_LOR_PAIR_BOOL, GxB_PLUS_ONEB_INT64;

GrB_Info GrB_Semiring_new(GrB_Semiring *semiring, GrB_Monoid add, GrB_BinaryOp multiply);
GrB_Info GrB_Semiring_free(GrB_Semiring *semiring);

//------------------------------------------------------------------------------
// descriptors
//------------------------------------------------------------------------------

typedef enum
{
    GrB_OUTP = 0,
    GrB_MASK = 1,
    GrB_INP0 = 2,
    GrB_INP1 = 3,
    GxB_AxB_METHOD = 7090,
    GxB_SORT = 7091,
    GxB_COMPRESSION = 7092,
    GxB_IMPORT = 7093
}
GrB_Desc_Field;

typedef enum
{
    GrB_DEFAULT = 0,
    GxB_DEFAULT = 0,
    GrB_REPLACE = 1,
    GrB_COMP = 2,
    GrB_STRUCTURE = 4,
    GrB_COMP_STRUCTURE = 6,
    GrB_TRAN = 3,
    GxB_AxB_GUSTAVSON = 7081,
    GxB_AxB_DOT = 7083,
    GxB_AxB_HASH = 7084,
    GxB_AxB_SAXPY = 7085,
    GxB_SECURE_IMPORT = 7080
}
GrB_Desc_Value;

#define GxB_FAST_IMPORT ((int) GrB_DEFAULT)

// predefined descriptors: R = replace, S = structural mask, C = complemented
// mask, T0/T1 = transpose the first/second input
extern GrB_Descriptor
    GrB_DESC_T1, GrB_DESC_T0, GrB_DESC_T0T1,
    GrB_DESC_C, GrB_DESC_CT1, GrB_DESC_CT0, GrB_DESC_CT0T1,
    GrB_DESC_S, GrB_DESC_ST1, GrB_DESC_ST0, GrB_DESC_ST0T1,
    GrB_DESC_SC, GrB_DESC_SCT1, GrB_DESC_SCT0, GrB_DESC_SCT0T1,
    GrB_DESC_R, GrB_DESC_RT1, GrB_DESC_RT0, GrB_DESC_RT0T1,
    GrB_DESC_RC, GrB_DESC_RCT1, GrB_DESC_RCT0, GrB_DESC_RCT0T1,
    GrB_DESC_RS, GrB_DESC_RST1, GrB_DESC_RST0, GrB_DESC_RST0T1,
    GrB_DESC_RSC, GrB_DESC_RSCT1, GrB_DESC_RSCT0, GrB_DESC_RSCT0T1;

GrB_Info GrB_Descriptor_new(GrB_Descriptor *descriptor);
GrB_Info GrB_Descriptor_set(GrB_Descriptor desc, GrB_Desc_Field field, GrB_Desc_Value val);
GrB_Info GrB_Descriptor_free(GrB_Descriptor *descriptor);

//------------------------------------------------------------------------------
// matrices (CSR in HBM)
//------------------------------------------------------------------------------

GrB_Info GrB_Matrix_new(GrB_Matrix *A, GrB_Type type, GrB_Index nrows, GrB_Index ncols);
GrB_Info GrB_Matrix_dup(GrB_Matrix *C, const GrB_Matrix A);
GrB_Info GrB_Matrix_clear(GrB_Matrix A);
GrB_Info GrB_Matrix_free(GrB_Matrix *A);
GrB_Info GrB_Matrix_nrows(GrB_Index *nrows, const GrB_Matrix A);
GrB_Info GrB_Matrix_ncols(GrB_Index *ncols, const GrB_Matrix A);
GrB_Info GrB_Matrix_nvals(GrB_Index *nvals, const GrB_Matrix A);
GrB_Info GxB_Matrix_type(GrB_Type *type, const GrB_Matrix A);
GrB_Info GxB_Matrix_iso(bool *iso, const GrB_Matrix A);

GrB_Info GrB_Matrix_build_BOOL(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                               const bool *X, GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Matrix_build_INT32(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                                const int32_t *X, GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Matrix_build_INT64(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                                const int64_t *X, GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Matrix_build_FP64(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                               const double *X, GrB_Index nvals, const GrB_BinaryOp dup);

GrB_Info GrB_Matrix_extractTuples_BOOL(GrB_Index *I, GrB_Index *J, bool *X,
                                       GrB_Index *nvals, const GrB_Matrix A);
GrB_Info GrB_Matrix_extractTuples_INT32(GrB_Index *I, GrB_Index *J, int32_t *X,
                                        GrB_Index *nvals, const GrB_Matrix A);
GrB_Info GrB_Matrix_extractTuples_INT64(GrB_Index *I, GrB_Index *J, int64_t *X,
                                        GrB_Index *nvals, const GrB_Matrix A);
GrB_Info GrB_Matrix_extractTuples_FP64(GrB_Index *I, GrB_Index *J, double *X,
                                       GrB_Index *nvals, const GrB_Matrix A);

// Move CSR arrays into / out of a matrix.  The host arrays are malloc'd
// memory: pack copies them into HBM and frees them (setting *Ap etc. to
// NULL); unpack returns malloc'd host copies and leaves A empty.
GrB_Info GxB_Matrix_pack_CSR(GrB_Matrix A, GrB_Index **Ap, GrB_Index **Aj, void **Ax,
                             GrB_Index Ap_size, GrB_Index Aj_size, GrB_Index Ax_size,
                             bool iso, bool jumbled, const GrB_Descriptor desc);
GrB_Info GxB_Matrix_unpack_CSR(GrB_Matrix A, GrB_Index **Ap, GrB_Index **Aj, void **Ax,
                               GrB_Index *Ap_size, GrB_Index *Aj_size, GrB_Index *Ax_size,
                               bool *iso, bool *jumbled, const GrB_Descriptor desc);

// Serialize to / from a blob (Include/GraphBLAS.h:10917-10990 of the
// reference), in the reference's blob layout (Source/serialize/GB_serialize.c).
// Written: sparse by row, every array one uncompressed block.  Read: sparse,
// hypersparse, bitmap or full, by row or by column, uncompressed or LZ4 /
// LZ4HC blocks; ZSTD blocks give GrB_NOT_IMPLEMENTED.
GrB_Info GxB_Matrix_serialize(void **blob_handle, GrB_Index *blob_size_handle, GrB_Matrix A,
                              const GrB_Descriptor desc);
GrB_Info GrB_Matrix_serialize(void *blob, GrB_Index *blob_size_handle, GrB_Matrix A);
GrB_Info GrB_Matrix_serializeSize(GrB_Index *blob_size_handle, GrB_Matrix A);
GrB_Info GxB_Matrix_deserialize(GrB_Matrix *C, GrB_Type type, const void *blob,
                                GrB_Index blob_size, const GrB_Descriptor desc);
GrB_Info GrB_Matrix_deserialize(GrB_Matrix *C, GrB_Type type, const void *blob,
                                GrB_Index blob_size);

//------------------------------------------------------------------------------
// sparse vectors (Include/GraphBLAS.h:3623-3670, 3745, 4197 of the reference)
//------------------------------------------------------------------------------

GrB_Info GrB_Vector_new(GrB_Vector *v, GrB_Type type, GrB_Index n);
GrB_Info GrB_Vector_dup(GrB_Vector *w, const GrB_Vector u);
GrB_Info GrB_Vector_clear(GrB_Vector v);
GrB_Info GrB_Vector_size(GrB_Index *n, const GrB_Vector v);
GrB_Info GrB_Vector_nvals(GrB_Index *nvals, const GrB_Vector v);
GrB_Info GrB_Vector_free(GrB_Vector *v);
GrB_Info GrB_Vector_build_BOOL(GrB_Vector w, const GrB_Index *I, const bool *X,
                               GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Vector_build_INT32(GrB_Vector w, const GrB_Index *I, const int32_t *X,
                                GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Vector_build_INT64(GrB_Vector w, const GrB_Index *I, const int64_t *X,
                                GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Vector_build_FP64(GrB_Vector w, const GrB_Index *I, const double *X,
                               GrB_Index nvals, const GrB_BinaryOp dup);
GrB_Info GrB_Vector_extractTuples_BOOL(GrB_Index *I, bool *X, GrB_Index *nvals,
                                       const GrB_Vector v);
GrB_Info GrB_Vector_extractTuples_INT32(GrB_Index *I, int32_t *X, GrB_Index *nvals,
                                        const GrB_Vector v);
GrB_Info GrB_Vector_extractTuples_INT64(GrB_Index *I, int64_t *X, GrB_Index *nvals,
                                        const GrB_Vector v);
GrB_Info GrB_Vector_extractTuples_FP64(GrB_Index *I, double *X, GrB_Index *nvals,
                                       const GrB_Vector v);

//------------------------------------------------------------------------------
// C<Mask> = accum (C, A*B): the hot path
//------------------------------------------------------------------------------

GrB_Info GrB_mxm
(
    GrB_Matrix C,                   // input/output matrix for results
    const GrB_Matrix Mask,          // optional mask for C, unused if NULL
    const GrB_BinaryOp accum,       // optional accum for Z=accum(C,T)
    const GrB_Semiring semiring,    // defines '+' and '*' for A*B
    const GrB_Matrix A,             // first input:  matrix A
    const GrB_Matrix B,             // second input: matrix B
    const GrB_Descriptor desc       // descriptor for C, Mask, A, and B
);

// w'<mask'> = accum (w', u'*A) and w<mask> = accum (w, A*u)
// (Include/GraphBLAS.h:5760-5780 of the reference); they run on GrB_mxm's
// kernels with the vectors as 1 x n matrices
GrB_Info GrB_vxm(GrB_Vector w, const GrB_Vector mask, const GrB_BinaryOp accum,
                 const GrB_Semiring semiring, const GrB_Vector u, const GrB_Matrix A,
                 const GrB_Descriptor desc);
GrB_Info GrB_mxv(GrB_Vector w, const GrB_Vector mask, const GrB_BinaryOp accum,
                 const GrB_Semiring semiring, const GrB_Matrix A, const GrB_Vector u,
                 const GrB_Descriptor desc);

//------------------------------------------------------------------------------
// reduce a matrix to a scalar: c = accum (c, reduce (A))
//------------------------------------------------------------------------------

GrB_Info GrB_Matrix_reduce_INT64(int64_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Matrix A, const GrB_Descriptor desc);
GrB_Info GrB_Matrix_reduce_INT32(int32_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Matrix A, const GrB_Descriptor desc);
GrB_Info GrB_Matrix_reduce_FP64(double *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Matrix A, const GrB_Descriptor desc);
GrB_Info GrB_Matrix_reduce_BOOL(bool *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Matrix A, const GrB_Descriptor desc);
GrB_Info GrB_Vector_reduce_INT64(int64_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Vector u, const GrB_Descriptor desc);
GrB_Info GrB_Vector_reduce_INT32(int32_t *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                 const GrB_Vector u, const GrB_Descriptor desc);
GrB_Info GrB_Vector_reduce_FP64(double *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Vector u, const GrB_Descriptor desc);
GrB_Info GrB_Vector_reduce_BOOL(bool *c, const GrB_BinaryOp accum, const GrB_Monoid monoid,
                                const GrB_Vector u, const GrB_Descriptor desc);

//------------------------------------------------------------------------------
// errors and global options
//------------------------------------------------------------------------------

GrB_Info GrB_Matrix_error(const char **error, const GrB_Matrix A);
GrB_Info GrB_Vector_error(const char **error, const GrB_Vector v);
#if defined(__cplusplus)
#define GrB_error(error, object) GrB_Matrix_error(error, object)
#else
#define GrB_error(error, object) \
    _Generic((object), GrB_Vector: GrB_Vector_error, default: GrB_Matrix_error)(error, object)
#endif

//------------------------------------------------------------------------------
// storage formats: sparse (CSR), bitmap and full, all held by row
// (the reference's GxB_Option_Field values and sparsity bits,
// Include/GraphBLAS.h:1389-1495; pack/unpack at :9686-9775, :10110-10167)
//------------------------------------------------------------------------------

typedef enum
{
    GxB_HYPER_SWITCH = 7000,        // accepted and ignored (no hypersparse format)
    GxB_BITMAP_SWITCH = 7001,       // accepted and ignored
    GxB_FORMAT = 7002,              // GxB_BY_ROW only
    GxB_SPARSITY_STATUS = 7034,     // get: GxB_SPARSE, GxB_BITMAP or GxB_FULL
    GxB_SPARSITY_CONTROL = 7036     // set/get: allowed formats (sum of the bits below)
} GxB_Option_Field;

typedef enum { GxB_BY_ROW = 0, GxB_BY_COL = 1, GxB_NO_FORMAT = -1 } GxB_Format_Value;

// serialize compression methods (GrB_Matrix_serialize writes uncompressed
// blocks; deserialize reads uncompressed and LZ4 / LZ4HC blocks)
#define GxB_COMPRESSION_NONE -1
#define GxB_COMPRESSION_DEFAULT 0
#define GxB_COMPRESSION_LZ4 1000
#define GxB_COMPRESSION_LZ4HC 2000
#define GxB_COMPRESSION_ZSTD 3000

#define GxB_HYPERSPARSE 1   // treated as sparse
#define GxB_SPARSE      2
#define GxB_BITMAP      4
#define GxB_FULL        8
#define GxB_ANY_SPARSITY (GxB_HYPERSPARSE + GxB_SPARSE + GxB_BITMAP + GxB_FULL)
#define GxB_AUTO_SPARSITY GxB_ANY_SPARSITY

GrB_Info GxB_Matrix_Option_set(GrB_Matrix A, GxB_Option_Field field, ...);
GrB_Info GxB_Matrix_Option_get(GrB_Matrix A, GxB_Option_Field field, ...);
GrB_Info GxB_Matrix_Option_set_INT32(GrB_Matrix A, GxB_Option_Field field, int32_t value);
GrB_Info GxB_Matrix_Option_get_INT32(GrB_Matrix A, GxB_Option_Field field, int32_t *value);

GrB_Info GxB_Matrix_pack_BitmapR(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index Ab_size,
                                 GrB_Index Ax_size, bool iso, GrB_Index nvals,
                                 const GrB_Descriptor desc);
GrB_Info GxB_Matrix_pack_FullR(GrB_Matrix A, void **Ax, GrB_Index Ax_size, bool iso,
                               const GrB_Descriptor desc);
GrB_Info GxB_Matrix_unpack_BitmapR(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index *Ab_size,
                                   GrB_Index *Ax_size, bool *iso, GrB_Index *nvals,
                                   const GrB_Descriptor desc);
GrB_Info GxB_Matrix_unpack_FullR(GrB_Matrix A, void **Ax, GrB_Index *Ax_size, bool *iso,
                                 const GrB_Descriptor desc);

#define GxB_BURBLE 7019   // GxB_Option_Field value in the reference
GrB_Info GxB_Global_Option_set_INT32(int field, int32_t value);
GrB_Info GxB_Global_Option_get_INT32(int field, int32_t *value);

#ifdef __cplusplus
}
#endif
#endif
