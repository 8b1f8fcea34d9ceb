// grb_internal.h: object model of the host library behind GraphBLAS.h
#ifndef GRB_INTERNAL_H
#define GRB_INTERNAL_H

#include "GraphBLAS.h"
#include "gm_kernels.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GB_MAGIC 0x72657473786f62ULL      // "boxster"
#define GB_FREED 0x6c6c756e6c6c756eULL

// operator codes (the subset of the reference's GB_Opcode this library uses)
typedef enum
{
    GB_OP_PLUS = 1, GB_OP_MINUS, GB_OP_TIMES, GB_OP_MIN, GB_OP_MAX, GB_OP_FIRST,
    GB_OP_SECOND, GB_OP_ANY, GB_OP_PAIR, GB_OP_LOR, GB_OP_LAND,
    GB_OP_FIRSTI, GB_OP_FIRSTJ, GB_OP_SECONDI, GB_OP_SECONDJ
} GB_opcode;

struct GB_Type_opaque
{
    uint64_t magic;
    int32_t code;           // gm_type
    size_t size;
    const char *name;
};

struct GB_BinaryOp_opaque
{
    uint64_t magic;
    GB_opcode opcode;
    GrB_Type xtype, ytype, ztype;
    const char *name;
    bool builtin;
};

struct GB_Monoid_opaque
{
    uint64_t magic;
    GrB_BinaryOp op;
    const char *name;
    bool builtin;
};

struct GB_Semiring_opaque
{
    uint64_t magic;
    GrB_Monoid add;
    GrB_BinaryOp mult;
    const char *name;
    bool builtin;
};

struct GB_Descriptor_opaque
{
    uint64_t magic;
    int out;        // GrB_DEFAULT or GrB_REPLACE
    int mask;       // 0, GrB_COMP, GrB_STRUCTURE or both
    int in0, in1;   // GrB_DEFAULT or GrB_TRAN
    int axb;        // GxB_AxB_* or default
    int sort;
    bool builtin;
};

// A matrix in HBM, held by row in one of three formats (sparsity):
//   GxB_SPARSE  CSR in p/j/x; p == NULL means "no entries" (allocated on
//               first use)
//   GxB_BITMAP  b (one byte per entry position) and xd (nrows*ncols values)
//   GxB_FULL    xd only, every entry present
// For bitmap and full, p/j/x hold a CSR copy of the same content built on
// first use by an operation that reads CSR (GB_matrix_csr), or are NULL.
struct GB_Matrix_opaque
{
    uint64_t magic;
    GrB_Type type;
    int64_t nrows, ncols, nvals;
    int64_t *p, *j;     // device
    void *x;            // device (one value if iso)
    bool iso;
    int sparsity;           // GxB_SPARSE, GxB_BITMAP or GxB_FULL
    int sparsity_control;   // GxB_SPARSITY_CONTROL hint (default GxB_AUTO_SPARSITY)
    int8_t *b;              // device bitmap (GxB_BITMAP)
    void *xd;               // device values by row (GxB_BITMAP, GxB_FULL; one if iso)
    // derived forms of the current content (gm_csr_aux_build), built on first
    // use by GrB_mxm and dropped with the content (GB_matrix_free_content)
    gm_csr_aux *aux;
    bool aux_tried;
    char err[384];
};

// runtime helpers
extern int GB_burble;
bool GB_initialized(void);
GrB_Info GB_from_gm(int rc);
GrB_Info GB_error(GrB_Matrix C, GrB_Info info, const char *fmt, ...);
GrB_Info GB_check_matrix(const GrB_Matrix A);
GrB_Info GB_matrix_set_csr(GrB_Matrix A, const gm_csr_out *R);
void GB_matrix_free_content(GrB_Matrix A);
GrB_Info GB_matrix_view(gm_csr *v, GrB_Matrix A);
GrB_Info GB_matrix_csr(GrB_Matrix A);
void GB_matrix_dense_view(gm_dense *d, const GrB_Matrix A);
GrB_Info GB_matrix_set_dense(GrB_Matrix A, const gm_dense_out *R);
GrB_Info GB_matrix_conform(GrB_Matrix A);
bool GB_matrix_is_dense(const GrB_Matrix A);
void GB_matrix_aux(gm_csr *v, GrB_Matrix A);
GrB_Type GB_type_from_code(int32_t code);

// method selection (pure host logic, exported for tests)
typedef enum { GB_METHOD_NONE = 0, GB_METHOD_DOT3 = 1, GB_METHOD_SAXPY3 = 2, GB_METHOD_DOT2 = 3 } GB_method;
GB_method GB_mxm_select(bool has_mask, bool mask_comp, bool a_tran, bool b_tran, int axb);
bool GB_mxm_prefer_dot3(int64_t flops, int64_t nnz_m, int64_t nnz_a, int64_t m, int64_t nnz_b,
                        int64_t n);
bool GB_AxB_dot2_control(double anz, double bnz, double anvec, double bnvec, double avlen);
bool GB_semiring_codes(const GrB_Semiring s, gm_semiring *out);

#endif
