// grb_matrix.c: CSR matrices held in HBM (new/free/dup/build/extract/pack).
//
// Construction and extraction run on the host (sort + combine of tuples, as
// GrB_Matrix_build does on the CPU in the reference, Source/builder) and move
// the CSR arrays across PCIe once; every GrB_mxm then works on the
// device-resident arrays.

#include "grb_internal.h"
#include <stdarg.h>

GrB_Info GB_check_matrix(const GrB_Matrix A)
{
    if (!A) return GrB_NULL_POINTER;
    if (A->magic == GB_FREED) return GrB_INVALID_OBJECT;
    if (A->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    return GrB_SUCCESS;
}

// drops the CSR arrays (and their derived forms) only
static void GB_matrix_free_csr(GrB_Matrix A)
{
    gm_csr_aux_free(A->aux);
    A->aux = NULL;
    A->aux_tried = false;
    gm_free(A->p);
    gm_free(A->j);
    gm_free(A->x);
    A->p = NULL;
    A->j = NULL;
    A->x = NULL;
}

void GB_matrix_free_content(GrB_Matrix A)
{
    GB_matrix_free_csr(A);
    gm_free(A->b);
    gm_free(A->xd);
    A->b = NULL;
    A->xd = NULL;
    A->sparsity = GxB_SPARSE;
    A->nvals = 0;
    A->iso = false;
}

GrB_Info GB_matrix_set_csr(GrB_Matrix A, const gm_csr_out *R)
{
    GB_matrix_free_content(A);
    A->p = R->p;
    A->j = R->j;
    A->x = R->x;
    A->nvals = R->nvals;
    A->iso = R->iso != 0;
    return GrB_SUCCESS;
}

bool GB_matrix_is_dense(const GrB_Matrix A)
{
    return A->sparsity == GxB_BITMAP || A->sparsity == GxB_FULL;
}

void GB_matrix_dense_view(gm_dense *d, const GrB_Matrix A)
{
    d->nrows = A->nrows;
    d->ncols = A->ncols;
    d->b = A->sparsity == GxB_BITMAP ? A->b : NULL;
    d->x = A->xd;
    d->iso = A->iso ? 1 : 0;
    d->type = A->type->code;
}

// A takes a bitmap (R->b set) or full result; R's arrays move into A
GrB_Info GB_matrix_set_dense(GrB_Matrix A, const gm_dense_out *R)
{
    GB_matrix_free_content(A);
    A->b = R->b;
    A->xd = R->x;
    A->nvals = R->nvals;
    A->iso = R->iso != 0;
    A->sparsity = R->b ? GxB_BITMAP : GxB_FULL;
    return GrB_SUCCESS;
}

// the CSR form of a bitmap or full matrix, built once per content
GrB_Info GB_matrix_csr(GrB_Matrix A)
{
    if (!GB_matrix_is_dense(A) || A->p != NULL) return GrB_SUCCESS;
    gm_dense d;
    GB_matrix_dense_view(&d, A);
    gm_csr_out R;
    memset(&R, 0, sizeof(R));
    int rc = gm_dense_to_csr(&R, &d);
    if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
    A->p = R.p;
    A->j = R.j;
    A->x = R.x;
    return GrB_SUCCESS;
}

// Stores A in a format its GxB_SPARSITY_CONTROL allows (the reference's
// GB_conform, Source/convert/GB_conform.c): a matrix with every entry present
// goes full when full is allowed and its format is not; a format that is not
// allowed moves to bitmap, else sparse; with only GxB_FULL allowed, a matrix
// with entries missing is kept bitmap, as the reference does.
GrB_Info GB_matrix_conform(GrB_Matrix A)
{
    int ctrl = A->sparsity_control & GxB_ANY_SPARSITY;
    if (ctrl == 0) ctrl = GxB_AUTO_SPARSITY;
    const bool sparse_ok = (ctrl & (GxB_SPARSE | GxB_HYPERSPARSE)) != 0;
    const bool bitmap_ok = (ctrl & GxB_BITMAP) != 0;
    const bool full_ok = (ctrl & GxB_FULL) != 0;
    const bool all = A->nvals == A->nrows * A->ncols;
    int want = A->sparsity;
    if (A->sparsity == GxB_SPARSE && !sparse_ok)
        want = (full_ok && all) ? GxB_FULL : GxB_BITMAP;
    else if (A->sparsity == GxB_BITMAP && !bitmap_ok)
        want = (full_ok && all) ? GxB_FULL : (sparse_ok ? GxB_SPARSE : GxB_BITMAP);
    else if (A->sparsity == GxB_FULL && !full_ok)
        want = bitmap_ok ? GxB_BITMAP : GxB_SPARSE;
    if (want == A->sparsity) return GrB_SUCCESS;
    GrB_Info info;
    if (want == GxB_SPARSE)
    {
        if ((info = GB_matrix_csr(A)) != GrB_SUCCESS) return info;
        gm_free(A->b);
        gm_free(A->xd);
        A->b = NULL;
        A->xd = NULL;
        A->sparsity = GxB_SPARSE;
        return GrB_SUCCESS;
    }
    if (A->sparsity == GxB_FULL && want == GxB_BITMAP)
    {
        size_t mn = (size_t) (A->nrows * A->ncols);
        A->b = (int8_t *) gm_malloc(mn > 0 ? mn : 1);
        if (!A->b) return GB_error(A, GrB_OUT_OF_MEMORY, "conform: bitmap");
        int rc = gm_memset(A->b, 1, mn);
        if (rc == GM_SUCCESS) rc = gm_synchronize();
        if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
        A->sparsity = GxB_BITMAP;
        return GrB_SUCCESS;
    }
    if (A->sparsity == GxB_BITMAP && want == GxB_FULL)
    {
        gm_free(A->b);
        A->b = NULL;
        A->sparsity = GxB_FULL;
        return GrB_SUCCESS;
    }
    // sparse -> bitmap / full
    gm_csr v;
    if ((info = GB_matrix_view(&v, A)) != GrB_SUCCESS) return info;
    gm_dense_out R;
    memset(&R, 0, sizeof(R));
    int rc = gm_csr_to_dense(&R, &v, want == GxB_FULL);
    if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
    const int64_t nv = A->nvals;
    GB_matrix_free_csr(A);
    A->b = R.b;
    A->xd = R.x;
    A->nvals = nv;
    A->sparsity = want;
    return GrB_SUCCESS;
}

// Device view of A for the kernel layer; an empty matrix gets zeroed row
// pointers on first use.
GrB_Info GB_matrix_view(gm_csr *v, GrB_Matrix A)
{
    if (GB_matrix_is_dense(A))
    {
        GrB_Info info = GB_matrix_csr(A);
        if (info != GrB_SUCCESS) return info;
    }
    if (A->p == NULL)
    {
        size_t bytes = sizeof(int64_t) * (size_t) (A->nrows + 1);
        A->p = (int64_t *) gm_malloc(bytes);
        if (!A->p) return GB_error(A, GB_from_gm(gm_init()) == GrB_SUCCESS ? GrB_OUT_OF_MEMORY
                                                                          : GrB_PANIC,
                                   "%s", gm_last_error());
        int rc = gm_memset(A->p, 0, bytes);
        if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
        A->nvals = 0;
    }
    v->nrows = A->nrows;
    v->ncols = A->ncols;
    v->p = A->p;
    v->j = A->j;
    v->x = A->x;
    v->iso = A->iso ? 1 : 0;
    v->type = A->type->code;
    v->aux = A->aux;
    return GrB_SUCCESS;
}

// Builds A's derived forms once for its current content (a failure only
// means none: the kernels then read p/j/x) and puts them in the view v.
void GB_matrix_aux(gm_csr *v, GrB_Matrix A)
{
    if (!A->aux_tried)
    {
        A->aux_tried = true;
        gm_csr_aux *aux = NULL;
        if (gm_csr_aux_build(&aux, v) == GM_SUCCESS) A->aux = aux;
    }
    v->aux = A->aux;
}

GrB_Info GrB_Matrix_new(GrB_Matrix *A, GrB_Type type, GrB_Index nrows, GrB_Index ncols)
{
    if (!GB_initialized()) return GrB_PANIC;
    if (!A || !type) return GrB_NULL_POINTER;
    *A = NULL;
    if (type->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    if (nrows > GrB_INDEX_MAX || ncols > GrB_INDEX_MAX) return GrB_INVALID_VALUE;
    GrB_Matrix M = (GrB_Matrix) calloc(1, sizeof(struct GB_Matrix_opaque));
    if (!M) return GrB_OUT_OF_MEMORY;
    M->magic = GB_MAGIC;
    M->type = type;
    M->nrows = (int64_t) nrows;
    M->ncols = (int64_t) ncols;
    M->sparsity = GxB_SPARSE;
    M->sparsity_control = GxB_AUTO_SPARSITY;
    *A = M;
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_free(GrB_Matrix *A)
{
    if (!A || !*A) return GrB_SUCCESS;
    GrB_Matrix M = *A;
    if (M->magic == GB_MAGIC)
    {
        GB_matrix_free_content(M);
        M->magic = GB_FREED;
        free(M);
    }
    *A = NULL;
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_clear(GrB_Matrix A)
{
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    GB_matrix_free_content(A);
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_nrows(GrB_Index *nrows, const GrB_Matrix A)
{
    if (!nrows) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    *nrows = (GrB_Index) A->nrows;
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_ncols(GrB_Index *ncols, const GrB_Matrix A)
{
    if (!ncols) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    *ncols = (GrB_Index) A->ncols;
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_nvals(GrB_Index *nvals, const GrB_Matrix A)
{
    if (!nvals) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    *nvals = (GrB_Index) A->nvals;
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_type(GrB_Type *type, const GrB_Matrix A)
{
    if (!type) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    *type = A->type;
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_iso(bool *iso, const GrB_Matrix A)
{
    if (!iso) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    *iso = A->iso;
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_dup(GrB_Matrix *C, const GrB_Matrix A)
{
    if (!C) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    info = GrB_Matrix_new(C, A->type, (GrB_Index) A->nrows, (GrB_Index) A->ncols);
    if (info != GrB_SUCCESS) return info;
    GrB_Matrix D = *C;
    D->sparsity_control = A->sparsity_control;
    if (GB_matrix_is_dense(A))
    {
        // bitmap / full: the by-row arrays
        const size_t mn = (size_t) (A->nrows * A->ncols);
        const size_t xb = A->type->size * (A->iso ? 1 : (mn > 0 ? mn : 1));
        D->xd = gm_malloc(xb);
        D->b = A->sparsity == GxB_BITMAP ? (int8_t *) gm_malloc(mn > 0 ? mn : 1) : NULL;
        if (!D->xd || (A->sparsity == GxB_BITMAP && !D->b))
        {
            GrB_Matrix_free(C);
            return GrB_OUT_OF_MEMORY;
        }
        int rc = gm_memcpy_d2d(D->xd, A->xd, xb);
        if (rc == GM_SUCCESS && D->b) rc = gm_memcpy_d2d(D->b, A->b, mn);
        if (rc == GM_SUCCESS) rc = gm_synchronize();
        if (rc != GM_SUCCESS)
        {
            GrB_Matrix_free(C);
            return GB_from_gm(rc);
        }
        D->sparsity = A->sparsity;
        D->nvals = A->nvals;
        D->iso = A->iso;
        return GrB_SUCCESS;
    }
    if (A->p == NULL) return GrB_SUCCESS;
    size_t pb = sizeof(int64_t) * (size_t) (A->nrows + 1);
    size_t jb = sizeof(int64_t) * (size_t) (A->nvals > 0 ? A->nvals : 1);
    size_t xb = A->type->size * (size_t) (A->iso ? 1 : (A->nvals > 0 ? A->nvals : 1));
    D->p = (int64_t *) gm_malloc(pb);
    D->j = (int64_t *) gm_malloc(jb);
    D->x = A->x ? gm_malloc(xb) : NULL;
    if (!D->p || !D->j || (A->x && !D->x))
    {
        GrB_Matrix_free(C);
        return GrB_OUT_OF_MEMORY;
    }
    int rc = gm_memcpy_d2d(D->p, A->p, pb);
    if (rc == GM_SUCCESS && A->j) rc = gm_memcpy_d2d(D->j, A->j, jb);
    if (rc == GM_SUCCESS && A->x) rc = gm_memcpy_d2d(D->x, A->x, xb);
    if (rc == GM_SUCCESS) rc = gm_synchronize();
    if (rc != GM_SUCCESS)
    {
        GrB_Matrix_free(C);
        return GB_from_gm(rc);
    }
    D->nvals = A->nvals;
    D->iso = A->iso;
    return GrB_SUCCESS;
}

//------------------------------------------------------------------------------
// upload / download helpers
//------------------------------------------------------------------------------

// upload host CSR (sorted rows, no duplicates) into A
static GrB_Info GB_upload(GrB_Matrix A, const int64_t *Ap, const int64_t *Aj, const void *Ax,
                          int64_t nvals, bool iso)
{
    size_t pb = sizeof(int64_t) * (size_t) (A->nrows + 1);
    size_t jb = sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1);
    size_t xb = A->type->size * (size_t) (iso ? 1 : (nvals > 0 ? nvals : 1));
    int64_t *dp = (int64_t *) gm_malloc(pb);
    int64_t *dj = (int64_t *) gm_malloc(jb);
    void *dx = gm_malloc(xb);
    if (!dp || !dj || !dx)
    {
        gm_free(dp);
        gm_free(dj);
        gm_free(dx);
        int rc = gm_init();
        return GB_error(A, rc == GM_SUCCESS ? GrB_OUT_OF_MEMORY : GrB_PANIC, "%s", gm_last_error());
    }
    int rc = gm_memcpy_h2d(dp, Ap, pb);
    if (rc == GM_SUCCESS && nvals > 0) rc = gm_memcpy_h2d(dj, Aj, sizeof(int64_t) * (size_t) nvals);
    if (rc == GM_SUCCESS && Ax && (nvals > 0 || iso))
        rc = gm_memcpy_h2d(dx, Ax, A->type->size * (size_t) (iso ? 1 : nvals));
    if (rc == GM_SUCCESS && !Ax) rc = gm_memset(dx, 0, xb);
    if (rc != GM_SUCCESS)
    {
        gm_free(dp);
        gm_free(dj);
        gm_free(dx);
        return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
    }
    GB_matrix_free_content(A);
    A->p = dp;
    A->j = dj;
    A->x = dx;
    A->nvals = nvals;
    A->iso = iso;
    return GrB_SUCCESS;
}

// download A into malloc'd host arrays
static GrB_Info GB_download(GrB_Matrix A, int64_t **Ap, int64_t **Aj, void **Ax)
{
    GrB_Info info0 = GB_matrix_csr(A);
    if (info0 != GrB_SUCCESS) return info0;
    int64_t n = A->nrows, nz = A->nvals;
    size_t xsz = A->type->size * (size_t) (A->iso ? 1 : (nz > 0 ? nz : 1));
    int64_t *hp = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n + 1));
    int64_t *hj = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nz > 0 ? nz : 1));
    void *hx = malloc(xsz);
    if (!hp || !hj || !hx)
    {
        free(hp);
        free(hj);
        free(hx);
        return GrB_OUT_OF_MEMORY;
    }
    if (A->p == NULL)
    {
        memset(hp, 0, sizeof(int64_t) * (size_t) (n + 1));
        memset(hx, 0, xsz);
    }
    else
    {
        int rc = gm_memcpy_d2h(hp, A->p, sizeof(int64_t) * (size_t) (n + 1));
        if (rc == GM_SUCCESS && nz > 0) rc = gm_memcpy_d2h(hj, A->j, sizeof(int64_t) * (size_t) nz);
        if (rc == GM_SUCCESS && A->x && (nz > 0 || A->iso)) rc = gm_memcpy_d2h(hx, A->x, xsz);
        if (rc != GM_SUCCESS)
        {
            free(hp);
            free(hj);
            free(hx);
            return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
        }
    }
    *Ap = hp;
    *Aj = hj;
    *Ax = hx;
    return GrB_SUCCESS;
}

//------------------------------------------------------------------------------
// build from tuples
//------------------------------------------------------------------------------

typedef struct { uint64_t i, j, k; } GB_tuple;

static int GB_tuple_cmp(const void *a, const void *b)
{
    const GB_tuple *x = (const GB_tuple *) a, *y = (const GB_tuple *) b;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return x->k < y->k ? -1 : (x->k > y->k);   // keep input order of duplicates
}

// z = dup(x, y) for the built-in operators; false if not applicable
static bool GB_apply_dup(const GrB_BinaryOp op, int32_t code, void *z, const void *x, const void *y)
{
    // integer arithmetic wraps (computed in the unsigned type); bool PLUS,
    // MINUS and TIMES are OR, XOR and AND
#define GB_DUP_CASE(CODE, CT, UT, ADD, SUB, MUL) /* UT: the type ADD/SUB/MUL use */                         \
    if (code == CODE)                                                    \
    {                                                                    \
        CT a = *(const CT *) x, b = *(const CT *) y, r;                  \
        switch (op->opcode)                                              \
        {                                                                \
            case GB_OP_PLUS: r = (CT) (ADD); break;                      \
            case GB_OP_MINUS: r = (CT) (SUB); break;                     \
            case GB_OP_TIMES: r = (CT) (MUL); break;                     \
            case GB_OP_MIN: r = a < b ? a : b; break;                    \
            case GB_OP_MAX: r = a > b ? a : b; break;                    \
            case GB_OP_FIRST: case GB_OP_ANY: r = a; break;              \
            case GB_OP_SECOND: r = b; break;                             \
            case GB_OP_PAIR: r = (CT) 1; break;                          \
            case GB_OP_LOR: r = (CT) (a || b); break;                    \
            case GB_OP_LAND: r = (CT) (a && b); break;                   \
            default: return false;                                       \
        }                                                                \
        *(CT *) z = r;                                                   \
        return true;                                                     \
    }
    GB_DUP_CASE(GM_BOOL, bool, bool, a || b, a != b, a && b)
    GB_DUP_CASE(GM_INT32, int32_t, uint32_t, (uint32_t) a + (uint32_t) b, (uint32_t) a - (uint32_t) b,
                (uint32_t) a * (uint32_t) b)
    GB_DUP_CASE(GM_INT64, int64_t, uint64_t, (uint64_t) a + (uint64_t) b, (uint64_t) a - (uint64_t) b,
                (uint64_t) a * (uint64_t) b)
    GB_DUP_CASE(GM_FP64, double, double, a + b, a - b, a * b)
#undef GB_DUP_CASE
    return false;
}

static GrB_Info GB_build(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J, const void *X,
                         GrB_Index nvals, const GrB_BinaryOp dup, int32_t xcode)
{
    GrB_Info info = GB_check_matrix(C);
    if (info != GrB_SUCCESS) return info;
    if ((!I || !J || !X) && nvals > 0) return GrB_NULL_POINTER;
    if (C->nvals > 0) return GB_error(C, GrB_OUTPUT_NOT_EMPTY, "GrB_Matrix_build: C not empty");
    if (C->type->code != xcode)
        return GB_error(C, GrB_DOMAIN_MISMATCH, "GrB_Matrix_build: X type differs from C type");
    if (dup && dup->magic != GB_MAGIC) return GrB_UNINITIALIZED_OBJECT;
    const size_t xs = C->type->size;
    GB_tuple *t = (GB_tuple *) malloc(sizeof(GB_tuple) * (size_t) (nvals > 0 ? nvals : 1));
    int64_t *Ap = (int64_t *) calloc((size_t) C->nrows + 1, sizeof(int64_t));
    int64_t *Aj = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1));
    uint8_t *Ax = (uint8_t *) malloc(xs * (size_t) (nvals > 0 ? nvals : 1));
    if (!t || !Ap || !Aj || !Ax)
    {
        free(t); free(Ap); free(Aj); free(Ax);
        return GrB_OUT_OF_MEMORY;
    }
    for (GrB_Index k = 0; k < nvals; k++)
    {
        if (I[k] >= (GrB_Index) C->nrows || J[k] >= (GrB_Index) C->ncols)
        {
            free(t); free(Ap); free(Aj); free(Ax);
            return GB_error(C, GrB_INDEX_OUT_OF_BOUNDS, "GrB_Matrix_build: tuple %llu out of bounds",
                            (unsigned long long) k);
        }
        t[k].i = I[k];
        t[k].j = J[k];
        t[k].k = k;
    }
    qsort(t, (size_t) nvals, sizeof(GB_tuple), GB_tuple_cmp);
    int64_t nz = 0;
    for (GrB_Index k = 0; k < nvals; k++)
    {
        const uint8_t *xk = (const uint8_t *) X + xs * t[k].k;
        if (nz > 0 && k > 0 && t[k].i == t[k - 1].i && t[k].j == t[k - 1].j)
        {
            if (!dup)
            {
                free(t); free(Ap); free(Aj); free(Ax);
                return GB_error(C, GrB_INVALID_VALUE, "GrB_Matrix_build: duplicate without dup op");
            }
            if (!GB_apply_dup(dup, xcode, Ax + xs * (nz - 1), Ax + xs * (nz - 1), xk))
            {
                free(t); free(Ap); free(Aj); free(Ax);
                return GB_error(C, GrB_DOMAIN_MISMATCH, "GrB_Matrix_build: dup op not supported");
            }
            continue;
        }
        Aj[nz] = (int64_t) t[k].j;
        memcpy(Ax + xs * nz, xk, xs);
        Ap[t[k].i + 1]++;
        nz++;
    }
    for (int64_t i = 0; i < C->nrows; i++) Ap[i + 1] += Ap[i];
    free(t);
    info = GB_upload(C, Ap, Aj, Ax, nz, false);
    free(Ap);
    free(Aj);
    free(Ax);
    return info;
}

GrB_Info GrB_Matrix_build_BOOL(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                               const bool *X, GrB_Index nvals, const GrB_BinaryOp dup)
{
    return GB_build(C, I, J, X, nvals, dup, GM_BOOL);
}
GrB_Info GrB_Matrix_build_INT32(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                                const int32_t *X, GrB_Index nvals, const GrB_BinaryOp dup)
{
    return GB_build(C, I, J, X, nvals, dup, GM_INT32);
}
GrB_Info GrB_Matrix_build_INT64(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                                const int64_t *X, GrB_Index nvals, const GrB_BinaryOp dup)
{
    return GB_build(C, I, J, X, nvals, dup, GM_INT64);
}
GrB_Info GrB_Matrix_build_FP64(GrB_Matrix C, const GrB_Index *I, const GrB_Index *J,
                               const double *X, GrB_Index nvals, const GrB_BinaryOp dup)
{
    return GB_build(C, I, J, X, nvals, dup, GM_FP64);
}

//------------------------------------------------------------------------------
// extract tuples
//------------------------------------------------------------------------------

static GrB_Info GB_extract(GrB_Index *I, GrB_Index *J, void *X, GrB_Index *nvals,
                           const GrB_Matrix A, int32_t xcode)
{
    if (!nvals) return GrB_NULL_POINTER;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (X && A->type->code != xcode)
        return GB_error(A, GrB_DOMAIN_MISMATCH, "extractTuples: X type differs from A type");
    if (*nvals < (GrB_Index) A->nvals)
        return GB_error(A, GrB_INSUFFICIENT_SPACE, "extractTuples: arrays too small");
    int64_t *Ap, *Aj;
    void *Ax;
    info = GB_download(A, &Ap, &Aj, &Ax);
    if (info != GrB_SUCCESS) return info;
    const size_t xs = A->type->size;
    for (int64_t i = 0; i < A->nrows; i++)
    {
        for (int64_t p = Ap[i]; p < Ap[i + 1]; p++)
        {
            if (I) I[p] = (GrB_Index) i;
            if (J) J[p] = (GrB_Index) Aj[p];
            if (X) memcpy((uint8_t *) X + xs * p, (uint8_t *) Ax + xs * (A->iso ? 0 : p), xs);
        }
    }
    *nvals = (GrB_Index) A->nvals;
    free(Ap);
    free(Aj);
    free(Ax);
    return GrB_SUCCESS;
}

GrB_Info GrB_Matrix_extractTuples_BOOL(GrB_Index *I, GrB_Index *J, bool *X, GrB_Index *nvals,
                                       const GrB_Matrix A)
{
    return GB_extract(I, J, X, nvals, A, GM_BOOL);
}
GrB_Info GrB_Matrix_extractTuples_INT32(GrB_Index *I, GrB_Index *J, int32_t *X,
                                        GrB_Index *nvals, const GrB_Matrix A)
{
    return GB_extract(I, J, X, nvals, A, GM_INT32);
}
GrB_Info GrB_Matrix_extractTuples_INT64(GrB_Index *I, GrB_Index *J, int64_t *X,
                                        GrB_Index *nvals, const GrB_Matrix A)
{
    return GB_extract(I, J, X, nvals, A, GM_INT64);
}
GrB_Info GrB_Matrix_extractTuples_FP64(GrB_Index *I, GrB_Index *J, double *X,
                                       GrB_Index *nvals, const GrB_Matrix A)
{
    return GB_extract(I, J, X, nvals, A, GM_FP64);
}

//------------------------------------------------------------------------------
// pack / unpack CSR
//------------------------------------------------------------------------------

static int GB_i64_cmp(const void *a, const void *b)
{
    int64_t x = *(const int64_t *) a, y = *(const int64_t *) b;
    return (x > y) - (x < y);
}

GrB_Info GxB_Matrix_pack_CSR(GrB_Matrix A, GrB_Index **Ap, GrB_Index **Aj, void **Ax,
                             GrB_Index Ap_size, GrB_Index Aj_size, GrB_Index Ax_size,
                             bool iso, bool jumbled, const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!Ap || !Aj || !*Ap) return GrB_NULL_POINTER;
    const int64_t n = A->nrows;
    if (Ap_size < sizeof(int64_t) * (GrB_Index) (n + 1))
        return GB_error(A, GrB_INVALID_VALUE, "pack_CSR: Ap too small");
    int64_t *p = (int64_t *) *Ap;
    int64_t *j = (int64_t *) (*Aj);
    const int64_t nz = p[n];
    if (nz < 0 || Aj_size < sizeof(int64_t) * (GrB_Index) nz)
        return GB_error(A, GrB_INVALID_VALUE, "pack_CSR: Aj too small");
    if (nz > 0 && !j) return GrB_NULL_POINTER;
    const size_t xs = A->type->size;
    const void *x = Ax ? *Ax : NULL;
    if (x && Ax_size < xs * (GrB_Index) (iso ? 1 : nz))
        return GB_error(A, GrB_INVALID_VALUE, "pack_CSR: Ax too small");
    if (!x && (nz > 0 || iso)) return GrB_NULL_POINTER;
    if (jumbled)
    {
        // the kernels need sorted rows; sort each row (with its values)
        for (int64_t i = 0; i < n; i++)
        {
            int64_t len = p[i + 1] - p[i];
            if (len < 2) continue;
            if (iso)
            {
                qsort(j + p[i], (size_t) len, sizeof(int64_t), GB_i64_cmp);
            }
            else
            {
                // insertion sort keeps values paired; rows are short in practice
                uint8_t *xv = (uint8_t *) x;
                uint8_t tmp[16];
                for (int64_t a = p[i] + 1; a < p[i + 1]; a++)
                {
                    int64_t key = j[a];
                    memcpy(tmp, xv + xs * a, xs);
                    int64_t b = a - 1;
                    while (b >= p[i] && j[b] > key)
                    {
                        j[b + 1] = j[b];
                        memcpy(xv + xs * (b + 1), xv + xs * b, xs);
                        b--;
                    }
                    j[b + 1] = key;
                    memcpy(xv + xs * (b + 1), tmp, xs);
                }
            }
        }
    }
    info = GB_upload(A, p, j, x, nz, iso);
    if (info != GrB_SUCCESS) return info;
    free(*Ap);
    *Ap = NULL;
    free(*Aj);
    *Aj = NULL;
    if (Ax)
    {
        free(*Ax);
        *Ax = NULL;
    }
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_unpack_CSR(GrB_Matrix A, GrB_Index **Ap, GrB_Index **Aj, void **Ax,
                               GrB_Index *Ap_size, GrB_Index *Aj_size, GrB_Index *Ax_size,
                               bool *iso, bool *jumbled, const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!Ap || !Aj || !Ax || !Ap_size || !Aj_size || !Ax_size) return GrB_NULL_POINTER;
    int64_t *hp, *hj;
    void *hx;
    info = GB_download(A, &hp, &hj, &hx);
    if (info != GrB_SUCCESS) return info;
    *Ap = (GrB_Index *) hp;
    *Aj = (GrB_Index *) hj;
    *Ax = hx;
    *Ap_size = sizeof(int64_t) * (GrB_Index) (A->nrows + 1);
    *Aj_size = sizeof(int64_t) * (GrB_Index) (A->nvals > 0 ? A->nvals : 1);
    *Ax_size = A->type->size * (GrB_Index) (A->iso ? 1 : (A->nvals > 0 ? A->nvals : 1));
    if (iso) *iso = A->iso;
    if (jumbled) *jumbled = false;
    GB_matrix_free_content(A);
    return GrB_SUCCESS;
}

//------------------------------------------------------------------------------
// serialize / deserialize (Include/GraphBLAS.h:10917-10990 of the reference)
//------------------------------------------------------------------------------

// The blob is the reference's serialized matrix (its writer
// Source/serialize/GB_serialize.c:300-390, header size and field macros
// GB_serialize.h:106-140, reader GB_deserialize.c and
// GB_deserialize_from_blob.c), little-endian:
//   uint64 blob_size; int32 type code, version; int64 vlen, vdim, nvec,
//   nvec_nonempty, nvals, type size, then the byte lengths of the arrays p, h,
//   b, i, x; float hyper_switch, bitmap_switch; int32 sparsity_control,
//   4*sparsity + 2*iso + is_csc; int32 (number of blocks, method) for each of
//   p, h, b, i, x; for each array its blocks' cumulative end offsets (int64);
//   the blocks of p, h, b, i, x in that order; the matrix name and the
//   element type name, each ended by a nul byte.
// This library writes every matrix sparse by row with each array as one
// uncompressed block (method GxB_COMPRESSION_NONE, which the reference reads
// as is; the descriptor's compression request is not honoured), and reads
// sparse, hypersparse, bitmap and full blobs by row or by column whose blocks
// are uncompressed or LZ4 / LZ4HC (decoded here); ZSTD blocks (the
// reference's default) are reported as GrB_NOT_IMPLEMENTED.

#define GB_BLOB_HEADER (sizeof(uint64_t) + 11 * sizeof(int64_t) + 14 * sizeof(int32_t) + 2 * sizeof(float))
#define GB_BLOB_VERSION ((int32_t) ((GxB_IMPLEMENTATION_MAJOR * 1000 + GxB_IMPLEMENTATION_MINOR) * 1000 + \
                                    GxB_IMPLEMENTATION_SUB))

// the reference's element type names of the built-in types
// (Source/get_set/GB_code_name_get.c)
static const char *GB_blob_type_name(int32_t code)
{
    switch (code)
    {
        case GM_BOOL: return "GrB_BOOL";
        case GM_INT8: return "GrB_INT8";
        case GM_UINT8: return "GrB_UINT8";
        case GM_INT16: return "GrB_INT16";
        case GM_UINT16: return "GrB_UINT16";
        case GM_INT32: return "GrB_INT32";
        case GM_UINT32: return "GrB_UINT32";
        case GM_INT64: return "GrB_INT64";
        case GM_UINT64: return "GrB_UINT64";
        case GM_FP32: return "GrB_FP32";
        case GM_FP64: return "GrB_FP64";
        default: return "";
    }
}

static GrB_Index GB_blob_size(int64_t nrows, int64_t nvals, size_t xsize, bool iso, int32_t code)
{
    const size_t xn = xsize * (size_t) (iso ? 1 : nvals);
    size_t s = GB_BLOB_HEADER;
    s += sizeof(int64_t) * (1 + (nvals > 0 ? 1 : 0) + (xn > 0 ? 1 : 0));    // block ends
    s += sizeof(int64_t) * (size_t) (nrows + 1) + sizeof(int64_t) * (size_t) nvals + xn;
    s += 1 + strlen(GB_blob_type_name(code)) + 1;                           // names
    return (GrB_Index) s;
}

GrB_Info GrB_Matrix_serializeSize(GrB_Index *blob_size_handle, GrB_Matrix A)
{
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!blob_size_handle) return GrB_NULL_POINTER;
    *blob_size_handle = GB_blob_size(A->nrows, A->nvals, A->type->size, A->iso, A->type->code);
    return GrB_SUCCESS;
}

#define GB_PUT(v, T)                              \
    do {                                          \
        const T v_ = (T) (v);                     \
        memcpy(q, &v_, sizeof(T));                \
        q += sizeof(T);                           \
    } while (0)

GrB_Info GrB_Matrix_serialize(void *blob, GrB_Index *blob_size_handle, GrB_Matrix A)
{
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!blob || !blob_size_handle) return GrB_NULL_POINTER;
    const int32_t code = A->type->code;
    const GrB_Index need = GB_blob_size(A->nrows, A->nvals, A->type->size, A->iso, code);
    if (*blob_size_handle < need)
        return GB_error(A, GrB_INSUFFICIENT_SPACE, "serialize: blob of %llu bytes, %llu needed",
                        (unsigned long long) *blob_size_handle, (unsigned long long) need);
    int64_t *hp, *hj;
    void *hx;
    if ((info = GB_download(A, &hp, &hj, &hx)) != GrB_SUCCESS) return info;
    const int64_t n = A->nrows, nz = A->nvals;
    int64_t nonempty = 0;
    for (int64_t i = 0; i < n; i++) nonempty += (hp[i + 1] > hp[i]) ? 1 : 0;
    const int64_t p_len = (int64_t) sizeof(int64_t) * (n + 1);
    const int64_t i_len = (int64_t) sizeof(int64_t) * nz;
    const int64_t x_len = (int64_t) A->type->size * (A->iso ? 1 : nz);
    const int32_t p_nb = 1, i_nb = i_len > 0 ? 1 : 0, x_nb = x_len > 0 ? 1 : 0;
    uint8_t *q = (uint8_t *) blob;
    GB_PUT(need, uint64_t);
    GB_PUT(code, int32_t);
    GB_PUT(GB_BLOB_VERSION, int32_t);
    GB_PUT(A->ncols, int64_t);                  // vlen: the length of a row
    GB_PUT(n, int64_t);                         // vdim: the number of rows
    GB_PUT(n, int64_t);                         // nvec
    GB_PUT(nonempty, int64_t);
    GB_PUT(nz, int64_t);
    GB_PUT(A->type->size, int64_t);
    GB_PUT(p_len, int64_t);
    GB_PUT(0, int64_t);                         // h
    GB_PUT(0, int64_t);                         // b
    GB_PUT(i_len, int64_t);
    GB_PUT(x_len, int64_t);
    GB_PUT(0.0625f, float);                     // hyper_switch (the reference's default)
    GB_PUT(0.04f, float);                       // bitmap_switch
    GB_PUT(A->sparsity_control, int32_t);
    GB_PUT(4 * GxB_SPARSE + (A->iso ? 2 : 0), int32_t);    // by row
    GB_PUT(p_nb, int32_t); GB_PUT(GxB_COMPRESSION_NONE, int32_t);
    GB_PUT(0, int32_t);    GB_PUT(GxB_COMPRESSION_NONE, int32_t);
    GB_PUT(0, int32_t);    GB_PUT(GxB_COMPRESSION_NONE, int32_t);
    GB_PUT(i_nb, int32_t); GB_PUT(GxB_COMPRESSION_NONE, int32_t);
    GB_PUT(x_nb, int32_t); GB_PUT(GxB_COMPRESSION_NONE, int32_t);
    GB_PUT(p_len, int64_t);
    if (i_nb) GB_PUT(i_len, int64_t);
    if (x_nb) GB_PUT(x_len, int64_t);
    memcpy(q, hp, (size_t) p_len);
    q += p_len;
    if (i_len > 0) memcpy(q, hj, (size_t) i_len);
    q += i_len;
    if (x_len > 0) memcpy(q, hx, (size_t) x_len);
    q += x_len;
    *q++ = 0;                                   // no matrix name
    const char *tn = GB_blob_type_name(code);
    memcpy(q, tn, strlen(tn) + 1);
    q += strlen(tn) + 1;
    free(hp);
    free(hj);
    free(hx);
    *blob_size_handle = (GrB_Index) (q - (uint8_t *) blob);
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_serialize(void **blob_handle, GrB_Index *blob_size_handle, GrB_Matrix A,
                              const GrB_Descriptor desc)
{
    (void) desc;        // compression: every array is written as one uncompressed block
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!blob_handle || !blob_size_handle) return GrB_NULL_POINTER;
    GrB_Index size = GB_blob_size(A->nrows, A->nvals, A->type->size, A->iso, A->type->code);
    void *blob = malloc((size_t) size);
    if (!blob) return GrB_OUT_OF_MEMORY;
    info = GrB_Matrix_serialize(blob, &size, A);
    if (info != GrB_SUCCESS)
    {
        free(blob);
        return info;
    }
    *blob_handle = blob;
    *blob_size_handle = size;
    return GrB_SUCCESS;
}

// One LZ4 block (the LZ4 block format: sequences of a token, literals and a
// back-reference of 4 or more bytes; the last sequence has literals only)
// into dst[0, cap); returns the decoded length, or -1 for a malformed block.
static int64_t GB_lz4_decode(const uint8_t *src, size_t n, uint8_t *dst, size_t cap)
{
    size_t ip = 0, op = 0;
    while (ip < n)
    {
        const unsigned token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15)
        {
            unsigned b;
            do
            {
                if (ip >= n) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip || lit > cap - op) return -1;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == n) break;                     // the block ends with literals
        if (n - ip < 2) return -1;
        const size_t off = (size_t) src[ip] | ((size_t) src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        size_t ml = token & 15;
        if (ml == 15)
        {
            unsigned b;
            do
            {
                if (ip >= n) return -1;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += 4;
        if (ml > cap - op) return -1;
        for (size_t k = 0; k < ml; k++, op++) dst[op] = dst[op - off];   // may overlap
    }
    return (int64_t) op;
}

// Reads one array of len bytes from the blob at *s (its nb block ends in
// ends[]), decoding its blocks; *out is malloc'd (NULL for len 0).
static GrB_Info GB_blob_array(uint8_t **out, int64_t len, const uint8_t *blob, size_t size,
                              const int64_t *ends, int32_t nb, int32_t method, size_t *s)
{
    *out = NULL;
    if (len < 0 || nb < 0) return GrB_INVALID_OBJECT;
    if (len == 0)
    {
        if (nb != 0 && !(nb == 1 && ends[0] == 0)) return GrB_INVALID_OBJECT;
        return GrB_SUCCESS;
    }
    if (nb == 0 || ends[nb - 1] <= 0 || *s + (size_t) ends[nb - 1] > size ||
        (method < 0 && ends[nb - 1] != len) || len / 255 > ends[nb - 1] + 16)
        return GrB_INVALID_OBJECT;
    uint8_t *x = (uint8_t *) malloc((size_t) len);
    if (!x) return GrB_OUT_OF_MEMORY;
    int64_t prev = 0;
    for (int32_t k = 0; k < nb; k++)
    {
        // block k: bytes ends[k-1] .. ends[k] of the array's area in the blob,
        // decoding to the k-th of nb equal parts of the array (GB_PARTITION)
        const int64_t d0 = (k == 0) ? 0 : (int64_t) (((double) k * (double) len) / (double) nb);
        const int64_t d1 = (k == nb - 1) ? len : (int64_t) (((double) (k + 1) * (double) len) / (double) nb);
        const int64_t e = ends[k];
        if (e <= prev || d1 <= d0 || *s + (size_t) e > size)
        {
            free(x);
            return GrB_INVALID_OBJECT;
        }
        const uint8_t *src = blob + *s + prev;
        const size_t ssz = (size_t) (e - prev), dsz = (size_t) (d1 - d0);
        if (method < 0)
        {
            if (nb != 1 || (int64_t) ssz != len)
            {
                free(x);
                return GrB_INVALID_OBJECT;
            }
            memcpy(x, src, ssz);
        }
        else if (method >= 1000 && method < 3000)
        {
            // an LZ4 block expands at most 255-fold
            if (dsz > 255 * ssz + 16 || GB_lz4_decode(src, ssz, x + d0, dsz) != (int64_t) dsz)
            {
                free(x);
                return GrB_INVALID_OBJECT;
            }
        }
        else
        {
            free(x);
            return GrB_NOT_IMPLEMENTED;         // ZSTD (method 0 and 3000 up)
        }
        prev = e;
    }
    *s += (size_t) prev;
    *out = x;
    return GrB_SUCCESS;
}

#define GB_GET(v, T)                                  \
    T v;                                              \
    memcpy(&v, blob + s, sizeof(T));                  \
    s += sizeof(T)

// A blob in host CSR form, by row: p (nrows+1), j, x (one value if iso)
typedef struct
{
    GrB_Type type;
    int64_t nrows, ncols, nvals;
    bool iso;
    int64_t *p, *j;
    uint8_t *x;
} GB_blob_csr;

static void GB_blob_csr_free(GB_blob_csr *c)
{
    free(c->p);
    free(c->j);
    free(c->x);
    c->p = c->j = NULL;
    c->x = NULL;
}

// Parses and checks a blob into host CSR (no device access).
static GrB_Info GB_blob_parse(const uint8_t *blob, size_t size, GB_blob_csr *c)
{
    memset(c, 0, sizeof(*c));
    if (!blob) return GrB_NULL_POINTER;
    if (size < GB_BLOB_HEADER) return GrB_INVALID_OBJECT;
    size_t s = 0;
    GB_GET(bsize, uint64_t);
    GB_GET(code, int32_t);
    GB_GET(version, int32_t);
    GB_GET(vlen, int64_t);
    GB_GET(vdim, int64_t);
    GB_GET(nvec, int64_t);
    GB_GET(nvec_nonempty, int64_t);
    GB_GET(nvals, int64_t);
    GB_GET(typesize, int64_t);
    GB_GET(p_len, int64_t);
    GB_GET(h_len, int64_t);
    GB_GET(b_len, int64_t);
    GB_GET(i_len, int64_t);
    GB_GET(x_len, int64_t);
    GB_GET(hyper_switch, float);
    GB_GET(bitmap_switch, float);
    GB_GET(sparsity_control, int32_t);
    GB_GET(sic, int32_t);
    int32_t nb[5], meth[5];
    for (int a = 0; a < 5; a++)
    {
        memcpy(&nb[a], blob + s, sizeof(int32_t));
        s += sizeof(int32_t);
        memcpy(&meth[a], blob + s, sizeof(int32_t));
        s += sizeof(int32_t);
    }
    (void) version; (void) nvec_nonempty; (void) hyper_switch; (void) bitmap_switch;
    (void) sparsity_control;
    if (bsize != (uint64_t) size || code < GM_BOOL || code > 14) return GrB_INVALID_OBJECT;
    GrB_Type t = GB_type_from_code(code);
    if (!t) return GrB_NOT_IMPLEMENTED;         // complex and user-defined types
    if ((int64_t) t->size != typesize) return GrB_INVALID_OBJECT;
    const int sparsity = sic / 4;
    const bool iso = (sic & 2) != 0, csc = (sic & 1) != 0;
    const int64_t lim = (int64_t) 1 << 60;
    // every size below times an element size stays under 2^63
    if (vlen < 0 || vdim < 0 || vlen >= lim || vdim >= lim || nvals < 0 || nvals >= lim || nvec < 0 ||
        nvec > vdim)
        return GrB_INVALID_OBJECT;
    int64_t nball = 0;
    for (int a = 0; a < 5; a++)
    {
        if (nb[a] < 0) return GrB_INVALID_OBJECT;
        nball += nb[a];
    }
    if (s + sizeof(int64_t) * (size_t) nball > size) return GrB_INVALID_OBJECT;
    int64_t *ends[5];
    int64_t *endbuf = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nball > 0 ? nball : 1));
    if (!endbuf) return GrB_OUT_OF_MEMORY;
    memcpy(endbuf, blob + s, sizeof(int64_t) * (size_t) nball);
    s += sizeof(int64_t) * (size_t) nball;
    for (int a = 0, o = 0; a < 5; o += nb[a], a++) ends[a] = endbuf + o;
    const int64_t lens[5] = {p_len, h_len, b_len, i_len, x_len};
    uint8_t *arr[5] = {NULL, NULL, NULL, NULL, NULL};
    GrB_Info info = GrB_SUCCESS;
    for (int a = 0; a < 5 && info == GrB_SUCCESS; a++)
        info = GB_blob_array(&arr[a], lens[a], blob, size, ends[a], nb[a], meth[a], &s);
    free(endbuf);
    // vectors (rows when by row, columns when by column) as pointers and
    // indices: (vp[vdim+1], vi[nvals]) and the values' positions vx[] (or
    // NULL for the stored order)
    int64_t *vp = NULL, *vi = NULL, *vx = NULL;
    const int64_t nx = iso ? 1 : nvals;
    if (info == GrB_SUCCESS && (sparsity == GxB_SPARSE || sparsity == GxB_HYPERSPARSE))
    {
        const int64_t *P = (const int64_t *) arr[0], *H = (const int64_t *) arr[1];
        const int64_t *I = (const int64_t *) arr[3];
        const int64_t pv = (sparsity == GxB_HYPERSPARSE) ? nvec : vdim;
        if (p_len != (int64_t) sizeof(int64_t) * (pv + 1) || i_len != (int64_t) sizeof(int64_t) * nvals ||
            (sparsity == GxB_HYPERSPARSE && h_len != (int64_t) sizeof(int64_t) * nvec) ||
            (sparsity == GxB_SPARSE && h_len != 0) || b_len != 0 ||
            x_len != typesize * nx || P == NULL || P[0] != 0 || P[pv] != nvals)
            info = GrB_INVALID_OBJECT;
        vp = (int64_t *) calloc((size_t) vdim + 1, sizeof(int64_t));
        vi = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1));
        if (info == GrB_SUCCESS && (!vp || !vi)) info = GrB_OUT_OF_MEMORY;
        for (int64_t k = 0; info == GrB_SUCCESS && k < pv; k++)
        {
            const int64_t v = (sparsity == GxB_HYPERSPARSE) ? H[k] : k;
            if (P[k + 1] < P[k] || v < 0 || v >= vdim ||
                (sparsity == GxB_HYPERSPARSE && k > 0 && v <= H[k - 1]))
            {
                info = GrB_INVALID_OBJECT;
                break;
            }
            vp[v + 1] = P[k + 1] - P[k];
            for (int64_t e = P[k]; e < P[k + 1]; e++)
                if (I[e] < 0 || I[e] >= vlen || (e > P[k] && I[e] <= I[e - 1])) info = GrB_INVALID_OBJECT;
        }
        if (info == GrB_SUCCESS)
        {
            for (int64_t v = 0; v < vdim; v++) vp[v + 1] += vp[v];
            if (nvals > 0) memcpy(vi, I, sizeof(int64_t) * (size_t) nvals);
        }
    }
    else if (info == GrB_SUCCESS && (sparsity == GxB_BITMAP || sparsity == GxB_FULL))
    {
        // every position of a bitmap that holds an entry, or of a full matrix
        const double cells = (double) vlen * (double) vdim;
        const int64_t nc = (int64_t) cells;
        if (cells >= (double) lim || p_len != 0 || h_len != 0 || i_len != 0 ||
            (sparsity == GxB_BITMAP && b_len != nc) || (sparsity == GxB_FULL && b_len != 0) ||
            x_len != typesize * (iso ? 1 : nc) || (sparsity == GxB_FULL && nvals != nc))
            info = GrB_INVALID_OBJECT;
        int64_t cnt = 0;
        if (info == GrB_SUCCESS && sparsity == GxB_BITMAP)
            for (int64_t q = 0; q < nc; q++) cnt += arr[2][q] ? 1 : 0;
        if (info == GrB_SUCCESS && sparsity == GxB_FULL) cnt = nc;
        if (info == GrB_SUCCESS && cnt != nvals) info = GrB_INVALID_OBJECT;
        vp = (int64_t *) calloc((size_t) vdim + 1, sizeof(int64_t));
        vi = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1));
        vx = iso ? NULL : (int64_t *) malloc(sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1));
        if (info == GrB_SUCCESS && (!vp || !vi || (!iso && !vx))) info = GrB_OUT_OF_MEMORY;
        int64_t e = 0;
        for (int64_t v = 0; info == GrB_SUCCESS && v < vdim; v++)
        {
            for (int64_t k = 0; k < vlen; k++)
            {
                const int64_t q = v * vlen + k;
                if (sparsity == GxB_BITMAP && !arr[2][q]) continue;
                vi[e] = k;
                if (vx) vx[e] = q;
                e++;
            }
            vp[v + 1] = e;
        }
    }
    else if (info == GrB_SUCCESS)
    {
        info = GrB_INVALID_OBJECT;
    }
    // to CSR by row: the vectors are rows, or (by column) transpose them
    if (info == GrB_SUCCESS)
    {
        const int64_t nrows = csc ? vlen : vdim, ncols = csc ? vdim : vlen;
        int64_t *p = (int64_t *) calloc((size_t) nrows + 1, sizeof(int64_t));
        int64_t *j = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nvals > 0 ? nvals : 1));
        uint8_t *x = (uint8_t *) malloc((size_t) typesize * (size_t) (nx > 0 ? nx : 1));
        if (!p || !j || !x)
        {
            free(p);
            free(j);
            free(x);
            info = GrB_OUT_OF_MEMORY;
        }
        else
        {
            const uint8_t *X = arr[4];
            if (!csc)
            {
                memcpy(p, vp, sizeof(int64_t) * (size_t) (nrows + 1));
                if (nvals > 0) memcpy(j, vi, sizeof(int64_t) * (size_t) nvals);
                for (int64_t e = 0; e < nx; e++)
                    memcpy(x + (size_t) e * (size_t) typesize,
                           X + (size_t) (vx ? vx[e] : e) * (size_t) typesize, (size_t) typesize);
            }
            else
            {
                // counting sort of the column vectors' entries by row; a row's
                // entries come in ascending column order
                for (int64_t e = 0; e < nvals; e++) p[vi[e] + 1]++;
                for (int64_t r = 0; r < nrows; r++) p[r + 1] += p[r];
                int64_t *next = (int64_t *) malloc(sizeof(int64_t) * (size_t) (nrows > 0 ? nrows : 1));
                if (!next)
                {
                    info = GrB_OUT_OF_MEMORY;
                }
                else
                {
                    memcpy(next, p, sizeof(int64_t) * (size_t) nrows);
                    for (int64_t col = 0; col < vdim; col++)
                        for (int64_t e = vp[col]; e < vp[col + 1]; e++)
                        {
                            const int64_t d = next[vi[e]]++;
                            j[d] = col;
                            if (!iso)
                                memcpy(x + (size_t) d * (size_t) typesize,
                                       X + (size_t) (vx ? vx[e] : e) * (size_t) typesize, (size_t) typesize);
                        }
                    if (iso && nx > 0) memcpy(x, X, (size_t) typesize);
                    free(next);
                }
            }
            if (info == GrB_SUCCESS)
            {
                c->type = t;
                c->nrows = nrows;
                c->ncols = ncols;
                c->nvals = nvals;
                c->iso = iso;
                c->p = p;
                c->j = j;
                c->x = x;
            }
            else
            {
                free(p);
                free(j);
                free(x);
            }
        }
    }
    free(vp);
    free(vi);
    free(vx);
    for (int a = 0; a < 5; a++) free(arr[a]);
    return info;
}

GrB_Info GrB_Matrix_deserialize(GrB_Matrix *C, GrB_Type type, const void *blob, GrB_Index blob_size)
{
    if (!C) return GrB_NULL_POINTER;
    *C = NULL;
    GB_blob_csr c;
    GrB_Info info = GB_blob_parse((const uint8_t *) blob, (size_t) blob_size, &c);
    if (info != GrB_SUCCESS) return info;
    if (type && type != c.type)
    {
        GB_blob_csr_free(&c);
        return GrB_DOMAIN_MISMATCH;
    }
    GrB_Matrix A;
    if ((info = GrB_Matrix_new(&A, c.type, (GrB_Index) c.nrows, (GrB_Index) c.ncols)) != GrB_SUCCESS)
    {
        GB_blob_csr_free(&c);
        return info;
    }
    info = GB_upload(A, c.p, c.j, c.x, c.nvals, c.iso);
    GB_blob_csr_free(&c);
    if (info != GrB_SUCCESS)
    {
        GrB_Matrix_free(&A);
        return info;
    }
    *C = A;
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_deserialize(GrB_Matrix *C, GrB_Type type, const void *blob,
                                GrB_Index blob_size, const GrB_Descriptor desc)
{
    (void) desc;
    return GrB_Matrix_deserialize(C, type, blob, blob_size);
}

//------------------------------------------------------------------------------
// storage options (Include/GraphBLAS.h:5156-5160 of the reference)
//------------------------------------------------------------------------------

GrB_Info GxB_Matrix_Option_set_INT32(GrB_Matrix A, GxB_Option_Field field, int32_t value)
{
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    switch (field)
    {
        case GxB_SPARSITY_CONTROL:
            A->sparsity_control = (value & GxB_ANY_SPARSITY) ? (value & GxB_ANY_SPARSITY) : GxB_AUTO_SPARSITY;
            return GB_matrix_conform(A);
        case GxB_FORMAT:
            if (value != GxB_BY_ROW)
                return GB_error(A, GrB_NOT_IMPLEMENTED, "GxB_FORMAT: matrices are held by row only");
            return GrB_SUCCESS;
        default:
            return GB_error(A, GrB_INVALID_VALUE, "option %d not settable as an int", (int) field);
    }
}

GrB_Info GxB_Matrix_Option_get_INT32(GrB_Matrix A, GxB_Option_Field field, int32_t *value)
{
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!value) return GrB_NULL_POINTER;
    switch (field)
    {
        case GxB_SPARSITY_STATUS: *value = A->sparsity; return GrB_SUCCESS;
        case GxB_SPARSITY_CONTROL: *value = A->sparsity_control; return GrB_SUCCESS;
        case GxB_FORMAT: *value = GxB_BY_ROW; return GrB_SUCCESS;
        default: return GB_error(A, GrB_INVALID_VALUE, "option %d has no int value", (int) field);
    }
}

GrB_Info GxB_Matrix_Option_set(GrB_Matrix A, GxB_Option_Field field, ...)
{
    va_list ap;
    va_start(ap, field);
    GrB_Info info;
    if (field == GxB_HYPER_SWITCH || field == GxB_BITMAP_SWITCH)
    {
        (void) va_arg(ap, double);          // no hypersparse format; the hint is accepted
        info = GB_check_matrix(A);
    }
    else info = GxB_Matrix_Option_set_INT32(A, field, va_arg(ap, int));
    va_end(ap);
    return info;
}

GrB_Info GxB_Matrix_Option_get(GrB_Matrix A, GxB_Option_Field field, ...)
{
    va_list ap;
    va_start(ap, field);
    GrB_Info info;
    if (field == GxB_HYPER_SWITCH || field == GxB_BITMAP_SWITCH)
    {
        double *d = va_arg(ap, double *);
        info = GB_check_matrix(A);
        if (info == GrB_SUCCESS && d) *d = 0;
    }
    else if (field == GxB_FORMAT)
    {
        GxB_Format_Value *f = va_arg(ap, GxB_Format_Value *);
        info = GB_check_matrix(A);
        if (info == GrB_SUCCESS && !f) info = GrB_NULL_POINTER;
        if (info == GrB_SUCCESS) *f = GxB_BY_ROW;
    }
    else info = GxB_Matrix_Option_get_INT32(A, field, va_arg(ap, int32_t *));
    va_end(ap);
    return info;
}

//------------------------------------------------------------------------------
// pack / unpack of bitmap and full matrices held by row
//------------------------------------------------------------------------------

// moves host arrays Ab (NULL: full) and Ax into A on the device
static GrB_Info GB_pack_dense(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index Ab_size, GrB_Index Ax_size,
                              bool iso, int64_t nvals)
{
    const int64_t mn = A->nrows * A->ncols;
    const size_t xs = A->type->size;
    if (!Ax || !*Ax) return GrB_NULL_POINTER;
    if (Ab && !*Ab && mn > 0) return GrB_NULL_POINTER;
    if (Ab && Ab_size < (GrB_Index) mn) return GB_error(A, GrB_INVALID_VALUE, "pack: Ab too small");
    if (Ax_size < (GrB_Index) (xs * (size_t) (iso ? 1 : mn)))
        return GB_error(A, GrB_INVALID_VALUE, "pack: Ax too small");
    if (Ab)
    {
        int64_t cnt = 0;
        for (int64_t t = 0; t < mn; t++) cnt += (*Ab)[t] != 0;
        if (cnt != nvals) return GB_error(A, GrB_INVALID_VALUE, "pack_BitmapR: nvals does not match Ab");
    }
    const size_t xb = xs * (size_t) (iso ? 1 : (mn > 0 ? mn : 1));
    int8_t *db = Ab ? (int8_t *) gm_malloc(mn > 0 ? (size_t) mn : 1) : NULL;
    void *dx = gm_malloc(xb);
    if (!dx || (Ab && !db))
    {
        gm_free(db);
        gm_free(dx);
        int rc = gm_init();
        return GB_error(A, rc == GM_SUCCESS ? GrB_OUT_OF_MEMORY : GrB_PANIC, "%s", gm_last_error());
    }
    int rc = GM_SUCCESS;
    if (db && mn > 0) rc = gm_memcpy_h2d(db, *Ab, (size_t) mn);
    if (rc == GM_SUCCESS && (mn > 0 || iso)) rc = gm_memcpy_h2d(dx, *Ax, xs * (size_t) (iso ? 1 : mn));
    if (rc != GM_SUCCESS)
    {
        gm_free(db);
        gm_free(dx);
        return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
    }
    GB_matrix_free_content(A);
    A->b = db;
    A->xd = dx;
    A->nvals = nvals;
    A->iso = iso;
    A->sparsity = Ab ? GxB_BITMAP : GxB_FULL;
    if (Ab)
    {
        free(*Ab);
        *Ab = NULL;
    }
    free(*Ax);
    *Ax = NULL;
    return GB_matrix_conform(A);
}

GrB_Info GxB_Matrix_pack_BitmapR(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index Ab_size,
                                 GrB_Index Ax_size, bool iso, GrB_Index nvals, const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!Ab) return GrB_NULL_POINTER;
    return GB_pack_dense(A, Ab, Ax, Ab_size, Ax_size, iso, (int64_t) nvals);
}

GrB_Info GxB_Matrix_pack_FullR(GrB_Matrix A, void **Ax, GrB_Index Ax_size, bool iso,
                               const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    return GB_pack_dense(A, NULL, Ax, 0, Ax_size, iso, A->nrows * A->ncols);
}

// moves A out as malloc'd host arrays in bitmap (Ab != NULL) or full form
static GrB_Info GB_unpack_dense(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index *Ab_size,
                                GrB_Index *Ax_size, bool *iso, GrB_Index *nvals)
{
    const int64_t mn = A->nrows * A->ncols;
    const bool full = (Ab == NULL);
    if (full && A->nvals != mn) return GB_error(A, GrB_INVALID_VALUE, "unpack_FullR: entries are missing");
    gm_dense_out R;
    memset(&R, 0, sizeof(R));
    const int8_t *sb;
    const void *sx;
    if (GB_matrix_is_dense(A))
    {
        sb = A->sparsity == GxB_BITMAP ? A->b : NULL;
        sx = A->xd;
    }
    else
    {
        gm_csr v;
        GrB_Info info = GB_matrix_view(&v, A);
        if (info != GrB_SUCCESS) return info;
        int rc = gm_csr_to_dense(&R, &v, full);
        if (rc != GM_SUCCESS) return GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
        sb = R.b;
        sx = R.x;
    }
    const size_t xs = A->type->size;
    const size_t xb = xs * (size_t) (A->iso ? 1 : (mn > 0 ? mn : 1));
    int8_t *hb = full ? NULL : (int8_t *) malloc(mn > 0 ? (size_t) mn : 1);
    void *hx = malloc(xb);
    GrB_Info info = GrB_SUCCESS;
    if (!hx || (!full && !hb)) info = GrB_OUT_OF_MEMORY;
    int rc = GM_SUCCESS;
    if (info == GrB_SUCCESS && !full && mn > 0)
    {
        if (sb) rc = gm_memcpy_d2h(hb, sb, (size_t) mn);
        else memset(hb, 1, (size_t) mn);            // full: every entry present
    }
    if (info == GrB_SUCCESS && rc == GM_SUCCESS && (mn > 0 || A->iso))
        rc = gm_memcpy_d2h(hx, sx, xs * (size_t) (A->iso ? 1 : mn));
    gm_free(R.b);
    gm_free(R.x);
    if (info == GrB_SUCCESS && rc != GM_SUCCESS) info = GB_error(A, GB_from_gm(rc), "%s", gm_last_error());
    if (info != GrB_SUCCESS)
    {
        free(hb);
        free(hx);
        return info;
    }
    if (!full)
    {
        *Ab = hb;
        *Ab_size = (GrB_Index) (mn > 0 ? mn : 1);
    }
    *Ax = hx;
    *Ax_size = (GrB_Index) xb;
    if (iso) *iso = A->iso;
    if (nvals) *nvals = (GrB_Index) A->nvals;
    GB_matrix_free_content(A);
    return GrB_SUCCESS;
}

GrB_Info GxB_Matrix_unpack_BitmapR(GrB_Matrix A, int8_t **Ab, void **Ax, GrB_Index *Ab_size,
                                   GrB_Index *Ax_size, bool *iso, GrB_Index *nvals,
                                   const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!Ab || !Ax || !Ab_size || !Ax_size) return GrB_NULL_POINTER;
    return GB_unpack_dense(A, Ab, Ax, Ab_size, Ax_size, iso, nvals);
}

GrB_Info GxB_Matrix_unpack_FullR(GrB_Matrix A, void **Ax, GrB_Index *Ax_size, bool *iso,
                                 const GrB_Descriptor desc)
{
    (void) desc;
    GrB_Info info = GB_check_matrix(A);
    if (info != GrB_SUCCESS) return info;
    if (!Ax || !Ax_size) return GrB_NULL_POINTER;
    return GB_unpack_dense(A, NULL, Ax, NULL, Ax_size, iso, NULL);
}
