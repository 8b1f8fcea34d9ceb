//------------------------------------------------------------------------------
// gm_kernels.h: C-ABI of the MI355X (gfx950) kernel layer behind GrB_mxm
//------------------------------------------------------------------------------

// This is the thin layer between the C host library (GrB_mxm and friends,
// include/GraphBLAS.h) and the hand-written HIP kernels.  Every entry point is
// plain C: device pointers, sizes and small POD structs.  All device memory
// named here lives in HBM of the current HIP device; the library gives each
// host thread that calls it its own non-blocking stream (default priority)
// and shares one caching device arena between them.
//
// Reference interfaces replaced (SuiteSparse:GraphBLAS v9.3.1):
//   gm_AxB_dot3_branch /         <- GB_cuda_AxB_dot3_branch / GB_cuda_AxB_dot3
//   gm_AxB_dot3                     with the same argument list, flipxy included
//                                   (Source/gateway/GB_cuda_gateway.h:237-256,
//                                   called from Source/mxm/GB_AxB_dot.c:189-197);
//   gm_dot3_branch / gm_dot3        the same pair with tuning knobs in place of
//                                   flipxy
//   gm_set_stream                <- the per-call stream of CUDA/GB_cuda_AxB_dot3.cpp
//                                   :52-54, owned by the caller instead
//   gm_saxpy3_branch / gm_saxpy3 <- GB_AxB_saxpy3 (Source/mxm/GB_AxB_saxpy3.c
//                                   :119-700; the reference has no GPU saxpy3
//                                   hook, so this mirrors the dot3 gateway pair)
//   gm_reduce_scalar             <- GB_reduce_to_scalar
//                                   (Source/reduce/template/
//                                   GB_reduce_to_scalar_template.c:40-59)
//   gm_init / gm_finalize        <- GB_cuda_init (CUDA/GB_cuda_init.c) and the
//                                   rmm pool set-up in Source/init/GB_init.c:84-95
//   gm_malloc / gm_free          <- rmm_wrap_malloc / rmm_wrap_free
//                                   (rmm_wrap/rmm_wrap.h)
//
// Return codes are GrB_Info values (Include/GraphBLAS.h:358-381): 0 success,
// 1 (GrB_NO_VALUE) "declined, try another implementation" as in
// Source/mxm/GB_AxB_dot3.c:277-336, negative values are errors.

#ifndef GM_KERNELS_H
#define GM_KERNELS_H

#include <stdint.h>
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

// status codes (GrB_Info numbering)
#define GM_SUCCESS            0
#define GM_NO_VALUE           1
#define GM_INVALID_VALUE     (-3)
#define GM_NOT_IMPLEMENTED   (-8)
#define GM_PANIC             (-101)
#define GM_OUT_OF_MEMORY     (-102)

// type codes: same order as the reference's GB_Type_code
// (Source/builtin/include/GB_opaque.h, GB_BOOL_code .. GB_FP64_code)
typedef enum
{
    GM_BOOL = 1, GM_INT8 = 2, GM_UINT8 = 3, GM_INT16 = 4, GM_UINT16 = 5,
    GM_INT32 = 6, GM_UINT32 = 7, GM_INT64 = 8, GM_UINT64 = 9,
    GM_FP32 = 10, GM_FP64 = 11
} gm_type;

// monoid ("add") and multiplier codes of a built-in semiring, as the
// reference decodes them in GB_AxB_semiring_builtin
// (Source/mxm/GB_AxB_semiring_builtin.c:20-123)
typedef enum { GM_ADD_PLUS = 1, GM_ADD_MIN = 2, GM_ADD_MAX = 3, GM_ADD_ANY = 4 } gm_add_op;
typedef enum
{
    GM_MULT_TIMES = 1,  // z = x*y
    GM_MULT_PLUS = 2,   // z = x+y
    GM_MULT_PAIR = 3,   // z = 1 (values ignored; "ONEB")
    GM_MULT_K = 4       // z = k, the inner index (SECONDI on A*B, FIRSTJ after
                        // the reference's swap; values ignored)
} gm_mult_op;

typedef struct
{
    int32_t add_op;     // gm_add_op
    int32_t mult_op;    // gm_mult_op
    int32_t ztype;      // gm_type of the result (and of x,y for valued ops)
} gm_semiring;

// Derived forms of a CSR operand (opaque; gm_csr_aux_build below).  Built
// once and kept with the operand while its content is unchanged, as the
// reference keeps a matrix's hyper-hash A->Y (Source/builtin/include/GB_matrix.h:243-306).
typedef struct gm_csr_aux gm_csr_aux;

// read-only CSR operand in device memory; rows sorted, no duplicates
typedef struct
{
    int64_t nrows, ncols;
    const int64_t *p;   // nrows+1 row pointers
    const int64_t *j;   // p[nrows] column indices
    const void *x;      // values (NULL: pattern only / all true)
    int32_t iso;        // 1: x holds one value used for every entry
    int32_t type;       // gm_type of x
    const gm_csr_aux *aux;  // NULL, or derived forms of this same content that
                            // kernels may read in place of p/j/x
} gm_csr;

// CSR result; p/j/x are allocated by the library with gm_malloc and owned by
// the caller afterwards (release with gm_free).  Rows are sorted.
typedef struct
{
    int64_t nrows, ncols;
    int64_t nvals;      // entries in the result
    int64_t *p;         // nrows+1
    int64_t *j;         // capacity >= nvals
    void *x;            // values (1 value if iso)
    int32_t iso;
    int32_t type;
} gm_csr_out;

// Bitmap or full matrix held by rows in device memory: entry (i,j) at
// i*ncols + j; b[i*ncols + j] != 0 marks a present entry, b == NULL means
// full (every entry present); x holds nrows*ncols values (one if iso).  The
// reference's GxB_BITMAP / GxB_FULL formats (Include/GraphBLAS.h:1491-1495)
// by row (GxB_Matrix_pack_BitmapR / pack_FullR, :9686-9775).
typedef struct
{
    int64_t nrows, ncols;
    const int8_t *b;
    const void *x;
    int32_t iso;
    int32_t type;
} gm_dense;

// bitmap/full result; b and x allocated with gm_malloc and owned by the caller
typedef struct
{
    int64_t nrows, ncols;
    int64_t nvals;      // present entries
    int8_t *b;          // NULL: full
    void *x;
    int32_t iso;
    int32_t type;
} gm_dense_out;

// format conversions on the device: the present entries of D as sorted CSR;
// A as bitmap (full = false) or full (full = true: every entry must be
// present, else GM_INVALID_VALUE)
int gm_dense_to_csr(gm_csr_out *R, const gm_dense *D);
int gm_csr_to_dense(gm_dense_out *D, const gm_csr *A, bool full);

// Products with a bitmap or full result (the reference's dense-result saxpy
// methods): C = A*B, C<M> = A*B or C<!M> = A*B with C bitmap (GB_AxB_saxbit,
// Source/mxm/GB_AxB_saxbit.c), or, when Cin is given, C = Cin (+) A*B with C
// full and the semiring's monoid as the accumulator (GB_AxB_saxpy4 for A
// sparse and B bitmap/full, GB_AxB_saxpy5 for A bitmap/full and B sparse,
// Source/mxm/GB_AxB_saxpy4.c, GB_AxB_saxpy5.c; chosen at
// Source/mxm/GB_AxB_saxpy.c:89-140).  Every operand is either sparse (the
// gm_csr) or bitmap/full (the gm_dense): exactly one of As/Ad and of Bs/Bd
// is given, at most one of Ms/Md (none: no mask).  Terms are applied for k
// ascending, so fp64 sums equal the reference's sequential sums; ANY keeps
// the last writer.  Returns GM_NO_VALUE for a case it does not handle
// (Cin with a mask, an ANY monoid, an iso or bitmap Cin, mixed types).
int gm_saxbit(gm_dense_out *C, const gm_csr *Ms, const gm_dense *Md, bool mask_comp, bool mask_struct,
              const gm_csr *As, const gm_dense *Ad, const gm_csr *Bs, const gm_dense *Bd,
              const gm_semiring *sr, const gm_dense *Cin);

// Builds the derived forms of X that apply to it (*aux = NULL when none does):
//   row records: every row of at most 16 entries with 8-byte values stored as
//   a fixed-stride run of 16-byte (column, value) records (stride: a power of
//   two >= the longest row; built only while the padding stays under half of
//   nnz), the row's length in the top byte of its first record's column.  A
//   saxpy3 gather of row k then reads one aligned run at k*stride and needs
//   neither the row pointers nor a second cache line.
// The forms describe X's content at the time of the call: free them (and
// build again) when the content changes.  Synchronises the stream, so the
// forms can be used from any thread's stream afterwards.
int gm_csr_aux_build(gm_csr_aux **aux, const gm_csr *X);
void gm_csr_aux_free(gm_csr_aux *aux);

// dot3 tuning knobs (0 = default).  Work per mask entry is
// min(|A(i,:)|+|B(j,:)|, s*ceil(log2(l+1))) with s,l the short/long list.
typedef struct
{
    int64_t thread_max;   // work <= thread_max: one thread per entry (8)
    int64_t team16_max;   // work <= team16_max: 16 lanes per entry (512)
    int64_t wave_max;     // work <= wave_max: one wavefront per entry (4096)
                          // larger: one 256-thread workgroup per entry
    int64_t bitmap_min_len;  // rows at least this long may get a bitmap (128);
                             // < 0: no bitmaps (merge / search only)
    int64_t order;        // order of the team16 / wave entries: 0 auto (grouped
                          // when nnz(M) >= 2^20), 1 mask order, 2 grouped by
                          // the row whose probe index they read
    int64_t row_min;      // mask rows with at least row_min entries whose A row
                          // spans < 2^17 columns are done whole against an LDS
                          // bitmap of that row (-1 = never by default: measured
                          // slower than the bins on the batch and RMAT 14/22)
} gm_dot3_params;

// saxpy3 tuning knobs (0 = default).  f = flops of a row; t = its table
// size bound (f, or min(f, |M(i,:)|) for C<M>, at most ncols).
typedef struct
{
    int64_t wave_small_max;   // f <= : "small" row (64): single-pass fused kernel,
                              //        or one wavefront with a 128-slot table
    int64_t wave_large_max;   // t <= : one wavefront, 1024-slot LDS table (512)
    int64_t block_max;        // t <= : one workgroup, 4096-slot LDS table (2048)
                              // larger: workgroup per row, global hash table
    int64_t mode;             // 0 auto, 1 prefer the single-pass fused kernel
                              // for small rows, 2 symbolic + numeric only
    int64_t table;            // rows with a table: 0 the reference's Hash /
                              // Gustavson rule, 1 hash (GxB_AxB_HASH), 2 dense
                              // Gustavson workspace where it fits
                              // (GxB_AxB_GUSTAVSON); not a tuning default
} gm_saxpy3_params;

//------------------------------------------------------------------------------
// runtime
//------------------------------------------------------------------------------

int gm_init(void);                  // idempotent; selects the current device
void gm_finalize(void);             // releases the arena and the stream
const char *gm_last_error(void);    // text of the last failure on this thread
int gm_device_count(void);

typedef struct
{
    char name[64];
    char arch[32];          // e.g. "gfx950"
    int32_t cu_count;
    int32_t lds_per_block;
    int64_t hbm_bytes;
    double hbm_peak_gbs;    // 2 x memory clock x bus width as HIP reports them; on
                            // MI355X this reads 4096, half the published 8 TB/s
                            // HBM3E rate, so rooflines use the published figure
} gm_device_info;
int gm_get_device_info(gm_device_info *info);

void *gm_malloc(size_t bytes);      // device memory from the caching arena
void gm_free(void *p);              // return to the arena (NULL ok)
int gm_arena_trim(void);            // give cached blocks back to the driver
int64_t gm_arena_bytes_in_use(void);
int64_t gm_arena_bytes_cached(void);   // freed blocks the arena holds for reuse
// host-forced allocation failure (tests): the n-th following gm_malloc
// fails; n < 0 disables.  Mirrors the malloc_debug count-down of
// Source/global/GB_Global.c:818-850.
void gm_debug_fail_after(int64_t n);
// gm_dot3 runs a mask of 2^32 or more entries as row slices below that and
// joins their results; tests lower the slice size to v entries (v <= 0
// restores 2^32).  No counterpart in the reference (its CPU kernels index
// with 64-bit integers throughout).
void gm_debug_dot3_slice_max(int64_t v);

// The stream this thread's library calls run on.  gm_set_stream(s) makes
// every later call of this thread enqueue its work on the caller's stream s
// (a hipStream_t of the current device), ordered with the caller's own work;
// gm_set_stream(NULL) restores the thread's library stream.  Calls that
// return host results (counts, nvals) still wait for their own work.  The
// reference's CUDA gateway makes a stream per call instead
// (CUDA/GB_cuda_AxB_dot3.cpp:52-54); a caller-owned stream lets a
// reference-side caller order the product with its other GPU work.
int gm_set_stream(void *stream);
void *gm_get_stream(void);

int gm_memcpy_h2d(void *dst, const void *src, size_t bytes);
int gm_memcpy_d2h(void *dst, const void *src, size_t bytes);
int gm_memcpy_d2d(void *dst, const void *src, size_t bytes);
int gm_memset(void *dst, int value, size_t bytes);
int gm_synchronize(void);

// Process-wide tuned values for the fields of gm_dot3_params /
// gm_saxpy3_params left 0 (and for calls that pass NULL, as GrB_mxm does);
// NULL restores the built-in values.  gm_load_tuning reads "key value" lines
// (dot3.thread_max, dot3.team16_max, dot3.wave_max, dot3.bitmap_min_len,
// dot3.order, dot3.row_min, saxpy3.wave_small_max, saxpy3.wave_large_max,
// saxpy3.block_max, saxpy3.mode; '#' starts a comment line) as written by
// the autotuner (graphblas_amd/tune.py); gm_init loads the file named by the
// environment variable GM_TUNING_FILE.  No counterpart in the reference.
int gm_dot3_set_defaults(const gm_dot3_params *p);
int gm_saxpy3_set_defaults(const gm_saxpy3_params *p);
int gm_load_tuning(const char *path);

// per-kernel timing with HIP events on the library stream
void gm_prof_enable(int on);
void gm_prof_reset(void);
// fills up to cap records; returns the number of distinct kernels timed
typedef struct { char name[48]; int64_t launches; double total_ms; } gm_prof_record;
int gm_prof_get(gm_prof_record *rec, int cap);

//------------------------------------------------------------------------------
// dot3: C<M> = A*B' with C taking the pattern of M (then zombies removed)
//------------------------------------------------------------------------------

// In CSR terms: for every entry (i,j) of M whose mask value allows it,
// C(i,j) = (+)_k A(i,k) (x) B(j,k) over the k present in both rows.
// The branch is a host-only test of shapes, types and the semiring.  The
// kernels number mask entries and rows with 32 bits: gm_dot3 declines
// (GM_NO_VALUE) a mask with 2^32 or more entries, and GrB_mxm, which knows
// nnz(M) on the host, sends such a product to saxpy3 on B' instead.
bool gm_dot3_branch(const gm_csr *M, bool mask_struct, const gm_csr *A,
                    const gm_csr *B, const gm_semiring *sr);
int gm_dot3(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
            const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params);

// The reference's gateway pair with its argument list
// (GB_cuda_AxB_dot3_branch / GB_cuda_AxB_dot3,
// Source/gateway/GB_cuda_gateway.h:237-256, called at
// Source/mxm/GB_AxB_dot.c:189-197): flipxy asks for z = mult(B(j,k), A(i,k))
// in place of mult(A(i,k), B(j,k)).  Every multiplier of the built-in
// semirings here is unchanged by the swap -- TIMES and PLUS commute (also
// bit for bit in IEEE fp64), PAIR is a constant, and the positional
// multiplier is the inner index k whichever operand comes first (the
// reference's GB_flip_binop turns FIRSTJ into SECONDI, both k here,
// Source/binaryop/GB_flip_binop.c:351-373) -- so flipxy selects the same
// kernels; a multiplier that does not commute would be declined.
bool gm_AxB_dot3_branch(const gm_csr *M, bool mask_struct, const gm_csr *A, const gm_csr *B,
                        const gm_semiring *sr, bool flipxy);
int gm_AxB_dot3(gm_csr_out *C, const gm_csr *M, bool mask_struct, const gm_csr *A,
                const gm_csr *B, const gm_semiring *sr, bool flipxy);

// dot2: C = A*B' or C<!M> = A*B' (M may be NULL; a mask that is not
// complemented goes to gm_dot3) with every dot product of a nonempty row of A
// and a nonempty row of B computed, for small, dense results: the reference's
// GB_AxB_dot2 (Source/mxm/GB_AxB_dot2.c), chosen by GB_AxB_dot2_control
// (Source/mxm/GB_AxB_dot2_control.c).  The pairs become a structural mask of
// dot3, so C comes back sparse and compacted instead of as a bitmap.
int gm_dot2(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct, const gm_csr *A,
            const gm_csr *B, const gm_semiring *sr, const gm_dot3_params *params);

// *count = rows of A with at least one entry (the reference's
// GB_nvec_nonempty, Source/hyper/GB_nvec_nonempty.c:17), for the dot2 rule
int gm_nonempty_rows(int64_t *count, const gm_csr *A);

// The list lengths a dot3 call C<M> = A*B' reads under the two models of its
// algorithmic bytes: *min_sum = sum over M's entries (i,j) of
// min(|A(i,:)|, |B(j,:)|) (the shorter list probed into the longer one's
// index) and *len_sum = sum of |A(i,:)| + |B(j,:)| (both lists merged, as
// the reference's GB_AxB_dot_cij.c:625-649 reads them).  Measurement only:
// no counterpart in the reference.
int gm_dot3_work(int64_t *min_sum, int64_t *len_sum, const gm_csr *M, const gm_csr *A,
                 const gm_csr *B);

//------------------------------------------------------------------------------
// saxpy3: C = A*B, C<M> = A*B, C<!M> = A*B (row-wise Gustavson / hash)
//------------------------------------------------------------------------------

bool gm_saxpy3_branch(const gm_csr *M, bool mask_comp, bool mask_struct,
                      const gm_csr *A, const gm_csr *B, const gm_semiring *sr);
int gm_saxpy3(gm_csr_out *C, const gm_csr *M, bool mask_comp, bool mask_struct,
              const gm_csr *A, const gm_csr *B, const gm_semiring *sr,
              const gm_saxpy3_params *params);

// *flops = multiplies of C=A*B without a mask: the sum over A's entries
// (i,k) of |B(k,:)| (the total of GB_AxB_saxpy3_flopcount,
// Source/mxm/GB_AxB_saxpy3_flopcount.c, taken over A's rows).  GrB_mxm uses
// it to choose between saxpy3 and dot3 for C<M>=A*B.
int gm_flopcount(int64_t *flops, const gm_csr *A, const gm_csr *B);

//------------------------------------------------------------------------------
// steps around the products: transpose, accumulate and mask
//------------------------------------------------------------------------------

// T = A' (CSR of the transpose, rows sorted; an iso A gives an iso T).
// Replaces GB_transpose (Source/transpose/GB_transpose.c) where the
// orientation rules of GB_AxB_meta.c:683-694,718-723 need an explicit one.
int gm_transpose(gm_csr_out *T, const gm_csr *A);

// O = C<M, replace> = accum (C, T), the reference's GB_accum_mask
// (Source/mask/GB_accum_mask.c) for sparse operands of one type (bool,
// int32, int64, fp64).  C may be NULL (empty); M may be NULL (no mask).
// accum_op: 0 none (C<M> = T), else the reference's opcode order
// 1 PLUS, 2 MINUS, 3 TIMES, 4 MIN, 5 MAX, 6 FIRST, 7 SECOND, 8 ANY, 9 PAIR,
// 10 LOR, 11 LAND.  Outside the mask C's entries stay unless replace.
int gm_accum_mask(gm_csr_out *O, const gm_csr *C, const gm_csr *M, bool mask_comp,
                  bool mask_struct, bool replace, const gm_csr *T, int32_t accum_op);

//------------------------------------------------------------------------------
// batches: many independent requests in one launch sequence (no counterpart
// in the reference; for callers with many small products)
//------------------------------------------------------------------------------

// C_r<M_r> = A_r * B_r' for r < nreq, one semiring.  Requests are stacked
// block-diagonally; C is the stacked result: rows [row_base[r],
// row_base[r+1]) are request r's, with its column indices offset by
// col_base[r].  row_base and col_base are caller-allocated host arrays of
// nreq+1 entries, filled here.
typedef struct { gm_csr M, A, B; } gm_dot3_request;
int gm_batch_dot3(gm_csr_out *C, int64_t *row_base, int64_t *col_base,
                  const gm_dot3_request *reqs, int64_t nreq, bool mask_struct,
                  const gm_semiring *sr, const gm_dot3_params *params);

// C_r<M_r> = A_r * B (M_r optional, all or none) with one shared B, e.g. many
// frontiers against one served graph.  Requests are stacked by rows.
typedef struct { gm_csr M, A; } gm_saxpy3_request;   // M.p == NULL: no mask
int gm_batch_saxpy3(gm_csr_out *C, int64_t *row_base, const gm_saxpy3_request *reqs,
                    int64_t nreq, bool mask_comp, bool mask_struct, const gm_csr *B,
                    const gm_semiring *sr, const gm_saxpy3_params *params);

// materialise the per-request results of a stacked C as separate CSR
// matrices (out[r], released with gm_free); col_base may be NULL
int gm_batch_split(gm_csr_out *out, const gm_csr_out *C, const int64_t *row_base,
                   const int64_t *col_base, int64_t nreq);

//------------------------------------------------------------------------------
// served frontier expansion (no counterpart in the reference; for callers
// that receive many small BFS-step requests against one resident graph)
//------------------------------------------------------------------------------

// A request is a breadth-first search of one or more levels on the server's
// graph A (square, in HBM): F has one row per source (F(r, src[r]) = 1, iso)
// and V is the request's visited set (k rows; V = F when none is given).
// Each level is C<!V> = F*A; with GxB_ANY_SECONDI_INT64, C(r,i) is the
// parent of every newly reached vertex i, the reference's BFS step
// (GraphBLAS/@GrB/bfs.m:140,162).  F and V stay in HBM between levels: the
// next frontier is C and V grows by C (V's values are the parents: the
// source for its own row, -1 for vertices the caller marked visited).
// Worker threads, each on its own HIP stream, admit pending requests in
// arrival order while the batch's workspace estimate stays within the byte
// budget (the lesser of mem_budget and half the HBM still free) and its flop
// bound (sum of deg over the frontier entries) within flop_budget, and run
// the admitted level of every request as one stacked saxpy3 call.  A request
// with levels left goes back to the front of the queue and meets the
// requests that arrived meanwhile in the next batch.  A batch that runs out
// of device memory is queued again and the next admissions take at most
// half as many requests; a single request that does not fit fails with
// GM_OUT_OF_MEMORY.
typedef struct
{
    int32_t workers;         // worker threads (2); 0: none, the caller runs each
                             // batch on its own thread with gm_server_step
    int32_t max_batch_reqs;  // requests per batch at most (256)
    int64_t flop_budget;     // flop bound per batch (2^24); a larger request runs alone
    int64_t mem_budget;      // workspace bytes per batch (0: a quarter of HBM)
} gm_server_config;

typedef struct
{
    int64_t max_steps;       // levels to run: 1 (the default), or <= 0 until the
                             // frontier is empty
    int32_t keep_visited;    // keep V for gm_server_fetch_visited
    int32_t pad;
} gm_server_request_opts;

typedef struct
{
    int32_t status;          // GM_SUCCESS or the batch's error code
    int32_t pad;
    int64_t k;               // rows (sources) of the request
    int64_t nvals;           // entries of the last level's result
    int64_t flops;           // flop bounds of the levels run
    double latency_us;       // arrival -> result complete
    double queue_us;         // arrival -> its first batch started
    int64_t batch_id;        // batch that computed its last level
    int64_t batch_size;      // requests in that batch
    int64_t steps;           // levels run
    int64_t vnvals;          // entries of V (when kept)
} gm_server_result;

typedef struct
{
    int64_t completed, batches, pending, max_in_flight;
    double mean_batch;       // requests per batch
    double flops;            // flop bound of the levels completed
    int64_t levels;          // request levels completed
    int64_t oom_splits;      // batches queued again after running out of memory
} gm_server_counters;

typedef struct gm_server gm_server;

// graph and sr are captured (graph's device arrays must outlive the server);
// cfg may be NULL.  Returns NULL on error (gm_last_error).
gm_server *gm_server_create(const gm_csr *graph, const gm_semiring *sr,
                            const gm_server_config *cfg);
// sources: k host vertex ids; vis_p (k+1) / vis_j: optional host CSR of the
// visited set (both NULL: V = F; vis_p[0] = 0, rows sorted, columns in
// [0, n)), uploaded to HBM here once.  opts may be NULL (one level).
// Returns the request id (> 0) or an error code (< 0).
int64_t gm_server_submit_bfs(gm_server *s, const int64_t *sources, int64_t k,
                             const int64_t *vis_p, const int64_t *vis_j,
                             const gm_server_request_opts *opts);
// one level (gm_server_submit_bfs with opts NULL)
int64_t gm_server_submit(gm_server *s, const int64_t *sources, int64_t k,
                         const int64_t *vis_p, const int64_t *vis_j);
// a server without worker threads: admits one batch and runs its level on
// the calling thread; returns the batch's request count (0: none pending)
int64_t gm_server_step(gm_server *s);
// blocks until the request is complete; fills *out (without worker threads
// a request that is not complete is an error)
int gm_server_wait(gm_server *s, int64_t id, gm_server_result *out);
// copies the request's last level's rows (k+1 row pointers from 0, nvals
// indices and values; Cj / Cx may be NULL) to host memory
int gm_server_fetch(gm_server *s, int64_t id, int64_t *Cp, int64_t *Cj, void *Cx);
// copies the request's visited set (k+1 row pointers, vnvals indices and
// parents) when it was submitted with keep_visited
int gm_server_fetch_visited(gm_server *s, int64_t id, int64_t *Vp, int64_t *Vj, void *Vx);
// forgets the request (its HBM results go with the last request that holds them)
int gm_server_release(gm_server *s, int64_t id);
int gm_server_stats(gm_server *s, gm_server_counters *st);
// drains pending requests, joins the workers, frees everything
void gm_server_destroy(gm_server *s);

//------------------------------------------------------------------------------
// reduce a value array to a scalar with a monoid (device -> host scalar)
//------------------------------------------------------------------------------

// result: host pointer to one value of `type`; x: device array of n values
// (iso: x[0] repeated n times).
int gm_reduce_scalar(void *result, int32_t add_op, int32_t type, const void *x,
                     int64_t n, int32_t iso);

// Allocation-free form of gm_reduce_scalar.  No counterpart in the reference:
// it is for callers that cannot allocate device or pinned memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// result: DEVICE pointer to one value of `type`, written on `stream`
// (NULL: the calling thread's library stream); nothing is copied to the host
// and nothing is waited for, so read it after synchronising the stream.
// The first pass's partials live in a fixed 8 KB device array per GPU (one
// value per workgroup of a fixed grid, so every n fits; the array is held
// until the process ends).  gm_init() must have been called (it looks up the
// array's address); otherwise GM_PANIC.  Calls that overlap on one GPU (on
// different streams) race on the array.  Same results as gm_reduce_scalar
// (the same kernels combine the same partials in the same order).
int gm_reduce_scalar_noalloc(void *result, int32_t add_op, int32_t type, const void *x,
                             int64_t n, int32_t iso, void *stream);

#ifdef __cplusplus
}
#endif
#endif
