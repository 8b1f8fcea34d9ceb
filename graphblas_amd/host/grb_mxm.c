// grb_mxm.c: GrB_mxm, C<Mask> = accum (C, A*B), on the gfx950 kernels.
//
// Follows the reference's call chain for CSR operands:
//   GrB_mxm      Source/mxm/GrB_mxm.c:18-67   (checks, descriptor decode)
//   GB_mxm       Source/mxm/GB_mxm.c:24-222   (types, dimensions, transplant)
//   GB_AxB_meta  Source/mxm/GB_AxB_meta.c:41-875 (orientation, method choice)
// With every matrix stored by row, the reference's CSR folding and swap rule
// (GB_AxB_meta.c:214-447) end in two kernels:
//   C<M> = A*B'  with M present and not complemented  -> dot3
//                (GB_AxB_dot3_control, Source/mxm/GB_mxm.h:233-241)
//   C = A*B, C<M> = A*B, C<!M> = A*B                   -> saxpy3
//                (GB_AxB_saxpy_sparsity.c:56-83, default policy)
// An A' operand, and a B' operand that does not go to dot3, are transposed
// on the device first (gm_transpose; the reference's explicit transposes of
// GB_AxB_meta.c:683-694,718-723).  The product T then reaches C by
// transplant, or through gm_accum_mask (GB_accum_mask) when there is an
// accumulator or a mask over a non-empty C without GrB_REPLACE.  A result
// typecast returns GrB_NOT_IMPLEMENTED: there is no CPU path behind this
// library.

#include "grb_internal.h"

GB_method GB_mxm_select(bool has_mask, bool mask_comp, bool a_tran, bool b_tran, int axb)
{
    (void) a_tran;      // an A' operand is transposed first; the choice is the same
    // C<M> = A*B' -> dot3 (GB_AxB_dot3_control, Source/mxm/GB_mxm.h:233-241);
    // C<M> = A*B with the GxB_AxB_DOT hint -> dot3 on a transposed B (the
    // user override of GB_AxB_meta_adotb_control.c:88-92); every other shape
    // (after transposing B' when needed) -> saxpy3
    if (has_mask && !mask_comp && (b_tran || axb == GxB_AxB_DOT)) return GB_METHOD_DOT3;
    return GB_METHOD_SAXPY3;
}

// C<M>=A*B with a plain mask and no method hint: the reference runs masked
// saxpy3 here (GB_AxB_meta.c:562), whose work is the flop count of A*B.  On
// this device the masked dot product on B' gives the same entries with work
// about nnz(M) * (mean |A(i,:)| + mean |B(:,j)|) plus a sort of B for the
// transpose, so dot3 is taken when that is well below the flop count (the
// triangle count C<L>=L*L, for example).  Each C(i,j) sums the same terms
// A(i,k)*B(k,j) either way, so integer and boolean results are identical;
// fp64 sums may round differently (dot3's teams add strided per-lane partials
// and combine them by a tree, saxpy3 adds in k order).
bool GB_mxm_prefer_dot3(int64_t flops, int64_t nnz_m, int64_t nnz_a, int64_t m, int64_t nnz_b,
                        int64_t n)
{
    if (nnz_m <= 0 || flops <= 0) return false;
    const double mean_a = m > 0 ? (double) nnz_a / (double) m : 0.0;
    const double mean_b = n > 0 ? (double) nnz_b / (double) n : 0.0;
    const double dot_work = (double) nnz_m * (mean_a + mean_b) + 16.0 * (double) nnz_b;
    return (double) flops > 4.0 * dot_work;
}

// C = A*B' (or C<!M> = A*B') as every dot product of the small result:
// the rule of the reference's GB_AxB_dot2_control
// (Source/mxm/GB_AxB_dot2_control.c:37-92), with A's rows in place of its
// vectors.  anvec / bnvec: nonempty rows of A and of B (the rows of B'
// columns), avlen: the inner dimension.
bool GB_AxB_dot2_control(double anz, double bnz, double anvec, double bnvec, double avlen)
{
    const double cnz = anvec * bnvec;                       // the pairs to compute
    const double row_degree = anz / (avlen > 1 ? avlen : 1);
    const double col_degree = anz / (anvec > 1 ? anvec : 1);
    if (cnz > anz + bnz) return false;                      // C large: saxpy
    if (anz + bnz > 10000 * cnz || cnz <= 100) return true;  // C tiny: dot
    if (row_degree < 0.125 && col_degree > 1200) return true;
    return false;
}

// dot2 pairs above this count go to saxpy3 even when GxB_AxB_DOT asks for a
// dot product (the structural mask of the pairs would not fit)
#define GB_DOT2_MAX_PAIRS ((double) (1ll << 30))

// the gm_csr view of a device result (no copy)
static gm_csr GB_view_of(const gm_csr_out *R)
{
    gm_csr v;
    memset(&v, 0, sizeof(v));
    v.nrows = R->nrows;
    v.ncols = R->ncols;
    v.p = R->p;
    v.j = R->j;
    v.x = R->x;
    v.iso = R->iso;
    v.type = R->type;
    return v;
}

static void GB_free_out(gm_csr_out *R)
{
    gm_free(R->p);
    gm_free(R->j);
    gm_free(R->x);
    memset(R, 0, sizeof(*R));
}

static const char *GB_sr_name(const gm_semiring *s, char *buf, size_t len)
{
    static const char *adds[] = {"?", "plus", "min", "max", "any"};
    static const char *mults[] = {"?", "times", "plus", "pair", "k"};
    static const char *types[] = {"?", "bool", "int8", "uint8", "int16", "uint16", "int32",
                                  "uint32", "int64", "uint64", "fp32", "fp64"};
    snprintf(buf, len, "%s_%s_%s", adds[(s->add_op >= 0 && s->add_op < 5) ? s->add_op : 0],
             mults[(s->mult_op >= 0 && s->mult_op < 5) ? s->mult_op : 0],
             types[(s->ztype >= 0 && s->ztype < 12) ? s->ztype : 0]);
    return buf;
}

GrB_Info GrB_mxm(GrB_Matrix C, const GrB_Matrix Mask, const GrB_BinaryOp accum,
                 const GrB_Semiring semiring, const GrB_Matrix A, const GrB_Matrix B,
                 const GrB_Descriptor desc)
{
    if (!GB_initialized()) return GrB_PANIC;
    GrB_Info info;
    if ((info = GB_check_matrix(C)) != GrB_SUCCESS) return info;
    if (!semiring) return GB_error(C, GrB_NULL_POINTER, "GrB_mxm: semiring is NULL");
    if (semiring->magic != GB_MAGIC)
        return GB_error(C, GrB_UNINITIALIZED_OBJECT, "GrB_mxm: semiring not initialised");
    if ((info = GB_check_matrix(A)) != GrB_SUCCESS) return GB_error(C, info, "GrB_mxm: A invalid");
    if ((info = GB_check_matrix(B)) != GrB_SUCCESS) return GB_error(C, info, "GrB_mxm: B invalid");
    if (Mask && (info = GB_check_matrix(Mask)) != GrB_SUCCESS)
        return GB_error(C, info, "GrB_mxm: Mask invalid");
    if (accum && accum->magic != GB_MAGIC)
        return GB_error(C, GrB_UNINITIALIZED_OBJECT, "GrB_mxm: accum not initialised");
    if (desc && desc->magic != GB_MAGIC)
        return GB_error(C, GrB_UNINITIALIZED_OBJECT, "GrB_mxm: descriptor not initialised");

    // descriptor decode (GrB_mxm.c:43-47)
    const bool C_replace = desc && desc->out == GrB_REPLACE;
    const bool mask_comp = desc && (desc->mask & GrB_COMP);
    const bool mask_struct = desc && (desc->mask & GrB_STRUCTURE);
    const bool a_tran = desc && desc->in0 == GrB_TRAN;
    const bool b_tran = desc && desc->in1 == GrB_TRAN;
    const int axb = desc ? desc->axb : 0;

    // dimensions (GB_mxm.c:85-99)
    const int64_t am = a_tran ? A->ncols : A->nrows, an = a_tran ? A->nrows : A->ncols;
    const int64_t bm = b_tran ? B->ncols : B->nrows, bn = b_tran ? B->nrows : B->ncols;
    if (an != bm || C->nrows != am || C->ncols != bn)
        return GB_error(C, GrB_DIMENSION_MISMATCH,
                        "GrB_mxm: dimensions (%lld x %lld) = (%lld x %lld) * (%lld x %lld)",
                        (long long) C->nrows, (long long) C->ncols, (long long) am,
                        (long long) an, (long long) bm, (long long) bn);
    if (Mask && (Mask->nrows != C->nrows || Mask->ncols != C->ncols))
        return GB_error(C, GrB_DIMENSION_MISMATCH, "GrB_mxm: Mask is %lld x %lld, C is %lld x %lld",
                        (long long) Mask->nrows, (long long) Mask->ncols, (long long) C->nrows,
                        (long long) C->ncols);

    // types (GB_mxm.c:65-82): the semiring result must be C's type here
    gm_semiring sr;
    if (!GB_semiring_codes(semiring, &sr))
        return GB_error(C, GrB_NOT_IMPLEMENTED, "GrB_mxm: semiring %s not supported",
                        semiring->name);
    if (C->type->code != sr.ztype)
        return GB_error(C, GrB_NOT_IMPLEMENTED,
                        "GrB_mxm: typecast of the %s result to C's type %s is not supported",
                        semiring->name, C->type->name);
    const bool valued = (sr.mult_op == GM_MULT_TIMES || sr.mult_op == GM_MULT_PLUS);
    if (valued && (A->type->code != sr.ztype || B->type->code != sr.ztype))
        return GB_error(C, GrB_NOT_IMPLEMENTED,
                        "GrB_mxm: A and B must have the semiring's type %s", C->type->name);
    // accumulator: a built-in operator on C's type (no typecast)
    int32_t accum_op = 0;
    if (accum)
    {
        if (accum->ztype != C->type || accum->xtype != C->type || accum->ytype != C->type ||
            accum->opcode < GB_OP_PLUS || accum->opcode > GB_OP_LAND)
            return GB_error(C, GrB_NOT_IMPLEMENTED,
                            "GrB_mxm: accum %s on C of type %s is not supported", accum->name,
                            C->type->name);
        accum_op = (int32_t) accum->opcode;
    }
    // T reaches C through GB_accum_mask when C's old entries matter
    const bool post = C->nvals > 0 && (accum != NULL || (Mask != NULL && !C_replace));

    // bitmap / full operands without transposes or a method hint
    // (GB_AxB_saxpy.c:89-140, GB_AxB_saxpy_sparsity.c:177-234): C += A*B
    // into a full C with the monoid as accum (saxpy4 / saxpy5), or a bitmap C
    // (saxbit) when the sparsity rule gives one
    if ((GB_matrix_is_dense(A) || GB_matrix_is_dense(B)) && !a_tran && !b_tran && axb == GxB_DEFAULT)
    {
        const bool ad = GB_matrix_is_dense(A), bd = GB_matrix_is_dense(B);
        const GrB_BinaryOp mon = semiring->add->op;
        const bool in_place = C->sparsity == GxB_FULL && !C->iso && !Mask && accum != NULL &&
                              accum->opcode == mon->opcode && accum->ztype == mon->ztype &&
                              ad != bd && sr.add_op != GM_ADD_ANY;
        bool bitmap = false;
        if (!in_place && !post && (!Mask || mask_comp || GB_matrix_is_dense(Mask)))
        {
            if (ad && bd) bitmap = true;
            else if (bd)
            {
                // {S} * {B,F}: bitmap unless A has few nonempty rows
                gm_csr va0;
                if ((info = GB_matrix_view(&va0, A)) != GrB_SUCCESS) return GB_error(C, info, "%s", A->err);
                int64_t anvec = 0;
                int rc0 = gm_nonempty_rows(&anvec, &va0);
                if (rc0 != GM_SUCCESS) return GB_error(C, GB_from_gm(rc0), "GrB_mxm: %s", gm_last_error());
                bitmap = anvec >= C->nrows / 4;
            }
            else bitmap = !(B->nvals < C->ncols / 20);   // {B,F} * {S}: sparse if B is very sparse
        }
        if (in_place || bitmap)
        {
            gm_csr as, bs, ms;
            gm_dense adn, bdn, mdn, cdn;
            const gm_csr *pas = NULL, *pbs = NULL, *pms = NULL;
            const gm_dense *pad = NULL, *pbd = NULL, *pmd = NULL;
            if (ad) { GB_matrix_dense_view(&adn, A); pad = &adn; }
            else
            {
                if ((info = GB_matrix_view(&as, A)) != GrB_SUCCESS) return GB_error(C, info, "%s", A->err);
                pas = &as;
            }
            if (bd) { GB_matrix_dense_view(&bdn, B); pbd = &bdn; }
            else
            {
                if ((info = GB_matrix_view(&bs, B)) != GrB_SUCCESS) return GB_error(C, info, "%s", B->err);
                pbs = &bs;
            }
            if (Mask && GB_matrix_is_dense(Mask)) { GB_matrix_dense_view(&mdn, Mask); pmd = &mdn; }
            else if (Mask)
            {
                if ((info = GB_matrix_view(&ms, Mask)) != GrB_SUCCESS)
                    return GB_error(C, info, "%s", Mask->err);
                pms = &ms;
            }
            const bool mstruct0 = mask_struct || (Mask && !GB_matrix_is_dense(Mask) && Mask->x == NULL);
            if (in_place) GB_matrix_dense_view(&cdn, C);
            if (GB_burble)
                printf("burble: GrB_mxm: %s (A %s, B %s)\n",
                       in_place ? (ad ? "C+=A*B saxpy5" : "C+=A*B saxpy4") : "C bitmap saxbit",
                       ad ? "bitmap/full" : "sparse", bd ? "bitmap/full" : "sparse");
            gm_dense_out D;
            memset(&D, 0, sizeof(D));
            int rc1 = gm_saxbit(&D, pms, pmd, mask_comp, mstruct0, pas, pad, pbs, pbd, &sr,
                                in_place ? &cdn : NULL);
            if (rc1 != GM_SUCCESS) return GB_error(C, GB_from_gm(rc1), "GrB_mxm: %s", gm_last_error());
            GB_matrix_set_dense(C, &D);
            C->err[0] = '\0';
            return GB_matrix_conform(C);
        }
    }

    GB_method method = GB_mxm_select(Mask != NULL, mask_comp, a_tran, b_tran, axb);

    gm_csr va, vb, vm, vc;
    if ((info = GB_matrix_view(&va, A)) != GrB_SUCCESS) return GB_error(C, info, "%s", A->err);
    if ((info = GB_matrix_view(&vb, B)) != GrB_SUCCESS) return GB_error(C, info, "%s", B->err);
    if (Mask && (info = GB_matrix_view(&vm, Mask)) != GrB_SUCCESS)
        return GB_error(C, info, "%s", Mask->err);
    // a valued boolean mask of an iso matrix with a NULL value array counts
    // as all true (structural)
    const bool mstruct = mask_struct || (Mask && Mask->x == NULL);

    gm_csr_out R, AT, BT;
    memset(&R, 0, sizeof(R));
    memset(&AT, 0, sizeof(AT));
    memset(&BT, 0, sizeof(BT));
    int rc = GM_SUCCESS;
    char nm[64];
    if (a_tran)
    {
        if (GB_burble) printf("burble: GrB_mxm: transpose A (%lld entries)\n", (long long) A->nvals);
        rc = gm_transpose(&AT, &va);
        if (rc != GM_SUCCESS) return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
        va = GB_view_of(&AT);
    }
    // (not for an ANY monoid with a positional multiplier, whose witness k
    // differs between the methods: the first match in a dot product, the
    // last writer in saxpy3)
    if (method == GB_METHOD_SAXPY3 && Mask && !mask_comp && !b_tran && axb == GxB_DEFAULT &&
        !(sr.add_op == GM_ADD_ANY && sr.mult_op == GM_MULT_K))
    {
        gm_csr bt_shape = vb;       // B' by its dimensions only, for the dot3 case check
        bt_shape.nrows = vb.ncols;
        bt_shape.ncols = vb.nrows;
        if (gm_dot3_branch(&vm, mstruct, &va, &bt_shape, &sr))
        {
            int64_t flops = 0;
            rc = gm_flopcount(&flops, &va, &vb);
            if (rc != GM_SUCCESS)
            {
                GB_free_out(&AT);
                return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
            }
            const int64_t nnz_a = a_tran ? AT.nvals : A->nvals;
            if (GB_mxm_prefer_dot3(flops, Mask->nvals, nnz_a, va.nrows, B->nvals, vb.ncols))
            {
                if (GB_burble)
                    printf("burble: GrB_mxm: C<M>=A*B with %lld flops: dot3 on B'\n", (long long) flops);
                method = GB_METHOD_DOT3;
            }
        }
    }
    if (method == GB_METHOD_SAXPY3 && (!Mask || mask_comp) && (b_tran || axb == GxB_AxB_DOT) &&
        (axb == GxB_DEFAULT || axb == GxB_AxB_DOT))
    {
        // C = A*B' or C<!M> = A*B': every dot product of a small, dense
        // result (GB_AxB_meta_adotb_control.c:57-92), or on request
        int64_t anvec = 0, bnvec = 0;
        rc = gm_nonempty_rows(&anvec, &va);
        if (rc == GM_SUCCESS && b_tran) rc = gm_nonempty_rows(&bnvec, &vb);
        if (!b_tran) bnvec = vb.ncols;      // B' rows: at most B's columns
        if (rc != GM_SUCCESS)
        {
            GB_free_out(&AT);
            return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
        }
        const double anz = (double) (a_tran ? AT.nvals : A->nvals), bnz = (double) B->nvals;
        const double pairs = (double) anvec * (double) bnvec;
        const bool dot2 = (axb == GxB_AxB_DOT) ? pairs <= GB_DOT2_MAX_PAIRS
                        : GB_AxB_dot2_control(anz, bnz, (double) anvec, (double) bnvec, (double) va.ncols);
        if (dot2)
        {
            if (GB_burble)
                printf("burble: GrB_mxm: C%s=A*B' with %lld x %lld nonempty rows: dot2\n",
                       Mask ? "<!M>" : "", (long long) anvec, (long long) bnvec);
            method = GB_METHOD_DOT2;
        }
    }
    if (method == GB_METHOD_DOT3 && (Mask->nvals >= ((int64_t) 1 << 32) || Mask->nrows >= ((int64_t) 1 << 32)))
    {
        // the dot3 kernels number mask entries and rows with 32 bits: a larger
        // mask takes saxpy3 on the rows of B' (the same entries, by the flop
        // order of C<M>=A*(B')')
        if (GB_burble)
            printf("burble: GrB_mxm: mask of %lld entries beyond dot3's 2^32: saxpy3\n",
                   (long long) Mask->nvals);
        method = GB_METHOD_SAXPY3;
    }
    if (b_tran != (method == GB_METHOD_DOT3 || method == GB_METHOD_DOT2))
    {
        // saxpy3 needs B by rows, dot3 needs the rows of B' (B's columns)
        if (GB_burble) printf("burble: GrB_mxm: transpose B (%lld entries)\n", (long long) B->nvals);
        rc = gm_transpose(&BT, &vb);
        if (rc != GM_SUCCESS)
        {
            GB_free_out(&AT);
            return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
        }
        vb = GB_view_of(&BT);
    }
    if (method == GB_METHOD_DOT3)
    {
        if (GB_burble)
            printf("burble: GrB_mxm C<%sM>=A*B' %s: dot3 (masked dot product), M nnz %lld\n",
                   mstruct ? "S:" : "", GB_sr_name(&sr, nm, sizeof(nm)),
                   (long long) Mask->nvals);
        rc = gm_dot3(&R, &vm, mstruct, &va, &vb, &sr, NULL);
    }
    else if (method == GB_METHOD_DOT2)
    {
        rc = gm_dot2(&R, Mask ? &vm : NULL, mask_comp, mstruct, &va, &vb, &sr, NULL);
    }
    else
    {
        // the descriptor's Hash / Gustavson hint picks the rows' tables
        // (Include/GraphBLAS.h:532-534 of the reference); otherwise its rule
        gm_saxpy3_params sp;
        memset(&sp, 0, sizeof(sp));
        sp.table = (axb == GxB_AxB_HASH) ? 1 : (axb == GxB_AxB_GUSTAVSON) ? 2 : 0;
        if (GB_burble)
            printf("burble: GrB_mxm C%s=A*B %s: saxpy3 (%s), A nnz %lld B nnz %lld\n",
                   Mask ? (mask_comp ? "<!M>" : "<M>") : "", GB_sr_name(&sr, nm, sizeof(nm)),
                   sp.table == 1 ? "hash" : sp.table == 2 ? "gustavson" : "hash/gustavson",
                   (long long) (a_tran ? AT.nvals : A->nvals),
                   (long long) (b_tran ? BT.nvals : B->nvals));
        // B's row records (built once per content of B) when the fused kernel
        // gathers its 8-byte values row by row, about once per row or more
        if (!b_tran && !B->iso && B->x && B->type->size == 8 && sr.mult_op != GM_MULT_PAIR &&
            sr.mult_op != GM_MULT_K && 2 * (a_tran ? AT.nvals : A->nvals) >= B->nrows)
            GB_matrix_aux(&vb, B);
        rc = gm_saxpy3(&R, Mask ? &vm : NULL, mask_comp, mstruct, &va, &vb, &sr, &sp);
    }
    GB_free_out(&AT);
    GB_free_out(&BT);
    if (rc != GM_SUCCESS) return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
    if (post)
    {
        // C<M, replace> = accum (C, T) on the device (GB_mxm.c:201-214 -> GB_accum_mask)
        if (GB_burble)
            printf("burble: GrB_mxm: accum/mask into C (%lld entries, accum op %d)\n",
                   (long long) C->nvals, (int) accum_op);
        gm_csr_out O;
        memset(&O, 0, sizeof(O));
        R.type = C->type->code;
        if (R.iso && R.x == NULL) R.iso = 0;
        const gm_csr vt = GB_view_of(&R);
        if ((info = GB_matrix_view(&vc, C)) != GrB_SUCCESS)
        {
            GB_free_out(&R);
            return GB_error(C, info, "%s", C->err);
        }
        rc = gm_accum_mask(&O, &vc, Mask ? &vm : NULL, mask_comp, mstruct, C_replace, &vt, accum_op);
        GB_free_out(&R);
        if (rc != GM_SUCCESS) return GB_error(C, GB_from_gm(rc), "GrB_mxm: %s", gm_last_error());
        R = O;
    }
    // transplant T into C (GB_mxm.c:168-200), then C's format hint
    GB_matrix_set_csr(C, &R);
    C->err[0] = '\0';
    return GB_matrix_conform(C);
}
