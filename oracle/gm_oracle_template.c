// gm_oracle_template.c: per-semiring body of the CPU oracle.  TEST
// INFRASTRUCTURE ONLY -- included by gm_oracle.c once per semiring with:
//
//   SR           name suffix
//   ZT           C/Z type
//   XT, YT       A and B value types (ignored when PATTERN_A / PATTERN_B)
//   PATTERN_A/B  1 if the multiplier ignores the A / B value
//   MULT(z,a,b,i,k,j)   z = a (x) b  (positional ops use i,k,j)
//   UPDATE(z,t)          z = z (+) t   (monoid update)
//   IS_ANY        1 for the ANY monoid
//   PLUS_PAIR_WIDE 1 for plus_pair on int64/fp64 ("cij != 0" existence test,
//                  GB_AxB_dot_cij.h:40-49); small-int plus_pair uses a
//                  cij_exists flag (GB_AxB_dot_cij.h:53-58)
//   TERMINAL(z)   1 if z is the terminal value of the monoid (or 0)

#define GM_CAT2(a, b) a##b
#define GM_CAT(a, b) GM_CAT2(a, b)

// C(i,j) = A(i,:) . B(j,:) for one mask entry.  Follows the sparse x sparse
// branch of GB_AxB_dot_cij.c:516-652: no-overlap early exit (:528-537),
// binary trim of the longer list when it is more than 8x longer
// (:538-616), otherwise a linear merge (:625-649).  GB_DOT semantics from
// GB_AxB_dot_cij.h:30-128: ANY takes the first match and stops; other
// monoids do MULT on the first match and MULTADD afterwards, stopping early
// at the terminal value.  Returns 1 if C(i,j) exists.
static int GM_CAT(dot_cij_, SR)(int64_t i, int64_t j,
    const int64_t *Ai, const void *Ax_, int A_iso, int64_t pA, int64_t pA_end,
    const int64_t *Bi, const void *Bx_, int B_iso, int64_t pB, int64_t pB_end,
    ZT *cij_out)
{
    const XT *Ax = (const XT *) Ax_;
    const YT *Bx = (const YT *) Bx_;
    (void) Ax; (void) Bx; (void) A_iso; (void) B_iso; (void) i; (void) j;
    int64_t ainz = pA_end - pA, bjnz = pB_end - pB;
    if (ainz == 0 || bjnz == 0) return 0;
    if (Ai[pA_end - 1] < Bi[pB] || Bi[pB_end - 1] < Ai[pA]) return 0;
    int exists = 0;
    ZT cij = 0;
    const int trimA = ainz > 8 * bjnz;
    const int trimB = bjnz > 8 * ainz;
    while (pA < pA_end && pB < pB_end)
    {
        int64_t ia = Ai[pA], ib = Bi[pB];
        if (ia < ib)
        {
            if (trimA)
            {
                // discard A entries below ib by binary search
                int64_t lo = pA + 1, hi = pA_end - 1;
                while (lo < hi)
                {
                    int64_t mid = (lo + hi) / 2;
                    if (Ai[mid] < ib) lo = mid + 1; else hi = mid;
                }
                if (lo == pA_end - 1 && Ai[lo] < ib) lo = pA_end;
                pA = lo;
            }
            else pA++;
        }
        else if (ib < ia)
        {
            if (trimB)
            {
                int64_t lo = pB + 1, hi = pB_end - 1;
                while (lo < hi)
                {
                    int64_t mid = (lo + hi) / 2;
                    if (Bi[mid] < ia) lo = mid + 1; else hi = mid;
                }
                if (lo == pB_end - 1 && Bi[lo] < ia) lo = pB_end;
                pB = lo;
            }
            else pB++;
        }
        else
        {
            const int64_t k = ia;
            (void) k;
            XT a = 0; YT b = 0;
#if !PATTERN_A
            a = Ax[A_iso ? 0 : pA];
#endif
#if !PATTERN_B
            b = Bx[B_iso ? 0 : pB];
#endif
            (void) a; (void) b;
#if IS_ANY
            MULT(cij, a, b, i, k, j);
            exists = 1;
            break;
#else
            if (exists)
            {
                ZT t;
                MULT(t, a, b, i, k, j);
                UPDATE(cij, t);
            }
            else
            {
                MULT(cij, a, b, i, k, j);
                exists = 1;
            }
            if (TERMINAL(cij)) break;
#endif
            pA++;
            pB++;
        }
    }
#if PLUS_PAIR_WIDE
    exists = (cij != 0);
#endif
    *cij_out = cij;
    return exists;
}

// dot3: C<M> = A*B' with C taking M's pattern, then zombies removed.
// Restates GB_AxB_dot3_template.c:111-165 (mask value test via GB_MCAST,
// dot product, zombie on an empty result) and the B(:,j)-empty shortcut at
// :86-99.  The zombie prune that GB_wait would do later is folded in here:
// the output is compact CSR.  Returns nnz(C).
static int64_t GM_CAT(dot3_, SR)(const or_csr *M, int mask_struct,
    const or_csr *A, const or_csr *B, int64_t *Cp, int64_t *Ci, void *Cx_)
{
    ZT *Cx = (ZT *) Cx_;
    const int64_t n = M->nrows;
    const int64_t mnz = M->p[n];
    ZT *W = (ZT *) malloc(sizeof(ZT) * (size_t) (mnz + 1));
    uint8_t *live = (uint8_t *) malloc((size_t) mnz + 1);
    if (!W || !live) { free(W); free(live); return -1; }
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++)
    {
        for (int64_t p = M->p[i]; p < M->p[i + 1]; p++)
        {
            int64_t j = M->j[p];
            live[p] = 0;
            if (!mask_struct && !or_mask_value(M, p)) continue;
            ZT cij;
            if (GM_CAT(dot_cij_, SR)(i, j, A->j, A->x, A->iso, A->p[i], A->p[i + 1],
                                     B->j, B->x, B->iso, B->p[j], B->p[j + 1], &cij))
            {
                live[p] = 1;
                W[p] = cij;
            }
        }
    }
    int64_t nz = 0;
    for (int64_t i = 0; i < n; i++)
    {
        Cp[i] = nz;
        for (int64_t p = M->p[i]; p < M->p[i + 1]; p++)
        {
            if (live[p]) { Ci[nz] = M->j[p]; Cx[nz] = W[p]; nz++; }
        }
    }
    Cp[n] = nz;
    free(W);
    free(live);
    return nz;
}

// saxpy3 for one row of C: C(i,:)<M(i,:)> = sum_k A(i,k) (x) B(k,:).
// Accumulation order follows the coarse tasks of the reference: k ascending
// over A(i,:), then j ascending over B(k,:) (GB_AxB_saxpy3_coarseHash_phase5.c
// :22-93, GB_AxB_saxpy3_coarseGus_noM_phase5.c:11-119): the first product
// for C(i,j) is stored with MULT, later ones combined with the monoid update
// (GB_AxB_saxpy3_template.h:393-530; for ANY the update is "z = t", so the
// last writer wins).  Hf is a dense flag workspace of size ncols (Gustavson
// form); mark values avoid clearing it between rows (mark = 2*row+1 in the
// reference, GB_AxB_saxpy3_template.c:456).
//   Hf[j] == mark        : C(i,j) seen
//   Hf[j] == mark+1      : M(i,j) is true (C<M>) / forbidden (C<!M>)
// phase 0 counts (Cx == NULL); phase 1 fills with sorted rows.
static int64_t GM_CAT(saxpy3_row_, SR)(int64_t i, const or_csr *M, int mask_comp,
    int mask_struct, const or_csr *A, const or_csr *B, int64_t *Hf, ZT *Hx,
    int64_t *Hlist, int64_t mark, int64_t *Ci, ZT *Cx)
{
    const XT *Ax = (const XT *) A->x;
    const YT *Bx = (const YT *) B->x;
    (void) Ax; (void) Bx;
    const int64_t mflag = mark + 1;
    if (M)
    {
        for (int64_t p = M->p[i]; p < M->p[i + 1]; p++)
        {
            if (mask_struct || or_mask_value(M, p)) Hf[M->j[p]] = mflag;
        }
    }
    int64_t cnt = 0;
    for (int64_t pA = A->p[i]; pA < A->p[i + 1]; pA++)
    {
        const int64_t k = A->j[pA];
        XT a = 0;
#if !PATTERN_A
        a = Ax[A->iso ? 0 : pA];
#endif
        (void) a;
        for (int64_t pB = B->p[k]; pB < B->p[k + 1]; pB++)
        {
            const int64_t j = B->j[pB];
            const int64_t f = Hf[j];
            int allowed;
            if (!M) allowed = 1;
            else if (mask_comp) allowed = (f != mflag);
            else allowed = (f == mflag || f == mark);
            if (!allowed) continue;
            YT b = 0;
#if !PATTERN_B
            b = Bx[B->iso ? 0 : pB];
#endif
            (void) b;
            ZT t;
            MULT(t, a, b, i, k, j);
            if (f == mark)
            {
                if (Hx) { UPDATE(Hx[j], t); }
            }
            else
            {
                Hf[j] = mark;
                if (Hx) Hx[j] = t;
                Hlist[cnt++] = j;
            }
        }
    }
    if (Cx)
    {
        or_sort_i64(Hlist, cnt);
        for (int64_t t = 0; t < cnt; t++) { Ci[t] = Hlist[t]; Cx[t] = Hx[Hlist[t]]; }
    }
    return cnt;
}

// saxpy3 driver: symbolic (count) then numeric (fill) over all rows.
// Mirrors the symbolic/numeric split of GB_AxB_saxpy3.c:591-678 with
// a Cp cumsum between the phases (GB_AxB_saxpy3_cumsum.c:132-133).
static int64_t GM_CAT(saxpy3_, SR)(const or_csr *M, int mask_comp, int mask_struct,
    const or_csr *A, const or_csr *B, int64_t *Cp, int64_t *Ci, void *Cx_, int phase)
{
    const int64_t m = A->nrows, n = B->ncols;
    int nth = omp_get_max_threads();
    int64_t *Hf_all = (int64_t *) malloc(sizeof(int64_t) * (size_t) n * (size_t) nth);
    ZT *Hx_all = (ZT *) malloc(sizeof(ZT) * (size_t) n * (size_t) nth);
    int64_t *Hl_all = (int64_t *) malloc(sizeof(int64_t) * (size_t) n * (size_t) nth);
    if (!Hf_all || !Hx_all || !Hl_all) { free(Hf_all); free(Hx_all); free(Hl_all); return -1; }
    for (size_t t = 0; t < (size_t) n * (size_t) nth; t++) Hf_all[t] = -1;
    int64_t total = 0;
    #pragma omp parallel
    {
        int tid = omp_get_thread_num();
        int64_t *Hf = Hf_all + (size_t) tid * n;
        ZT *Hx = Hx_all + (size_t) tid * n;
        int64_t *Hl = Hl_all + (size_t) tid * n;
        #pragma omp for schedule(dynamic, 64) reduction(+:total)
        for (int64_t i = 0; i < m; i++)
        {
            int64_t mark = 2 * i;     // unique per row, never -1
            if (phase == 0)
            {
                int64_t c = GM_CAT(saxpy3_row_, SR)(i, M, mask_comp, mask_struct, A, B,
                                                    Hf, NULL, Hl, mark, NULL, NULL);
                Cp[i] = c;
                total += c;
            }
            else
            {
                GM_CAT(saxpy3_row_, SR)(i, M, mask_comp, mask_struct, A, B, Hf, Hx, Hl,
                                        mark, Ci + Cp[i], ((ZT *) Cx_) + Cp[i]);
            }
        }
    }
    free(Hf_all);
    free(Hx_all);
    free(Hl_all);
    return total;
}

#undef SR
#undef ZT
#undef XT
#undef YT
#undef PATTERN_A
#undef PATTERN_B
#undef MULT
#undef UPDATE
#undef IS_ANY
#undef PLUS_PAIR_WIDE
#undef TERMINAL
