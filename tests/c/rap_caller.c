/*
 * A C caller of the drop-in entry points, written the way the reference's
 * test driver uses them (code/test/Sp3MM_test.c:254, 318-346): host spmat
 * operands, a CONFIG, an SpMM function pointer for the pair form, results
 * released with freeSpmat, and the 1-based _1 set through the Fortran index
 * shift (code/src/include/sparseUtilsGeneric.h:64-72).
 *
 * Problem: the 5-point Laplacian on an n x n grid with 2 x 2 aggregates, P the
 * 0/1 aggregation matrix and R = P^T.  Then R*A*P has the closed form
 * RAP[I][J] = sum of A[i][j] over i in aggregate I, j in aggregate J, which
 * the caller accumulates densely and compares with every product.
 *
 * Build: gcc -std=c11 -I include tests/c/rap_caller.c -L sp3mm4amg_amd
 *        -lsp3mm_amd -Wl,-rpath,<abs>/sp3mm4amg_amd -lm
 * Exit status 0 and "rap_caller ok" on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sp3mm_amd.h"

static spmat *csr_new(idx_t M, idx_t N, idx_t nz)
{
    spmat *m = calloc(1, sizeof *m);
    m->M = M;
    m->N = N;
    m->NZ = nz;
    m->IRP = malloc((M + 1) * sizeof(idx_t));
    m->JA = malloc((nz ? nz : 1) * sizeof(idx_t));
    m->AS = malloc((nz ? nz : 1) * sizeof(double));
    return m;
}

static spmat *csr_copy(const spmat *a)
{
    spmat *m = csr_new(a->M, a->N, a->NZ);
    memcpy(m->IRP, a->IRP, (a->M + 1) * sizeof(idx_t));
    memcpy(m->JA, a->JA, a->NZ * sizeof(idx_t));
    memcpy(m->AS, a->AS, a->NZ * sizeof(double));
    return m;
}

/* 5-point Laplacian, Dirichlet, rows in column order */
static spmat *laplacian(int n)
{
    spmat *a = csr_new((idx_t)n * n, (idx_t)n * n, (idx_t)5 * n * n);
    idx_t k = 0;
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            const idx_t r = (idx_t)y * n + x;
            a->IRP[r] = k;
            if (y > 0) { a->JA[k] = r - n; a->AS[k++] = -1.0; }
            if (x > 0) { a->JA[k] = r - 1; a->AS[k++] = -1.0; }
            a->JA[k] = r;
            a->AS[k++] = 4.0;
            if (x < n - 1) { a->JA[k] = r + 1; a->AS[k++] = -1.0; }
            if (y < n - 1) { a->JA[k] = r + n; a->AS[k++] = -1.0; }
        }
    a->IRP[a->M] = k;
    a->NZ = k;
    return a;
}

static idx_t aggregate(int n, idx_t r) { return (idx_t)((r / n) / 2) * (n / 2) + (idx_t)((r % n) / 2); }

/* P: one 1.0 per fine row, in the column of its aggregate */
static spmat *aggregation(int n)
{
    const idx_t M = (idx_t)n * n, nc = (idx_t)(n / 2) * (n / 2);
    spmat *p = csr_new(M, nc, M);
    for (idx_t r = 0; r < M; r++) {
        p->IRP[r] = r;
        p->JA[r] = aggregate(n, r);
        p->AS[r] = 1.0;
    }
    p->IRP[M] = M;
    return p;
}

/* host transpose of a 0-based CSR matrix (rows of the result in column order) */
static spmat *transposed(const spmat *a)
{
    spmat *t = csr_new(a->N, a->M, a->NZ);
    idx_t *cnt = calloc(a->N + 1, sizeof(idx_t));
    for (idx_t e = 0; e < a->NZ; e++) cnt[a->JA[e] + 1]++;
    for (idx_t c = 0; c < a->N; c++) cnt[c + 1] += cnt[c];
    memcpy(t->IRP, cnt, (a->N + 1) * sizeof(idx_t));
    for (idx_t r = 0; r < a->M; r++)
        for (idx_t e = a->IRP[r]; e < a->IRP[r + 1]; e++) {
            const idx_t pos = cnt[a->JA[e]]++;
            t->JA[pos] = r;
            t->AS[pos] = a->AS[e];
        }
    free(cnt);
    return t;
}

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            fprintf(stderr, "rap_caller: " __VA_ARGS__); \
            fprintf(stderr, "\n");                        \
            failures++;                                   \
        }                                                 \
    } while (0)

/* every row of c against the dense closed form (pattern and values) */
static void check_against(const spmat *c, const double *dense, idx_t nc, const char *what)
{
    if (!c) {
        CHECK(0, "%s returned NULL", what);
        return;
    }
    CHECK(c->M == nc && c->N == nc, "%s: shape %lu x %lu", what, c->M, c->N);
    if (c->M != nc) return;
    idx_t want_nz = 0;
    for (idx_t i = 0; i < nc * nc; i++) want_nz += dense[i] != 0.0;
    CHECK(c->NZ == want_nz && c->IRP[nc] == want_nz, "%s: %lu nonzeros, want %lu", what, c->NZ, want_nz);
    for (idx_t r = 0; r < nc; r++) {
        idx_t e = c->IRP[r];
        for (idx_t j = 0; j < nc; j++) {
            const double v = dense[r * nc + j];
            if (v == 0.0) continue;
            if (e >= c->IRP[r + 1] || c->JA[e] != j) {
                CHECK(0, "%s: row %lu misses column %lu", what, r, j);
                return;
            }
            CHECK(fabs(c->AS[e] - v) <= 1e-12 * fabs(v), "%s: (%lu, %lu) = %g, want %g", what, r, j, c->AS[e], v);
            e++;
        }
        CHECK(e == c->IRP[r + 1], "%s: row %lu has extra entries", what, r);
    }
}

int main(void)
{
    const int n = 48;
    spmat *A = laplacian(n), *P = aggregation(n), *R = transposed(P);
    const idx_t nc = P->N;
    double *dense = calloc(nc * nc, sizeof(double));
    for (idx_t i = 0; i < A->M; i++)
        for (idx_t e = A->IRP[i]; e < A->IRP[i + 1]; e++)
            dense[aggregate(n, i) * nc + aggregate(n, A->JA[e])] += A->AS[e];

    CONFIG cfg = {1, 1, IDXMAP, 0, NULL};

    /* sizing: the upper bound of R*A (Sp3MM_test.c sizes with it) */
    idx_t *ub = spMMSizeUpperbound_0(R, A);
    CHECK(ub != NULL, "spMMSizeUpperbound_0 returned NULL");
    if (ub) {
        idx_t want = 0;
        for (idx_t e = 0; e < R->NZ; e++) want += A->IRP[R->JA[e] + 1] - A->IRP[R->JA[e]];
        CHECK(ub[R->M] == want, "upper bound total %lu, want %lu", ub[R->M], want);
        free(ub);
    }

    /* the pair form with the library's SpMMs as the function pointer */
    SPMM_INTERF spmms[2] = {&spmmSerial_0, &spmmRowByRow_SymbNum_0};
    for (int i = 0; i < 2; i++) {
        spmat *c = sp3mmRowByRowPair_0(R, A, P, &cfg, spmms[i]);
        check_against(c, dense, nc, i ? "pair (SymbNum)" : "pair (serial)");
        if (c) freeSpmat(c);
    }
    /* direct forwarding */
    spmat *cm = sp3mmRowByRowMerged_0(R, A, P, &cfg, NULL);
    check_against(cm, dense, nc, "merged");
    printf("device time of the last product: %.3f ms\n", ElapsedInternal * 1e3);
    if (cm) freeSpmat(cm);

    /* the 1-based set, as a Fortran caller passes its data */
    spmat *R1 = csr_copy(R), *A1 = csr_copy(A), *P1 = csr_copy(P);
    C_FortranShiftIdxs(R1);
    C_FortranShiftIdxs(A1);
    C_FortranShiftIdxs(P1);
    CHECK(R1->IRP[0] == 1 && R1->JA[0] == R->JA[0] + 1, "C_FortranShiftIdxs did not shift to base 1");
    spmat *c1 = sp3mmRowByRowMerged_1(R1, A1, P1, &cfg, NULL);
    CHECK(c1 && c1->IRP[0] == 1, "sp3mmRowByRowMerged_1: result not 1-based");
    if (c1) {
        Fortran_C_ShiftIdxs(c1);
        check_against(c1, dense, nc, "merged, 1-based");
        freeSpmat(c1);
    }
    spmat *c2 = sp3mmRowByRowPair_1(R1, A1, P1, &cfg, &spmmRowByRow_1);
    if (c2) {
        Fortran_C_ShiftIdxs(c2);
        check_against(c2, dense, nc, "pair, 1-based");
        freeSpmat(c2);
    } else {
        CHECK(0, "sp3mmRowByRowPair_1 returned NULL");
    }

    freeSpmat(R1);
    freeSpmat(A1);
    freeSpmat(P1);
    freeSpmat(R);
    freeSpmat(A);
    freeSpmat(P);
    free(dense);
    if (failures) {
        fprintf(stderr, "rap_caller: %d check(s) failed\n", failures);
        return 1;
    }
    printf("rap_caller ok (%lu coarse rows)\n", nc);
    return 0;
}
