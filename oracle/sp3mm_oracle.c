/*
 * sp3mm_oracle.c -- CPU oracle for the sparse triple product R*A*P on sorted CSR.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, never by the product path.
 *
 * It restates the reference library's CPU algorithms (Sp3MM4AMG, C99+OpenMP):
 *   - upper-bound row sizing          code/src/include/SpMMUtilsGeneric.h:112-174
 *   - bitmap symbolic row lengths      code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:164-231,281-417
 *   - Gustavson numeric with a dense accumulator, first-touch index append and
 *     per-row sort                     code/src/include/SpMMUtilsGeneric.h:39-103,230-241
 *                                      code/src/Sp3MM_CSR_OMP_Num_Generic.c:93-147
 *   - pair product (R*A)*P            code/src/Sp3MM_CSR_OMP_UB_Generic.c:456-494
 *   - direct row forwarding            code/src/Sp3MM_CSR_OMP_UB_Generic.c:498-573
 *   - 0/1-based index shift            code/src/include/sparseUtilsGeneric.h:64-72
 *
 * Arithmetic order is the reference's: for output row r the contributions of
 * the left row's entries are added in stored (sorted-column) order, each one
 * scanning the right row in stored order, with a separate multiply and add
 * (the reference is built with gcc -O2 for x86-64, which does not contract to
 * FMA; this file is built with -ffp-contract=off for the same reason).  The
 * direct form forwards the intermediate row in first-touch order, as the
 * reference does.
 *
 * Two accumulator-reset policies produce bit-identical outputs:
 *   faithful=1  clears the full-width dense accumulator for every row, as the
 *               reference does (code/src/include/sparseMatrix.h:106-110);
 *               this is the timed CPU baseline;
 *   faithful=0  clears only the touched slots; this is the fast checker.
 *
 * Indices are int64 on both sides of this interface; `base` (0 or 1) is the
 * index base of the inputs and of the returned matrix.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

typedef struct {
    int64_t M, N, NZ;
    const int64_t *irp;   /* [M+1], base-shifted */
    const int64_t *ja;    /* [NZ],  base-shifted */
    const double  *as;    /* [NZ] (may be NULL for symbolic-only calls) */
} ocsr_in;

typedef struct {
    int64_t M, N, NZ;
    int64_t *irp;
    int64_t *ja;
    double  *as;
} ocsr_out;

/* ---- per-thread dense accumulator (sparseMatrix.h:72-77 semantics) ----- */
typedef struct {
    int64_t  width;
    double  *v;        /* dense values                         */
    int64_t *touched;  /* first-touch ordered column list        */
    uint64_t *bits;    /* presence bitmap, one bit per column   */
    int64_t  len;      /* number of touched columns              */
} dacc;

static int dacc_init(dacc *a, int64_t width)
{
    a->width = width;
    a->len = 0;
    a->v = calloc((size_t)(width ? width : 1), sizeof(double));
    a->touched = calloc((size_t)(width ? width : 1), sizeof(int64_t));
    a->bits = calloc((size_t)((width + 63) / 64 + 1), sizeof(uint64_t));
    return (a->v && a->touched && a->bits) ? 0 : -1;
}

static void dacc_free(dacc *a)
{
    free(a->v); free(a->touched); free(a->bits);
    a->v = NULL; a->touched = NULL; a->bits = NULL;
}

/* returns 1 when column j is touched for the first time in this row */
static inline int dacc_mark(dacc *a, int64_t j)
{
    uint64_t m = (uint64_t)1 << (j & 63);
    uint64_t *w = a->bits + (j >> 6);
    if (*w & m) return 0;
    *w |= m;
    a->touched[a->len++] = j;
    return 1;
}

/* scale-and-accumulate one right-hand row: v[j] += b_kj * s  (SpMMUtilsGeneric.h:39-55) */
static inline void dacc_axpy_row(dacc *a, double s, const ocsr_in *B, int64_t k, int base)
{
    int64_t lo = B->irp[k] - base, hi = B->irp[k + 1] - base;
    for (int64_t t = lo; t < hi; t++) {
        int64_t j = B->ja[t] - base;
        double prod = B->as[t] * s;   /* rounded product; built without FP contraction */
        a->v[j] += prod;
        dacc_mark(a, j);
    }
}

static void dacc_reset(dacc *a, int faithful)
{
    if (faithful) {            /* full-width clear, sparseMatrix.h:106-110 */
        memset(a->v, 0, (size_t)a->width * sizeof(double));
        memset(a->touched, 0, (size_t)a->width * sizeof(int64_t));
        memset(a->bits, 0, (size_t)((a->width + 63) / 64 + 1) * sizeof(uint64_t));
    } else {                   /* clear only what the row touched */
        for (int64_t t = 0; t < a->len; t++) {
            int64_t j = a->touched[t];
            a->v[j] = 0.0;
            a->bits[j >> 6] = 0;
        }
    }
    a->len = 0;
}

static int cmp_i64(const void *x, const void *y)
{
    int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
    return (a > b) - (a < b);
}

static int threads_for(int64_t rows)
{
    int t = omp_get_max_threads();
    if (rows < 64) t = 1;
    return t < 1 ? 1 : t;
}

/* ------------------------------------------------------------------------ */
/* Upper bound: ub[r] = sum_{k in A_r} nnz(B_k); ub[M] = total.              */
/* SpMMUtilsGeneric.h:112-139                                                */
int orc_ub(const ocsr_in *A, const ocsr_in *B, int base, int64_t *ub)
{
    int64_t total = 0;
    #pragma omp parallel for schedule(static) reduction(+:total)
    for (int64_t r = 0; r < A->M; r++) {
        int64_t s = 0;
        for (int64_t t = A->irp[r] - base; t < A->irp[r + 1] - base; t++) {
            int64_t k = A->ja[t] - base;
            s += B->irp[k + 1] - B->irp[k];
        }
        ub[r] = s;
        total += s;
    }
    ub[A->M] = total;
    return 0;
}

/* Upper bound split over gridCols uniform column ranges of B's columns.
 * out[r*gridCols+g] = sum_{k in A_r} nnz(B_k restricted to range g); out[M*gridCols]=total.
 * Column ranges give the first (N % gridCols) ranges one extra column
 * (macros.h:50-55); per-row range offsets as in sparseUtilsGeneric.c:284-307.   */
int orc_ub_colparts(const ocsr_in *A, const ocsr_in *B, int base, int gridCols, int64_t *out)
{
    if (gridCols < 1) return -1;
    int64_t cb = B->N / gridCols, cr = B->N % gridCols;
    int64_t total = 0;
    #pragma omp parallel for schedule(static) reduction(+:total)
    for (int64_t r = 0; r < A->M; r++) {
        for (int g = 0; g < gridCols; g++) out[r * gridCols + g] = 0;
        for (int64_t t = A->irp[r] - base; t < A->irp[r + 1] - base; t++) {
            int64_t k = A->ja[t] - base;
            int64_t lo = B->irp[k] - base, hi = B->irp[k + 1] - base;
            int64_t p = lo;
            for (int g = 0; g < gridCols; g++) {
                int64_t gend = (int64_t)(g + 1) * cb + (g + 1 < cr ? g + 1 : cr);
                int64_t q = p;
                while (q < hi && B->ja[q] - base < gend) q++;
                if (g == gridCols - 1) q = hi;
                out[r * gridCols + g] += q - p;
                total += q - p;
                p = q;
            }
        }
    }
    out[A->M * (int64_t)gridCols] = total;
    return 0;
}

/* Symbolic row lengths of A*B with a presence bitmap (SymbStep_Generic.c:164-231). */
int orc_symb(const ocsr_in *A, const ocsr_in *B, int base, int64_t *rowlens)
{
    int nt = threads_for(A->M);
    int64_t total = 0;
    int err = 0;
    #pragma omp parallel num_threads(nt) reduction(+:total)
    {
        dacc acc;
        int ok = dacc_init(&acc, B->N) == 0;
        if (!ok) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(static)
        for (int64_t r = 0; r < A->M; r++) {
            if (!ok) continue;
            for (int64_t t = A->irp[r] - base; t < A->irp[r + 1] - base; t++) {
                int64_t k = A->ja[t] - base;
                for (int64_t u = B->irp[k] - base; u < B->irp[k + 1] - base; u++)
                    dacc_mark(&acc, B->ja[u] - base);
            }
            rowlens[r] = acc.len;
            total += acc.len;
            dacc_reset(&acc, 0);
        }
        dacc_free(&acc);
    }
    rowlens[A->M] = total;
    return err ? -1 : 0;
}

/* Direct triple symbolic: row lengths of R*A*P without forming R*A.
 * Built from the mathematics (the reference's own version of this step,
 * SymbStep_Generic.c:423-579, is not wired to any numeric phase).            */
int orc_symb3(const ocsr_in *R, const ocsr_in *A, const ocsr_in *P, int base, int64_t *rowlens)
{
    int nt = threads_for(R->M);
    int64_t total = 0;
    int err = 0;
    #pragma omp parallel num_threads(nt) reduction(+:total)
    {
        dacc a1, a2;
        int ok = (dacc_init(&a1, A->N) == 0) & (dacc_init(&a2, P->N) == 0);
        if (!ok) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(static)
        for (int64_t r = 0; r < R->M; r++) {
            if (!ok) continue;
            for (int64_t t = R->irp[r] - base; t < R->irp[r + 1] - base; t++) {
                int64_t k = R->ja[t] - base;
                for (int64_t u = A->irp[k] - base; u < A->irp[k + 1] - base; u++)
                    dacc_mark(&a1, A->ja[u] - base);
            }
            for (int64_t t = 0; t < a1.len; t++) {
                int64_t j = a1.touched[t];
                for (int64_t u = P->irp[j] - base; u < P->irp[j + 1] - base; u++)
                    dacc_mark(&a2, P->ja[u] - base);
            }
            rowlens[r] = a2.len;
            total += a2.len;
            dacc_reset(&a1, 0);
            dacc_reset(&a2, 0);
        }
        dacc_free(&a1);
        dacc_free(&a2);
    }
    rowlens[R->M] = total;
    return err ? -1 : 0;
}

static int out_alloc(ocsr_out *C, int64_t M, int64_t N)
{
    C->M = M; C->N = N; C->NZ = 0;
    C->irp = calloc((size_t)M + 1, sizeof(int64_t));
    C->ja = NULL; C->as = NULL;
    return C->irp ? 0 : -1;
}

void orc_free(ocsr_out *C)
{
    if (!C) return;
    free(C->irp); free(C->ja); free(C->as);
    C->irp = NULL; C->ja = NULL; C->as = NULL;
}

/* sort the touched list of one row and write it at C row offset (SpMMUtilsGeneric.h:230-241) */
static void emit_sorted(dacc *acc, ocsr_out *C, int64_t off)
{
    qsort(acc->touched, (size_t)acc->len, sizeof(int64_t), cmp_i64);
    for (int64_t t = 0; t < acc->len; t++) {
        int64_t j = acc->touched[t];
        C->ja[off + t] = j;
        C->as[off + t] = acc->v[j];
    }
}

static void shift_out(ocsr_out *C, int base)
{
    if (!base) return;
    for (int64_t r = 0; r <= C->M; r++) C->irp[r] += base;
    for (int64_t t = 0; t < C->NZ; t++) C->ja[t] += base;
}

/* Symbolic + numeric Gustavson SpMM C = A*B (Num_Generic.c:93-147). */
int orc_spmm(const ocsr_in *A, const ocsr_in *B, int base, int faithful, ocsr_out *C)
{
    if (A->N != B->M) return -2;
    if (out_alloc(C, A->M, B->N)) return -1;
    int64_t *lens = malloc(((size_t)A->M + 1) * sizeof(int64_t));
    if (!lens) return -1;
    if (orc_symb(A, B, base, lens)) { free(lens); return -1; }
    int64_t s = 0;
    for (int64_t r = 0; r < A->M; r++) { C->irp[r] = s; s += lens[r]; }
    C->irp[A->M] = s;
    C->NZ = s;
    free(lens);
    C->ja = malloc((size_t)(s ? s : 1) * sizeof(int64_t));
    C->as = malloc((size_t)(s ? s : 1) * sizeof(double));
    if (!C->ja || !C->as) return -1;
    int nt = threads_for(A->M);
    int err = 0;
    #pragma omp parallel num_threads(nt)
    {
        dacc acc;
        int ok = dacc_init(&acc, B->N) == 0;
        if (!ok) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < A->M; r++) {
            if (!ok) continue;
            dacc_reset(&acc, faithful);
            for (int64_t t = A->irp[r] - base; t < A->irp[r + 1] - base; t++)
                dacc_axpy_row(&acc, A->as[t], B, A->ja[t] - base, base);
            emit_sorted(&acc, C, C->irp[r]);
        }
        dacc_free(&acc);
    }
    if (err) return -1;
    shift_out(C, base);
    return 0;
}

/* Pair form: (R*A)*P, two SpMMs (UB_Generic.c:456-494). */
int orc_sp3mm_pair(const ocsr_in *R, const ocsr_in *A, const ocsr_in *P, int base,
                   int faithful, ocsr_out *C)
{
    ocsr_out RA;
    memset(&RA, 0, sizeof(RA));
    int rc = orc_spmm(R, A, base, faithful, &RA);
    if (rc) { orc_free(&RA); return rc; }
    ocsr_in ra = { RA.M, RA.N, RA.NZ, RA.irp, RA.ja, RA.as };
    rc = orc_spmm(&ra, P, base, faithful, C);
    orc_free(&RA);
    return rc;
}

/* Direct form: row r of R*A is accumulated densely, then forwarded into P in
 * first-touch order of its columns (UB_Generic.c:533-550).  Output rows are
 * sized by the direct symbolic pass and written sorted.                      */
int orc_sp3mm_merged(const ocsr_in *R, const ocsr_in *A, const ocsr_in *P, int base,
                     int faithful, ocsr_out *C)
{
    if (R->N != A->M || A->N != P->M) return -2;
    if (out_alloc(C, R->M, P->N)) return -1;
    int64_t *lens = malloc(((size_t)R->M + 1) * sizeof(int64_t));
    if (!lens) return -1;
    if (orc_symb3(R, A, P, base, lens)) { free(lens); return -1; }
    int64_t s = 0;
    for (int64_t r = 0; r < R->M; r++) { C->irp[r] = s; s += lens[r]; }
    C->irp[R->M] = s;
    C->NZ = s;
    free(lens);
    C->ja = malloc((size_t)(s ? s : 1) * sizeof(int64_t));
    C->as = malloc((size_t)(s ? s : 1) * sizeof(double));
    if (!C->ja || !C->as) return -1;
    int nt = threads_for(R->M);
    int err = 0;
    #pragma omp parallel num_threads(nt)
    {
        dacc a1, a2;
        int ok = (dacc_init(&a1, A->N) == 0) & (dacc_init(&a2, P->N) == 0);
        if (!ok) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < R->M; r++) {
            if (!ok) continue;
            for (int64_t t = R->irp[r] - base; t < R->irp[r + 1] - base; t++)
                dacc_axpy_row(&a1, R->as[t], A, R->ja[t] - base, base);
            for (int64_t t = 0; t < a1.len; t++) {
                int64_t c = a1.touched[t];
                dacc_axpy_row(&a2, a1.v[c], P, c, base);
            }
            emit_sorted(&a2, C, C->irp[r]);
            dacc_reset(&a1, faithful);
            dacc_reset(&a2, faithful);
        }
        dacc_free(&a1);
        dacc_free(&a2);
    }
    if (err) return -1;
    shift_out(C, base);
    return 0;
}

/* Extended-precision value check: the pair product accumulated in long double
 * (x87 80-bit).  Pattern identical to orc_sp3mm_pair; values rounded once to
 * double at the end.  Used to bound the rounding error of every FP64 path.   */
int orc_sp3mm_pair_ld(const ocsr_in *R, const ocsr_in *A, const ocsr_in *P, int base,
                      ocsr_out *C)
{
    if (R->N != A->M || A->N != P->M) return -2;
    if (out_alloc(C, R->M, P->N)) return -1;
    int64_t *lens = malloc(((size_t)R->M + 1) * sizeof(int64_t));
    if (!lens) return -1;
    if (orc_symb3(R, A, P, base, lens)) { free(lens); return -1; }
    int64_t s = 0;
    for (int64_t r = 0; r < R->M; r++) { C->irp[r] = s; s += lens[r]; }
    C->irp[R->M] = s;
    C->NZ = s;
    free(lens);
    C->ja = malloc((size_t)(s ? s : 1) * sizeof(int64_t));
    C->as = malloc((size_t)(s ? s : 1) * sizeof(double));
    if (!C->ja || !C->as) return -1;
    int nt = threads_for(R->M);
    int err = 0;
    #pragma omp parallel num_threads(nt)
    {
        long double *v1 = calloc((size_t)A->N + 1, sizeof(long double));
        long double *v2 = calloc((size_t)P->N + 1, sizeof(long double));
        dacc m1, m2;   /* presence tracking only */
        int ok = v1 && v2 && (dacc_init(&m1, A->N) == 0) & (dacc_init(&m2, P->N) == 0);
        if (!ok) {
            #pragma omp atomic write
            err = 1;
        }
        #pragma omp for schedule(dynamic, 64)
        for (int64_t r = 0; r < R->M; r++) {
            if (!ok) continue;
            for (int64_t t = R->irp[r] - base; t < R->irp[r + 1] - base; t++) {
                int64_t k = R->ja[t] - base;
                long double rv = R->as[t];
                for (int64_t u = A->irp[k] - base; u < A->irp[k + 1] - base; u++) {
                    int64_t j = A->ja[u] - base;
                    v1[j] += rv * (long double)A->as[u];
                    dacc_mark(&m1, j);
                }
            }
            for (int64_t t = 0; t < m1.len; t++) {
                int64_t j = m1.touched[t];
                for (int64_t u = P->irp[j] - base; u < P->irp[j + 1] - base; u++) {
                    int64_t c = P->ja[u] - base;
                    v2[c] += v1[j] * (long double)P->as[u];
                    dacc_mark(&m2, c);
                }
            }
            qsort(m2.touched, (size_t)m2.len, sizeof(int64_t), cmp_i64);
            int64_t off = C->irp[r];
            for (int64_t t = 0; t < m2.len; t++) {
                int64_t c = m2.touched[t];
                C->ja[off + t] = c;
                C->as[off + t] = (double)v2[c];
                v2[c] = 0.0L;
            }
            for (int64_t t = 0; t < m1.len; t++) v1[m1.touched[t]] = 0.0L;
            dacc_reset(&m1, 0);
            dacc_reset(&m2, 0);
        }
        free(v1); free(v2);
        if (ok) { dacc_free(&m1); dacc_free(&m2); }
    }
    if (err) return -1;
    shift_out(C, base);
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
double orc_wtime(void) { return omp_get_wtime(); }
