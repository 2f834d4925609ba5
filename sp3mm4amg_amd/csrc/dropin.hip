// dropin.hip -- the Sp3MM4AMG entry points (_0 and _1 variants) on the GPU.
//
// Conventions kept from the reference: borrowed host spmat inputs are not
// modified; results are calloc/malloc'd host memory released by the caller with
// freeSpmat (code/src/include/sparseMatrix.h:154-168); errors print a message
// on stderr and return NULL (code/src/include/macros.h:85-86); the _1 variants
// take and return 1-based IRP/JA (code/src/Fortran_C_indexing.md).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "engine.hpp"

using namespace sp3mm;

extern "C" {
extern double Start, End, Elapsed, ElapsedInternal;
}

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void report(const char *fn)
{
    fprintf(stderr, "%s: %s\n", fn, sp3mm_last_error());
}

// RAII holder of a device copy of a host spmat
struct DevMat {
    sp3mm_dmat m;
    bool ok = false;
    DevMat(const spmat *h, int base, const char *fn)
    {
        memset(&m, 0, sizeof m);
        if (!h || (!h->IRP) || (h->NZ && (!h->JA || !h->AS))) {
            fprintf(stderr, "%s: null matrix argument\n", fn);
            return;
        }
        ok = sp3mm_dmat_upload(&m, (int64_t)h->M, (int64_t)h->N, (int64_t)h->NZ,
                               reinterpret_cast<const int64_t *>(h->IRP),
                               reinterpret_cast<const int64_t *>(h->JA), h->AS, SP3MM_F64, base) == 0;
        if (!ok) report(fn);
    }
    ~DevMat() { sp3mm_dmat_free(&m); }
};

spmat *to_host(const sp3mm_dmat &d, int base, const char *fn)
{
    spmat *o = static_cast<spmat *>(calloc(1, sizeof(spmat)));
    if (!o) {
        fprintf(stderr, "%s: spmat calloc failed\n", fn);
        return nullptr;
    }
    o->M = (idx_t)d.M;
    o->N = (idx_t)d.N;
    o->NZ = (idx_t)d.nnz;
    o->IRP = static_cast<idx_t *>(malloc(((size_t)d.M + 1) * sizeof(idx_t)));
    o->JA = static_cast<idx_t *>(malloc((size_t)(d.nnz ? d.nnz : 1) * sizeof(idx_t)));
    o->AS = static_cast<double *>(malloc((size_t)(d.nnz ? d.nnz : 1) * sizeof(double)));
    if (!o->IRP || !o->JA || !o->AS) {
        fprintf(stderr, "%s: output malloc failed\n", fn);
        freeSpmat(o);
        return nullptr;
    }
    if (sp3mm_dmat_download(&d, reinterpret_cast<int64_t *>(o->IRP), reinterpret_cast<int64_t *>(o->JA),
                            o->AS, base)) {
        report(fn);
        freeSpmat(o);
        return nullptr;
    }
    return o;
}

idx_t *rowlens_to_host(const std::vector<int64_t> &v)
{
    idx_t *o = static_cast<idx_t *>(malloc(v.size() * sizeof(idx_t)));
    if (o) memcpy(o, v.data(), v.size() * sizeof(idx_t));
    return o;
}

// C = A*B on the GPU for every SpMM entry point
spmat *spmm_gpu(spmat *A, spmat *B, int base, const char *fn)
{
    const double t0 = now_s();
    if (!A || !B) {
        fprintf(stderr, "%s: null matrix argument\n", fn);
        return nullptr;
    }
    if (A->N != B->M) {
        fprintf(stderr, "%s: invalid sizes %lu x %lu * %lu x %lu\n", fn, A->M, A->N, B->M, B->N);
        return nullptr;
    }
    DevMat a(A, base, fn), b(B, base, fn);
    if (!a.ok || !b.ok) return nullptr;
    sp3mm_dmat c;
    sp3mm_stats st{};
    if (sp3mm_spgemm(&a.m, &b.m, &c, &st)) {
        report(fn);
        return nullptr;
    }
    spmat *o = to_host(c, base, fn);
    sp3mm_dmat_free(&c);
    Start = t0;
    End = now_s();
    Elapsed = End - t0;
    ElapsedInternal = st.ms_total * 1e-3;
    return o;
}

spmat *rap_gpu(spmat *R, spmat *AC, spmat *P, int base, int path, const char *fn)
{
    const double t0 = now_s();
    if (!R || !AC || !P) {
        fprintf(stderr, "%s: null matrix argument\n", fn);
        return nullptr;
    }
    if (R->N != AC->M || AC->N != P->M) {
        fprintf(stderr, "%s: invalid sizes in R <-> AC <-> P\n", fn);
        return nullptr;
    }
    DevMat r(R, base, fn), a(AC, base, fn), p(P, base, fn);
    if (!r.ok || !a.ok || !p.ok) return nullptr;
    sp3mm_dmat c;
    sp3mm_opts o{};
    o.path = path;
    sp3mm_stats st{};
    if (sp3mm_rap(&r.m, &a.m, &p.m, &c, &o, &st)) {
        report(fn);
        return nullptr;
    }
    spmat *out = to_host(c, base, fn);
    sp3mm_dmat_free(&c);
    Start = t0;
    End = now_s();
    Elapsed = End - t0;
    ElapsedInternal = st.ms_total * 1e-3;
    return out;
}

}  // namespace

extern "C" {

// --- helpers (code/src/include/sparseMatrix.h:154-214, sparseUtilsGeneric.h:64-72)
void freeSpmatInternal(spmat *mat)
{
    if (!mat) return;
    free(mat->AS);
    free(mat->JA);
    free(mat->IRP);
    mat->AS = nullptr;
    mat->JA = nullptr;
    mat->IRP = nullptr;
}

void freeSpmat(spmat *mat)
{
    if (!mat) return;
    freeSpmatInternal(mat);
    free(mat);
}

int allocSpMatrixInternal(idx_t rows, idx_t cols, spmat *mat)
{
    mat->M = rows;
    mat->N = cols;
    if (!(mat->IRP = static_cast<idx_t *>(calloc(rows + 1, sizeof(idx_t))))) {
        fprintf(stderr, "allocSpMatrixInternal: no host memory for %llu row offsets\n",
                (unsigned long long)(rows + 1));
        freeSpmatInternal(mat);
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

spmat *allocSpMatrix(idx_t rows, idx_t cols)
{
    spmat *mat = static_cast<spmat *>(calloc(1, sizeof(spmat)));
    if (!mat) {
        fprintf(stderr, "allocSpMatrix: no host memory for the matrix header\n");
        return nullptr;
    }
    if (allocSpMatrixInternal(rows, cols, mat)) {
        free(mat);
        return nullptr;
    }
    return mat;
}

void C_FortranShiftIdxs(spmat *m)
{
    for (idx_t r = 0; r < m->M + 1; r++) m->IRP[r]++;
    for (idx_t i = 0; i < m->NZ; i++) m->JA[i]++;
}

void Fortran_C_ShiftIdxs(spmat *m)
{
    for (idx_t r = 0; r < m->M + 1; r++) m->IRP[r]--;
    for (idx_t i = 0; i < m->NZ; i++) m->JA[i]--;
}

// --- SpMM entry points
#define SP3MM_SPMM_PAIR(name)                                                   \
    spmat *name##_0(spmat *A, spmat *B, CONFIG *) { return spmm_gpu(A, B, 0, #name "_0"); } \
    spmat *name##_1(spmat *A, spmat *B, CONFIG *) { return spmm_gpu(A, B, 1, #name "_1"); }

SP3MM_SPMM_PAIR(spmmSerial)
SP3MM_SPMM_PAIR(spmmRowByRow)
SP3MM_SPMM_PAIR(spmmRowByRow1DBlocks)
SP3MM_SPMM_PAIR(spmmRowByRow2DBlocks)
SP3MM_SPMM_PAIR(spmmRowByRow2DBlocksAllocated)
SP3MM_SPMM_PAIR(spmmRowByRow_SymbNum)
SP3MM_SPMM_PAIR(spmmRowByRow1DBlocks_SymbNum)
SP3MM_SPMM_PAIR(spmmRowByRow2DBlocks_SymbNum)
#undef SP3MM_SPMM_PAIR

}  // extern "C"

namespace {

bool is_ours(SPMM_INTERF f)
{
    const SPMM_INTERF mine[] = {
        spmmSerial_0, spmmSerial_1, spmmRowByRow_0, spmmRowByRow_1, spmmRowByRow1DBlocks_0,
        spmmRowByRow1DBlocks_1, spmmRowByRow2DBlocks_0, spmmRowByRow2DBlocks_1,
        spmmRowByRow2DBlocksAllocated_0, spmmRowByRow2DBlocksAllocated_1, spmmRowByRow_SymbNum_0,
        spmmRowByRow_SymbNum_1, spmmRowByRow1DBlocks_SymbNum_0, spmmRowByRow1DBlocks_SymbNum_1,
        spmmRowByRow2DBlocks_SymbNum_0, spmmRowByRow2DBlocks_SymbNum_1};
    for (auto m : mine)
        if (m == f) return true;
    return false;
}

// caller-supplied SpMM: compose on the host exactly as the reference does
// (code/src/Sp3MM_CSR_OMP_UB_Generic.c:473-493)
spmat *pair_with(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm)
{
    spmat *RAC = spmm(R, AC, cfg);
    if (!RAC) return nullptr;
    spmat *out = spmm(RAC, P, cfg);
    freeSpmat(RAC);
    return out;
}

std::vector<int64_t> symbolic_lens(spmat *a, spmat *b, spmat *c, int base, const char *fn, bool &ok)
{
    ok = false;
    std::vector<int64_t> lens;
    if (!a || !b || (a->N != b->M) || (c && b->N != c->M)) {
        fprintf(stderr, "%s: invalid operands\n", fn);
        return lens;
    }
    DevMat da(a, base, fn), db(b, base, fn);
    if (!da.ok || !db.ok) return lens;
    lens.resize(a->M + 1);
    if (c) {
        DevMat dc(c, base, fn);
        if (!dc.ok) return lens;
        if (sp3mm_symbolic(&da.m, &db.m, &dc.m, lens.data())) {
            report(fn);
            return lens;
        }
    } else if (sp3mm_symbolic(&da.m, &db.m, nullptr, lens.data())) {
        report(fn);
        return lens;
    }
    ok = true;
    return lens;
}

// row lengths + (optionally) sorted row indices in one allocation
idx_t *symb_out(spmat *a, spmat *b, spmat *c, int base, idx_t ***outIdxs, const char *fn)
{
    bool ok;
    if (!outIdxs) {
        std::vector<int64_t> lens = symbolic_lens(a, b, c, base, fn, ok);
        return ok ? rowlens_to_host(lens) : nullptr;
    }
    spmat *prod = c ? rap_gpu(a, b, c, base, SP3MM_PATH_DIRECT, fn) : spmm_gpu(a, b, base, fn);
    if (!prod) return nullptr;
    std::vector<int64_t> lens(prod->M + 1);
    for (idx_t r = 0; r < prod->M; r++) lens[r] = (int64_t)(prod->IRP[r + 1] - prod->IRP[r]);
    lens[prod->M] = (int64_t)prod->NZ;
    idx_t **rows = static_cast<idx_t **>(malloc((prod->M ? prod->M : 1) * sizeof(idx_t *)));
    idx_t *rl = rowlens_to_host(lens);
    if (!rows || !rl) {
        free(rows);
        free(rl);
        freeSpmat(prod);
        fprintf(stderr, "%s: malloc failed\n", fn);
        return nullptr;
    }
    idx_t *all = prod->JA;  // keeps the (base-shifted) column indices
    prod->JA = nullptr;
    for (idx_t r = 0; r < prod->M; r++) rows[r] = all + (prod->IRP[r] - base);
    if (prod->M == 0) rows[0] = all;
    freeSpmat(prod);
    *outIdxs = rows;
    return rl;
}

// Row lengths of A*B and, per row, the nonzeros in each of gridCols column
// shares of B's N columns: share g starts at g*(N/gridCols) + min(g, N %
// gridCols) and the first N % gridCols shares hold one column more (the
// reference's uniform shares, macros.h:49-52).  outIdxs as SpMM_Symb_OutIdxs__.
idx_t *symb_colparts(spmat *a, spmat *b, int base, idx_t ***outIdxs, unsigned short gridCols,
                     idx_t **rowColPartsLens, const char *fn)
{
    if (!a || !b || !rowColPartsLens || gridCols == 0) {
        fprintf(stderr, "%s: null argument or no column partitions\n", fn);
        return nullptr;
    }
    idx_t **rows = nullptr;
    idx_t *rl = symb_out(a, b, nullptr, base, &rows, fn);
    if (!rl) return nullptr;
    const idx_t M = a->M, N = b->N, div = N / gridCols, rem = N % gridCols, wide = rem * (div + 1);
    idx_t *parts = static_cast<idx_t *>(calloc((size_t)(M ? M : 1) * gridCols, sizeof(idx_t)));
    if (!parts) {
        fprintf(stderr, "%s: malloc failed\n", fn);
        free(rl);
        if (M) free(rows[0]);
        free(rows);
        return nullptr;
    }
    for (idx_t r = 0; r < M; r++)
        for (idx_t j = 0; j < rl[r]; j++) {
            const idx_t c = rows[r][j] - base;
            const idx_t g = c < wide ? c / (div + 1) : rem + (c - wide) / div;
            parts[r * gridCols + g]++;
        }
    *rowColPartsLens = parts;
    if (outIdxs) {
        *outIdxs = rows;
    } else {
        free(rows[0]);
        free(rows);
    }
    return rl;
}

}  // namespace

extern "C" {

// --- triple products
// With no SpMM of the caller's own, the pair entry points compute R*AC*P with
// the direct kernels: the same pattern (structural products are associative)
// and FP64 values within the same summation-order tolerance, in about half
// the time of two SpGEMMs with the intermediate in HBM (sp3mm_rap with
// SP3MM_PATH_PAIR keeps that form for callers who want it).
spmat *sp3mmRowByRowPair_0(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm)
{
    if (spmm && !is_ours(spmm)) return pair_with(R, AC, P, cfg, spmm);
    return rap_gpu(R, AC, P, 0, SP3MM_PATH_DIRECT, "sp3mmRowByRowPair_0");
}
spmat *sp3mmRowByRowPair_1(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm)
{
    if (spmm && !is_ours(spmm)) return pair_with(R, AC, P, cfg, spmm);
    return rap_gpu(R, AC, P, 1, SP3MM_PATH_DIRECT, "sp3mmRowByRowPair_1");
}
spmat *sp3mmRowByRowMerged_0(spmat *R, spmat *AC, spmat *P, CONFIG *, SPMM_INTERF)
{
    return rap_gpu(R, AC, P, 0, SP3MM_PATH_DIRECT, "sp3mmRowByRowMerged_0");
}
spmat *sp3mmRowByRowMerged_1(spmat *R, spmat *AC, spmat *P, CONFIG *, SPMM_INTERF)
{
    return rap_gpu(R, AC, P, 1, SP3MM_PATH_DIRECT, "sp3mmRowByRowMerged_1");
}

// --- sizing and symbolic phases
static idx_t *ub_gpu(spmat *A, spmat *B, int base, const char *fn)
{
    if (!A || !B || A->N != B->M) {
        fprintf(stderr, "%s: invalid operands\n", fn);
        return nullptr;
    }
    DevMat a(A, base, fn), b(B, base, fn);
    if (!a.ok || !b.ok) return nullptr;
    idx_t *out = static_cast<idx_t *>(calloc(A->M + 1, sizeof(idx_t)));
    if (!out) {
        fprintf(stderr, "%s: rowSizes calloc errd\n", fn);
        return nullptr;
    }
    if (sp3mm_ub(&a.m, &b.m, reinterpret_cast<int64_t *>(out))) {
        report(fn);
        free(out);
        return nullptr;
    }
    return out;
}
idx_t *spMMSizeUpperbound_0(spmat *A, spmat *B) { return ub_gpu(A, B, 0, "spMMSizeUpperbound_0"); }
idx_t *spMMSizeUpperbound_1(spmat *A, spmat *B) { return ub_gpu(A, B, 1, "spMMSizeUpperbound_1"); }

static idx_t *ub_colparts_gpu(spmat *A, spmat *B, unsigned short gridCols, idx_t *boff, int base,
                              const char *fn)
{
    if (!A || !B || !boff || gridCols == 0 || A->N != B->M) {
        fprintf(stderr, "%s: invalid operands\n", fn);
        return nullptr;
    }
    DevMat a(A, base, fn);
    if (!a.ok) return nullptr;
    idx_t *out = static_cast<idx_t *>(calloc(A->M * gridCols + 1, sizeof(idx_t)));
    if (!out) {
        fprintf(stderr, "%s: rowPartsSizes calloc errd\n", fn);
        return nullptr;
    }
    if (ub_colparts(a.m, reinterpret_cast<const int64_t *>(boff), (int64_t)(B->M * gridCols + 1), gridCols,
                    reinterpret_cast<int64_t *>(out))) {
        report(fn);
        free(out);
        return nullptr;
    }
    return out;
}
idx_t *spMMSizeUpperboundColParts_0(spmat *A, spmat *B, unsigned short gridCols, idx_t *bColPartOffsets)
{
    return ub_colparts_gpu(A, B, gridCols, bColPartOffsets, 0, "spMMSizeUpperboundColParts_0");
}
idx_t *spMMSizeUpperboundColParts_1(spmat *A, spmat *B, unsigned short gridCols, idx_t *bColPartOffsets)
{
    return ub_colparts_gpu(A, B, gridCols, bColPartOffsets, 1, "spMMSizeUpperboundColParts_1");
}

idx_t *SpMM_Symb___0(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b)
{
    return symb_out(a, b, nullptr, 0, nullptr, "SpMM_Symb___0");
}
idx_t *SpMM_Symb___1(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b)
{
    return symb_out(a, b, nullptr, 1, nullptr, "SpMM_Symb___1");
}
idx_t *SpMM_Symb_OutIdxs__0(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, idx_t ***outIdxs)
{
    return symb_out(a, b, nullptr, 0, outIdxs, "SpMM_Symb_OutIdxs__0");
}
idx_t *SpMM_Symb_OutIdxs__1(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, idx_t ***outIdxs)
{
    return symb_out(a, b, nullptr, 1, outIdxs, "SpMM_Symb_OutIdxs__1");
}
idx_t *SpMM_Symb__ColParts_0(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, unsigned short gridCols,
                             idx_t **rowColPartsLens)
{
    return symb_colparts(a, b, 0, nullptr, gridCols, rowColPartsLens, "SpMM_Symb__ColParts_0");
}
idx_t *SpMM_Symb__ColParts_1(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, unsigned short gridCols,
                             idx_t **rowColPartsLens)
{
    return symb_colparts(a, b, 1, nullptr, gridCols, rowColPartsLens, "SpMM_Symb__ColParts_1");
}
idx_t *SpMM_Symb_OutIdxs_ColParts_0(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, idx_t ***outIdxs,
                                    unsigned short gridCols, idx_t **rowColPartsLens)
{
    return symb_colparts(a, b, 0, outIdxs, gridCols, rowColPartsLens, "SpMM_Symb_OutIdxs_ColParts_0");
}
idx_t *SpMM_Symb_OutIdxs_ColParts_1(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, idx_t ***outIdxs,
                                    unsigned short gridCols, idx_t **rowColPartsLens)
{
    return symb_colparts(a, b, 1, outIdxs, gridCols, rowColPartsLens, "SpMM_Symb_OutIdxs_ColParts_1");
}
idx_t *Sp3MM_Symb_OutIdxs_0(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, spmat *c, idx_t ***outIdxs)
{
    return symb_out(a, b, c, 0, outIdxs, "Sp3MM_Symb_OutIdxs_0");
}
idx_t *Sp3MM_Symb_OutIdxs_1(ROW_MMSYM_IMPL_MODE, spmat *a, spmat *b, spmat *c, idx_t ***outIdxs)
{
    return symb_out(a, b, c, 1, outIdxs, "Sp3MM_Symb_OutIdxs_1");
}

}  // extern "C"
