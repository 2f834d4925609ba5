/*
 * sp3mm_amd.h -- C ABI of the MI355X (gfx950) Galerkin triple product R*A*P.
 *
 * Part 1 is a drop-in for the Sp3MM4AMG C library: the same structs with the
 * same byte layout, and the same entry-point names and signatures, in the 0-based
 * (_0) and 1-based Fortran (_1) variants.  Every entry point computes on the
 * GPU; host arrays are borrowed inputs, outputs are malloc'd host memory that
 * the caller frees with free()/freeSpmat(), and errors return NULL with a
 * message on stderr, as in the reference.
 *
 * Part 2 is a device-resident interface (no counterpart in the reference) for
 * callers that keep matrices in HBM: uploads, products on device-resident CSR,
 * downloads, a batched form and run statistics.
 *
 * Reference paths below are relative to the Sp3MM4AMG repository.
 */
#ifndef SP3MM_AMD_H
#define SP3MM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================= */
/* Part 1: reference-compatible types and entry points                     */
/* ======================================================================= */

/* code/src/include/config.h:51 -- 64-bit unsigned index type */
typedef unsigned long idx_t;

/* code/src/include/config.h:36-40 -- symbolic row implementation selector.
 * Both values select the same GPU symbolic phase (LDS hash / bitmap). */
typedef enum { RBTREE, IDXMAP } ROW_MMSYM_IMPL_MODE;

/* code/src/include/config.h:42-49 -- run configuration.  gridRows/gridCols,
 * threadNum and chunkDistrbFunc steer the reference's OpenMP loops; they are
 * accepted and ignored (chunkDistrbFunc is never called). */
typedef struct {
    unsigned short gridRows;
    unsigned short gridCols;
    ROW_MMSYM_IMPL_MODE symbMMRowImplID;
    unsigned int threadNum;
    void *chunkDistrbFunc;
} CONFIG;

/* code/src/include/sparseMatrix.h:41-53 -- CSR matrix (built without ROWLENS) */
typedef struct {
    idx_t NZ, M, N;
    double *AS;
    idx_t *JA;
    idx_t *IRP;
    idx_t MAX_ROW_NZ;
} spmat;

/* code/src/include/Sp3MM_CSR_OMP_Multi.h:50-53 */
typedef spmat *(*SPMM_INTERF)(spmat *, spmat *, CONFIG *);
typedef spmat *(*SP3MM_INTERF)(spmat *, spmat *, spmat *, CONFIG *, SPMM_INTERF);

/* code/src/Sp3MM_CSR_OMP_Multi.c:54 -- audit timers (seconds).  ElapsedInternal
 * receives the device compute time of the last product (uploads/downloads
 * excluded); Elapsed the wall time of the last call. */
extern double Start, End, Elapsed, ElapsedInternal;

/* --- triple products ---------------------------------------------------- */
/* code/src/Sp3MM_CSR_OMP_UB_Generic.c:456-494 (declared in
 * code/src/include/Sp3MM_CSR_OMP_UB_Generic.h:44-50): (R*AC)*P as two SpMMs.
 * When spmm is NULL or one of this library's SpMM entry points the product
 * runs on the GPU with the direct kernels (the same pattern, FP64 values
 * within the summation-order tolerance; sp3mm_rap with SP3MM_PATH_PAIR gives
 * the two-SpGEMM form); any other function pointer is called twice on the
 * host, as in the reference. */
spmat *sp3mmRowByRowPair_0(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm);
spmat *sp3mmRowByRowPair_1(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm);

/* code/src/Sp3MM_CSR_OMP_UB_Generic.c:498-573 (header :52-57): direct row
 * forwarding; each row of R*AC is built in LDS and forwarded into P without
 * an intermediate matrix in HBM.  spmm is ignored, as in the reference. */
spmat *sp3mmRowByRowMerged_0(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm);
spmat *sp3mmRowByRowMerged_1(spmat *R, spmat *AC, spmat *P, CONFIG *cfg, SPMM_INTERF spmm);

/* --- SpMM C = A*B --------------------------------------------------------- */
/* All return the same sorted CSR product; in the reference they differ only in
 * their OpenMP decomposition:
 *   spmmSerial_              code/src/Sp3MM_CSR_OMP_UB_Generic.c:39-53
 *   spmmRowByRow_            code/src/Sp3MM_CSR_OMP_UB_Generic.c:56-124
 *   spmmRowByRow1DBlocks_    code/src/Sp3MM_CSR_OMP_UB_Generic.c:126-201
 *   spmmRowByRow2DBlocks_    code/src/Sp3MM_CSR_OMP_UB_Generic.c:205-323
 *   spmmRowByRow2DBlocksAllocated_  code/src/Sp3MM_CSR_OMP_UB_Generic.c:325-454
 *   spmmRowByRow_SymbNum_    code/src/Sp3MM_CSR_OMP_Num_Generic.c:93-147
 *   spmmRowByRow1DBlocks_SymbNum_   code/src/Sp3MM_CSR_OMP_Num_Generic.c:148-210
 *   spmmRowByRow2DBlocks_SymbNum_   code/src/Sp3MM_CSR_OMP_Num_Generic.c:211-239
 *       (a stub that always returns NULL in the reference; here it computes the
 *        product like the others). */
spmat *spmmSerial_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmSerial_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow1DBlocks_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow1DBlocks_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocks_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocks_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocksAllocated_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocksAllocated_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow_SymbNum_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow_SymbNum_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow1DBlocks_SymbNum_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow1DBlocks_SymbNum_1(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocks_SymbNum_0(spmat *A, spmat *B, CONFIG *cfg);
spmat *spmmRowByRow2DBlocks_SymbNum_1(spmat *A, spmat *B, CONFIG *cfg);

/* --- sizing and symbolic phases (arrays are malloc'd; caller frees) ----- */
/* code/src/include/SpMMUtilsGeneric.h:112-139: ub[r] = sum_{k in A_r} nnz(B_k),
 * ub[A->M] = total; A->M+1 entries. */
idx_t *spMMSizeUpperbound_0(spmat *A, spmat *B);
idx_t *spMMSizeUpperbound_1(spmat *A, spmat *B);

/* code/src/include/SpMMUtilsGeneric.h:146-174: the same bound split over
 * gridCols column ranges of B.  bColPartOffsets holds, for every row of B and
 * every range, the index into B->JA where the range starts (B->M*gridCols+1
 * entries, as built by colsOffsetsPartitioningUnifRanges_,
 * code/src/commons/sparseUtilsGeneric.c:284-307); A->M*gridCols+1 entries out. */
idx_t *spMMSizeUpperboundColParts_0(spmat *A, spmat *B, unsigned short gridCols,
                                    idx_t *bColPartOffsets);
idx_t *spMMSizeUpperboundColParts_1(spmat *A, spmat *B, unsigned short gridCols,
                                    idx_t *bColPartOffsets);

/* code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:281-417: exact row lengths of A*B
 * plus the total in the last entry (A->M+1 entries).  Both modes are served by
 * the same GPU symbolic kernels. */
idx_t *SpMM_Symb___0(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b);
idx_t *SpMM_Symb___1(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b);

/* code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:281-417 built with OUT_IDXS
 * (code/src/Sp3MM_CSR_OMP_SymbStep_Multi.c:59-92): row lengths as above and,
 * when outIdxs != NULL, *outIdxs = malloc'd array of a->M pointers to each
 * row's sorted column indices.  All rows live in one allocation starting at
 * (*outIdxs)[0]; the caller frees (*outIdxs)[0] and then *outIdxs. */
idx_t *SpMM_Symb_OutIdxs__0(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, idx_t ***outIdxs);
idx_t *SpMM_Symb_OutIdxs__1(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, idx_t ***outIdxs);

/* code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:281-417 built with COL_PARTS (and
 * OUT_IDXS for the second pair): row lengths and indices as above, and
 * *rowColPartsLens = malloc'd a->M x gridCols array (row-major) of each row's
 * nonzeros per column share.  The shares split b's N columns uniformly as the
 * reference's column partitioning does (macros.h:49-52: share g starts at
 * g*(N/gridCols) + min(g, N%gridCols)).  Deviation: the reference sizes the
 * shares from the row's own length and counts into uninitialised memory
 * (SymbStep_Generic.c:196-220, 329-334), so its counts are not defined. */
idx_t *SpMM_Symb__ColParts_0(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, unsigned short gridCols,
                             idx_t **rowColPartsLens);
idx_t *SpMM_Symb__ColParts_1(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, unsigned short gridCols,
                             idx_t **rowColPartsLens);
idx_t *SpMM_Symb_OutIdxs_ColParts_0(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, idx_t ***outIdxs,
                                    unsigned short gridCols, idx_t **rowColPartsLens);
idx_t *SpMM_Symb_OutIdxs_ColParts_1(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, idx_t ***outIdxs,
                                    unsigned short gridCols, idx_t **rowColPartsLens);

/* code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:453-579: exact row lengths of
 * a*b*c by direct row forwarding, plus the total (a->M+1 entries), and the
 * sorted row indices as for SpMM_Symb_OutIdxs__ when outIdxs != NULL. */
idx_t *Sp3MM_Symb_OutIdxs_0(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, spmat *c,
                            idx_t ***outIdxs);
idx_t *Sp3MM_Symb_OutIdxs_1(ROW_MMSYM_IMPL_MODE mode, spmat *a, spmat *b, spmat *c,
                            idx_t ***outIdxs);

/* --- helpers the reference's callers link against ------------------------ */
/* code/src/include/sparseMatrix.h:188-214 */
spmat *allocSpMatrix(idx_t rows, idx_t cols);
int allocSpMatrixInternal(idx_t rows, idx_t cols, spmat *mat);
/* code/src/include/sparseMatrix.h:154-168 */
void freeSpmatInternal(spmat *mat);
void freeSpmat(spmat *mat);
/* code/src/include/sparseUtilsGeneric.h:64-72 */
void C_FortranShiftIdxs(spmat *m);
void Fortran_C_ShiftIdxs(spmat *m);

/* ======================================================================= */
/* Part 2: device-resident interface (no counterpart in the reference)     */
/* ======================================================================= */

enum { SP3MM_F64 = 0, SP3MM_F32 = 1 };
enum { SP3MM_PATH_PAIR = 0, SP3MM_PATH_DIRECT = 1 };
/* FP64: FP64 operands, FP64 arithmetic.
 * MIXED: FP32 operands and multiplies; products formed exactly as FP32 pairs
 *   (TwoProd) and accumulated as FP32 pairs (compensated, TwoSum); diagonal
 *   and row-sum entries accumulated in FP64; rows with a cancelling
 *   off-diagonal entry recomputed in FP64 (see sp3mm_opts.guard).
 * MIXED64: FP32 operands; every FP32 x FP32 product is exact in FP64 and is
 *   accumulated in FP64 (FP32 storage, FP64 sums; no guard needed). */
enum { SP3MM_PREC_FP64 = 0, SP3MM_PREC_MIXED = 1, SP3MM_PREC_MIXED64 = 2 };

/* A sorted CSR matrix in HBM, 0-based.  rp has M+1 int64 offsets; rp32 holds
 * the same offsets as int32 whenever nnz < 2^31 (it is what the gathers read);
 * col holds int32 column indices (N < 2^31); val holds doubles (SP3MM_F64) or
 * floats (SP3MM_F32). */
typedef struct {
    int64_t M, N, nnz;
    int64_t *rp;
    int32_t *rp32;
    int32_t *col;
    void *val;
    int32_t vtype;
    int32_t owned;
} sp3mm_dmat;

typedef struct {
    int32_t path;        /* SP3MM_PATH_PAIR or SP3MM_PATH_DIRECT */
    int32_t precision;   /* SP3MM_PREC_FP64, or SP3MM_PREC_MIXED / _MIXED64 (direct
                            path, FP32 operands, FP64 output) */
    double guard;        /* MIXED: a row holding an off-diagonal entry with
                            |value| < guard * (|R||A||P|)_ij is recomputed in FP64;
                            diagonal entries are always summed in FP64.
                            <= 0 selects the default 1e-2 */
    int32_t timing;      /* 1: time each kernel with HIP events (adds syncs) */
    int32_t reserved;
    double *rowsum;      /* MIXED, optional: device array of R->M doubles that
                            receives the FP64 row sums of R*A*P */
} sp3mm_opts;

typedef struct {
    double ms_total;       /* device time of the whole product                 */
    double ms_ub;          /* upper bounds and binning                          */
    double ms_symbolic;    /* symbolic phase(s)                                */
    double ms_scan;        /* row-offset scans                                  */
    double ms_numeric;     /* numeric phase(s)                                  */
    double ms_main;        /* the dominant kernel (numeric of the last product) */
    int64_t mac1, mac2;    /* multiply-adds of R*A and of (RA)*P; -1: not
                              counted (a direct product whose rows all fit one
                              bin by their longest-row bound is not measured
                              row by row)                                      */
    int64_t nnz_mid;       /* nnz of R*A (pair path) or its symbolic count      */
    int64_t nnz_out;
    int64_t guarded_rows;  /* MIXED: rows recomputed in FP64                    */
    int64_t guarded_entries; /* direct path: rows computed off the hash tables
                                (long rows and rows whose tables filled)       */
    int32_t main_kernel_launches;
    int32_t spilled_rows;  /* direct path: of those, rows whose column windows
                              exceeded LDS (global dense accumulators)          */
    int32_t reassociated;  /* direct path: 1 if the product ran as R*(A*P)
                              (most rows beyond the one-wavefront tables and
                              an output wider than an LDS window)               */
    int32_t reserved2;
} sp3mm_stats;

/* Select (and initialise) the HIP device; called implicitly with device 0. */
int sp3mm_init(int device);
/* Text of the last error of the calling thread ("" if none). */
const char *sp3mm_last_error(void);
/* Block until all queued device work of the library has finished. */
int sp3mm_sync(void);
/* Queue the library's device work on the caller's HIP stream (a hipStream_t
 * created by the caller, not the null stream; NULL restores the library's own
 * stream).  Work already queued on the previous stream finishes first.  Calls
 * that must read sizes back still wait for that stream; sp3mm_batch_create
 * and sp3mm_vcycle_create capture on it. */
int sp3mm_set_stream(void *stream);
void *sp3mm_get_stream(void);

/* Upload host CSR (int64 row offsets and columns in index base `base`,
 * values of type vtype) after validating shape, sortedness and ranges. */
int sp3mm_dmat_upload(sp3mm_dmat *out, int64_t M, int64_t N, int64_t nnz,
                      const int64_t *irp, const int64_t *ja, const void *val,
                      int32_t vtype, int32_t base);
/* Download to host arrays of M+1, nnz, nnz entries (values as stored). */
int sp3mm_dmat_download(const sp3mm_dmat *m, int64_t *irp, int64_t *ja, void *val,
                        int32_t base);
/* The same for CSR already in device memory (device pointers; e.g. a
 * PSBLAS-side matrix in the reference's spmat layout, 1-based with base 1):
 * import validates and converts on the GPU into a library-owned matrix,
 * export writes a matrix back into caller-owned device arrays (M+1, nnz, nnz
 * entries).  No host copy of the arrays is made. */
int sp3mm_dmat_import(sp3mm_dmat *out, int64_t M, int64_t N, int64_t nnz,
                      const int64_t *d_irp, const int64_t *d_ja, const void *d_val,
                      int32_t vtype, int32_t base);
int sp3mm_dmat_export(const sp3mm_dmat *m, int64_t *d_irp, int64_t *d_ja, void *d_val,
                      int32_t base);
/* Convert values to another type in place (allocates a new value array). */
int sp3mm_dmat_convert(sp3mm_dmat *m, int32_t vtype);
void sp3mm_dmat_free(sp3mm_dmat *m);

/* out = R*A*P on device-resident operands. */
int sp3mm_rap(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
              sp3mm_dmat *out, const sp3mm_opts *opts, sp3mm_stats *stats);
/* Bytes of the fixed device scratch array behind sp3mm_rap_noalloc (one copy
 * per GPU the library runs on, held until the process ends). */
#define SP3MM_NOALLOC_SCRATCH_BYTES (256ll << 20)

/* out = R*A*P, FP64, direct path (as sp3mm_rap with SP3MM_PATH_DIRECT and
 * SP3MM_PREC_FP64), with no device allocation at all: no hipMalloc/hipFree,
 * no pool, no HIP events.  It has no counterpart in the reference; it is for
 * callers that cannot allocate device memory (code around HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The caller passes device
 * arrays for the result: out_rp (R->M + 1 int64 offsets, 0-based), out_col
 * (int32) and out_val (double) of `cap` entries each.  Every scratch buffer of
 * the product comes from a fixed device array of SP3MM_NOALLOC_SCRATCH_BYTES
 * bytes (roughly 64 bytes per row of R plus 12 bytes per hash-table slot of
 * every row); a product that needs more is refused with an error that says
 * so.  On success *nnz receives the number of nonzeros written.  If the
 * product has more than `cap` nonzeros the call fails and *nnz receives the
 * capacity a repeated call needs (else -1).  The scratch array is shared by
 * every call on a GPU and keeps the last call's contents; calls return after
 * their device work has finished and are serialised by the library, so calls
 * that overlap on one GPU (e.g. from several processes' streams sharing a
 * device by other means) race and are not supported. */
int sp3mm_rap_noalloc(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                      int64_t *out_rp, int32_t *out_col, double *out_val, int64_t cap,
                      int64_t *nnz, sp3mm_stats *stats);
/* out = A*B. */
int sp3mm_spgemm(const sp3mm_dmat *A, const sp3mm_dmat *B, sp3mm_dmat *out,
                 sp3mm_stats *stats);
/* ub (host, A->M+1 int64) = row upper bounds of A*B, last entry the total. */
int sp3mm_ub(const sp3mm_dmat *A, const sp3mm_dmat *B, int64_t *ub);
/* rowlens (host, A->M+1 int64) = exact row lengths of A*B (C != NULL: of A*B*C). */
int sp3mm_symbolic(const sp3mm_dmat *A, const sp3mm_dmat *B, const sp3mm_dmat *C,
                   int64_t *rowlens);

/* Batched products: `count` independent triples (R[i], A[i], P[i]) -- levels
 * of an AMG hierarchy and/or unrelated problems of uneven sizes -- packed
 * block-diagonally into one product.  create() packs the operands and runs the
 * product once (sizing bins, scratch and output); with use_graph the whole
 * launch sequence is captured into one HIP graph.  run() recomputes every
 * product from the packed operands (repack() refreshes them from the
 * caller's matrices: values may change, patterns may not, and a pattern
 * change that alters the sizes the batch was built for is reported as an
 * error) with one graph launch or a plain launch sequence, and
 * returns the device time in *ms.  result() copies product i into a new
 * device matrix.  Direct path; FP64 operands, or FP32 operands with
 * opts->precision == SP3MM_PREC_MIXED or SP3MM_PREC_MIXED64. */
typedef struct sp3mm_batch sp3mm_batch;
int sp3mm_batch_create(int32_t count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                       const sp3mm_opts *opts, int32_t use_graph, sp3mm_batch **out);
int sp3mm_batch_run(sp3mm_batch *b, int32_t use_graph, double *ms);
/* Copy the current entries of the operands into the batch (the batch works on
 * packed copies made by create()): the same count, shapes and nonzero counts
 * as at creation, values and patterns as they are now.  The next run()
 * recomputes every product from them; a pattern change that alters a
 * product's sizes from those the batch was built for is reported by run(). */
int sp3mm_batch_repack(sp3mm_batch *b, int32_t count, const sp3mm_dmat *R, const sp3mm_dmat *A,
                       const sp3mm_dmat *P);
int sp3mm_batch_result(sp3mm_batch *b, int32_t i, sp3mm_dmat *out);
int sp3mm_batch_info(sp3mm_batch *b, double *ms_pack, double *ms_build, int64_t *nnz, int32_t *has_graph);
void sp3mm_batch_free(sp3mm_batch *b);

/* Setup operations of an aggregation AMG level (no counterpart in the
 * reference, whose caller supplies P and R), so that a hierarchy can be built
 * on the GPU level after level:
 *   sp3mm_transpose: AT = A^T (sorted CSR, values of A's type);
 *   sp3mm_sa_prolongator: P = (I - omega D^-1 A) P0 for square FP64 A, where
 *     P0 is the 0/1 aggregation matrix of agg (host array of A->M aggregate
 *     ids in [0, nagg)); D = diag(A) (diag_mode 0) or |diag(A)| + 1 (1). */
int sp3mm_transpose(const sp3mm_dmat *A, sp3mm_dmat *AT);
int sp3mm_sa_prolongator(const sp3mm_dmat *A, const int64_t *agg, int64_t nagg, double omega,
                         int32_t diag_mode, sp3mm_dmat *P);

/* A whole smoothed-aggregation hierarchy as one replayable plan (no
 * counterpart in the reference, whose caller builds the hierarchy): level l
 * takes its operator A_l and the host array aggs[l] of A_l->M aggregate ids in
 * [0, naggs[l]), and forms P_l = (I - omega D^-1 A_l) P0_l, R_l = P_l^T and
 * A_{l+1} = R_l A_l P_l, all in HBM.  Creation builds every level once and,
 * with use_graph, captures the chain as one HIP graph.  sp3mm_hier_update
 * recomputes every level from A's current values (A must keep its pattern,
 * which is checked against a fingerprint taken at creation; its values may
 * change in place); ms receives its device time.  sp3mm_hier_level gives
 * views (owned = 0, valid until sp3mm_hier_free) of R_l, P_l and A_{l+1}. */
typedef struct sp3mm_hier sp3mm_hier;
int sp3mm_hier_create(const sp3mm_dmat *A, int32_t nlevels, const int64_t *const *aggs, const int64_t *naggs,
                      double omega, int32_t diag_mode, int32_t use_graph, sp3mm_hier **out);
int sp3mm_hier_update(sp3mm_hier *h, int32_t use_graph, double *ms);
int sp3mm_hier_level(sp3mm_hier *h, int32_t level, sp3mm_dmat *R, sp3mm_dmat *P, sp3mm_dmat *Ac);
int sp3mm_hier_info(sp3mm_hier *h, int32_t *nlevels, double *ms_build, int32_t *has_graph);
void sp3mm_hier_free(sp3mm_hier *h);

/* Consumers of a level (no counterpart in the reference, whose products feed
 * the caller's own solver), so that a V-cycle over a device hierarchy runs on
 * the GPU; FP64 operators, device vectors (x and y distinct):
 *   sp3mm_spmv: y = alpha A x + beta y (beta == 0: y is not read);
 *   sp3mm_residual: r = b - A x for square A (r may be b);
 *   sp3mm_jacobi: `sweeps` damped Jacobi sweeps on A x = b,
 *     x <- x + omega D^-1 (b - A x), in place; work holds A->M doubles;
 *   sp3mm_axpby: y = alpha x + beta y on n entries (a zero coefficient skips
 *     the read of its vector: alpha = beta = 0 clears y);
 *   sp3mm_dot: *out = x . y, summed in an order that does not depend on the
 *     launch (the same inputs give the same bits).
 * spmv, residual, jacobi and axpby are queued on the library's stream in
 * call order and return without waiting; sp3mm_dot, sp3mm_copy_to_host and
 * sp3mm_sync wait.  A Jacobi sweep that meets a zero diagonal entry is
 * reported (once) by the next call that waits. */
int sp3mm_spmv(const sp3mm_dmat *A, const double *x, double *y, double alpha, double beta);
int sp3mm_residual(const sp3mm_dmat *A, const double *b, const double *x, double *r);
int sp3mm_jacobi(const sp3mm_dmat *A, const double *b, double *x, double *work, double omega,
                 int32_t sweeps);
int sp3mm_axpby(int64_t n, double alpha, const double *x, double beta, double *y);
int sp3mm_dot(int64_t n, const double *x, const double *y, double *out);

/* V(nu1, nu2) cycle over a device hierarchy (no counterpart in the
 * reference): level l (0 = finest, nlevels >= 1 levels above the coarsest)
 * smooths with nu1 damped Jacobi sweeps, restricts the residual with R[l],
 * recurses, prolongates with P[l], adds the correction and smooths with nu2
 * sweeps; the coarsest operator gets coarse_sweeps sweeps from zero.  The
 * matrices are referenced, not copied: they must outlive the handle.  With
 * use_graph one cycle is captured as a HIP graph and replayed.
 * solve: b and x are device vectors of A[0]->M entries; x holds the initial
 * guess and receives the solution.  Cycles until ||b - A x|| <= rtol ||b||
 * or maxit cycles; *cycles receives the count, hist (maxit + 1 entries, or
 * NULL) the relative residual before every cycle and after the last, *ms the
 * device time of the cycles.  Calls on one handle must not overlap. */
typedef struct sp3mm_vcycle sp3mm_vcycle;
int sp3mm_vcycle_create(int32_t nlevels, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                        const sp3mm_dmat *coarse, int32_t nu1, int32_t nu2, double omega,
                        int32_t coarse_sweeps, int32_t use_graph, sp3mm_vcycle **out);
int sp3mm_vcycle_solve(sp3mm_vcycle *v, const double *b, double *x, double rtol, int32_t maxit,
                       int32_t *cycles, double *hist, double *ms);
void sp3mm_vcycle_free(sp3mm_vcycle *v);

/* Scheduling of the one-pass direct kernels.  These parameters pick table sizes,
 * group shapes and the spread of work over lanes; results are the same up to
 * FP64 summation order. */
typedef struct {
    int32_t u1, u2;      /* left entries each lane holds in stage 1 (2 or 4) / stage 2
                            (2, 3 or 4) */
    int32_t subwave_h;   /* largest stage-1 table run by 32-lane groups (two rows per
                            wavefront); 0, or a power of two in [64, 4096] */
    int32_t slog1_cap;   /* largest lane sub-group (log2) per left entry in stage 1, 0..8
                            (the sub-group never exceeds the row's group) */
    int32_t slog2_cap;   /* the same for stage 2 */
    int32_t h1_shift;    /* stage-1 table = bin upper bound >> h1_shift, -1..3 (-1: twice
                            the bound; a row whose table fills is redone by the
                            long-row kernels) */
    int32_t h2_div;      /* stage-2 table = max(64, stage-1 table / h2_div): 1, 2, 4, 8 */
    int32_t xcd_map;     /* 1: the workgroups of each XCD (workgroup index mod 8) take a
                            contiguous part of a bin's rows, 0: rows interleaved over all
                            workgroups */
    uint32_t hash1, hash2; /* multipliers of the stage-1 / stage-2 LDS hash tables of the
                              direct kernels (and of the pair path's two SpGEMMs; slot =
                              top bits of key * multiplier); odd, 0: the default
                              0x9E3779B1.  The autotuner scores candidates on sample
                              rows of the operands and takes the ones that leave the
                              fewest keys off their home slots */
} sp3mm_tuning;
int sp3mm_get_tuning(sp3mm_tuning *t);
/* NULL restores the defaults. */
int sp3mm_set_tuning(const sp3mm_tuning *t);
/* Autotuner: times the direct product R*A*P (opts->precision as in sp3mm_rap)
 * under candidate schedules, `reps` products each, and installs the fastest
 * for later calls; *best receives it, *ms_best and *ms_default the median
 * device times of the chosen and of the starting schedule. */
int sp3mm_autotune(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                   const sp3mm_opts *opts, int32_t reps, sp3mm_tuning *best,
                   double *ms_best, double *ms_default);
/* The same on a batch's packed product (with the batch's precision); the
 * schedule found is installed and the batch (and its graph) rebuilt under
 * it. */
int sp3mm_batch_autotune(sp3mm_batch *b, int32_t reps, sp3mm_tuning *best,
                         double *ms_best, double *ms_default);

/* Scores of the candidate hash multipliers the autotuner ranks (no
 * counterpart in the reference; exposed for inspection and tests).  For
 * tables of H1 and H2 slots (powers of two in [64, 4096]) and the rows
 * rows[0..nrows) of R (host array): cand[c] receives candidate c's
 * multiplier, score1[c] / score2[c] the summed distances of the rows'
 * distinct stage-1 keys (columns of the rows of A that R's row selects) /
 * stage-2 keys (columns of the rows of P that those select) from their home
 * slots under linear probing, nkeys[2i] and nkeys[2i+1] row i's distinct key
 * counts (INT32_MAX beyond 1024).  A row with more than 1024 distinct keys at
 * a stage, or more than 7/8 of the stage's table, adds nothing to that
 * stage's scores.  Returns the number of candidates (cap must hold them), or
 * -1. */
int32_t sp3mm_hash_scores(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int32_t H1, int32_t H2,
                          const int32_t *rows, int32_t nrows, int32_t cap, uint32_t *cand, uint64_t *score1,
                          uint64_t *score2, int32_t *nkeys);

/* Device memory from the library's pool (e.g. for sp3mm_opts.rowsum), its
 * release, and blocking copies between host and device. */
void *sp3mm_device_alloc(int64_t bytes);
void sp3mm_device_free(void *p);
int sp3mm_copy_to_host(void *dst, const void *src, int64_t bytes);
int sp3mm_copy_to_device(void *dst, const void *src, int64_t bytes);

/* Compute units, LDS bytes per workgroup on the device, LDS bytes the bin
 * tables are allowed to use. */
int sp3mm_device_info(int32_t *num_cu, int64_t *lds_per_block, int64_t *lds_used);

/* Time a device-to-device copy of `bytes` bytes (HBM bandwidth probe); returns ms. */
double sp3mm_copy_probe(int64_t bytes, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif /* SP3MM_AMD_H */
