// kernels.hpp -- host-side declarations of the gfx950 kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sp3mm {

// Device view of one CSR operand.  Left operands are read through rp (int64),
// right operands through rp32 (their nnz must be < 2^31).
struct SpOperand {
    const int64_t *rp;
    const int32_t *rp32;
    const int32_t *col;
    const void *val;
    int64_t M, N;
};

constexpr int kMaxBins = 14;
constexpr int kWaveUbRow = 64;

// Default multiplier of the LDS hash tables (slot = top bits of key * mul):
// 2^32 / golden ratio, rounded to odd.
constexpr unsigned kHashMul = 0x9E3779B1u;

// Row goes to the first bin b with measure <= v[b]; above v[nb-1] -> bin nb.
struct BinThresholds {
    int nb;
    long long v[kMaxBins];
};

// binning scratch: bin_scratch_ints(M) ints of device memory
int64_t bin_scratch_ints(int64_t M);
// Upper bounds of A*B's rows and bin lists.  counts[] accumulates the bin
// populations.  merge_lim > 0: bins of at most merge_lim rows fold into the
// next larger non-empty hash bin while the lists are written, and mcounts[]
// receives the merged populations (the lists are laid out by them).
// ctl != null: the first kernel clears ctl_words words of the control block
// `ctl` (the caller need not), and with ctl_host != null the scatter kernel
// also copies them (populations and totals final by then) into coherent host
// memory at ctl_host and then writes seq to *flag (host-mapped).  totals[0]
// receives the sum of the bounds, totals[1] their maximum and, with
// span_lim > 0 (ctl given), totals[2] the number of rows of A whose columns
// spread over more than span_lim (last column minus first).  pred_count !=
// null: the scatter also writes the (merged) population of bin `pred` there,
// or 0 if reassoc_possible and at least half of the rows are in bins >=
// big_from (or long) and at least half are spread wider than span_lim.
// One-bin mode of the binning (direct plans): k_row_maxlen writes the
// longest row of A and of B per workgroup into part (2 * kMaxLenBlocks ints);
// when their product is at most `bound`, every row is binned with measure `u`
// without measuring it, the per-row bounds are not summed (totals[0] stays
// 0), and word `word` of the control block is 1.
constexpr int kMaxLenBlocks = 1024;
struct OneBin {
    const int *part;  // 2 * kMaxLenBlocks ints
    int nblk;         // kMaxLenBlocks
    long long bound, u;
    int word;
};
void launch_ub_bin(const SpOperand &A, const SpOperand &B, int64_t *ub, int32_t *rownnz,
                   int32_t *lists, int *counts, const BinThresholds &thr,
                   unsigned long long *totals, int *scratch, hipStream_t st, int merge_lim = 0,
                   int *mcounts = nullptr, const void *ctl = nullptr, void *ctl_host = nullptr, int ctl_words = 0,
                   unsigned long long *flag = nullptr, unsigned long long seq = 0, int span_lim = 0,
                   int *pred_count = nullptr, int pred = -1, int big_from = 0, bool reassoc_possible = false,
                   const OneBin *one = nullptr, bool wave_ub = false);
// wave_ub: the bounds are measured a wavefront per row of A first (rows of
// A averaging kWaveUbRow entries or more), then binned
void launch_ub_colparts(const int64_t *arp, const int32_t *acol, int64_t M, const int64_t *boff, int G,
                        int64_t *out, unsigned long long *total, hipStream_t st);
void launch_rebin(const int32_t *m1, const int32_t *m2, int64_t M, int32_t *lists, int *counts,
                  const BinThresholds &thr, int *scratch, hipStream_t st);

size_t symb_lds(int G, int H);
size_t num_lds(int G, int H);
size_t direct_lds(int G, int H1, int H2, bool num);

void launch_symb(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                 const int *count, int rows_host, int32_t *rownnz, int32_t *retry, int *nretry,
                 int *err, int max_blocks, hipStream_t st);
void launch_num(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                const int *count, int rows_host, const int64_t *crp, int32_t *ccol, double *cval,
                int *err, int max_blocks, hipStream_t st);
void launch_direct(int G, bool num, int H1, int H2, int slog1, int slog2, const SpOperand &R,
                   const SpOperand &A, const SpOperand &P, const int32_t *rows, const int *count,
                   int rows_host, int32_t *n1out, int32_t *rownnz, const int64_t *crp, int32_t *ccol,
                   double *cval, int32_t *retry, int *nretry, int *err, int max_blocks,
                   hipStream_t st, bool f32in = false);

// per-workgroup dense workspaces of the long-row kernels (nw* = words per bitmap,
// arrays of values/list entries hold nw*32 slots per workgroup)
struct DenseWs {
    unsigned *bitsA;
    double *valsA;
    int32_t *listc;
    double *listv;
    unsigned *bitsP;
    double *valsP;
    int32_t *cur;  // numeric pass, windowed second stage: one cursor per first-stage entry
    int nwA, nwP;
};

int dense_words(int64_t ncols);
// Long rows of A*B: one 1024-thread workgroup per row.  Bitmap (and, in the
// numeric pass, values) over B's columns in LDS when they fit in max_lds, else
// in the HBM workspace `bits`/`vals`; a numeric pass whose values do not fit
// runs in LDS column windows instead (k_spgemm_window) and needs no workspace.
bool spgemm_dense_in_lds(bool num, int nw, size_t max_lds);
int spgemm_window_cols(size_t max_lds);
void launch_dense_spgemm(bool num, int blocks, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                         const int *count, unsigned *bits, double *vals, int nw, int slog, int32_t *rownnz,
                         const int64_t *crp, int32_t *ccol, double *cval, int *err, hipStream_t st,
                         size_t max_lds = 0);
// long rows whose column windows fit in `lds` bytes of LDS (see k_window_direct);
// the count pass (num = false) lists the rows that do not fit in `spill`
int window_cap(size_t lds);
void launch_window_direct(bool num, int blocks, size_t lds, const SpOperand &R, const SpOperand &A,
                          const SpOperand &P, const int32_t *rows, const int *count, int slog1, int32_t *spill,
                          int *nspill, int32_t *rownnz, const int64_t *crp, int32_t *ccol, double *cval,
                          double *rowsum, int64_t out_cap, int *err, hipStream_t st, bool f32in,
                          int32_t *slotn = nullptr, int32_t *scol = nullptr, double *sval = nullptr,
                          int cap_rows = 0x7fffffff);
void launch_dense_direct(bool num, int blocks, const SpOperand &R, const SpOperand &A, const SpOperand &P,
                         const int32_t *rows, const int *count, const DenseWs &ws, int slog1, int slog2,
                         int32_t *n1out, int32_t *rownnz, const int64_t *crp, int32_t *ccol, double *cval,
                         int *err, hipStream_t st, bool f32in = false, double *rowsum = nullptr,
                         size_t max_lds = 0);

size_t fused_lds(int G, int H1, int H2);
// Scheduling parameters of the one-pass direct kernels.  They decide table
// sizes, group shapes and how work is spread over lanes; results are the same
// up to FP64 summation order.
struct Tuning {
    int u1 = 4, u2 = 4;     // left entries held per lane in stage 1 (2, 4) / stage 2 (2, 3, 4)
    int subwave_h = 256;    // largest stage-1 table run by 32-lane groups (0: none)
    int slog1_cap = 3;      // largest lane sub-group (log2) per left entry, stage 1
    int slog2_cap = 1;      // the same, stage 2
    int h1_shift = 0;       // stage-1 table = bin upper bound >> h1_shift, -1: twice the bound
                            // (a full table -> retry)
    int h2_div = 4;         // stage-2 table = max(64, H1 / h2_div)
    int xcd_map = 1;        // 1: each XCD (block % 8) works through its own contiguous part of
                            // a bin's row list (rows that share A and P rows share one L2)
    unsigned hash1 = kHashMul, hash2 = kHashMul;  // multipliers of the stage-1 / stage-2 table hashes (odd)
};

void launch_direct_fused(int G, int H1, int H2, int slog1, int slog2, const SpOperand &R, const SpOperand &A,
                         const SpOperand &P, const int32_t *rows, const int *count, int rows_host,
                         int32_t *rownnz, int32_t *slotn, int32_t *scol, double *sval, int32_t *retry,
                         int *nretry, int *err, int max_blocks, hipStream_t st, bool f32in,
                         const Tuning &tu);
// Scores of candidate hash multipliers for the direct kernel's tables of H1 and
// H2 slots (both at most kScoreTable) on sample rows of R (see k_hash_score):
// score1[c] / score2[c] (zeroed here) receive the summed home-slot distances,
// nkeys (2 * nrows) every sample row's distinct stage-1 and stage-2 key
// counts.  The launch takes hash_score_lds(H1, H2) bytes of LDS per workgroup.
size_t hash_score_lds(int H1, int H2);
constexpr int kScoreTable = 4096;
void launch_hash_score(const SpOperand &R, const SpOperand &A, const SpOperand &P, const int32_t *rows, int nrows,
                       const unsigned *cand, int ncand, int H1, int H2, unsigned long long *score1,
                       unsigned long long *score2, int *nkeys, hipStream_t st);
// one-pass SpGEMM rows into scratch slots of H entries (see k_spgemm_fused)
size_t spgemm_fused_lds(int G, int H);
void launch_spgemm_fused(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                         const int *count, int rows_host, int32_t *rownnz, int32_t *slotn, int32_t *scol,
                         double *sval, int *err, int max_blocks, hipStream_t st, int xcd_map,
                         unsigned hmul = kHashMul);
// xdst (optional): the kernel's workgroup 0 first copies xwords words from
// xsrc to xdst (a control block into host memory, read after the stream ends)
void launch_gather_rows(const int32_t *rows, const int *count, int rows_host, const int32_t *slotn, int cap,
                        const int32_t *scol, const double *sval, const int64_t *crp, int32_t *ccol, double *cval,
                        int64_t out_cap, int *err, int max_blocks, hipStream_t st, const void *xsrc = nullptr,
                        void *xdst = nullptr, int xwords = 0);
// populations of the bins a plan launched (bins 0..n-1, n <= kMaxBins + 1)
struct BinCounts {
    int n;
    int c[kMaxBins + 1];
};
void launch_check_nnz(const int64_t *nnz, int64_t want, const int *nretry, int want_retry, const int *nspill,
                      int want_spill, int *err, hipStream_t st, const int *counts = nullptr,
                      const BinCounts *want_counts = nullptr);

size_t mixed_lds(int G, int H1, int H2);
void launch_direct_mixed(int G, int H1, int H2, int slog1, int slog2, float guard, const SpOperand &R,
                         const SpOperand &A, const SpOperand &P, const int32_t *rows, const int *count,
                         int rows_host, int32_t *rownnz, int32_t *slotn, int32_t *scol, double *sval,
                         double *rowsum, int32_t *retry, int *nretry, int32_t *fix, int *nfix, int *err,
                         int max_blocks, hipStream_t st, const Tuning &tu);

struct PackItem {
    const int64_t *rp;
    const int32_t *col;
    const void *val;
    int64_t M, nnz, roff, nzoff;
    int32_t coloff;
    int32_t pad;
};
void launch_pack(const PackItem *items, int count, int vtype, int64_t *rp, int32_t *col, void *val, hipStream_t st);
void launch_extract(const int64_t *rp, const int32_t *col, const double *val, int64_t r0, int64_t M, int32_t coloff,
                    int64_t *orp, int32_t *orp32, int32_t *ocol, double *oval, hipStream_t st);

int64_t scan_tiles(int64_t n);
void launch_scan(const int32_t *in, int64_t n, int64_t *out, int32_t *out32, long long *tile_sums,
                 int64_t *total, hipStream_t st);

void launch_rp_convert(const int64_t *in, int64_t M, int64_t nnz, int base, int64_t *rp, int32_t *rp32,
                       int *err, hipStream_t st);
void launch_col_convert(const int64_t *in, const int64_t *rp, int64_t M, int64_t N, int base, int32_t *col,
                        int *err, hipStream_t st);
void launch_col_widen(const int32_t *col, int64_t nnz, int base, int64_t *out, hipStream_t st);
void launch_rp_shift(const int64_t *rp, int64_t n, int base, int64_t *out, hipStream_t st);
void launch_f64_to_f32(const double *in, int64_t n, float *out, hipStream_t st);
void launch_f32_to_f64(const float *in, int64_t n, double *out, hipStream_t st);
void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
// order-sensitive 64-bit fingerprint of a CSR pattern (offsets and columns) into *out
void launch_fingerprint(const int64_t *rp, int64_t M, const int32_t *col, int64_t nnz, unsigned long long *out,
                        hipStream_t st);
// *out = OR of the n device words *words[i] (one wavefront)
void launch_or_words(const int *const *words, int n, int *out, hipStream_t st);
// setup.hip: transpose and smoothed-aggregation prolongator
void launch_col_count(const int32_t *col, int64_t nnz, int32_t *cnt, hipStream_t st);
void launch_transpose_scatter(const int64_t *rp, const int32_t *col, const void *val, bool f32, int64_t M,
                              int64_t nnz, const int64_t *trp, int32_t *cursor, int32_t *tcol, void *tval,
                              int32_t *tsrc, hipStream_t st);
void launch_sort_rows(const int64_t *rp, int64_t M, const int32_t *icol, const void *ival, const int32_t *isrc,
                      bool f32, int32_t *ocol, void *oval, int32_t *osrc, hipStream_t st);
// out[j] = src[perm[j]] for n values (FP32 or FP64)
void launch_gather_vals(const void *src, const int32_t *perm, int64_t n, bool f32, void *out, hipStream_t st);
void launch_sa_count(const int64_t *rp, const int32_t *col, int64_t M, int64_t nnz, int32_t *len, hipStream_t st);
void launch_sa_fill(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, double omega,
                    int mode, const int64_t *srp, int32_t *scol, double *sval, int *err, hipStream_t st);
void launch_p0_fill(const int64_t *agg, int64_t M, int64_t nagg, int64_t *rp, int32_t *rp32, int32_t *col,
                    double *val, int *err, hipStream_t st);
void launch_spmv(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, const double *x,
                 const double *z, double *y, double alpha, double beta, hipStream_t st);
// all sweeps in one workgroup when the system fits in max_lds bytes of LDS
// (returns false, launching nothing, otherwise)
bool launch_jacobi_small(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz,
                         const double *b, double *x, double omega, int sweeps, int *err, size_t max_lds,
                         hipStream_t st);
void launch_axpby(int64_t n, double alpha, const double *x, double beta, double *y, hipStream_t st);
void launch_dot(int64_t n, const double *x, const double *y, double *part, double *out, hipStream_t st);
int dot_parts();
void launch_jacobi(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, const double *b,
                   const double *xi, double *xo, double omega, int *err, hipStream_t st);
void launch_export_words(const void *src, void *host_dst, int words, hipStream_t st);
void launch_copy_probe(const void *in, void *out, int64_t bytes, hipStream_t st);

}  // namespace sp3mm
