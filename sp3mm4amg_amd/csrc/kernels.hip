// kernels.hip -- gfx950 kernels of the Galerkin triple product R*A*P on sorted CSR.
//
// Phases (one launch sequence per product):
//   k_ub_bin      per-row upper bound ub[r] = sum_{k in A_r} nnz(B_k)
//                 (code/src/include/SpMMUtilsGeneric.h:112-139), rows scattered
//                 into per-bin lists by a size measure; empty rows finished here
//   k_symb<G>     exact row lengths with an LDS key hash (replaces the bitmap /
//                 RB-tree symbolic, code/src/Sp3MM_CSR_OMP_SymbStep_Generic.c:84-231)
//   k_scan_*      exclusive scan of row lengths into row offsets (replaces the
//                 serial IRP prefix sum, code/src/Sp3MM_CSR_OMP_Num_Generic.c:40-47)
//   k_num<G>      Gustavson numeric: products accumulated in an LDS hash
//                 (ds_add_f64), compacted and rank-sorted into the output row
//                 (replaces the dense accumulator + qsort,
//                 code/src/include/SpMMUtilsGeneric.h:39-103,230-241)
//   k_direct<G,NUM>  row r of R*A built in one LDS hash and forwarded into P
//                 through a second LDS hash without leaving the CU (replaces
//                 code/src/Sp3MM_CSR_OMP_UB_Generic.c:533-550)
//
// G is the number of threads cooperating on one output row: 64 (a wavefront;
// a 256-thread workgroup then works on 4 rows at once, each wavefront on its
// own LDS tables) or 256 (the whole workgroup on one row).  Table sizes H are
// runtime powers of two chosen by the host from the bin of the row.
#include "device_common.hpp"
#include "kernels.hpp"
#include "expand.hpp"
#include <algorithm>
#include <climits>
#include <stdlib.h>
#include <string.h>

namespace sp3mm {

// ---------------------------------------------------------------------------
// upper bounds + binning
// ---------------------------------------------------------------------------
// Binning in three short kernels with no contended global atomics:
//   k_bin_hist    one thread per row: the row's measure (its upper bound
//                 ub[r] = sum_{k in A_r} nnz(B_k), code/src/include/SpMMUtilsGeneric.h:112-139,
//                 or a given exact length), its bin, a per-workgroup histogram
//                 (LDS atomics) stored bin-major, and per-workgroup ub sum/max;
//   k_bin_offsets one workgroup per bin: exclusive scan of the bin's histogram over
//                 the workgroups -> each workgroup's base inside the bin, bin
//                 populations, ub totals;
//   k_bin_scatter one thread per row: list[bin][base + rank] = row.
// Rows with a zero measure get rownnz = 0 and no bin.
constexpr int kBinThreads = 256;

__device__ __forceinline__ int bin_of(long long u, const BinThresholds &thr)
{
    if (u <= 0) return -1;
    for (int i = 0; i < thr.nb; i++)
        if (u <= thr.v[i]) return i;
    return thr.nb;
}

// kspan receives the spread (last - first) of the row's columns (0 for the
// given measures)
__device__ __forceinline__ long long row_measure(int64_t r, const int64_t *__restrict__ arp,
                                                 const int32_t *__restrict__ acol, const int32_t *__restrict__ brp,
                                                 const int32_t *__restrict__ meas1, const int32_t *__restrict__ meas2,
                                                 int &kspan)
{
    kspan = 0;
    if (meas1) {
        int u = meas1[r];
        if (meas2) u = max(u, meas2[r]);
        return u;
    }
    const int64_t a0 = arp[r], a1 = arp[r + 1];
    if (a1 > a0) kspan = acol[a1 - 1] - acol[a0];
    long long u = 0;
    // batches of kB entries: all their columns, then all their row bounds, are
    // loaded back to back (two memory round trips per batch); entries past
    // the row read its first entry and count nothing
    constexpr int kB = 32;
    for (int64_t e = a0; e < a1; e += kB) {
        int k[kB];
#pragma unroll
        for (int j = 0; j < kB; j++) k[j] = acol[e + j < a1 ? e + j : a0];
        int len[kB];
#pragma unroll
        for (int j = 0; j < kB; j++) len[j] = brp[k[j] + 1] - brp[k[j]];
#pragma unroll
        for (int j = 0; j < kB; j++) u += e + j < a1 ? len[j] : 0;
    }
    return u;
}

// Longest row of A (rows 0..M-1) and of B: per-workgroup maxima into
// part[2b] and part[2b + 1] (kMaxLenBlocks workgroups), reduced by every
// workgroup of k_bin_hist.  Their product bounds every row's ub.
template <typename IA, typename IB>
__global__ void __launch_bounds__(256) k_row_maxlen(const IA *__restrict__ arp, int64_t M,
                                                     const IB *__restrict__ brp, int64_t BM,
                                                     int *__restrict__ part)
{
    __shared__ long long red[2][256 / 64];
    long long ma = 0, mb = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // four rows per thread and step, their offsets loaded together
    for (int64_t i = i0; i < M; i += 4 * stride) {
        int64_t d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t k = i + u * stride;
            d[u] = k < M ? (int64_t)(arp[k + 1] - arp[k]) : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) ma = max(ma, (long long)d[u]);
    }
    for (int64_t i = i0; i < BM; i += 4 * stride) {
        int64_t d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t k = i + u * stride;
            d[u] = k < BM ? (int64_t)(brp[k + 1] - brp[k]) : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) mb = max(mb, (long long)d[u]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        ma = max(ma, (long long)__shfl_xor(ma, o));
        mb = max(mb, (long long)__shfl_xor(mb, o));
    }
    if (lane() == 0) {
        red[0][threadIdx.x / 64] = ma;
        red[1][threadIdx.x / 64] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / 64; w++) {
            ma = max(ma, red[0][w]);
            mb = max(mb, red[1][w]);
        }
        part[2 * blockIdx.x] = (int)min(ma, (long long)INT32_MAX);
        part[2 * blockIdx.x + 1] = (int)min(mb, (long long)INT32_MAX);
    }
}

// Upper bounds ub(r) = sum of nnz(B_k) over the entries k of row r of A, one
// wavefront per row (four entries per lane in flight): for operators whose
// rows run to hundreds of entries, where one thread per row would walk its
// row alone (on the 216-row level of a 96^3 hierarchy, rows of up to 1004
// entries, that took 67 us).
__global__ void __launch_bounds__(256) k_row_ub_wave(const int64_t *__restrict__ arp,
                                                      const int32_t *__restrict__ acol,
                                                      const int32_t *__restrict__ brp, int64_t M,
                                                      int64_t *__restrict__ ub)
{
    const int64_t r = (int64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
    if (r >= M) return;  // whole wavefronts leave together
    const int l = lane();
    const int64_t a0 = arp[r], a1 = arp[r + 1];
    long long s = 0;
    for (int64_t e = a0 + l; e < a1; e += 4 * kWave) {
        int k[4];
#pragma unroll
        for (int u = 0; u < 4; u++) k[u] = acol[e + u * kWave < a1 ? e + u * kWave : a0];
        int len[4];
#pragma unroll
        for (int u = 0; u < 4; u++) len[u] = brp[k[u] + 1] - brp[k[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) s += e + u * kWave < a1 ? len[u] : 0;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) ub[r] = s;
}

__global__ void __launch_bounds__(kBinThreads) k_bin_hist(const int64_t *__restrict__ arp, const int32_t *__restrict__ acol,
                                                           const int32_t *__restrict__ brp, const int32_t *__restrict__ meas1,
                                                           const int32_t *__restrict__ meas2, int64_t M,
                                                           int64_t *__restrict__ ub, int32_t *__restrict__ rownnz,
                                                           int *__restrict__ hist, unsigned long long *__restrict__ part,
                                                           BinThresholds thr, uint32_t *__restrict__ zero, int zero_words,
                                                           int span_lim, OneBin ob, int pre_ub)
{
    __shared__ int h[kMaxBins + 1];
    // one-bin mode: the longest rows bound every row's ub by ob.bound, so
    // every row goes to the bin of ob.u without measuring it
    bool one = false;
    if (ob.part) {
        static_assert(kMaxLenBlocks % kBinThreads == 0, "whole partials per thread");
        __shared__ int pm[2][kBinThreads / 64];
        // every thread loads its kMaxLenBlocks / kBinThreads partials at once
        int a[kMaxLenBlocks / kBinThreads], b[kMaxLenBlocks / kBinThreads];
#pragma unroll
        for (int u = 0; u < kMaxLenBlocks / kBinThreads; u++) {
            const int i = u * kBinThreads + threadIdx.x;
            a[u] = ob.part[2 * i];
            b[u] = ob.part[2 * i + 1];
        }
        int ma = 0, mb = 0;
#pragma unroll
        for (int u = 0; u < kMaxLenBlocks / kBinThreads; u++) {
            ma = max(ma, a[u]);
            mb = max(mb, b[u]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            ma = max(ma, __shfl_xor(ma, o));
            mb = max(mb, __shfl_xor(mb, o));
        }
        if (lane() == 0) {
            pm[0][threadIdx.x / 64] = ma;
            pm[1][threadIdx.x / 64] = mb;
        }
        __syncthreads();
        for (int w = 0; w < kBinThreads / 64; w++) {
            ma = max(ma, pm[0][w]);
            mb = max(mb, pm[1][w]);
        }
        one = (long long)ma * mb <= ob.bound;
    }
    // workgroup 0 clears the caller's control block (zero_words words), which
    // the kernels after this one accumulate into; word ob.word (if any)
    // receives the one-bin decision
    if (zero && blockIdx.x == 0)
        for (int i = threadIdx.x; i < zero_words; i += blockDim.x) zero[i] = i == ob.word && one ? 1u : 0u;
    __shared__ unsigned long long red[3][kBinThreads / 64];
    const int nbin = thr.nb + 1;
    if ((int)threadIdx.x < nbin) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
    long long u = 0;
    int kspan = 0;
    if (one) {
        // every row of the workgroup in the bin of ob.u; no bound stored
        // (the scatter lists the rows in row order without reading them)
        if (threadIdx.x == 0) h[bin_of(ob.u, thr)] = (int)min((int64_t)kBinThreads, M - blockIdx.x * (int64_t)kBinThreads);
    } else if (r < M) {
        if (pre_ub) {  // k_row_ub_wave measured the row
            u = ub[r];
            const int64_t a0 = arp[r], a1 = arp[r + 1];
            kspan = a1 > a0 ? acol[a1 - 1] - acol[a0] : 0;
        } else {
            u = row_measure(r, arp, acol, brp, meas1, meas2, kspan);
            if (ub) ub[r] = u;
        }
        if (u <= 0 && rownnz) rownnz[r] = 0;
        const int b = bin_of(u, thr);
        if (b >= 0) atomicAdd(h + b, 1);
    }
    // rows whose columns spread over more than span_lim (unstructured rows);
    // in one-bin mode the bounds are not known and not summed
    unsigned long long s = (unsigned long long)(u > 0 && !one ? u : 0), mx = s;
    unsigned long long wide = span_lim > 0 && kspan > span_lim ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        wide += __shfl_xor(wide, o);
        const unsigned long long t = __shfl_xor(mx, o);
        mx = t > mx ? t : mx;
    }
    if (lane() == 0) {
        red[0][threadIdx.x / 64] = s;
        red[1][threadIdx.x / 64] = mx;
        red[2][threadIdx.x / 64] = wide;
    }
    __syncthreads();
    if ((int)threadIdx.x < nbin) hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
    if (threadIdx.x == 0) {
        unsigned long long ts = 0, tm = 0, tw = 0;
        for (int i = 0; i < kBinThreads / 64; i++) {
            ts += red[0][i];
            tm = red[1][i] > tm ? red[1][i] : tm;
            tw += red[2][i];
        }
        part[3 * (int64_t)blockIdx.x] = ts;
        part[3 * (int64_t)blockIdx.x + 1] = tm;
        part[3 * (int64_t)blockIdx.x + 2] = tw;
    }
}

// one workgroup of 1024 threads
__global__ void __launch_bounds__(1024) k_bin_offsets(int *__restrict__ hist, int nblk, int nbin,
                                                       const unsigned long long *__restrict__ part,
                                                       int *__restrict__ counts,
                                                       unsigned long long *__restrict__ totals)
{
    __shared__ long long sh[1024 / 64];
    const int b = blockIdx.x;  // workgroup b < nbin scans bin b; the last one sums the bounds
    if (b < nbin) {
        int *hb = hist + (int64_t)b * nblk;
        long long carry = 0;
        for (int i0 = 0; i0 < nblk; i0 += 1024) {
            const int i = i0 + threadIdx.x;
            const long long v = i < nblk ? hb[i] : 0;
            long long total;
            const long long ex = block_exclusive_scan(v, sh, total);
            if (i < nblk) hb[i] = (int)(ex + carry);
            carry += total;
        }
        if (threadIdx.x == 0) counts[b] += (int)carry;
    } else if (totals) {
        unsigned long long s = 0, m = 0, wd = 0;
        for (int i = threadIdx.x; i < nblk; i += 1024) {
            s += part[3 * (int64_t)i];
            m = part[3 * (int64_t)i + 1] > m ? part[3 * (int64_t)i + 1] : m;
            wd += part[3 * (int64_t)i + 2];
        }
        for (int o = 32; o > 0; o >>= 1) {
            s += __shfl_xor(s, o);
            wd += __shfl_xor(wd, o);
            const unsigned long long t = __shfl_xor(m, o);
            m = t > m ? t : m;
        }
        __shared__ unsigned long long rs[16], rm[16], rw[16];
        if (lane() == 0) {
            rs[threadIdx.x / 64] = s;
            rm[threadIdx.x / 64] = m;
            rw[threadIdx.x / 64] = wd;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long ts = 0, tm = 0, tw = 0;
            for (int w = 0; w < 16; w++) {
                ts += rs[w];
                tm = rm[w] > tm ? rm[w] : tm;
                tw += rw[w];
            }
            totals[0] += ts;
            totals[1] = tm > totals[1] ? tm : totals[1];
            totals[2] += tw;
        }
    }
}

// Bins of at most `lim` rows fold into the next larger non-empty hash bin
// (a launch of a handful of rows costs a whole row's latency, while appended
// rows fill groups of the larger launch; chains of small bins move up
// together).  From the bin populations raw[0..nb]: dest[b] is the bin whose
// list takes bin b's rows, starting at position start[b]; cnt[b] the merged
// populations.  The long bin nb never merges.
__device__ __forceinline__ void merge_plan(const int *raw, int nb, int lim, int *dest, int *start, int *cnt)
{
    for (int b = 0; b <= nb; b++) {
        cnt[b] = raw[b];
        dest[b] = b;
        start[b] = 0;
    }
    for (int b = 0; b < nb; b++) {
        if (cnt[b] == 0 || cnt[b] > lim) continue;
        int to = b + 1;
        while (to < nb && cnt[to] == 0) to++;
        if (to >= nb) continue;
        for (int c = 0; c <= b; c++)
            if (dest[c] == b) {
                dest[c] = to;
                start[c] += cnt[to];
            }
        cnt[to] += cnt[b];
        cnt[b] = 0;
    }
}

// merge_lim > 0: rows land in their merged bins (merge_plan over the
// populations `raw` that k_bin_offsets produced), and workgroup 0 writes the
// merged populations to `mcounts`
// ex (optional): workgroup 0 also exports the control block ex.src (bin
// populations final, merged ones included) into coherent host memory and
// then raises the host flag, so the host learns the populations while the
// scatter and the work queued behind it still run.
struct CtlExport {
    const uint32_t *src;
    uint32_t *host;
    int words;
    unsigned long long *flag;
    unsigned long long seq;
    bool clear;    // k_bin_hist clears src (the control block) first
    int span_lim;  // > 0: totals[2] counts the rows whose columns spread wider
    // the count a launch queued before the read-back (bin `pred`) reads:
    // *pred_count = the merged population of bin pred, or 0 when the product
    // is to be reassociated (reassoc_possible, and at least half of the rows
    // in bins >= big_from or long, and at least half of them wide)
    int *pred_count;
    int pred, big_from;
    bool reassoc_possible;
    const unsigned long long *totals;
    int onebin_word;  // >= 0: src[onebin_word] != 0 lists every row in row order
};

__global__ void __launch_bounds__(kBinThreads) k_bin_scatter(const int64_t *__restrict__ ub, const int32_t *__restrict__ meas1,
                                                              const int32_t *__restrict__ meas2, int64_t M,
                                                              const int *__restrict__ hist, int32_t *__restrict__ lists,
                                                              BinThresholds thr, const int *__restrict__ raw,
                                                              int merge_lim, int *__restrict__ mcounts, CtlExport ex)
{
    __shared__ int h[kMaxBins + 1];
    // the merge plan in LDS (indexed by bin: no private-memory arrays)
    __shared__ int mdest[kMaxBins + 1], mstart[kMaxBins + 1], cnt[kMaxBins + 1];
    __shared__ int sone_sh;
    const int nbin = thr.nb + 1;
    if ((int)threadIdx.x < nbin) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        const bool one = ex.src && ex.onebin_word >= 0 && ex.pred >= 0 && ex.src[ex.onebin_word] != 0u;
        sone_sh = one;
    }
    __syncthreads();
    // in one-bin mode only workgroup 0 needs the plan (for the populations
    // it publishes); the others list their rows in row order
    if (threadIdx.x == 0 && (sone_sh == 0 || blockIdx.x == 0)) {
        if (merge_lim > 0) {
            merge_plan(raw, thr.nb, merge_lim, mdest, mstart, cnt);
        } else {
            for (int b = 0; b < nbin; b++) {
                mdest[b] = b;
                mstart[b] = 0;
            }
        }
        if (merge_lim > 0 && blockIdx.x == 0)
            for (int b = 0; b < nbin; b++) mcounts[b] = cnt[b];
        if (ex.pred_count && blockIdx.x == 0) {
            bool reassoc = false;
            if (ex.reassoc_possible) {
                long long big = 0;
                for (int b = ex.big_from; b < nbin; b++) big += merge_lim > 0 ? cnt[b] : raw[b];
                reassoc = 2 * big >= M && 2 * (long long)ex.totals[2] >= M;
            }
            *ex.pred_count = reassoc || ex.pred < 0 ? 0 : (merge_lim > 0 ? cnt[ex.pred] : raw[ex.pred]);
        }
    }
    __syncthreads();
    if (ex.host && blockIdx.x == 0) {
        for (int i = threadIdx.x; i < ex.words; i += blockDim.x)
            __hip_atomic_store(ex.host + i, ex.src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence_system();
            __hip_atomic_store(ex.flag, ex.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    const bool sone = sone_sh != 0;
    const int64_t r = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
    if (r >= M) return;
    if (sone) {  // one-bin mode: bin ex.pred (never merged) lists every row in row order
        lists[(int64_t)ex.pred * M + r] = (int32_t)r;
        return;
    }
    long long u;
    if (meas1) {
        int v = meas1[r];
        if (meas2) v = max(v, meas2[r]);
        u = v;
    } else {
        u = ub[r];
    }
    const int b = bin_of(u, thr);
    if (b < 0) return;
    const int rank = atomicAdd(h + b, 1);
    lists[(int64_t)mdest[b] * M + mstart[b] + hist[(int64_t)b * gridDim.x + blockIdx.x] + rank] = (int32_t)r;
}

// Upper bound split over column ranges of B: boff[k*G+g] .. boff[k*G+g+1] are
// the entries of B's row k in range g (code/src/include/SpMMUtilsGeneric.h:146-174).
__global__ void k_ub_colparts(const int64_t *__restrict__ arp, const int32_t *__restrict__ acol, int64_t M,
                              const int64_t *__restrict__ boff, int G, int64_t *__restrict__ out,
                              unsigned long long *__restrict__ total)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long s = 0;
    if (r < M) {
        for (int g = 0; g < G; g++) out[r * G + g] = 0;
        for (int64_t e = arp[r]; e < arp[r + 1]; e++) {
            const int64_t k = acol[e];
            for (int g = 0; g < G; g++) {
                const int64_t len = boff[k * G + g + 1] - boff[k * G + g];
                out[r * G + g] += len;
                s += (unsigned long long)len;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane() == 0 && s) atomicAdd(total, s);
}

// ---------------------------------------------------------------------------
// product expansion helpers
// ---------------------------------------------------------------------------
// Expand one left row (entries [e0,e1) of lcol/lval) against right operand B:
// every product (b_col, lval*b_val) goes to the LDS table.
//
// Lanes form sub-groups of S = 2^slog lanes; a sub-group owns one left entry at
// a time and its lanes walk that entry's right row with stride S (S is chosen
// near the typical right-row length, so one step covers most rows).  Every lane
// holds U left entries at once (entries g, g + ngrp, ... of the pass), loads all
// their row bounds, then loads one right entry per left entry and inserts the U
// products: a pass over U*ngrp left entries costs three dependent global round
// trips.  The first probe of every insert is one compare-and-swap issued back
// to back with the others; only an insert whose first slot holds another key
// enters the probing loop, so the common path has no divergent loops.
template <int G, bool NUM, typename LV, typename BV = double, int U = 4>
__device__ __forceinline__ void expand_into(const int32_t *lcol, const LV *lval,
                                            int64_t e0, int64_t e1, const int32_t *__restrict__ brp,
                                            const int32_t *__restrict__ bcol,
                                            const BV *__restrict__ bval, int slog, int *keys,
                                            double *vals, unsigned mask, int shift, int &fresh_cnt,
                                            bool &full)
{
    constexpr int kLogG = G >= 256 ? 8 : G >= 128 ? 7 : G >= 64 ? 6 : G >= 32 ? 5 : 4;
    if (slog > kLogG) slog = kLogG;  // a sub-group never spans more than the group
    const int t = Group<G>::tid();
    const int S = 1 << slog;
    const int sub = t & (S - 1);
    const int ngrp = G >> slog;
    for (int64_t eb = e0 + (t >> slog); eb < e1; eb += (int64_t)ngrp * U) {
        int q[U], qe[U];
        double lv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t e = eb + (int64_t)u * ngrp;
            const bool ok = e < e1;
            const int k = ok ? lcol[e] : 0;
            lv[u] = 0.0;
            if constexpr (NUM) lv[u] = ok ? (double)lval[e] : 0.0;
            q[u] = ok ? brp[k] + sub : 0;
            qe[u] = ok ? brp[k + 1] : 0;
        }
        while (true) {
            bool more = false;
#pragma unroll
            for (int u = 0; u < U; u++) more |= q[u] < qe[u];
            if (!more) break;
            int c[U], old[U];
            unsigned sl[U];
            double w[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = q[u] < qe[u];
                c[u] = ok ? bcol[q[u]] : kEmpty;
                w[u] = 0.0;
                if constexpr (NUM) w[u] = ok ? (double)bval[q[u]] : 0.0;
                q[u] += S;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                sl[u] = hslot(c[u], shift) & mask;
                old[u] = c[u] != kEmpty ? atomicCAS(keys + sl[u], kEmpty, c[u]) : kEmpty;
            }
            bool pend = false;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool hit = c[u] != kEmpty && (old[u] == kEmpty || old[u] == c[u]);
                pend |= c[u] != kEmpty && !hit;
                if (hit) {
                    fresh_cnt += old[u] == kEmpty;
                    if constexpr (NUM) atomicAdd(vals + sl[u], lv[u] * w[u]);
                }
            }
            if (pend) {
                for (int u = 0; u < U; u++) {
                    if (c[u] == kEmpty || old[u] == kEmpty || old[u] == c[u]) continue;
                    unsigned s2 = sl[u];
                    bool done = false;
                    for (unsigned probe = 1; probe <= mask; probe++) {
                        s2 = (s2 + 1) & mask;
                        const int o = atomicCAS(keys + s2, kEmpty, c[u]);
                        if (o == kEmpty || o == c[u]) {
                            fresh_cnt += o == kEmpty;
                            if constexpr (NUM) atomicAdd(vals + s2, lv[u] * w[u]);
                            done = true;
                            break;
                        }
                    }
                    if (!done) full = true;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// pairwise SpMM C = A*B: symbolic
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_symb(SpOperand A, SpOperand B, const int32_t *__restrict__ rows,
                                               const int *__restrict__ count, int H, int slog,
                                               int32_t *__restrict__ rownnz,
                                               int32_t *__restrict__ retry, int *__restrict__ nretry,
                                               int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;  // groups per block
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    int *keys = reinterpret_cast<int *>(smem) + g * H;
    long long *red = reinterpret_cast<long long *>(smem + (size_t)GPB * H * 4) + g * Group<G>::kWaves;
    const unsigned mask = (unsigned)H - 1;
    const int shift = 32 - ilog2((unsigned)H);
    const int n = *count;
    for (int i = blockIdx.x * GPB + g; i < n; i += gridDim.x * GPB) {
        const int r = rows[i];
        for (int s = t; s < H; s += G) keys[s] = kEmpty;
        Group<G>::sync();
        int cnt = 0;
        bool full = false;
        expand_into<G, false, double, double>(A.col, (const double *)nullptr, A.rp[r], A.rp[r + 1], B.rp32,
                                              B.col, (const double *)nullptr, slog, keys, nullptr, mask, shift,
                                              cnt, full);
        const long long tot = group_sum<G>(cnt, red);
        const long long anyfull = group_max<G>(full ? 1 : 0, red);
        if (t == 0) {
            if (anyfull) {
                rownnz[r] = -1;
                if (retry) retry[atomicAdd(nretry, 1)] = r;
                else atomicOr(err, kErrHashFull);
            } else {
                rownnz[r] = (int32_t)tot;
            }
        }
        Group<G>::sync();
    }
}

// ---------------------------------------------------------------------------
// pairwise SpMM C = A*B: numeric
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(256) k_num(SpOperand A, SpOperand B, const int32_t *__restrict__ rows,
                                              const int *__restrict__ count, int H, int slog,
                                              const int64_t *__restrict__ crp,
                                              int32_t *__restrict__ ccol, double *__restrict__ cval,
                                              int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    double *vals = reinterpret_cast<double *>(smem) + g * H;
    int *keys = reinterpret_cast<int *>(smem + (size_t)GPB * H * 8) + g * H;
    int *ired = reinterpret_cast<int *>(smem + (size_t)GPB * H * 12) + g * (Group<G>::kWaves + 1);
    const unsigned mask = (unsigned)H - 1;
    const int shift = 32 - ilog2((unsigned)H);
    const int n = *count;
    const double *aval = static_cast<const double *>(A.val);
    const double *bval = static_cast<const double *>(B.val);
    for (int i = blockIdx.x * GPB + g; i < n; i += gridDim.x * GPB) {
        const int r = rows[i];
        for (int s = t; s < H; s += G) {
            keys[s] = kEmpty;
            vals[s] = 0.0;
        }
        Group<G>::sync();
        int cnt = 0;
        bool full = false;
        expand_into<G, true, double>(A.col, aval, A.rp[r], A.rp[r + 1], B.rp32, B.col, bval, slog,
                                     keys, vals, mask, shift, cnt, full);
        if (full) atomicOr(err, kErrHashFull);
        Group<G>::sync();
        const int nn = compact_table<G, double>(keys, vals, H, ired);
        const int64_t off = crp[r];
        const int64_t len = crp[r + 1] - off;
        if (nn != len) {
            if (t == 0) atomicOr(err, kErrOverflow);
        } else {
            for (int j = t; j < nn; j += G) {
                const int k = keys[j];
                const int rk = key_rank(keys, nn, k);
                ccol[off + rk] = k;
                cval[off + rk] = vals[j];
            }
        }
        Group<G>::sync();
    }
}

// ---------------------------------------------------------------------------
// direct triple product: row r of R*A in table 1, forwarded into P in table 2
// ---------------------------------------------------------------------------
template <int G, bool NUM, typename VT = double>
__global__ void __launch_bounds__(256) k_direct(SpOperand R, SpOperand A, SpOperand P,
                                                 const int32_t *__restrict__ rows,
                                                 const int *__restrict__ count, int H1, int H2,
                                                 int slog1, int slog2,
                                                 int32_t *__restrict__ n1out,
                                                 int32_t *__restrict__ rownnz,
                                                 const int64_t *__restrict__ crp,
                                                 int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                 int32_t *__restrict__ retry, int *__restrict__ nretry,
                                                 int *__restrict__ err, int cap_rows)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    unsigned char *p = smem;
    double *v1 = nullptr, *v2 = nullptr;
    if constexpr (NUM) {
        v1 = reinterpret_cast<double *>(p) + g * H1;
        p += (size_t)GPB * H1 * 8;
        v2 = reinterpret_cast<double *>(p) + g * H2;
        p += (size_t)GPB * H2 * 8;
    }
    int *k1 = reinterpret_cast<int *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    int *k2 = reinterpret_cast<int *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    long long *red = reinterpret_cast<long long *>(p) + g * Group<G>::kWaves;
    p += (size_t)GPB * Group<G>::kWaves * 8;
    int *ired = reinterpret_cast<int *>(p) + g * (Group<G>::kWaves + 1);

    const unsigned m1 = (unsigned)H1 - 1, m2 = (unsigned)H2 - 1;
    const int sh1 = 32 - ilog2((unsigned)H1), sh2 = 32 - ilog2((unsigned)H2);
    // the row list was sized for cap_rows rows (a larger count is reported)
    const int n = clamp_count(count, cap_rows, err);
    const VT *rval = static_cast<const VT *>(R.val);
    const VT *aval = static_cast<const VT *>(A.val);
    const VT *pval = static_cast<const VT *>(P.val);
    for (int i = blockIdx.x * GPB + g; i < n; i += gridDim.x * GPB) {
        const int r = rows[i];
        for (int s = t; s < H1; s += G) {
            k1[s] = kEmpty;
            if constexpr (NUM) v1[s] = 0.0;
        }
        for (int s = t; s < H2; s += G) {
            k2[s] = kEmpty;
            if constexpr (NUM) v2[s] = 0.0;
        }
        Group<G>::sync();
        int c1 = 0;
        bool full = false;
        expand_into<G, NUM, VT, VT>(R.col, rval, R.rp[r], R.rp[r + 1], A.rp32, A.col, aval, slog1,
                                    k1, v1, m1, sh1, c1, full);
        Group<G>::sync();
        const int nn1 = compact_table<G, double>(k1, v1, H1, ired);
        int c2 = 0;
        expand_into<G, NUM, double, VT>(k1, v1, 0, nn1, P.rp32, P.col, pval, slog2, k2, v2, m2, sh2, c2,
                                    full);
        if constexpr (!NUM) {
            const long long tot = group_sum<G>(c2, red);
            const long long anyfull = group_max<G>(full ? 1 : 0, red);
            if (t == 0) {
                if (n1out) n1out[r] = nn1;
                if (anyfull) {
                    rownnz[r] = -1;
                    if (retry) retry[atomicAdd(nretry, 1)] = r;
                    else atomicOr(err, kErrHashFull);
                } else {
                    rownnz[r] = (int32_t)tot;
                }
            }
        } else {
            if (full) atomicOr(err, kErrHashFull);
            Group<G>::sync();
            const int nn2 = compact_table<G, double>(k2, v2, H2, ired);
            const int64_t off = crp[r];
            const int64_t len = crp[r + 1] - off;
            if (nn2 != len) {
                if (t == 0) atomicOr(err, kErrOverflow);
            } else {
                for (int j = t; j < nn2; j += G) {
                    const int k = k2[j];
                    const int rk = key_rank(k2, nn2, k);
                    ccol[off + rk] = k;
                    cval[off + rk] = v2[j];
                }
            }
        }
        Group<G>::sync();
    }
}

// ---------------------------------------------------------------------------
// direct triple product in ONE pass over the operands
// ---------------------------------------------------------------------------
// Row r of R*A is accumulated in table 1 (H1 >= the row's upper bound, so it
// cannot fill), compacted, and forwarded into P through table 2 of H2 slots.
// The finished row is rank-sorted and written to its slot of `cap` = H2
// entries in a scratch area (slot i of the bin list -> scratch + i*cap); its
// length goes to rownnz[r] and slotn[i].  A row whose table 2 fills is marked
// slotn[i] = -1 and queued on `retry` for the long-row kernels.  After the
// scan of rownnz, k_gather_rows moves every slot to its final place.  Table 1
// hashes with the multiplier hm1, table 2 with hm2.
template <int G, int U1, int U2, typename VT = double>
__global__ void __launch_bounds__(256, 6) k_direct_fused(SpOperand R, SpOperand A, SpOperand P,
                                                       const int32_t *__restrict__ rows,
                                                       const int *__restrict__ count, int H1, int H2,
                                                       int slog1, int slog2, int32_t *__restrict__ rownnz,
                                                       int32_t *__restrict__ slotn,
                                                       int32_t *__restrict__ scol, double *__restrict__ sval,
                                                       int32_t *__restrict__ retry, int *__restrict__ nretry,
                                                       int *__restrict__ err, int nxcd, int cap_rows, unsigned hm1,
                                                       unsigned hm2)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    unsigned char *p = smem;
    // per group: values [table 1 | table 2 | G dummy slots], keys the same;
    // the dummy slots (one per lane, see expand_flat) serve both tables
    const int W = H1 + H2 + G;
    double *v1 = reinterpret_cast<double *>(p) + g * W;
    double *v2 = v1 + H1;
    p += (size_t)GPB * W * 8;
    int *k1 = reinterpret_cast<int *>(p) + g * W;
    int *k2 = k1 + H1;
    p += (size_t)GPB * W * 4;
    long long *red = reinterpret_cast<long long *>(p) + g * Group<G>::kWaves;
    p += (size_t)GPB * Group<G>::kWaves * 8;
    int *ired = reinterpret_cast<int *>(p) + g * (Group<G>::kWaves + 4);
    int *fullflag = ired + Group<G>::kWaves + 1;  // a table of this row filled
    // keys claimed in table 1 and table 2 (multi-wavefront groups, whose
    // rows are given up past 7/8 of a table)
    int *claimed1 = ired + Group<G>::kWaves + 2, *claimed2 = ired + Group<G>::kWaves + 3;
    const int lim1 = H1 - H1 / 8, lim2 = H2 - H2 / 8;

    const unsigned m1 = (unsigned)H1 - 1, m2 = (unsigned)H2 - 1;
    const int sh1 = 32 - ilog2((unsigned)H1), sh2 = 32 - ilog2((unsigned)H2);
    // the scratch slots hold cap_rows rows: a larger count (a replay whose
    // bins differ from the plan's) is reported and never written past
    const int n = clamp_count(count, cap_rows, err);
    const VT *rval = static_cast<const VT *>(R.val);
    const VT *aval = static_cast<const VT *>(A.val);
    const VT *pval = static_cast<const VT *>(P.val);
    const RowWalk w = row_walk<GPB>(n, g, nxcd);
    const int stride = w.stride, end = w.end;
    int i = w.i;
    // The row index and R's row bounds of the group's next row are loaded while
    // the current row is processed (two dependent loads off the critical path).
    int r = i < end ? rows[i] : 0;
    int64_t rb = 0, re = 0;
    if (i < end) {
        rb = R.rp[r];
        re = R.rp[r + 1];
    }
    for (; i < end; i += stride) {
        const int rn = i + stride < end ? rows[i + stride] : 0;
        for (int s = t; s < W; s += G) {
            k1[s] = kEmpty;
            v1[s] = 0.0;
        }
        if (t == 0) {
            *fullflag = 0;
            *claimed1 = 0;
            *claimed2 = 0;
        }
        Group<G>::sync();
        bool full1 = false, full2 = false;
        // multi-wavefront groups (large tables): a row that claims more than
        // 7/8 of a table is given up, and the retry path redoes it
        if constexpr (G > kWave)
            expand_flat_crowded<G, VT, VT, U1>(R.col, rval, rb, re, A.rp32, A.col, aval, slog1, k1, v1, m1, sh1,
                                               (unsigned)(H1 + H2 + t), full1, fullflag, claimed1, lim1, hm1);
        else
            expand_flat<G, VT, VT, U1>(R.col, rval, rb, re, A.rp32, A.col, aval, slog1, k1, v1, m1, sh1,
                                       (unsigned)(H1 + H2 + t), full1, fullflag, hm1);
        const int64_t rbn = R.rp[rn], ren = R.rp[rn + 1];
        Group<G>::sync();
        if constexpr (G > kWave) full1 = full1 || *claimed1 > lim1;
        const int nn1 = compact_table<G, double>(k1, v1, H1, ired);
        if (!full1 && !__hip_atomic_load(fullflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
            if constexpr (G > kWave)
                expand_flat_crowded<G, double, VT, U2>(k1, v1, 0, nn1, P.rp32, P.col, pval, slog2, k2, v2, m2, sh2,
                                                       (unsigned)(H2 + t), full2, fullflag, claimed2, lim2, hm2);
            else
                expand_flat<G, double, VT, U2>(k1, v1, 0, nn1, P.rp32, P.col, pval, slog2, k2, v2, m2, sh2,
                                               (unsigned)(H2 + t), full2, fullflag, hm2);
        }
        if constexpr (G > kWave) {
            Group<G>::sync();  // every wavefront's claims are counted
            full2 = full2 || *claimed2 > lim2;
        }
        const long long anyfull = group_max<G>((full1 || full2) ? 1 : 0, red);
        if (anyfull) {
            if (t == 0) {
                slotn[i] = -1;
                rownnz[r] = 0;
                retry[atomicAdd(nretry, 1)] = r;
            }
        } else {
            Group<G>::sync();
            const int nn2 = compact_table<G, double>(k2, v2, H2, ired);
            const int64_t off = (int64_t)i * H2;
            // streaming stores: the slots are read once, by the gather, and
            // should not evict the A and P rows other groups are gathering
            for (int j = t; j < nn2; j += G) {
                const int k = k2[j];
                const int rk = key_rank(k2, nn2, k);
                __builtin_nontemporal_store(k, scol + off + rk);
                __builtin_nontemporal_store(v2[j], sval + off + rk);
            }
            if (t == 0) {
                slotn[i] = nn2;
                rownnz[r] = nn2;
            }
        }
        Group<G>::sync();
        r = rn;
        rb = rbn;
        re = ren;
    }
}

// ---------------------------------------------------------------------------
// hash multipliers for the direct kernel's two tables
// ---------------------------------------------------------------------------
// The stage-1 keys of row r of R*A*P are the columns of the rows of A that
// row r of R selects, the stage-2 keys the columns of the rows of P that
// those keys select.  In a mesh-ordered operator the key sets of most rows
// are translates of one stencil, and whether a multiplier sends such a set to
// distinct slots depends on the multiplier and the strides, not on the row:
// some multipliers leave no key off its home slot, where others make every
// pass of the direct kernel probe.  A workgroup takes one sample row: it
// collects the row's distinct stage-1 and stage-2 keys (dedup tables with the
// default multiplier, compacted), then each wavefront takes candidates
// c = w, w + 4, ...: it inserts both key sets by linear probing into tables of
// H1 and H2 slots, as the direct kernel does, and adds the keys' distances
// from their home slots to score1[c] and score2[c].  For distinct keys under
// linear probing that sum does not depend on the insert order, so the scores
// are deterministic.  A row with more distinct keys at a stage than its
// dedup table takes (kScoreDedup / 2), or than 7/8 of H1 (H2), is skipped for
// that stage by every candidate.  nkeys[2i], nkeys[2i+1] receive sample row
// i's distinct stage-1 and stage-2 key counts.
constexpr int kScoreDedup = 2048;
__global__ void __launch_bounds__(256) k_hash_score(SpOperand R, SpOperand A, SpOperand P,
                                                   const int32_t *__restrict__ rows, const unsigned *__restrict__ cand,
                                                   int ncand, int H1, int H2,
                                                   unsigned long long *__restrict__ score1,
                                                   unsigned long long *__restrict__ score2,
                                                   int *__restrict__ nkeys)
{
    __shared__ int d1[kScoreDedup], d2[kScoreDedup];
    __shared__ int red[Group<256>::kWaves + 1];
    __shared__ int over;
    extern __shared__ int tabs[];  // per wavefront: H1 + H2 slots
    const int t = threadIdx.x, w = t / kWave, l = t % kWave;
    const int r = rows[blockIdx.x];
    constexpr unsigned dmask = kScoreDedup - 1;
    constexpr int dshift = 32 - 11;  // log2(kScoreDedup) == 11
    static_assert(kScoreDedup == 2048, "dshift");
    for (int i = t; i < kScoreDedup; i += 256) {
        d1[i] = kEmpty;
        d2[i] = kEmpty;
    }
    if (t == 0) over = 0;
    __syncthreads();
    // stage-1 keys: every lane takes entries of R's row and walks their rows of A
    const int64_t rb = R.rp[r], re = R.rp[r + 1];
    for (int64_t e = rb + t; e < re; e += 256) {
        const int k = R.col[e];
        for (int q = A.rp32[k]; q < A.rp32[k + 1]; q++) {
            int fresh = 0;
            if (hash_insert(d1, dmask, dshift, A.col[q], &fresh) < 0) over = 1;
        }
    }
    __syncthreads();
    const int n1 = over ? kScoreDedup : compact_table<256, double>(d1, nullptr, kScoreDedup, red);
    __syncthreads();
    // stage-2 keys: the columns of the rows of P that the stage-1 keys select
    if (n1 <= kScoreDedup / 2) {
        for (int i = t; i < n1; i += 256) {
            const int k = d1[i];
            for (int q = P.rp32[k]; q < P.rp32[k + 1]; q++) {
                int fresh = 0;
                if (hash_insert(d2, dmask, dshift, P.col[q], &fresh) < 0) over = 1;
            }
        }
    }
    __syncthreads();
    const int n2 = (over || n1 > kScoreDedup / 2) ? kScoreDedup : compact_table<256, double>(d2, nullptr, kScoreDedup, red);
    // the row's distinct key counts (INT32_MAX: beyond the dedup tables)
    if (t == 0) {
        nkeys[2 * blockIdx.x] = n1 <= kScoreDedup / 2 ? n1 : INT32_MAX;
        nkeys[2 * blockIdx.x + 1] = n2 <= kScoreDedup / 2 ? n2 : INT32_MAX;
    }
    const bool use1 = n1 <= kScoreDedup / 2 && n1 <= H1 - H1 / 8;
    const bool use2 = n2 <= kScoreDedup / 2 && n2 <= H2 - H2 / 8;
    int *t1 = tabs + w * (H1 + H2), *t2 = t1 + H1;
    const unsigned m1 = (unsigned)H1 - 1, m2 = (unsigned)H2 - 1;
    const int sh1 = 32 - ilog2((unsigned)H1), sh2 = 32 - ilog2((unsigned)H2);
    for (int c = w; c < ncand; c += Group<256>::kWaves) {
        const unsigned mul = cand[c];
        for (int i = l; i < H1 + H2; i += kWave) t1[i] = kEmpty;
        wave_sync();
        unsigned long long dist1 = 0, dist2 = 0;
        // keys are distinct: a swap either claims an empty slot or moves on
        if (use1)
            for (int i = l; i < n1; i += kWave) {
                const int k = d1[i];
                const unsigned home = hslot(k, sh1, mul) & m1;
                unsigned s = home;
                for (unsigned probe = 0; probe <= m1; probe++) {
                    if (atomicCAS(t1 + s, kEmpty, k) == kEmpty) break;
                    s = (s + 1) & m1;
                }
                dist1 += (s - home) & m1;
            }
        if (use2)
            for (int i = l; i < n2; i += kWave) {
                const int k = d2[i];
                const unsigned home = hslot(k, sh2, mul) & m2;
                unsigned s = home;
                for (unsigned probe = 0; probe <= m2; probe++) {
                    if (atomicCAS(t2 + s, kEmpty, k) == kEmpty) break;
                    s = (s + 1) & m2;
                }
                dist2 += (s - home) & m2;
            }
        dist1 = (unsigned long long)group_sum<kWave>((long long)dist1, nullptr);
        dist2 = (unsigned long long)group_sum<kWave>((long long)dist2, nullptr);
        if (l == 0) {
            if (dist1) atomicAdd(score1 + c, dist1);
            if (dist2) atomicAdd(score2 + c, dist2);
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// SpGEMM C = A*B in one pass
// ---------------------------------------------------------------------------
// Row r of A*B is accumulated in an LDS table of H >= its upper bound slots
// (branch-free expansion, so it cannot fill; home slots from the multiplier
// hmul), compacted, rank-sorted and
// written to its scratch slot of H entries (slot i of the bin list -> scratch
// + i*H); its length goes to rownnz[r] and slotn[i].  After the scan of
// rownnz, k_gather_rows moves every slot to its final place.
template <int G, int U>
__global__ void __launch_bounds__(256) k_spgemm_fused(SpOperand A, SpOperand B, const int32_t *__restrict__ rows,
                                                       const int *__restrict__ count, int H, int slog,
                                                       int32_t *__restrict__ rownnz, int32_t *__restrict__ slotn,
                                                       int32_t *__restrict__ scol, double *__restrict__ sval,
                                                       int *__restrict__ err, int nxcd, unsigned hmul)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    const int W = H + G;  // H slots + one dummy slot per lane
    unsigned char *p = smem;
    double *v = reinterpret_cast<double *>(p) + g * W;
    p += (size_t)GPB * W * 8;
    int *k = reinterpret_cast<int *>(p) + g * W;
    p += (size_t)GPB * W * 4;
    int *ired = reinterpret_cast<int *>(p) + g * (Group<G>::kWaves + 2);
    int *fullflag = ired + Group<G>::kWaves + 1;
    const unsigned m = (unsigned)H - 1;
    const int sh = 32 - ilog2((unsigned)H);
    const int n = *count;
    const double *aval = static_cast<const double *>(A.val);
    const double *bval = static_cast<const double *>(B.val);
    const RowWalk w = row_walk<GPB>(n, g, nxcd);
    const int stride = w.stride, end = w.end;
    int i = w.i;
    int r = i < end ? rows[i] : 0;
    int64_t rb = 0, re = 0;
    if (i < end) {
        rb = A.rp[r];
        re = A.rp[r + 1];
    }
    for (; i < end; i += stride) {
        const int rn = i + stride < end ? rows[i + stride] : 0;
        for (int s = t; s < W; s += G) {
            k[s] = kEmpty;
            v[s] = 0.0;
        }
        if (t == 0) *fullflag = 0;
        Group<G>::sync();
        bool full = false;
        expand_flat<G, double, double, U>(A.col, aval, rb, re, B.rp32, B.col, bval, slog, k, v, m, sh,
                                          (unsigned)(H + t), full, fullflag, hmul);
        const int64_t rbn = A.rp[rn], ren = A.rp[rn + 1];
        if (full) atomicOr(err, kErrHashFull);  // H >= the row's upper bound: cannot happen
        Group<G>::sync();
        const int nn = compact_table<G, double>(k, v, H, ired);
        const int64_t off = (int64_t)i * H;
        if (nn > (G > kWave ? 512 : 256)) {
            // long rows: sort in LDS (padded to a power of two, at most H)
            // and write in order.  Ranking costs n^2/G comparisons per lane;
            // the bitonic network (log2 N)^2/2 stages, each a dependent LDS
            // round trip plus a barrier, which pays from about 256 keys
            // (one-wavefront groups) or 512 (multi-wavefront groups) on
            int N = 1;
            while (N < nn) N <<= 1;
            for (int j = nn + t; j < N; j += G) k[j] = INT32_MAX;
            Group<G>::sync();
            bitonic_sort<G, double>(k, v, N, t);
            for (int j = t; j < nn; j += G) {
                scol[off + j] = k[j];
                sval[off + j] = v[j];
            }
        } else {
            for (int j = t; j < nn; j += G) {
                const int key = k[j];
                const int rk = key_rank(k, nn, key);
                scol[off + rk] = key;
                sval[off + rk] = v[j];
            }
        }
        if (t == 0) {
            slotn[i] = nn;
            rownnz[r] = nn;
        }
        Group<G>::sync();
        r = rn;
        rb = rbn;
        re = ren;
    }
}

size_t spgemm_fused_lds(int G, int H)
{
    const size_t gpb = 256 / G;
    return gpb * ((size_t)H + G) * 12 + gpb * (group_waves(G) + 2) * 4;
}

static unsigned grid_for(int rows, int G, int max_blocks);

void launch_spgemm_fused(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                         const int *count, int rows_host, int32_t *rownnz, int32_t *slotn, int32_t *scol,
                         double *sval, int *err, int max_blocks, hipStream_t st, int xcd_map, unsigned hmul)
{
    const unsigned grid = grid_for(rows_host, G, max_blocks);
    const size_t lds = spgemm_fused_lds(G, H);
    slog = std::min(slog, ilog2((unsigned)G));
    const int nxcd = xcd_map && grid >= (unsigned)kXcds ? kXcds : 1;
    hmul |= 1u;
#define SP3MM_SPG_ARGS A, B, rows, count, H, slog, rownnz, slotn, scol, sval, err, nxcd, hmul
    if (G == 32) hipLaunchKernelGGL((k_spgemm_fused<32, 4>), dim3(grid), dim3(256), lds, st, SP3MM_SPG_ARGS);
    else if (G == 64) hipLaunchKernelGGL((k_spgemm_fused<64, 4>), dim3(grid), dim3(256), lds, st, SP3MM_SPG_ARGS);
    else hipLaunchKernelGGL((k_spgemm_fused<256, 4>), dim3(grid), dim3(256), lds, st, SP3MM_SPG_ARGS);
#undef SP3MM_SPG_ARGS
}

// Move the rows of one bin from their scratch slots to the output, L lanes per
// row (several rows per wavefront, so their metadata loads overlap); slots
// marked -1 belong to the long-row path.
template <int L>
__global__ void __launch_bounds__(256) k_gather_rows(const int32_t *__restrict__ rows, const int *__restrict__ count,
                                                      const int32_t *__restrict__ slotn, int cap,
                                                      const int32_t *__restrict__ scol,
                                                      const double *__restrict__ sval,
                                                      const int64_t *__restrict__ crp, int32_t *__restrict__ ccol,
                                                      double *__restrict__ cval, int64_t out_cap,
                                                      int *__restrict__ err, int cap_rows,
                                                      const uint32_t *__restrict__ xsrc, uint32_t *xdst, int xwords)
{
    // xdst (optional): workgroup 0 first copies xwords words from xsrc (a
    // control block the kernels before this one completed) to host memory
    if (xdst && blockIdx.x == 0)
        for (int i = threadIdx.x; i < xwords; i += blockDim.x) xdst[i] = xsrc[i];
    const int gpb = blockDim.x / L;
    const int hl = threadIdx.x & (L - 1);
    const int n = clamp_count(count, cap_rows, err);
    for (int i = blockIdx.x * gpb + threadIdx.x / L; i < n; i += gridDim.x * gpb) {
        const int len = slotn[i];
        if (len <= 0) continue;
        const int64_t src = (int64_t)i * cap, dst = crp[rows[i]];
        if (dst + len > out_cap) {  // output smaller than this replay needs
            if (hl == 0) atomicOr(err, kErrOverflow);
            continue;
        }
        for (int j = hl; j < len; j += L) {
            ccol[dst + j] = scol[src + j];
            cval[dst + j] = sval[src + j];
        }
    }
}

// ---------------------------------------------------------------------------
// exclusive scan of int32 row lengths into int64 offsets (3 phases)
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int32_t *__restrict__ in, int64_t n,
                                                              int64_t *__restrict__ out,
                                                              long long *__restrict__ tile_sums)
{
    __shared__ long long sh[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long s = 0;
    for (int j = 0; j < kScanItems; j++) {
        const int64_t idx = base + j;
        v[j] = idx < n ? (long long)in[idx] : 0;
        s += v[j];
    }
    long long total;
    long long ex = block_exclusive_scan(s, sh, total);
    for (int j = 0; j < kScanItems; j++) {
        const int64_t idx = base + j;
        if (idx < n) out[idx] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// one workgroup scans the tile sums (up to kScanTile tiles per pass, looped)
__global__ void __launch_bounds__(kScanThreads) k_scan_top(long long *__restrict__ tile_sums, int64_t ntiles,
                                                            int64_t *__restrict__ total_out)
{
    __shared__ long long sh[kScanThreads / 64];
    long long carry = 0;
    for (int64_t t0 = 0; t0 < ntiles; t0 += kScanTile) {
        const int64_t base = t0 + (int64_t)threadIdx.x * kScanItems;
        long long v[kScanItems];
        long long s = 0;
        for (int j = 0; j < kScanItems; j++) {
            const int64_t idx = base + j;
            v[j] = idx < ntiles ? tile_sums[idx] : 0;
            s += v[j];
        }
        long long total;
        long long ex = block_exclusive_scan(s, sh, total) + carry;
        for (int j = 0; j < kScanItems; j++) {
            const int64_t idx = base + j;
            if (idx < ntiles) tile_sums[idx] = ex;
            ex += v[j];
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(int64_t *__restrict__ out, int64_t n, const long long *__restrict__ tile_off,
                           const int64_t *__restrict__ total, int32_t *__restrict__ out32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = *total;
    if (i < n) {
        const long long v = out[i] + tile_off[i / kScanTile];
        out[i] = v;
        if (out32 && tot < 2147483647LL) out32[i] = (int32_t)v;
    }
    if (i == n) {
        out[n] = tot;
        if (out32 && tot < 2147483647LL) out32[n] = (int32_t)tot;
    }
}

// The second phase of a short scan (at most kScanShortTiles tiles) without
// the tile-sum scan kernel: every workgroup sums the raw tile sums before its
// rows' tile itself (and all of them for the total), then adds.
constexpr int kScanShortTiles = 2048;
__global__ void __launch_bounds__(256) k_scan_add_short(int64_t *__restrict__ out, int64_t n,
                                                         const long long *__restrict__ tile_sums, int ntiles,
                                                         int64_t *__restrict__ total_out, int32_t *__restrict__ out32)
{
    static_assert(kScanTile % 256 == 0, "a workgroup's rows lie in one tile");
    __shared__ long long sp[256 / 64], st[256 / 64];
    const int64_t i0 = (int64_t)blockIdx.x * 256;
    const int tile = (int)(i0 / kScanTile);
    long long pre = 0, tot = 0;
    for (int j = threadIdx.x; j < ntiles; j += 256) {
        const long long v = tile_sums[j];
        tot += v;
        pre += j < tile ? v : 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        pre += __shfl_xor(pre, o);
        tot += __shfl_xor(tot, o);
    }
    if (lane() == 0) {
        sp[threadIdx.x / 64] = pre;
        st[threadIdx.x / 64] = tot;
    }
    __syncthreads();
    pre = sp[0] + sp[1] + sp[2] + sp[3];
    tot = st[0] + st[1] + st[2] + st[3];
    const int64_t i = i0 + threadIdx.x;
    if (i < n) {
        const long long v = out[i] + pre;
        out[i] = v;
        if (out32 && tot < 2147483647LL) out32[i] = (int32_t)v;
    }
    if (i == n) {
        out[n] = tot;
        if (out32 && tot < 2147483647LL) out32[n] = (int32_t)tot;
        *total_out = tot;
    }
}

// ---------------------------------------------------------------------------
// long rows: one workgroup per row on a dense accumulator in HBM
// ---------------------------------------------------------------------------
// Each persistent workgroup owns one workspace: a presence bitmap of 32-bit
// words and (numeric) a dense FP64 value array over the output columns.
// Products set bits with global atomicOr and add values with global FP64
// atomics; sorted emission walks the bitmap word by word (1024 words per step,
// one per thread, a workgroup scan giving output positions) and clears every
// bit and value it emits, so the workspace is clean for the next row.  1024
// threads (16 wavefronts) per workgroup: a workgroup whose LDS workspace
// fills a CU still has the waves to hide the latency of its row gathers.
constexpr int kDense = 1024;

__device__ __forceinline__ unsigned ld_word(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double ld_val(const double *p)
{
    const unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}

template <bool NUM, typename LV, typename BV = double>
__device__ __forceinline__ void dense_expand(const int32_t *lcol, const LV *lval, int64_t e0, int64_t e1,
                                             const int32_t *__restrict__ brp,
                                             const int32_t *__restrict__ bcol,
                                             const BV *__restrict__ bval, int slog, unsigned *bits,
                                             double *vals, long long &fresh, double *psum = nullptr)
{
    const int t = threadIdx.x;
    const int sub = t & ((1 << slog) - 1);
    const int step_e = kDense >> slog;
    const int step_b = 1 << slog;
    for (int64_t e = e0 + (t >> slog); e < e1; e += step_e) {
        const int k = lcol[e];
        double lv = 0.0;
        if constexpr (NUM) lv = (double)lval[e];
        const int b0 = brp[k], b1 = brp[k + 1];
        for (int q = b0 + sub; q < b1; q += step_b) {
            const int c = bcol[q];
            const unsigned bit = 1u << (c & 31);
            const unsigned old = atomicOr(bits + (c >> 5), bit);
            fresh += (old & bit) ? 0 : 1;
            if constexpr (NUM) {
                const double pr = lv * (double)bval[q];
                unsafeAtomicAdd(vals + c, pr);
                if (psum) *psum += pr;
            }
        }
    }
}

// index (0..31) of the k-th set bit of x (k < popcount(x))
__device__ __forceinline__ int nth_set_bit(unsigned x, int k)
{
    int b = 0;
    for (int s = 16; s > 0; s >>= 1) {
        const unsigned low = (b + s) >= 32 ? x : (x & ((1u << (b + s)) - 1u));
        if (__popc(low) <= k) b += s;
    }
    return b;
}

// Sorted emission of the set bits (and values) of a workspace of nw words;
// returns the count.  Entries beyond `cap` are counted but not written.
// Every thread owns one word per pass; a wavefront's 64 words hold a run of
// consecutive outputs, which the wavefront writes 64 entries per store
// instruction: output j of the run goes to lane j % 64, which finds the word
// holding it (binary search over the lanes' exclusive counts) and the bit
// inside the word.  Emitted bits and values are cleared.
// A workspace (LDS, or global memory) belongs to one workgroup, so the
// fences between its phases are workgroup-scope: a device-scope fence
// would also write back the XCD's L2 at every emission.
__device__ __forceinline__ int dense_emit(unsigned *bits, double *vals, int nw, int32_t *outc, double *outv,
                                          long long cap, long long *sh, int c0 = 0)
{
    const int t = threadIdx.x, l = lane();
    long long base = 0;
    for (int w0 = 0; w0 < nw; w0 += kDense) {
        const int w = w0 + t;
        const unsigned x = w < nw ? ld_word(bits + w) : 0u;
        const int pc = __popc(x);
        long long total;
        const long long pos = base + block_exclusive_scan(pc, sh, total);
        const long long wfirst = __shfl(pos, 0);
        const int wcount = (int)(__shfl(pos + pc, 63) - wfirst);
        const int ex = (int)(pos - wfirst);  // this lane's first output within the wavefront's run
        const int wbase = w - l;             // word of lane 0
        for (int j0 = 0; j0 < wcount; j0 += 64) {
            const int j = j0 + l;
            // the last lane m with ex(m) <= j holds output j
            int m = 0;
            for (int st = 32; st > 0; st >>= 1) {
                const int e = __shfl(ex, m + st);
                if (m + st < 64 && e <= j) m += st;
            }
            const unsigned xm = __shfl(x, m);
            const int em = __shfl(ex, m);
            if (j < wcount) {
                const int c = (wbase + m) * 32 + nth_set_bit(xm, j - em);
                const long long o = wfirst + j;
                const bool in = o < cap;
                if (outc && in) outc[o] = c0 + c;
                if (vals) {
                    const double v = ld_val(vals + c);
                    vals[c] = 0.0;
                    if (outv && in) outv[o] = v;
                }
            }
        }
        if (x) bits[w] = 0u;
        base += total;
    }
    __threadfence_block();
    __syncthreads();
    return (int)base;
}

__device__ __forceinline__ long long block_sum_dense(long long v, long long *sh)
{
    long long total;
    (void)block_exclusive_scan(v, sh, total);
    return total;
}

// pairwise SpMM rows that are too long for the LDS tables
// LDS: B's column space fits in LDS (bitmap, and values in the numeric pass);
// the workspace is zeroed once per workgroup and kept clear as above.
template <bool NUM, bool LDS = false>
__global__ void __launch_bounds__(kDense) k_dense_spgemm(SpOperand A, SpOperand B, const int32_t *__restrict__ rows,
                                                          const int *__restrict__ count, unsigned *ws_bits,
                                                          double *ws_vals, int nw, int slog,
                                                          int32_t *__restrict__ rownnz,
                                                          const int64_t *__restrict__ crp,
                                                          int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                          int *__restrict__ err)
{
    __shared__ long long sh[kDense / 64];
    unsigned *bits;
    double *vals;
    if constexpr (LDS) {
        extern __shared__ __align__(16) unsigned char dsm[];
        vals = NUM ? reinterpret_cast<double *>(dsm) : nullptr;
        bits = reinterpret_cast<unsigned *>(dsm + (NUM ? (size_t)nw * 32 * 8 : 0));
        for (int w = threadIdx.x; w < nw; w += kDense) bits[w] = 0u;
        if constexpr (NUM)
            for (int s = threadIdx.x; s < nw * 32; s += kDense) vals[s] = 0.0;
        __syncthreads();
    } else {
        bits = ws_bits + (size_t)blockIdx.x * nw;
        vals = NUM ? ws_vals + (size_t)blockIdx.x * ((size_t)nw * 32) : nullptr;
    }
    const double *aval = static_cast<const double *>(A.val);
    const double *bval = static_cast<const double *>(B.val);
    const int n = *count;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int r = rows[i];
        long long fresh = 0;
        dense_expand<NUM, double>(A.col, aval, A.rp[r], A.rp[r + 1], B.rp32, B.col, bval, slog, bits, vals,
                                  fresh);
        __threadfence_block();
        __syncthreads();
        if constexpr (!NUM) {
            const long long tot = block_sum_dense(fresh, sh);
            for (int w = threadIdx.x; w < nw; w += kDense) bits[w] = 0u;
            if (threadIdx.x == 0) rownnz[r] = (int32_t)tot;
            __threadfence_block();
            __syncthreads();
        } else {
            const int64_t off = crp[r];
            const int got = dense_emit(bits, vals, nw, ccol + off, cval + off, crp[r + 1] - off, sh);
            if (threadIdx.x == 0 && got != crp[r + 1] - off) atomicOr(err, kErrOverflow);
        }
    }
}

// Long rows of pairwise SpMM C = A*B, walked window by window over B's
// columns with every left entry's position kept between windows.
//
// Per row, the left entries are staged once in LDS (for the first
// kWinEntries of a row): cursor into the right row, its end, and (numeric
// pass) the left value; an entry beyond that re-derives its cursor by binary
// search and its row bounds from global memory.  A window then costs one
// round of loads per kWinStep positions of every entry: sub-groups of
// kWinSub lanes each own one entry, every lane loads kWinVec positions
// (cursor + sub + v*kWinSub) at once, the positions whose column falls inside
// the window form a prefix of the sub-group's kWinVec*kWinSub positions (the
// right row is sorted), and their count advances the cursor; a step with a
// position past the window ends the entry's slice.  Each sub-group loads its
// next round before running this round's atomics, so a round's loads wait
// behind the atomics instead of after them.
//
// Count pass (NUM false): presence bits only, one window spanning all of B's
// columns (they fit in LDS as bits), the row length to rownnz[r].  Numeric
// pass: windows of `wcols` columns with FP64 values (native ds_add_f64) and
// presence bits in LDS, each window's part emitted in column order at
// crp[r] + (entries of the earlier windows).  One 1024-thread workgroup per
// row, persistent over the row list.
constexpr int kWinEntries = 1024;
constexpr int kWinPiece = 256;
constexpr int kWinSub = 16;
constexpr int kWinVec = 4;
constexpr int kWinStep = kWinSub * kWinVec;

__device__ __forceinline__ int lower_col(const int32_t *__restrict__ col, int lo, int hi, int c)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// LDS bytes of k_spgemm_window<NUM> for windows of wcols columns
__host__ __device__ constexpr size_t spgemm_window_lds(bool num, int wcols)
{
    return (num ? (size_t)wcols * 8 : 0) + (size_t)(wcols + 31) / 32 * 4 + (size_t)kWinEntries * (num ? 16 : 8);
}

template <bool NUM>
__global__ void __launch_bounds__(kDense) k_spgemm_window(SpOperand A, SpOperand B, const int32_t *__restrict__ rows,
                                                           const int *__restrict__ count, int wcols,
                                                           int32_t *__restrict__ rownnz,
                                                           const int64_t *__restrict__ crp,
                                                           int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                           int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char dsm[];
    __shared__ long long sh[kDense / 64];
    const int nwords = (wcols + 31) / 32;
    double *wv = reinterpret_cast<double *>(dsm);                                    // NUM: wcols values
    unsigned *wbits = reinterpret_cast<unsigned *>(dsm + (NUM ? (size_t)wcols * 8 : 0));  // nwords words
    unsigned char *ep = dsm + (NUM ? (size_t)wcols * 8 : 0) + (size_t)nwords * 4;
    double *elv = reinterpret_cast<double *>(ep);                                    // NUM: kWinEntries values
    int *ecur = reinterpret_cast<int *>(ep + (NUM ? (size_t)kWinEntries * 8 : 0));   // kWinEntries cursors
    int *eend = ecur + kWinEntries;                                                  // kWinEntries row ends
    {
        unsigned *z = reinterpret_cast<unsigned *>(dsm);
        const int nz = (NUM ? wcols * 2 : 0) + nwords;
        for (int w = threadIdx.x; w < nz; w += kDense) z[w] = 0u;
        __syncthreads();
    }
    const int t = threadIdx.x;
    const int sub = t & (kWinSub - 1);
    constexpr int ng = kDense / kWinSub;
    const int sh0 = lane() & ~(kWinSub - 1);  // this sub-group's bits of a wavefront ballot
    const double *aval = static_cast<const double *>(A.val);
    const double *bval = static_cast<const double *>(B.val);
    const int n = *count;
    const int ncol = (int)B.N;
    const int span = NUM ? wcols : ncol;
    __shared__ int lohi[2 * (kDense / 64)];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int r = rows[i];
        const int64_t e0 = A.rp[r];
        const int na = (int)(A.rp[r + 1] - e0);
        // Stage the entries.  A right row longer than kWinPiece positions is
        // cut into pieces of at most kWinPiece, each staged with its own
        // cursor, so that no sub-group walks a long row alone while the others
        // wait at the window's barrier -- if all pieces of the row fit the
        // staging area; else the entries are staged whole (and those beyond
        // it read from global memory).  The numeric pass also finds the row's
        // column range [lo, hi] from the first and last column of every right
        // row and runs its windows over that range only.
        int pieces = 0;
        for (int e = t; e < na; e += kDense) {
            const int k = A.col[e0 + e];
            pieces += (B.rp32[k + 1] - B.rp32[k] + kWinPiece - 1) / kWinPiece;
        }
        const bool split = block_sum_dense(pieces, sh) <= kWinEntries;
        int lo = INT_MAX, hi = -1;
        long long pbase = 0;
        for (int e0c = 0; e0c < na; e0c += kDense) {
            const int e = e0c + t;
            int b0 = 0, b1 = 0, np = 0;
            double x = 0.0;
            if (e < na) {
                const int k = A.col[e0 + e];
                b0 = B.rp32[k];
                b1 = B.rp32[k + 1];
                if constexpr (NUM) x = aval[e0 + e];
                np = split ? (b1 - b0 + kWinPiece - 1) / kWinPiece : 1;
                if (NUM && b1 > b0) {
                    lo = min(lo, B.col[b0]);
                    hi = max(hi, B.col[b1 - 1]);
                }
            }
            long long tot;
            const long long first = pbase + (split ? block_exclusive_scan(np, sh, tot) : (long long)e0c + t);
            if (!split) tot = min(kDense, na - e0c);
            for (int j = 0; j < np; j++) {
                const long long v = first + j;
                if (v < kWinEntries) {
                    ecur[v] = b0 + j * kWinPiece;
                    eend[v] = split ? min(b1, b0 + (j + 1) * kWinPiece) : b1;
                    if constexpr (NUM) elv[v] = x;
                }
            }
            pbase += tot;
        }
        // staged units: pieces (split) or entries; with !split, entries at
        // and past kWinEntries come from global memory
        const int nunit = split ? (int)pbase : na;
        if constexpr (NUM) {
            for (int o = 32; o > 0; o >>= 1) {
                lo = min(lo, __shfl_xor(lo, o));
                hi = max(hi, __shfl_xor(hi, o));
            }
            if (lane() == 0) {
                lohi[t / 64] = lo;
                lohi[kDense / 64 + t / 64] = hi;
            }
        }
        __syncthreads();
        int c0 = 0, c1 = ncol;  // columns [c0, c1) in windows of `span`
        if constexpr (NUM) {
            lo = lohi[0];
            hi = lohi[kDense / 64];
            for (int w = 1; w < kDense / 64; w++) {
                lo = min(lo, lohi[w]);
                hi = max(hi, lohi[kDense / 64 + w]);
            }
            c0 = hi >= lo ? lo : 0;
            c1 = hi >= lo ? hi + 1 : 0;
        }
        int64_t off = 0;
        long long cap = 0;
        if constexpr (NUM) {
            off = crp[r];
            cap = crp[r + 1] - off;
        }
        long long base = 0;
        for (int w0 = c0; w0 < c1; w0 += span) {
            const int w1 = min(c1, w0 + span);
            // every sub-group runs the same number of rounds, so the ballots
            // below see whole sub-groups
            for (int eb = 0; eb < nunit; eb += ng) {
                const int e = eb + t / kWinSub;
                const bool ok = e < nunit;
                int q = 0, qe = 0;
                double lv = 0.0;
                if (ok) {
                    if (e < kWinEntries) {
                        q = ecur[e];
                        qe = eend[e];
                        if constexpr (NUM) lv = elv[e];
                    } else {
                        const int k = A.col[e0 + e];
                        qe = B.rp32[k + 1];
                        q = w0 == c0 ? B.rp32[k] : lower_col(B.col, B.rp32[k], qe, w0);
                        if constexpr (NUM) lv = aval[e0 + e];
                    }
                }
                bool more = ok && q < qe;
                // rounds of kWinStep positions, loaded one round ahead into
                // two register sets in turn: a sub-group's next round starts
                // kWinStep positions on when this round was all inside the
                // window, and its loads are in flight while this round's
                // atomics run
                int ca[kWinVec], cb[kWinVec];
                double va[kWinVec], vb[kWinVec];
                auto load = [&](int at, int *c, double *bv, unsigned &m) {
                    m = 0u;
#pragma unroll
                    for (int v = 0; v < kWinVec; v++) {
                        const int qq = at + v * kWinSub + sub;
                        const bool valid = more && qq < qe;
                        // unconditional loads (a valid index either way) and
                        // the validity kept apart, so that neither a branch
                        // nor a select on the loaded values comes before the
                        // next round's loads are issued
                        const int qs = valid ? qq : 0;
                        m |= valid ? 1u << v : 0u;
                        c[v] = B.col[qs];
                        bv[v] = 0.0;
                        if constexpr (NUM) bv[v] = bval[qs];
                    }
                };
                auto round = [&](const int *c, const double *bv, unsigned m) {
                    int cnt = 0;
#pragma unroll
                    for (int v = 0; v < kWinVec; v++) {
                        const bool in = ((m >> v) & 1u) && c[v] < w1;
                        if (in) {
                            const int cw = c[v] - w0;
                            atomicOr(wbits + (cw >> 5), 1u << (cw & 31));
                            if constexpr (NUM) atomicAdd(wv + cw, lv * bv[v]);
                        }
                        cnt += __popcll((ballot(in) >> sh0) & ((1ull << kWinSub) - 1));
                    }
                    q += cnt;
                    more = more && cnt == kWinStep;
                };
                unsigned ma, mb;
                if (ballot(more)) {  // some lane has entries: B is not empty
                    load(q, ca, va, ma);
                    for (;;) {
                        load(q + kWinStep, cb, vb, mb);
                        round(ca, va, ma);
                        if (!ballot(more)) break;
                        load(q + kWinStep, ca, va, ma);
                        round(cb, vb, mb);
                        if (!ballot(more)) break;
                    }
                }
                if (ok && sub == 0 && e < kWinEntries) ecur[e] = q;
            }
            __syncthreads();
            if constexpr (NUM) {
                base += dense_emit(wbits, wv, (w1 - w0 + 31) / 32, ccol + off + base, cval + off + base,
                                   cap - base, sh, w0);
            } else {
                long long pc = 0;
                for (int w = t; w < nwords; w += kDense) {
                    pc += __popc(wbits[w]);
                    wbits[w] = 0u;
                }
                base += block_sum_dense(pc, sh);
                __syncthreads();
            }
        }
        if constexpr (NUM) {
            if (t == 0 && base != cap) atomicOr(err, kErrOverflow);
        } else {
            if (t == 0) rownnz[r] = (int32_t)base;
        }
        __syncthreads();
    }
}

// direct triple product rows that are too long for the LDS tables:
// row of R*A densely over A's columns, compacted into a sorted list, then
// forwarded densely over P's columns.  L2: P's column space fits in LDS, and
// the second stage accumulates there (bitmap + values) instead of in global
// memory; L1: the first stage's bitmap over A's columns lives in LDS (its
// values stay in global memory).  LDS workspaces are zeroed once per
// workgroup, and emission clears what it emits.
template <bool NUM, typename VT = double, bool L1 = false, bool L2 = false>
__global__ void __launch_bounds__(kDense) k_dense_direct(SpOperand R, SpOperand A, SpOperand P,
                                                          const int32_t *__restrict__ rows,
                                                          const int *__restrict__ count, DenseWs ws,
                                                          int slog1, int slog2, int32_t *__restrict__ n1out,
                                                          int32_t *__restrict__ rownnz,
                                                          const int64_t *__restrict__ crp,
                                                          int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                          double *__restrict__ rowsum, int *__restrict__ err)
{
    __shared__ long long sh[kDense / 64];
    __shared__ double dsh[kDense / 64];
    const size_t b = blockIdx.x;
    extern __shared__ __align__(16) unsigned char dsm[];
    // LDS layout: [stage-2 values | stage-2 bitmap | stage-1 bitmap]
    const size_t l2bytes = L2 ? (NUM ? (size_t)ws.nwP * 32 * 8 : 0) + (size_t)ws.nwP * 4 : 0;
    unsigned *bitsA = ws.bitsA + b * ws.nwA;
    if constexpr (L1) {
        bitsA = reinterpret_cast<unsigned *>(dsm + l2bytes);
        for (int w = threadIdx.x; w < ws.nwA; w += kDense) bitsA[w] = 0u;
    }
    double *valsA = NUM ? ws.valsA + b * ((size_t)ws.nwA * 32) : nullptr;
    int32_t *listc = ws.listc + b * ((size_t)ws.nwA * 32);
    double *listv = NUM ? ws.listv + b * ((size_t)ws.nwA * 32) : nullptr;
    unsigned *bitsP;
    double *valsP;
    if constexpr (L2) {
        valsP = NUM ? reinterpret_cast<double *>(dsm) : nullptr;
        bitsP = reinterpret_cast<unsigned *>(dsm + (NUM ? (size_t)ws.nwP * 32 * 8 : 0));
        for (int w = threadIdx.x; w < ws.nwP; w += kDense) bitsP[w] = 0u;
        if constexpr (NUM)
            for (int s = threadIdx.x; s < ws.nwP * 32; s += kDense) valsP[s] = 0.0;
    } else {
        bitsP = ws.bitsP + b * ws.nwP;
        valsP = NUM ? ws.valsP + b * ((size_t)ws.nwP * 32) : nullptr;
    }
    if constexpr (L1 || L2) __syncthreads();
    const VT *rval = static_cast<const VT *>(R.val);
    const VT *aval = static_cast<const VT *>(A.val);
    const VT *pval = static_cast<const VT *>(P.val);
    const int n = *count;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int r = rows[i];
        long long f1 = 0, f2 = 0;
        dense_expand<NUM, VT, VT>(R.col, rval, R.rp[r], R.rp[r + 1], A.rp32, A.col, aval, slog1, bitsA, valsA,
                                  f1);
        __threadfence_block();
        __syncthreads();
        const int n1 = dense_emit(bitsA, valsA, ws.nwA, listc, listv, (long long)ws.nwA * 32, sh);
        double ps = 0.0;
        dense_expand<NUM, double, VT>(listc, listv, 0, n1, P.rp32, P.col, pval, slog2, bitsP, valsP, f2, &ps);
        __threadfence_block();
        __syncthreads();
        if (NUM && rowsum) {
            for (int o = 32; o > 0; o >>= 1) ps += __shfl_xor(ps, o);
            if (lane() == 0) dsh[threadIdx.x / 64] = ps;
            __syncthreads();
            if (threadIdx.x == 0) {
                double tot = 0.0;
                for (int w = 0; w < kDense / 64; w++) tot += dsh[w];
                rowsum[r] = tot;
            }
        }
        if constexpr (!NUM) {
            const long long tot = block_sum_dense(f2, sh);
            for (int w = threadIdx.x; w < ws.nwP; w += kDense) bitsP[w] = 0u;
            if (threadIdx.x == 0) {
                rownnz[r] = (int32_t)tot;
                if (n1out) n1out[r] = n1;
            }
            __threadfence_block();
            __syncthreads();
        } else {
            const int64_t off = crp[r];
            const int got = dense_emit(bitsP, valsP, ws.nwP, ccol + off, cval + off, crp[r + 1] - off, sh);
            if (threadIdx.x == 0 && got != crp[r + 1] - off) atomicOr(err, kErrOverflow);
        }
    }
}

// Numeric pass of the direct long-row kernel when P's column space exceeds
// LDS: the first stage as in k_dense_direct (its bitmap in LDS when L1), then
// the second stage one window of `wcols` of P's columns at a time, accumulated
// in LDS.  Every first-stage entry keeps a cursor into its (sorted) row of P,
// so each window consumes only its slice and P's rows are read once in all;
// each window's slice of the output row is emitted in column order straight
// after the previous one.  The window region reuses the first stage's bitmap
// space (clear once that stage is emitted); emission keeps it clear.
template <typename VT, bool L1>
__global__ void __launch_bounds__(kDense) k_dense_window(SpOperand R, SpOperand A, SpOperand P,
                                                          const int32_t *__restrict__ rows,
                                                          const int *__restrict__ count, DenseWs ws, int wcols,
                                                          int slog1, const int64_t *__restrict__ crp,
                                                          int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                          double *__restrict__ rowsum, int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char dsm[];
    __shared__ long long sh[kDense / 64];
    __shared__ double dsh[kDense / 64];
    const size_t b = blockIdx.x;
    double *wv = reinterpret_cast<double *>(dsm);                        // wcols values
    unsigned *wbits = reinterpret_cast<unsigned *>(dsm + (size_t)wcols * 8);  // wcols / 32 words
    unsigned *bitsA = L1 ? reinterpret_cast<unsigned *>(dsm) : ws.bitsA + b * ws.nwA;
    double *valsA = ws.valsA + b * ((size_t)ws.nwA * 32);
    int32_t *listc = ws.listc + b * ((size_t)ws.nwA * 32);
    double *listv = ws.listv + b * ((size_t)ws.nwA * 32);
    int32_t *cur = ws.cur + b * ((size_t)ws.nwA * 32);
    {
        // clear the LDS once: the larger of the two uses of the region
        const int nw = max(L1 ? ws.nwA : 0, wcols * 2 + wcols / 32);
        unsigned *z = reinterpret_cast<unsigned *>(dsm);
        for (int w = threadIdx.x; w < nw; w += kDense) z[w] = 0u;
        __syncthreads();
    }
    const VT *rval = static_cast<const VT *>(R.val);
    const VT *aval = static_cast<const VT *>(A.val);
    const VT *pval = static_cast<const VT *>(P.val);
    const int n = *count;
    const int ncol = (int)P.N;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int r = rows[i];
        long long f1 = 0;
        dense_expand<true, VT, VT>(R.col, rval, R.rp[r], R.rp[r + 1], A.rp32, A.col, aval, slog1, bitsA, valsA, f1);
        __threadfence_block();
        __syncthreads();
        const int n1 = dense_emit(bitsA, valsA, ws.nwA, listc, listv, (long long)ws.nwA * 32, sh);
        for (int e = threadIdx.x; e < n1; e += kDense) cur[e] = P.rp32[listc[e]];
        __threadfence_block();
        __syncthreads();
        const int64_t off = crp[r];
        const long long cap = crp[r + 1] - off;
        long long base = 0;
        double ps = 0.0;
        for (int w0 = 0; w0 < ncol; w0 += wcols) {
            const int w1 = min(ncol, w0 + wcols);
            for (int e = threadIdx.x; e < n1; e += kDense) {
                int q = cur[e];
                const int qe = P.rp32[listc[e] + 1];
                const double lv = listv[e];
                for (; q < qe; q++) {
                    const int c = P.col[q];
                    if (c >= w1) break;
                    const int cw = c - w0;
                    atomicOr(wbits + (cw >> 5), 1u << (cw & 31));
                    const double pr = lv * (double)pval[q];
                    atomicAdd(wv + cw, pr);
                    ps += pr;
                }
                cur[e] = q;
            }
            __syncthreads();
            base += dense_emit(wbits, wv, (w1 - w0 + 31) / 32, ccol + off + base, cval + off + base,
                               cap - base, sh, w0);
        }
        if (rowsum) {
            for (int o = 32; o > 0; o >>= 1) ps += __shfl_xor(ps, o);
            if (lane() == 0) dsh[threadIdx.x / 64] = ps;
            __syncthreads();
            if (threadIdx.x == 0) {
                double tot = 0.0;
                for (int w = 0; w < kDense / 64; w++) tot += dsh[w];
                rowsum[r] = tot;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && base != cap) atomicOr(err, kErrOverflow);
    }
}

// ---------------------------------------------------------------------------
// upload / download conversions and validation
// ---------------------------------------------------------------------------
// rp64 (host offsets, index base `base`) -> 0-based rp (+ rp32); checks monotone
__global__ void k_rp_convert(const int64_t *__restrict__ in, int64_t M, int64_t nnz, int base,
                             int64_t *__restrict__ rp, int32_t *__restrict__ rp32,
                             int *__restrict__ err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > M) return;
    const int64_t v = in[i] - base;
    rp[i] = v;
    if (rp32) rp32[i] = (int32_t)v;
    bool bad = v < 0 || v > nnz;
    if (i == 0) bad |= v != 0;
    if (i == M) bad |= v != nnz;
    if (i < M && in[i + 1] < in[i]) bad = true;
    if (bad && err) atomicOr(err, kErrBadInput);
}

// int64 columns (base) -> int32 0-based; checks range and strict increase per row
__global__ void k_col_convert(const int64_t *__restrict__ in, const int64_t *__restrict__ rp, int64_t M,
                              int64_t N, int base, int32_t *__restrict__ col, int *__restrict__ err)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const int64_t nnz = rp[M];
    int64_t e0 = rp[r], e1 = rp[r + 1];
    // offsets are validated separately; never walk outside [0, nnz) whatever they hold
    if (e0 < 0 || e1 > nnz || e0 > e1 || nnz < 0) {
        atomicOr(err, kErrBadInput);
        return;
    }
    int64_t prev = -1;
    bool bad = false;
    for (int64_t e = e0; e < e1; e++) {
        const int64_t c = in[e] - base;
        bad |= c < 0 || c >= N || c <= prev;
        prev = c;
        col[e] = (int32_t)c;
    }
    if (bad) atomicOr(err, kErrBadInput);
}

__global__ void k_col_widen(const int32_t *__restrict__ col, int64_t nnz, int base, int64_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) out[i] = (int64_t)col[i] + base;
}

__global__ void k_rp_shift(const int64_t *__restrict__ rp, int64_t n, int base, int64_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = rp[i] + base;
}

__global__ void k_f64_to_f32(const double *__restrict__ in, int64_t n, float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void k_f32_to_f64(const float *__restrict__ in, int64_t n, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// copies a small device block (the control counters) straight into pinned
// host memory, so reading it back needs no copy-engine transfer
__global__ void k_export_words(const uint32_t *__restrict__ src, uint32_t *dst, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_probe(const double4 *__restrict__ in, double4 *__restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

int64_t bin_scratch_ints(int64_t M) { return (kMaxBins + 1) * ((M + kBinThreads - 1) / kBinThreads + 1) + 6 * ((M + kBinThreads - 1) / kBinThreads + 1) + 2; }

static void launch_bins(const SpOperand *A, const SpOperand *B, const int32_t *m1, const int32_t *m2, int64_t M,
                        int64_t *ub, int32_t *rownnz, int32_t *lists, int *counts, const BinThresholds &thr,
                        unsigned long long *totals, int *scratch, hipStream_t st, int merge_lim = 0,
                        int *mcounts = nullptr, const CtlExport *ex = nullptr, const OneBin *one = nullptr,
                        bool wave_ub = false)
{
    if (M == 0) return;
    // long rows of A: their bounds measured a wavefront per row first (not
    // when the one-bin test may make measuring unnecessary)
    const bool pre = wave_ub && A && B && ub && !one;
    if (pre)
        hipLaunchKernelGGL(k_row_ub_wave, dim3(blocks_for(M, 256 / kWave)), dim3(256), 0, st, A->rp, A->col,
                           B->rp32, M, ub);
    OneBin ob{};
    ob.word = -1;
    if (one && A && B) {
        ob = *one;
        // the int32 offsets when the operands have them (half the bytes)
        int *part = const_cast<int *>(ob.part);
        if (A->rp32 && B->rp32)
            hipLaunchKernelGGL((k_row_maxlen<int32_t, int32_t>), dim3(kMaxLenBlocks), dim3(256), 0, st, A->rp32, M,
                               B->rp32, B->M, part);
        else
            hipLaunchKernelGGL((k_row_maxlen<int64_t, int64_t>), dim3(kMaxLenBlocks), dim3(256), 0, st, A->rp, M,
                               B->rp, B->M, part);
    }
    const int nblk = (int)((M + kBinThreads - 1) / kBinThreads);
    int *hist = scratch;
    unsigned long long *part = reinterpret_cast<unsigned long long *>(scratch + (int64_t)(kMaxBins + 1) * (nblk + 1) + 1) ;
    part = reinterpret_cast<unsigned long long *>((reinterpret_cast<uintptr_t>(part) + 7) & ~(uintptr_t)7);
    uint32_t *zero = ex && ex->clear ? const_cast<uint32_t *>(ex->src) : nullptr;
    hipLaunchKernelGGL(k_bin_hist, dim3(nblk), dim3(kBinThreads), 0, st, A ? A->rp : nullptr, A ? A->col : nullptr,
                       B ? B->rp32 : nullptr, m1, m2, M, ub, rownnz, hist, part, thr, zero, zero ? ex->words : 0,
                       ex ? ex->span_lim : 0, ob, pre ? 1 : 0);
    hipLaunchKernelGGL(k_bin_offsets, dim3(thr.nb + 2), dim3(1024), 0, st, hist, nblk, thr.nb + 1, part, counts,
                       totals);
    CtlExport e{};
    if (ex) e = *ex;
    hipLaunchKernelGGL(k_bin_scatter, dim3(nblk), dim3(kBinThreads), 0, st, ub, m1, m2, M, hist, lists, thr, counts,
                       merge_lim, mcounts, e);
}

void launch_ub_bin(const SpOperand &A, const SpOperand &B, int64_t *ub, int32_t *rownnz,
                   int32_t *lists, int *counts, const BinThresholds &thr,
                   unsigned long long *totals, int *scratch, hipStream_t st, int merge_lim, int *mcounts,
                   const void *ctl, void *ctl_host, int ctl_words, unsigned long long *flag, unsigned long long seq,
                   int span_lim, int *pred_count, int pred, int big_from, bool reassoc_possible, const OneBin *one,
                   bool wave_ub)
{
    CtlExport e{static_cast<const uint32_t *>(ctl), static_cast<uint32_t *>(ctl_host), ctl_words, flag, seq,
                ctl != nullptr, span_lim, pred_count, pred, big_from, reassoc_possible, totals,
                ctl && one ? one->word : -1};
    launch_bins(&A, &B, nullptr, nullptr, A.M, ub, rownnz, lists, counts, thr, totals, scratch, st, merge_lim,
                mcounts, ctl ? &e : nullptr, ctl ? one : nullptr, wave_ub);
}

void launch_ub_colparts(const int64_t *arp, const int32_t *acol, int64_t M, const int64_t *boff, int G,
                        int64_t *out, unsigned long long *total, hipStream_t st)
{
    if (M == 0) return;
    hipLaunchKernelGGL(k_ub_colparts, dim3(blocks_for(M, 256)), dim3(256), 0, st, arp, acol, M, boff, G, out,
                       total);
}

void launch_rebin(const int32_t *m1, const int32_t *m2, int64_t M, int32_t *lists, int *counts,
                  const BinThresholds &thr, int *scratch, hipStream_t st)
{
    launch_bins(nullptr, nullptr, m1, m2, M, nullptr, nullptr, lists, counts, thr, nullptr, scratch, st);
}

static unsigned grid_for(int rows, int G, int max_blocks)
{
    const int gpb = 256 / G;
    int64_t b = (rows + gpb - 1) / gpb;
    if (b > max_blocks) b = max_blocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

size_t symb_lds(int G, int H) { return (size_t)(256 / G) * H * 4 + (size_t)(256 / G) * group_waves(G) * 8; }
size_t num_lds(int G, int H) { return (size_t)(256 / G) * H * 12 + (size_t)(256 / G) * (group_waves(G) + 1) * 4; }
size_t direct_lds(int G, int H1, int H2, bool num)
{
    const size_t gpb = 256 / G;
    size_t b = gpb * ((size_t)H1 + H2) * (num ? 12 : 4);
    b += gpb * group_waves(G) * 8 + gpb * (group_waves(G) + 1) * 4;
    return b;
}

void launch_symb(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                 const int *count, int rows_host, int32_t *rownnz, int32_t *retry, int *nretry,
                 int *err, int max_blocks, hipStream_t st)
{
    const unsigned grid = grid_for(rows_host, G, max_blocks);
    const size_t lds = symb_lds(G, H);
    if (G == 64)
        hipLaunchKernelGGL(k_symb<64>, dim3(grid), dim3(256), lds, st, A, B, rows, count, H, slog,
                           rownnz, retry, nretry, err);
    else
        hipLaunchKernelGGL(k_symb<256>, dim3(grid), dim3(256), lds, st, A, B, rows, count, H, slog,
                           rownnz, retry, nretry, err);
}

void launch_num(int G, int H, int slog, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                const int *count, int rows_host, const int64_t *crp, int32_t *ccol, double *cval,
                int *err, int max_blocks, hipStream_t st)
{
    const unsigned grid = grid_for(rows_host, G, max_blocks);
    const size_t lds = num_lds(G, H);
    if (G == 64)
        hipLaunchKernelGGL(k_num<64>, dim3(grid), dim3(256), lds, st, A, B, rows, count, H, slog, crp,
                           ccol, cval, err);
    else
        hipLaunchKernelGGL(k_num<256>, dim3(grid), dim3(256), lds, st, A, B, rows, count, H, slog, crp,
                           ccol, cval, err);
}

void launch_direct(int G, bool num, int H1, int H2, int slog1, int slog2, const SpOperand &R,
                   const SpOperand &A, const SpOperand &P, const int32_t *rows, const int *count,
                   int rows_host, int32_t *n1out, int32_t *rownnz, const int64_t *crp, int32_t *ccol,
                   double *cval, int32_t *retry, int *nretry, int *err, int max_blocks,
                   hipStream_t st, bool f32in)
{
    if (G < 64) G = 64;  // the two-pass kernels run one row per wavefront or workgroup
    const unsigned grid = grid_for(rows_host, G, max_blocks);
    const size_t lds = direct_lds(G, H1, H2, num);
#define SP3MM_DIRECT_ARGS R, A, P, rows, count, H1, H2, slog1, slog2, n1out, rownnz, crp, ccol, cval, \
                          retry, nretry, err, rows_host
    if (f32in) {
        if (!num) return;  // the mixed path only recomputes numerics from FP32 inputs
        if (G == 64) hipLaunchKernelGGL((k_direct<64, true, float>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
        else hipLaunchKernelGGL((k_direct<256, true, float>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
    } else if (G == 64) {
        if (num) hipLaunchKernelGGL((k_direct<64, true>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
        else hipLaunchKernelGGL((k_direct<64, false>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
    } else {
        if (num) hipLaunchKernelGGL((k_direct<256, true>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
        else hipLaunchKernelGGL((k_direct<256, false>), dim3(grid), dim3(256), lds, st, SP3MM_DIRECT_ARGS);
    }
#undef SP3MM_DIRECT_ARGS
}

int dense_words(int64_t ncols) { return (int)((ncols + 31) / 32) + 1; }

// widest power-of-two column window whose values and presence bits, with the
// staged entries of k_spgemm_window, fit in `max_lds` bytes (0: none)
int spgemm_window_cols(size_t max_lds)
{
    for (int w = 1 << 20; w >= 1024; w >>= 1)
        if (spgemm_window_lds(true, w) <= max_lds) return w;
    return 0;
}

bool spgemm_dense_in_lds(bool num, int nw, size_t max_lds)
{
    return (size_t)nw * 4 + (num ? (size_t)nw * 32 * 8 : 0) <= max_lds;
}

void launch_dense_spgemm(bool num, int blocks, const SpOperand &A, const SpOperand &B, const int32_t *rows,
                         const int *count, unsigned *bits, double *vals, int nw, int slog, int32_t *rownnz,
                         const int64_t *crp, int32_t *ccol, double *cval, int *err, hipStream_t st,
                         size_t max_lds)
{
    const size_t lds = (size_t)nw * 4 + (num ? (size_t)nw * 32 * 8 : 0);
    const bool in_lds = lds <= max_lds;
    if (!num && spgemm_window_lds(false, (int)B.N) <= max_lds) {
        // count pass: presence bits over all of B's columns, staged entries
        hipLaunchKernelGGL(k_spgemm_window<false>, dim3(blocks), dim3(kDense), spgemm_window_lds(false, (int)B.N), st,
                           A, B, rows, count, (int)B.N, rownnz, crp, ccol, cval, err);
        return;
    }
    if (num && !in_lds) {
        // values beyond LDS: windows of B's columns in LDS, when one fits
        const int wcols = spgemm_window_cols(max_lds);
        if (wcols > 0) {
            hipLaunchKernelGGL(k_spgemm_window<true>, dim3(blocks), dim3(kDense), spgemm_window_lds(true, wcols), st,
                               A, B, rows, count, wcols, rownnz, crp, ccol, cval, err);
            return;
        }
    }
#define SP3MM_DSPG_ARGS A, B, rows, count, bits, vals, nw, slog, rownnz, crp, ccol, cval, err
    if (num && in_lds)
        hipLaunchKernelGGL((k_dense_spgemm<true, true>), dim3(blocks), dim3(kDense), lds, st, SP3MM_DSPG_ARGS);
    else if (num)
        hipLaunchKernelGGL((k_dense_spgemm<true, false>), dim3(blocks), dim3(kDense), 0, st, SP3MM_DSPG_ARGS);
    else if (in_lds)
        hipLaunchKernelGGL((k_dense_spgemm<false, true>), dim3(blocks), dim3(kDense), lds, st, SP3MM_DSPG_ARGS);
    else
        hipLaunchKernelGGL((k_dense_spgemm<false, false>), dim3(blocks), dim3(kDense), 0, st, SP3MM_DSPG_ARGS);
#undef SP3MM_DSPG_ARGS
}

// ---------------------------------------------------------------------------
// long rows with narrow column windows: dense accumulators in LDS
// ---------------------------------------------------------------------------
// Rows too long for the hash tables often still touch a narrow band of
// columns (coarse levels of a hierarchy, block-diagonal batches).  One
// workgroup per row: stage 1 finds the window [lo1, hi1] of the R*A row from
// the first and last column of every A row it references, clears that part of
// a dense LDS array (FP64 values + presence bits) and accumulates into it with
// native ds_add_f64; stage 2 finds the window of the forwarded row in P and
// does the same in a second LDS region.  The presence bits of stage 2, read in
// order, are the sorted output row.  A row whose windows exceed the LDS
// capacity `cap` (values) goes to `spill` (count pass, or slot mode), for the
// global dense kernel.  Workgroups of kWin lanes: with one row in LDS per
// workgroup, 16 wavefronts per CU keep enough loads in flight.
//
// Slot mode (numeric pass with `scol` set): one pass instead of count +
// numeric.  Row i of the list is emitted to its scratch slot scol/sval + i*cap
// (a row that fits the windows has at most cap entries) and its length goes
// to rownnz[r] and slotn[i]; k_gather_rows later moves the slots to their
// final place.
constexpr int kWin = 1024;

__device__ __forceinline__ void block_minmax(int &lo, int &hi, int *sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    __syncthreads();
    if (lane() == 0) {
        sh[threadIdx.x / 64] = lo;
        sh[kWin / 64 + threadIdx.x / 64] = hi;
    }
    __syncthreads();
    lo = sh[0];
    hi = sh[kWin / 64];
    for (int w = 1; w < kWin / 64; w++) {
        lo = min(lo, sh[w]);
        hi = max(hi, sh[kWin / 64 + w]);
    }
    __syncthreads();
}

constexpr int kWinSub2 = 8;  // lanes per forwarded entry in k_window_direct's stage 2

template <bool NUM, typename VT>
__global__ void __launch_bounds__(kWin) k_window_direct(SpOperand R, SpOperand A, SpOperand P,
                                                           const int32_t *__restrict__ rows,
                                                           const int *__restrict__ count, int cap, int slog1,
                                                           int32_t *__restrict__ spill, int *__restrict__ nspill,
                                                           int32_t *__restrict__ rownnz,
                                                           const int64_t *__restrict__ crp,
                                                           int32_t *__restrict__ ccol, double *__restrict__ cval,
                                                           double *__restrict__ rowsum, int64_t out_cap,
                                                           int *__restrict__ err, int32_t *__restrict__ slotn,
                                                           int32_t *__restrict__ scol, double *__restrict__ sval,
                                                           int cap_rows)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const bool slots = NUM && scol != nullptr;
    double *V = reinterpret_cast<double *>(smem);                  // cap values
    unsigned *B = reinterpret_cast<unsigned *>(smem + (size_t)cap * 8);  // cap/32 words
    __shared__ int ish[2 * (kWin / 64)];
    __shared__ long long lsh[kWin / 64];
    __shared__ double dsh[kWin / 64];
    const int t = threadIdx.x;
    const VT *rval = static_cast<const VT *>(R.val);
    const VT *aval = static_cast<const VT *>(A.val);
    const VT *pval = static_cast<const VT *>(P.val);
    const int S = 1 << slog1;
    // list entries i index the per-row slots: at most cap_rows of them
    const int n = clamp_count(count, cap_rows, err);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int r = rows[i];
        const int64_t e0 = R.rp[r], e1 = R.rp[r + 1];
        // window of the R*A row
        int lo1 = INT_MAX, hi1 = -1;
        for (int64_t e = e0 + t; e < e1; e += kWin) {
            const int k = R.col[e];
            const int a0 = A.rp32[k], a1 = A.rp32[k + 1];
            if (a1 > a0) {
                lo1 = min(lo1, A.col[a0]);
                hi1 = max(hi1, A.col[a1 - 1]);
            }
        }
        block_minmax(lo1, hi1, ish);
        const int span1 = hi1 >= lo1 ? hi1 - lo1 + 1 : 0;
        const int nw1 = (span1 + 31) / 32;
        if (nw1 * 32 > cap) {
            if ((!NUM || slots) && t == 0) {
                spill[atomicAdd(nspill, 1)] = r;
                rownnz[r] = 0;
                if (slots) slotn[i] = 0;
            }
            continue;
        }
        for (int w = t; w < nw1; w += kWin) B[w] = 0u;
        if constexpr (NUM)
            for (int j = t; j < span1; j += kWin) V[j] = 0.0;
        __syncthreads();
        // stage 1: sub-groups of S lanes per R entry walk its A row, four
        // entries per lane at a time (their loads in flight together)
        for (int64_t e = e0 + (t >> slog1); e < e1; e += (kWin >> slog1)) {
            const int k = R.col[e];
            double rv = 0.0;
            if constexpr (NUM) rv = (double)rval[e];
            const int a1 = A.rp32[k + 1];
            for (int q0 = A.rp32[k] + (t & (S - 1)); q0 < a1; q0 += 4 * S) {
                int jj[4];
                double av[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int q = q0 + u * S;
                    const bool ok = q < a1;
                    jj[u] = ok ? A.col[q] - lo1 : 0;
                    av[u] = 0.0;
                    if constexpr (NUM) av[u] = ok ? (double)aval[q] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (q0 + u * S >= a1) break;
                    const int j = jj[u];
                    if ((unsigned)j >= (unsigned)span1) {  // an unsorted A row
                        atomicOr(err, kErrBadInput);
                        continue;
                    }
                    const unsigned bit = 1u << (j & 31);
                    if (!(B[j >> 5] & bit)) atomicOr(B + (j >> 5), bit);  // set bits are never cleared here
                    if constexpr (NUM) atomicAdd(V + j, rv * av[u]);
                }
            }
        }
        __syncthreads();
        // window of the forwarded row in P
        int lo2 = INT_MAX, hi2 = -1;
        for (int j = t; j < span1; j += kWin) {
            if (B[j >> 5] & (1u << (j & 31))) {
                const int c = lo1 + j;
                const int p0 = P.rp32[c], p1 = P.rp32[c + 1];
                if (p1 > p0) {
                    lo2 = min(lo2, P.col[p0]);
                    hi2 = max(hi2, P.col[p1 - 1]);
                }
            }
        }
        block_minmax(lo2, hi2, ish);
        const int span2 = hi2 >= lo2 ? hi2 - lo2 + 1 : 0;
        const int nw2 = (span2 + 31) / 32;
        const int off2 = nw1 * 32;  // region 2 starts on a word boundary
        if (off2 + nw2 * 32 > cap) {
            if ((!NUM || slots) && t == 0) {
                spill[atomicAdd(nspill, 1)] = r;
                rownnz[r] = 0;
                if (slots) slotn[i] = 0;
            }
            continue;
        }
        unsigned *B2 = B + nw1;
        double *V2 = V + off2;
        for (int w = t; w < nw2; w += kWin) B2[w] = 0u;
        if constexpr (NUM)
            for (int j = t; j < span2; j += kWin) V2[j] = 0.0;
        __syncthreads();
        // stage 2: sub-groups of kWinSub2 lanes walk the P rows of the present
        // entries (positions strided over the sub-groups), two entries per
        // lane per step.  A lane per present entry walking its P row alone
        // made every load of a wavefront touch 64 rows (the 1728-row level
        // of a 96^3 hierarchy took 0.95 instead of 0.70 ms per product).
        double ps = 0.0;
        {
            constexpr int S2 = kWinSub2;
            const int sub = t & (S2 - 1), ngrp = kWin / S2;
            for (int j = t / S2; j < span1; j += ngrp) {
                if (!(B[j >> 5] & (1u << (j & 31)))) continue;
                const int c = lo1 + j;
                double x = 0.0;
                if constexpr (NUM) x = V[j];
                const int p1 = P.rp32[c + 1];
                for (int q0 = P.rp32[c] + sub; q0 < p1; q0 += 2 * S2) {
                    int jj[2];
                    double pv[2];
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        const bool ok = q0 + u * S2 < p1;
                        jj[u] = ok ? P.col[q0 + u * S2] - lo2 : 0;
                        pv[u] = 0.0;
                        if constexpr (NUM) pv[u] = ok ? (double)pval[q0 + u * S2] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        if (q0 + u * S2 >= p1) break;
                        const int j2 = jj[u];
                        if ((unsigned)j2 >= (unsigned)span2) {  // an unsorted P row
                            atomicOr(err, kErrBadInput);
                            continue;
                        }
                        const unsigned bit2 = 1u << (j2 & 31);
                        if (!(B2[j2 >> 5] & bit2)) atomicOr(B2 + (j2 >> 5), bit2);
                        if constexpr (NUM) {
                            const double pr = x * pv[u];
                            atomicAdd(V2 + j2, pr);
                            ps += pr;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if constexpr (!NUM) {
            long long cnt = 0;
            for (int w = t; w < nw2; w += kWin) cnt += __popc(B2[w]);
            long long total;
            (void)block_exclusive_scan(cnt, lsh, total);
            if (t == 0) rownnz[r] = (int32_t)total;
        } else {
            if (rowsum) {
                for (int o = 32; o > 0; o >>= 1) ps += __shfl_xor(ps, o);
                if (lane() == 0) dsh[t / 64] = ps;
                __syncthreads();
                if (t == 0) {
                    double tot = 0.0;
                    for (int w = 0; w < kWin / 64; w++) tot += dsh[w];
                    rowsum[r] = tot;
                }
            }
            // sorted emission of region 2 (to the output row, or to the slot)
            int32_t *oc;
            double *ov;
            int64_t len;
            if (slots) {
                oc = scol + (size_t)i * cap;
                ov = sval + (size_t)i * cap;
                len = cap;
            } else {
                const int64_t off = crp[r];
                len = crp[r + 1] - off;
                if (off + len > out_cap) {  // output allocated smaller than this row needs
                    if (t == 0) atomicOr(err, kErrOverflow);
                    __syncthreads();
                    continue;
                }
                oc = ccol + off;
                ov = cval + off;
            }
            long long base = 0;
            for (int w0 = 0; w0 < nw2; w0 += kWin) {
                const int w = w0 + t;
                unsigned x = w < nw2 ? B2[w] : 0u;
                long long total;
                long long pos = base + block_exclusive_scan(__popc(x), lsh, total);
                while (x) {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    const int j2 = w * 32 + b;
                    if (pos < len) {
                        oc[pos] = lo2 + j2;
                        ov[pos] = V2[j2];
                    }
                    pos++;
                }
                base += total;
            }
            if (slots) {
                if (t == 0) {
                    rownnz[r] = (int32_t)base;
                    slotn[i] = (int32_t)base;
                }
            } else if (t == 0 && base != len) {
                atomicOr(err, kErrOverflow);
            }
        }
        __syncthreads();
    }
}

// values the window kernel can hold in `lds` bytes (a multiple of 32)
int window_cap(size_t lds)
{
    const size_t c = lds / (8 + 4.0 / 32);
    return (int)(c / 32 * 32);
}

void launch_window_direct(bool num, int blocks, size_t lds, const SpOperand &R, const SpOperand &A,
                          const SpOperand &P, const int32_t *rows, const int *count, int slog1, int32_t *spill,
                          int *nspill, int32_t *rownnz, const int64_t *crp, int32_t *ccol, double *cval,
                          double *rowsum, int64_t out_cap, int *err, hipStream_t st, bool f32in,
                          int32_t *slotn, int32_t *scol, double *sval, int cap_rows)
{
    const int cap = window_cap(lds);
    const size_t bytes = (size_t)cap * 8 + (size_t)cap / 32 * 4;
    // at most 16 lanes per R entry: 64 entries of R's row in flight per
    // workgroup (64 lanes each, 16 in flight, made the 1728-row level of a
    // 96^3 hierarchy 6% slower)
    if (slog1 > 4) slog1 = 4;
#define SP3MM_WIN_ARGS R, A, P, rows, count, cap, slog1, spill, nspill, rownnz, crp, ccol, cval, rowsum, out_cap, \
                       err, slotn, scol, sval, cap_rows
    if (f32in) {
        if (num) hipLaunchKernelGGL((k_window_direct<true, float>), dim3(blocks), dim3(kWin), bytes, st, SP3MM_WIN_ARGS);
        else hipLaunchKernelGGL((k_window_direct<false, float>), dim3(blocks), dim3(kWin), bytes, st, SP3MM_WIN_ARGS);
    } else {
        if (num) hipLaunchKernelGGL((k_window_direct<true, double>), dim3(blocks), dim3(kWin), bytes, st, SP3MM_WIN_ARGS);
        else hipLaunchKernelGGL((k_window_direct<false, double>), dim3(blocks), dim3(kWin), bytes, st, SP3MM_WIN_ARGS);
    }
#undef SP3MM_WIN_ARGS
}

void launch_dense_direct(bool num, int blocks, const SpOperand &R, const SpOperand &A, const SpOperand &P,
                         const int32_t *rows, const int *count, const DenseWs &ws, int slog1, int slog2,
                         int32_t *n1out, int32_t *rownnz, const int64_t *crp, int32_t *ccol, double *cval,
                         int *err, hipStream_t st, bool f32in, double *rowsum, size_t max_lds)
{
    // LDS workspaces within max_lds bytes: stage 2 (P's column space: bitmap,
    // and values in the numeric pass) first, as it takes most of the
    // accumulations, then stage 1's bitmap over A's columns
    const size_t lds2 = (size_t)ws.nwP * 4 + (num ? (size_t)ws.nwP * 32 * 8 : 0);
    const size_t lds1 = (size_t)ws.nwA * 4;
    const bool l2 = lds2 <= max_lds;
    const bool l1 = (l2 ? lds2 : 0) + lds1 <= max_lds;
    const size_t lds = (l2 ? lds2 : 0) + (l1 ? lds1 : 0);
    if (num && !l2 && ws.cur) {
        // the second stage in LDS windows of P's columns (k_dense_window): the
        // widest power-of-two window whose values and bitmap fit, sharing the
        // region with the first stage's bitmap when that fits too
        int wcols = 1 << 20;
        while (wcols >= 1024 && (size_t)wcols * 8 + (size_t)wcols / 8 > max_lds) wcols >>= 1;
        if (wcols >= 1024) {
            const size_t wbytes = (size_t)wcols * 8 + (size_t)wcols / 8;
            const bool w1 = lds1 <= max_lds;
            const size_t wl = std::max(w1 ? lds1 : 0, wbytes);
#define SP3MM_WIN_LAUNCH(VTT)                                                                                   \
    do {                                                                                                         \
        if (w1)                                                                                                  \
            hipLaunchKernelGGL((k_dense_window<VTT, true>), dim3(blocks), dim3(kDense), wl, st, R, A, P, rows,   \
                               count, ws, wcols, slog1, crp, ccol, cval, rowsum, err);                           \
        else                                                                                                     \
            hipLaunchKernelGGL((k_dense_window<VTT, false>), dim3(blocks), dim3(kDense), wl, st, R, A, P, rows,  \
                               count, ws, wcols, slog1, crp, ccol, cval, rowsum, err);                           \
    } while (0)
            if (f32in) SP3MM_WIN_LAUNCH(float);
            else SP3MM_WIN_LAUNCH(double);
#undef SP3MM_WIN_LAUNCH
            return;
        }
    }
#define SP3MM_DENSE_ARGS R, A, P, rows, count, ws, slog1, slog2, n1out, rownnz, crp, ccol, cval, rowsum, err
#define SP3MM_DENSE_LAUNCH(NUMV, VTT)                                                                           \
    do {                                                                                                         \
        if (l1 && l2)                                                                                            \
            hipLaunchKernelGGL((k_dense_direct<NUMV, VTT, true, true>), dim3(blocks), dim3(kDense), lds, st,    \
                               SP3MM_DENSE_ARGS);                                                                \
        else if (l2)                                                                                             \
            hipLaunchKernelGGL((k_dense_direct<NUMV, VTT, false, true>), dim3(blocks), dim3(kDense), lds, st,   \
                               SP3MM_DENSE_ARGS);                                                                \
        else if (l1)                                                                                             \
            hipLaunchKernelGGL((k_dense_direct<NUMV, VTT, true, false>), dim3(blocks), dim3(kDense), lds, st,   \
                               SP3MM_DENSE_ARGS);                                                                \
        else                                                                                                     \
            hipLaunchKernelGGL((k_dense_direct<NUMV, VTT, false, false>), dim3(blocks), dim3(kDense), 0, st,    \
                               SP3MM_DENSE_ARGS);                                                                \
    } while (0)
    if (f32in) {
        if (num) SP3MM_DENSE_LAUNCH(true, float);
        else SP3MM_DENSE_LAUNCH(false, float);
    } else {
        if (num) SP3MM_DENSE_LAUNCH(true, double);
        else SP3MM_DENSE_LAUNCH(false, double);
    }
#undef SP3MM_DENSE_LAUNCH
#undef SP3MM_DENSE_ARGS
}

size_t fused_lds(int G, int H1, int H2)
{
    const size_t gpb = 256 / G;
    return gpb * ((size_t)H1 + H2 + (size_t)G) * 12 + gpb * group_waves(G) * 8 + gpb * (group_waves(G) + 4) * 4;
}

using FusedFn = void (*)(SpOperand, SpOperand, SpOperand, const int32_t *, const int *, int, int, int, int,
                        int32_t *, int32_t *, int32_t *, double *, int32_t *, int *, int *, int, int, unsigned,
                        unsigned);

// left entries held per lane: stage 1 U1 in {2, 4}, stage 2 U2 in {2, 3, 4}
template <int G, typename VT>
static FusedFn fused_variant(int u1, int u2)
{
    static const FusedFn tab[2][3] = {
        {k_direct_fused<G, 2, 2, VT>, k_direct_fused<G, 2, 3, VT>, k_direct_fused<G, 2, 4, VT>},
        {k_direct_fused<G, 4, 2, VT>, k_direct_fused<G, 4, 3, VT>, k_direct_fused<G, 4, 4, VT>},
    };
    return tab[u1 == 4 ? 1 : 0][u2 == 2 ? 0 : u2 == 3 ? 1 : 2];
}

void launch_direct_fused(int G, int H1, int H2, int slog1, int slog2, const SpOperand &R, const SpOperand &A,
                         const SpOperand &P, const int32_t *rows, const int *count, int rows_host,
                         int32_t *rownnz, int32_t *slotn, int32_t *scol, double *sval, int32_t *retry,
                         int *nretry, int *err, int max_blocks, hipStream_t st, bool f32in,
                         const Tuning &tu)
{
    const unsigned grid = grid_for(rows_host, G, max_blocks);
    const size_t lds = fused_lds(G, H1, H2);
    const int u1 = tu.u1, u2 = tu.u2;
    const int cap = ilog2((unsigned)G);
    slog1 = std::min(slog1, std::min(cap, std::max(0, tu.slog1_cap)));
    slog2 = std::min(slog2, std::min(cap, std::max(0, tu.slog2_cap)));
    const FusedFn fn = f32in ? (G == 32   ? fused_variant<32, float>(u1, u2)
                                : G == 64 ? fused_variant<64, float>(u1, u2)
                                          : fused_variant<256, float>(u1, u2))
                             : (G == 32   ? fused_variant<32, double>(u1, u2)
                                : G == 64 ? fused_variant<64, double>(u1, u2)
                                          : fused_variant<256, double>(u1, u2));
    const int nxcd = tu.xcd_map && grid >= (unsigned)kXcds ? kXcds : 1;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(256), lds, st, R, A, P, rows, count, H1, H2, slog1, slog2, rownnz,
                       slotn, scol, sval, retry, nretry, err, nxcd, rows_host, tu.hash1 | 1u, tu.hash2 | 1u);
}

size_t hash_score_lds(int H1, int H2)
{
    // dynamic: one pair of tables per wavefront; static: the two dedup tables and reductions
    return (size_t)Group<256>::kWaves * (H1 + H2) * sizeof(int) + 2 * kScoreDedup * sizeof(int) + 64;
}

void launch_hash_score(const SpOperand &R, const SpOperand &A, const SpOperand &P, const int32_t *rows, int nrows,
                       const unsigned *cand, int ncand, int H1, int H2, unsigned long long *score1,
                       unsigned long long *score2, int *nkeys, hipStream_t st)
{
    (void)hipMemsetAsync(score1, 0, sizeof(unsigned long long) * ncand, st);
    (void)hipMemsetAsync(score2, 0, sizeof(unsigned long long) * ncand, st);
    if (nrows <= 0) return;
    const size_t lds = (size_t)Group<256>::kWaves * (H1 + H2) * sizeof(int);
    hipLaunchKernelGGL(k_hash_score, dim3(nrows), dim3(256), lds, st, R, A, P, rows, cand, ncand, H1, H2, score1,
                       score2, nkeys);
}

void launch_gather_rows(const int32_t *rows, const int *count, int rows_host, const int32_t *slotn, int cap,
                        const int32_t *scol, const double *sval, const int64_t *crp, int32_t *ccol, double *cval,
                        int64_t out_cap, int *err, int max_blocks, hipStream_t st, const void *xsrc, void *xdst,
                        int xwords)
{
    const int L = cap <= 128 ? 16 : cap <= 4096 ? 32 : 256;
    unsigned grid = (unsigned)((rows_host + 256 / L - 1) / (256 / L));
    if (grid > (unsigned)max_blocks) grid = (unsigned)max_blocks;
    if (grid < 1) grid = 1;
    const auto *xs = static_cast<const uint32_t *>(xsrc);
    auto *xd = static_cast<uint32_t *>(xdst);
    if (L == 256)
        hipLaunchKernelGGL(k_gather_rows<256>, dim3(grid), dim3(256), 0, st, rows, count, slotn, cap, scol, sval,
                           crp, ccol, cval, out_cap, err, rows_host, xs, xd, xwords);
    else if (L == 16)
        hipLaunchKernelGGL(k_gather_rows<16>, dim3(grid), dim3(256), 0, st, rows, count, slotn, cap, scol, sval, crp,
                           ccol, cval, out_cap, err, rows_host, xs, xd, xwords);
    else
        hipLaunchKernelGGL(k_gather_rows<32>, dim3(grid), dim3(256), 0, st, rows, count, slotn, cap, scol, sval, crp,
                           ccol, cval, out_cap, err, rows_host, xs, xd, xwords);
}

// A replay computed the sizes its plan was built for: nnz, retried and
// spilled rows, and (counts != null) every bin's population, whose launches
// and scratch slots the plan fixed.  Anything else is reported.
__global__ void k_check_nnz(const int64_t *nnz, int64_t want, const int *nretry, int want_retry,
                            const int *nspill, int want_spill, const int *counts, BinCounts want_counts, int *err)
{
    if (threadIdx.x != 0) return;
    bool bad = *nnz != want || *nretry != want_retry || *nspill != want_spill;
    if (counts)
        for (int b = 0; b < want_counts.n; b++) bad = bad || counts[b] != want_counts.c[b];
    if (bad) atomicOr(err, kErrOverflow);
}

void launch_check_nnz(const int64_t *nnz, int64_t want, const int *nretry, int want_retry, const int *nspill,
                      int want_spill, int *err, hipStream_t st, const int *counts, const BinCounts *want_counts)
{
    BinCounts wc{};
    if (want_counts) wc = *want_counts;
    hipLaunchKernelGGL(k_check_nnz, dim3(1), dim3(64), 0, st, nnz, want, nretry, want_retry, nspill, want_spill,
                       want_counts ? counts : nullptr, wc, err);
}


// ---------------------------------------------------------------------------
// batches: block-diagonal packing of many operands, and extraction of one block
// ---------------------------------------------------------------------------
// blockIdx.y = problem; its rows get offset roff, its entries nzoff, its
// columns coloff.  The packed row offsets are written for rows [roff, roff+M).
template <typename VT>
__global__ void k_pack(const PackItem *__restrict__ items, int64_t *__restrict__ rp, int32_t *__restrict__ col,
                       VT *__restrict__ val)
{
    const PackItem it = items[blockIdx.y];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t r = t0; r < it.M; r += stride) rp[it.roff + r] = it.rp[r] + it.nzoff;
    const VT *v = static_cast<const VT *>(it.val);
    for (int64_t e = t0; e < it.nnz; e += stride) {
        col[it.nzoff + e] = it.col[e] + it.coloff;
        val[it.nzoff + e] = v[e];
    }
}

void launch_pack(const PackItem *items, int count, int vtype, int64_t *rp, int32_t *col, void *val, hipStream_t st)
{
    if (count <= 0) return;
    if (vtype == 1)
        hipLaunchKernelGGL(k_pack<float>, dim3(64, count), dim3(256), 0, st, items, rp, col, static_cast<float *>(val));
    else
        hipLaunchKernelGGL(k_pack<double>, dim3(64, count), dim3(256), 0, st, items, rp, col,
                           static_cast<double *>(val));
}

// rows [r0, r0+M) of a packed product, rebased to 0 and shifted back by coloff
__global__ void k_extract(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                          const double *__restrict__ val, int64_t r0, int64_t M, int32_t coloff,
                          int64_t *__restrict__ orp, int32_t *__restrict__ orp32, int32_t *__restrict__ ocol,
                          double *__restrict__ oval)
{
    const int64_t base = rp[r0];
    const int64_t nnz = rp[r0 + M] - base;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t r = t0; r <= M; r += stride) {
        const int64_t v = rp[r0 + r] - base;
        orp[r] = v;
        if (orp32) orp32[r] = (int32_t)v;
    }
    for (int64_t e = t0; e < nnz; e += stride) {
        ocol[e] = col[base + e] - coloff;
        oval[e] = val[base + e];
    }
}

void launch_extract(const int64_t *rp, const int32_t *col, const double *val, int64_t r0, int64_t M, int32_t coloff,
                    int64_t *orp, int32_t *orp32, int32_t *ocol, double *oval, hipStream_t st)
{
    hipLaunchKernelGGL(k_extract, dim3(256), dim3(256), 0, st, rp, col, val, r0, M, coloff, orp, orp32, ocol, oval);
}

void launch_scan(const int32_t *in, int64_t n, int64_t *out, int32_t *out32, long long *tile_sums,
                 int64_t *total, hipStream_t st)
{
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (n > 0)
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(kScanThreads), 0, st, in, n, out,
                           tile_sums);
    if (ntiles <= kScanShortTiles) {  // two kernels: the adding one sums the tile sums it needs itself
        hipLaunchKernelGGL(k_scan_add_short, dim3(blocks_for(n + 1, 256)), dim3(256), 0, st, out, n, tile_sums,
                           (int)ntiles, total, out32);
        return;
    }
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, tile_sums, ntiles, total);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n + 1, 256)), dim3(256), 0, st, out, n, tile_sums, total,
                       out32);
}

int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_rp_convert(const int64_t *in, int64_t M, int64_t nnz, int base, int64_t *rp, int32_t *rp32,
                       int *err, hipStream_t st)
{
    hipLaunchKernelGGL(k_rp_convert, dim3(blocks_for(M + 1, 256)), dim3(256), 0, st, in, M, nnz, base, rp,
                       rp32, err);
}

void launch_col_convert(const int64_t *in, const int64_t *rp, int64_t M, int64_t N, int base, int32_t *col,
                        int *err, hipStream_t st)
{
    if (M == 0) return;
    hipLaunchKernelGGL(k_col_convert, dim3(blocks_for(M, 256)), dim3(256), 0, st, in, rp, M, N, base, col,
                       err);
}

void launch_col_widen(const int32_t *col, int64_t nnz, int base, int64_t *out, hipStream_t st)
{
    if (nnz == 0) return;
    hipLaunchKernelGGL(k_col_widen, dim3(blocks_for(nnz, 256)), dim3(256), 0, st, col, nnz, base, out);
}

void launch_rp_shift(const int64_t *rp, int64_t n, int base, int64_t *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_rp_shift, dim3(blocks_for(n, 256)), dim3(256), 0, st, rp, n, base, out);
}

void launch_f64_to_f32(const double *in, int64_t n, float *out, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, st, in, n, out);
}

void launch_f32_to_f64(const float *in, int64_t n, double *out, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, st, in, n, out);
}

void launch_export_words(const void *src, void *host_dst, int words, hipStream_t st)
{
    hipLaunchKernelGGL(k_export_words, dim3(1), dim3(64), 0, st, static_cast<const uint32_t *>(src),
                       static_cast<uint32_t *>(host_dst), words);
}

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n, v);
}

void launch_copy_probe(const void *in, void *out, int64_t bytes, hipStream_t st)
{
    const int64_t n = bytes / 32;
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, st, (const double4 *)in, (double4 *)out, n);
}

}  // namespace sp3mm
