// mixed.hip -- mixed-precision direct triple product R*A*P (gfx950).
//
// FP32 storage and FP32 multiplies; every product is formed exactly as an
// FP32 pair with TwoProd (one FMA) and accumulated into its LDS slot as an
// FP32 pair (hi, lo): a returning FP32 atomic add on hi, TwoSum of the value
// it returned with the addend to recover that addition's rounding error, and
// a plain atomic add of the error (plus the addend's low part) on lo -- two
// atomics per product and no compare-and-swap loop.  Alongside every slot an
// FP32 sum of |terms| gives the magnitude scale (|R||A||P|)_ij of the entry.
//
// FP64 is used only where cancellation needs it:
//   * products landing on the diagonal (column == row) are also summed in an
//     FP64 LDS cell (ds_add_f64), which replaces the pair sum on output;
//   * the row sum of the coarse operator is accumulated in FP64 registers
//     (optional output);
//   * a row holding an off-diagonal entry with |value| < guard * scale is
//     queued for an FP64 recomputation (k_direct<G, true, float>) that
//     overwrites it after the gather.
// Structure, binning, scratch slots and gather are those of the FP64 one-pass
// kernel (k_direct_fused).
#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sp3mm {

#pragma clang fp contract(off)

struct FF {
    float h, l;
};

__device__ __forceinline__ FF two_sum(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ FF quick_two_sum(float a, float b)
{
    const float s = a + b;
    const float e = b - (s - a);
    return {s, e};
}

__device__ __forceinline__ FF two_prod(float a, float b)
{
    const float p = a * b;
    const float e = __fmaf_rn(a, b, -p);
    return {p, e};
}

// accurate pair addition (both halves summed with TwoSum, renormalised twice)
__device__ __forceinline__ FF ff_add(FF x, FF y)
{
    FF s = two_sum(x.h, y.h);
    const FF t = two_sum(x.l, y.l);
    s.l += t.h;
    s = quick_two_sum(s.h, s.l);
    s.l += t.l;
    return quick_two_sum(s.h, s.l);
}

// pair times float
__device__ __forceinline__ FF ff_mul(FF x, float b)
{
    FF p = two_prod(x.h, b);
    p.l = __fmaf_rn(x.l, b, p.l);
    return quick_two_sum(p.h, p.l);
}

// Adds the pair v to the slot (hi, lo) with two atomics and no retry loop:
// the returning FP32 atomic add on hi yields the value it was added to, so
// the rounding error of that addition is recovered exactly with TwoSum on the
// lane and goes, together with v.l, into lo with a plain atomic add.  The LDS
// adder rounds to nearest even like the lane's own addition, so the recovered
// error is the one the hardware made.
__device__ __forceinline__ void ff_atomic_accum(float *hi, float *lo, FF v)
{
    const float old = atomicAdd(hi, v.h);
    const FF s = two_sum(old, v.h);
    atomicAdd(lo, v.l + s.l);
}

// compaction of keys with three FP32 value arrays (pair hi, pair lo, scale),
// slot order kept
template <int G>
__device__ int compact3(int *keys, float *h, float *l, float *a, int H, int *red)
{
    const int t = Group<G>::tid();
    int base = 0;
    for (int s0 = 0; s0 < H; s0 += G) {
        const int s = s0 + t;
        const bool in = s < H;
        const int k = in ? keys[s] : kEmpty;
        const float hh = in ? h[s] : 0.f, ll = in ? l[s] : 0.f, aa = in ? a[s] : 0.f;
        const bool occ = k != kEmpty;
        const uint64_t m = Group<G>::part(ballot(occ));
        int pos = G >= kWave ? lanes_below(m) : __popcll(m & ((1ull << t) - 1));
        int tot = __popcll(m);
        if constexpr (G > kWave) {
            const int w = t / kWave;
            __syncthreads();
            if (lane() == 0) red[w] = tot;
            __syncthreads();
            int before = 0;
            tot = 0;
            for (int i = 0; i < Group<G>::kWaves; i++) {
                before += (i < w) ? red[i] : 0;
                tot += red[i];
            }
            pos += before;
        } else {
            wave_sync();
        }
        if (occ) {
            keys[base + pos] = k;
            h[base + pos] = hh;
            l[base + pos] = ll;
            a[base + pos] = aa;
        }
        base += tot;
        Group<G>::sync();
    }
    return base;
}

size_t mixed_lds(int G, int H1, int H2)
{
    const size_t gpb = 256 / G;
    return gpb * ((size_t)H1 + H2) * 16 + gpb * 8 + gpb * group_waves(G) * 8 + gpb * (group_waves(G) + 2) * 4;
}

// Expansion of one left row into an LDS table, as in the FP64 kernel (sub-groups
// of 2^slog lanes per left entry, U left entries per lane per pass, the first
// probe of every insert issued back to back); `loadl(e)` and `loadb(q)` fetch
// the left and right payloads and `acc(slot, left, right, col)` accumulates one
// product into an occupied slot.  Home slots come from the multiplier hmul.
template <int G, int U, typename LoadL, typename LoadB, typename Acc>
__device__ __forceinline__ void expand_mixed(const int32_t *lcol, int64_t e0, int64_t e1,
                                             const int32_t *__restrict__ brp, const int32_t *__restrict__ bcol,
                                             int slog, int *keys, unsigned mask, int shift, bool &full,
                                             int *fullflag, LoadL loadl, LoadB loadb, Acc acc, unsigned hmul)
{
    using LT = decltype(loadl(int64_t{}));
    using BT = decltype(loadb(int{}));
    constexpr int kLogG = G >= 256 ? 8 : G >= 128 ? 7 : G >= 64 ? 6 : G >= 32 ? 5 : 4;
    if (slog > kLogG) slog = kLogG;
    const int t = Group<G>::tid();
    const int S = 1 << slog;
    const int sub = t & (S - 1);
    const int ngrp = G >> slog;
    for (int64_t eb = e0 + (t >> slog); eb < e1; eb += (int64_t)ngrp * U) {
        int q[U], qe[U];
        LT lv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t e = eb + (int64_t)u * ngrp;
            const bool ok = e < e1;
            const int k = ok ? lcol[e] : 0;
            lv[u] = loadl(ok ? e : e0);
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
            BT w[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = q[u] < qe[u];
                c[u] = ok ? bcol[q[u]] : kEmpty;
                w[u] = loadb(ok ? q[u] : 0);
                q[u] += S;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                sl[u] = hslot(c[u], shift, hmul) & mask;
                old[u] = c[u] != kEmpty ? atomicCAS(keys + sl[u], kEmpty, c[u]) : kEmpty;
            }
            __builtin_amdgcn_sched_barrier(0);  // the U swaps stay in flight together
            bool pend = false;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool hit = c[u] != kEmpty && (old[u] == kEmpty || old[u] == c[u]);
                pend |= c[u] != kEmpty && !hit;
                if (hit) acc(sl[u], lv[u], w[u], c[u]);
            }
            if (pend) {
                for (int u = 0; u < U; u++) {
                    if (c[u] == kEmpty || old[u] == kEmpty || old[u] == c[u]) continue;
                    if (__hip_atomic_load(fullflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                        full = true;  // the row is redone elsewhere
                        continue;
                    }
                    unsigned s2 = sl[u];
                    bool done = false;
                    for (unsigned probe = 1; probe <= mask; probe++) {
                        s2 = (s2 + 1) & mask;
                        const int o = atomicCAS(keys + s2, kEmpty, c[u]);
                        if (o == kEmpty || o == c[u]) {
                            acc(s2, lv[u], w[u], c[u]);
                            done = true;
                            break;
                        }
                    }
                    if (!done) {
                        full = true;
                        __hip_atomic_store(fullflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
        }
    }
}

// stage-2 left payload: one entry of row r of R*A
struct Mid {
    FF x;       // value as a normalised FP32 pair
    float xa;   // magnitude scale (|R||A|)_rj
    double xd;  // value in FP64 (for the diagonal and the row sum)
};

template <int G>
__global__ void __launch_bounds__(256) k_direct_mixed(SpOperand R, SpOperand A, SpOperand P,
                                                       const int32_t *__restrict__ rows,
                                                       const int *__restrict__ count, int H1, int H2,
                                                       int slog1, int slog2, float guard,
                                                       int32_t *__restrict__ rownnz, int32_t *__restrict__ slotn,
                                                       int32_t *__restrict__ scol, double *__restrict__ sval,
                                                       double *__restrict__ rowsum,
                                                       int32_t *__restrict__ retry, int *__restrict__ nretry,
                                                       int32_t *__restrict__ fix, int *__restrict__ nfix,
                                                       int *__restrict__ err, int nxcd, int cap_rows,
                                                       unsigned hm1, unsigned hm2)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    constexpr int KW = Group<G>::kWaves;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    unsigned char *p = smem;
    double *dg = reinterpret_cast<double *>(p) + g;
    p += (size_t)GPB * 8;
    long long *red = reinterpret_cast<long long *>(p) + g * KW;
    p += (size_t)GPB * KW * 8;
    float *h1 = reinterpret_cast<float *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    float *l1 = reinterpret_cast<float *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    float *a1 = reinterpret_cast<float *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    int *k1 = reinterpret_cast<int *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    float *h2 = reinterpret_cast<float *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    float *l2 = reinterpret_cast<float *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    float *a2 = reinterpret_cast<float *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    int *k2 = reinterpret_cast<int *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    int *ired = reinterpret_cast<int *>(p) + g * (KW + 2);
    int *fullflag = ired + KW + 1;  // a table of this row filled

    const unsigned m1 = (unsigned)H1 - 1, m2 = (unsigned)H2 - 1;
    const int sh1 = 32 - ilog2((unsigned)H1), sh2 = 32 - ilog2((unsigned)H2);
    // the scratch slots hold cap_rows rows (a larger count is reported)
    const int n = clamp_count(count, cap_rows, err);
    const float *rv = static_cast<const float *>(R.val);
    const float *av = static_cast<const float *>(A.val);
    const float *pv = static_cast<const float *>(P.val);
    // the next row's index and bounds are loaded while this row is processed
    const RowWalk w = row_walk<GPB>(n, g, nxcd);
    const int stride = w.stride, end = w.end;
    int i = w.i;
    int r = i < end ? rows[i] : 0;
    int64_t rb = 0, re = 0;
    if (i < end) {
        rb = R.rp[r];
        re = R.rp[r + 1];
    }
    for (; i < end; i += stride) {
        const int rn = i + stride < end ? rows[i + stride] : 0;
        for (int s = t; s < H1; s += G) {
            k1[s] = kEmpty;
            h1[s] = 0.f;
            l1[s] = 0.f;
            a1[s] = 0.f;
        }
        for (int s = t; s < H2; s += G) {
            k2[s] = kEmpty;
            h2[s] = 0.f;
            l2[s] = 0.f;
            a2[s] = 0.f;
        }
        if (t == 0) {
            *dg = 0.0;
            *fullflag = 0;
        }
        Group<G>::sync();
        bool full1 = false, full2 = false;
        // stage 1: row r of R*A; every FP32 product is split exactly by TwoProd
        expand_mixed<G, 4>(
            R.col, rb, re, A.rp32, A.col, slog1, k1, m1, sh1, full1, fullflag,
            [&](int64_t e) { return rv[e]; }, [&](int q) { return av[q]; },
            [&](unsigned sl, float x, float w, int) {
                const FF pr = two_prod(x, w);
                ff_atomic_accum(h1 + sl, l1 + sl, pr);
                atomicAdd(a1 + sl, fabsf(pr.h));
            },
            hm1);
        const int64_t rbn = R.rp[rn], ren = R.rp[rn + 1];
        Group<G>::sync();
        const int nn1 = compact3<G>(k1, h1, l1, a1, H1, ired);
        // stage 2: forward into P; the diagonal and the row sum also in FP64
        double rs = 0.0;
        expand_mixed<G, 2>(
            k1, 0, (full1 || __hip_atomic_load(fullflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) ? 0 : nn1,
            P.rp32, P.col, slog2, k2, m2, sh2, full2, fullflag,
            [&](int64_t e) {
                Mid m;
                m.x = two_sum(h1[e], l1[e]);
                m.xa = a1[e];
                m.xd = (double)m.x.h + (double)m.x.l;
                return m;
            },
            [&](int q) { return pv[q]; },
            [&](unsigned sl, const Mid &m, float w, int c) {
                ff_atomic_accum(h2 + sl, l2 + sl, ff_mul(m.x, w));
                atomicAdd(a2 + sl, m.xa * fabsf(w));
                const double pd = m.xd * (double)w;
                rs += pd;
                if (c == r) atomicAdd(dg, pd);
            },
            hm2);
        const long long anyfull = group_max<G>((full1 || full2) ? 1 : 0, red);
        if (anyfull) {
            if (t == 0) {
                slotn[i] = -1;
                rownnz[r] = 0;
                retry[atomicAdd(nretry, 1)] = r;
            }
        } else {
            Group<G>::sync();
            const int nn2 = compact3<G>(k2, h2, l2, a2, H2, ired);
            const int64_t off = (int64_t)i * H2;
            bool cancel = false;
            for (int j = t; j < nn2; j += G) {
                const int k = k2[j];
                const FF x = two_sum(h2[j], l2[j]);
                double val = (double)x.h + (double)x.l;
                if (k == r) val = *dg;
                else cancel |= fabs(val) < (double)guard * (double)a2[j];
                const int rk = key_rank(k2, nn2, k);
                scol[off + rk] = k;
                sval[off + rk] = val;
            }
            const long long anycancel = group_max<G>(cancel ? 1 : 0, red);
            if (t == 0) {
                slotn[i] = nn2;
                rownnz[r] = nn2;
                if (anycancel) fix[atomicAdd(nfix, 1)] = r;
            }
        }
        if (rowsum && !anyfull) {
            double tot = rs;
            constexpr int top = G < kWave ? G / 2 : 32;
            for (int o = top; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
            if constexpr (G > kWave) {
                double *dred = reinterpret_cast<double *>(red);
                __syncthreads();
                if (lane() == 0) dred[t / kWave] = tot;
                __syncthreads();
                tot = 0.0;
                for (int w = 0; w < KW; w++) tot += dred[w];
            }
            if (t == 0) rowsum[r] = tot;
        }
        Group<G>::sync();
        r = rn;
        rb = rbn;
        re = ren;
    }
}

static unsigned grid_for_m(int rows, int G, int max_blocks)
{
    const int gpb = 256 / G;
    int64_t b = (rows + gpb - 1) / gpb;
    if (b > max_blocks) b = max_blocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

void launch_direct_mixed(int G, int H1, int H2, int slog1, int slog2, float guard, const SpOperand &R,
                         const SpOperand &A, const SpOperand &P, const int32_t *rows, const int *count,
                         int rows_host, int32_t *rownnz, int32_t *slotn, int32_t *scol, double *sval,
                         double *rowsum, int32_t *retry, int *nretry, int32_t *fix, int *nfix, int *err,
                         int max_blocks, hipStream_t st, const Tuning &tu)
{
    const unsigned grid = grid_for_m(rows_host, G, max_blocks);
    const size_t lds = mixed_lds(G, H1, H2);
    const int cap = ilog2((unsigned)G);
    slog1 = std::min(slog1, std::min(cap, std::max(0, tu.slog1_cap)));
    slog2 = std::min(slog2, std::min(cap, std::max(0, tu.slog2_cap)));
    auto fn = G == 32 ? k_direct_mixed<32> : G == 64 ? k_direct_mixed<64> : k_direct_mixed<256>;
    const int nxcd = tu.xcd_map && grid >= (unsigned)kXcds ? kXcds : 1;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(256), lds, st, R, A, P, rows, count, H1, H2, slog1, slog2, guard,
                       rownnz, slotn, scol, sval, rowsum, retry, nretry, fix, nfix, err, nxcd, rows_host,
                       tu.hash1 | 1u, tu.hash2 | 1u);
}

}  // namespace sp3mm
