// mixed.hip -- mixed-precision direct triple product R*A*P (gfx950).
//
// FP32 storage and FP32 multiplies; every product is formed exactly as an
// FP32 pair with TwoProd (one FMA) and accumulated into its LDS slot as an
// FP32 pair (hi, lo) with an error-free TwoSum-based addition, applied with a
// 64-bit compare-and-swap on the packed pair.  Alongside every slot an FP32
// sum of |terms| gives the magnitude scale (|R||A||P|)_ij of the entry.
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

__device__ __forceinline__ unsigned long long ff_pack(FF v)
{
    return (unsigned long long)__float_as_uint(v.h) | ((unsigned long long)__float_as_uint(v.l) << 32);
}

__device__ __forceinline__ FF ff_unpack(unsigned long long u)
{
    return {__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32))};
}

__device__ __forceinline__ void ff_atomic_add(unsigned long long *slot, FF v)
{
    unsigned long long old = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    unsigned long long assumed;
    do {
        assumed = old;
        old = atomicCAS(slot, assumed, ff_pack(ff_add(ff_unpack(assumed), v)));
    } while (old != assumed);
}

// compaction of keys with two value arrays (pair and scale), slot order kept
template <int G>
__device__ int compact2(int *keys, unsigned long long *v, float *a, int H, int *red)
{
    const int t = Group<G>::tid();
    int base = 0;
    for (int s0 = 0; s0 < H; s0 += G) {
        const int s = s0 + t;
        const int k = keys[s];
        const unsigned long long vv = v[s];
        const float aa = a[s];
        const bool occ = k != kEmpty;
        const uint64_t m = ballot(occ);
        int pos = lanes_below(m);
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
            v[base + pos] = vv;
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
    return gpb * ((size_t)H1 + H2) * 16 + gpb * 8 + gpb * (G / 64) * 8 + gpb * (G / 64 + 1) * 4;
}

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
                                                       int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int GPB = 256 / G;
    const int g = threadIdx.x / G;
    const int t = Group<G>::tid();
    unsigned char *p = smem;
    unsigned long long *v1 = reinterpret_cast<unsigned long long *>(p) + g * H1;
    p += (size_t)GPB * H1 * 8;
    unsigned long long *v2 = reinterpret_cast<unsigned long long *>(p) + g * H2;
    p += (size_t)GPB * H2 * 8;
    double *dg = reinterpret_cast<double *>(p) + g;
    p += (size_t)GPB * 8;
    long long *red = reinterpret_cast<long long *>(p) + g * Group<G>::kWaves;
    p += (size_t)GPB * Group<G>::kWaves * 8;
    float *a1 = reinterpret_cast<float *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    float *a2 = reinterpret_cast<float *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    int *k1 = reinterpret_cast<int *>(p) + g * H1;
    p += (size_t)GPB * H1 * 4;
    int *k2 = reinterpret_cast<int *>(p) + g * H2;
    p += (size_t)GPB * H2 * 4;
    int *ired = reinterpret_cast<int *>(p) + g * (Group<G>::kWaves + 1);

    const unsigned m1 = (unsigned)H1 - 1, m2 = (unsigned)H2 - 1;
    const int sh1 = 32 - ilog2((unsigned)H1), sh2 = 32 - ilog2((unsigned)H2);
    const int S1 = 1 << slog1, S2 = 1 << slog2;
    const int n = *count;
    const float *rv = static_cast<const float *>(R.val);
    const float *av = static_cast<const float *>(A.val);
    const float *pv = static_cast<const float *>(P.val);
    for (int i = blockIdx.x * GPB + g; i < n; i += gridDim.x * GPB) {
        const int r = rows[i];
        for (int s = t; s < H1; s += G) {
            k1[s] = kEmpty;
            v1[s] = 0ull;
            a1[s] = 0.f;
        }
        for (int s = t; s < H2; s += G) {
            k2[s] = kEmpty;
            v2[s] = 0ull;
            a2[s] = 0.f;
        }
        if (t == 0) *dg = 0.0;
        Group<G>::sync();
        bool full1 = false, full2 = false;
        // stage 1: row r of R*A as FP32 pairs
        for (int64_t e = R.rp[r] + (t >> slog1); e < R.rp[r + 1]; e += (G >> slog1)) {
            const int k = R.col[e];
            const float x = rv[e];
            const int b1 = A.rp32[k + 1];
            for (int q0 = A.rp32[k] + (t & (S1 - 1)); q0 < b1; q0 += S1 * 8) {
                int c[8];
                float w[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int q = q0 + u * S1;
                    c[u] = q < b1 ? A.col[q] : kEmpty;
                    w[u] = q < b1 ? av[q] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    if (c[u] == kEmpty) continue;
                    int fresh;
                    const int sl = hash_insert(k1, m1, sh1, c[u], &fresh);
                    if (sl < 0) { full1 = true; continue; }
                    ff_atomic_add(v1 + sl, two_prod(x, w[u]));
                    atomicAdd(a1 + sl, fabsf(x) * fabsf(w[u]));
                }
            }
        }
        Group<G>::sync();
        const int nn1 = compact2<G>(k1, v1, a1, H1, ired);
        // stage 2: forward into P; diagonal and row sum also in FP64
        double rs = 0.0;
        for (int e = t >> slog2; e < (full1 ? 0 : nn1); e += (G >> slog2)) {
            const int j = k1[e];
            const FF x = ff_unpack(v1[e]);
            const float xa = a1[e];
            const double xd = (double)x.h + (double)x.l;
            const int b1 = P.rp32[j + 1];
            for (int q0 = P.rp32[j] + (t & (S2 - 1)); q0 < b1; q0 += S2 * 8) {
                int c[8];
                float w[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int q = q0 + u * S2;
                    c[u] = q < b1 ? P.col[q] : kEmpty;
                    w[u] = q < b1 ? pv[q] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    if (c[u] == kEmpty) continue;
                    const double pd = xd * (double)w[u];
                    rs += pd;
                    int fresh;
                    const int sl = hash_insert(k2, m2, sh2, c[u], &fresh);
                    if (sl < 0) { full2 = true; continue; }
                    ff_atomic_add(v2 + sl, ff_mul(x, w[u]));
                    atomicAdd(a2 + sl, xa * fabsf(w[u]));
                    if (c[u] == r) atomicAdd(dg, pd);
                }
            }
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
            const int nn2 = compact2<G>(k2, v2, a2, H2, ired);
            const int64_t off = (int64_t)i * H2;
            bool cancel = false;
            for (int j = t; j < nn2; j += G) {
                const int k = k2[j];
                const FF x = ff_unpack(v2[j]);
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
            for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
            if constexpr (G > kWave) {
                double *dred = reinterpret_cast<double *>(red);
                __syncthreads();
                if (lane() == 0) dred[t / kWave] = tot;
                __syncthreads();
                tot = 0.0;
                for (int w = 0; w < Group<G>::kWaves; w++) tot += dred[w];
            }
            if (t == 0) rowsum[r] = tot;
        }
        Group<G>::sync();
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
                         int max_blocks, hipStream_t st)
{
    const unsigned grid = grid_for_m(rows_host, G, max_blocks);
    const size_t lds = mixed_lds(G, H1, H2);
    if (G == 64)
        hipLaunchKernelGGL(k_direct_mixed<64>, dim3(grid), dim3(256), lds, st, R, A, P, rows, count, H1, H2, slog1,
                           slog2, guard, rownnz, slotn, scol, sval, rowsum, retry, nretry, fix, nfix, err);
    else
        hipLaunchKernelGGL(k_direct_mixed<256>, dim3(grid), dim3(256), lds, st, R, A, P, rows, count, H1, H2,
                           slog1, slog2, guard, rownnz, slotn, scol, sval, rowsum, retry, nretry, fix, nfix, err);
}

}  // namespace sp3mm
