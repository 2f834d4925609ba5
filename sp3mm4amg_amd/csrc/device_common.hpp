// device_common.hpp -- wavefront/workgroup helpers and LDS hash tables for the
// gfx950 Sp3MM kernels.  CDNA wavefronts are 64 lanes; a "group" of G lanes
// works on one output row: a power-of-two part of a wavefront (G = 16 or 32,
// several rows per wavefront advancing in lockstep), one wavefront (G == 64)
// or a whole workgroup (G > 64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace sp3mm {

constexpr int kWave = 64;
// MI355X (gfx950) has 8 XCDs, each with its own L2; the workgroups of a
// launch are dispatched to them round-robin (workgroup b runs on XCD b % 8)
constexpr int kXcds = 8;
constexpr int kEmpty = -1;

// Error bits raised by kernels instead of hanging or writing out of bounds.
enum : int {
    kErrHashFull = 1,     // an LDS hash table ran out of slots (sizing bug)
    kErrOverflow = 2,     // an output row did not fit the space reserved for it
    kErrBadInput = 4,     // input validation failed
    kErrSpin = 8,         // a bounded wait gave up
};

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

// A row count read from device memory, bounded by the count the launch (and
// the buffers it indexes) was sized for; a larger count raises kErrOverflow
// once per launch and the extra rows are not processed.
__device__ __forceinline__ int clamp_count(const int *count, int cap, int *err)
{
    const int n = *count;
    if (n > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrOverflow);
        return cap;
    }
    return n;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of `mask` strictly below this lane
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G>
struct Group {
    static_assert((G >= 16 && G <= kWave && (G & (G - 1)) == 0) || G % kWave == 0,
                  "group must be a power-of-two part of a wavefront or whole wavefronts");
    // wavefronts per group (1 for sub-wavefront groups)
    static constexpr int kWaves = G >= kWave ? G / kWave : 1;
    __device__ static int tid() { return (int)(threadIdx.x % G); }
    // first lane of this group inside its wavefront
    __device__ static int lane0() { return G >= kWave ? 0 : (lane() & ~(G - 1)); }
    // this group's bits of a wavefront-wide ballot, shifted down to bit 0
    __device__ static uint64_t part(uint64_t m)
    {
        if constexpr (G >= kWave) return m;
        else return (m >> lane0()) & ((1ull << G) - 1);
    }
    __device__ static void sync()
    {
        if constexpr (G <= kWave) wave_sync();
        else __syncthreads();
    }
};

// wavefronts spanned by a group of G lanes (host and device)
__host__ __device__ constexpr int group_waves(int G) { return G >= kWave ? G / kWave : 1; }

// sum of an int over the group; `red` is kWaves ints of LDS private to the group
template <int G>
__device__ __forceinline__ long long group_sum(long long v, long long *red)
{
    constexpr int top = G < kWave ? G / 2 : 32;
    for (int o = top; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if constexpr (G <= kWave) {
        return v;
    } else {
        const int w = Group<G>::tid() / kWave;
        __syncthreads();
        if (lane() == 0) red[w] = v;
        __syncthreads();
        long long t = 0;
        for (int i = 0; i < Group<G>::kWaves; i++) t += red[i];
        __syncthreads();
        return t;
    }
}

template <int G>
__device__ __forceinline__ long long group_max(long long v, long long *red)
{
    constexpr int top = G < kWave ? G / 2 : 32;
    for (int o = top; o > 0; o >>= 1) {
        long long u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    if constexpr (G <= kWave) {
        return v;
    } else {
        const int w = Group<G>::tid() / kWave;
        __syncthreads();
        if (lane() == 0) red[w] = v;
        __syncthreads();
        long long t = red[0];
        for (int i = 1; i < Group<G>::kWaves; i++) t = red[i] > t ? red[i] : t;
        __syncthreads();
        return t;
    }
}

// ---------------------------------------------------------------------------
// LDS hash table with linear probing.  Keys are column indices (>= 0), the
// empty slot is -1.  `mask` = H-1 for a power-of-two table of H slots.  A probe
// sequence is bounded by H: a full table raises kErrHashFull instead of
// spinning forever.  A key's home slot is the top bits of key * mul (`mul`
// odd); the direct kernel's two tables take their multipliers from the
// schedule, which the autotuner fits to the operands (k_hash_score).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned hslot(int key, int shift, unsigned mul = kHashMul)
{
    return ((unsigned)key * mul) >> shift;
}

// returns slot index, or -1 when the table is full; *fresh = 1 for a new key.
// One compare-and-swap per probe: an empty slot is claimed, a slot already
// holding the key is found by the failed swap's return value.
__device__ __forceinline__ int hash_insert(int *keys, unsigned mask, int shift, int key,
                                           int *fresh)
{
    unsigned s = hslot(key, shift) & mask;
    for (unsigned probe = 0; probe <= mask; probe++) {
        const int old = atomicCAS(keys + s, kEmpty, key);
        if (old == kEmpty) { *fresh = 1; return (int)s; }
        if (old == key) { *fresh = 0; return (int)s; }
        s = (s + 1) & mask;
    }
    *fresh = 0;
    return -1;
}

// log2 of a power of two
__host__ __device__ __forceinline__ int ilog2(unsigned v)
{
    int r = 0;
    while ((1u << r) < v) r++;
    return r;
}

// In-place compaction of the occupied slots of a table of H slots to its
// front, preserving slot order.  Values (if any) move with keys.  Returns the
// number of occupied slots.  `red` holds kWaves+1 ints.
template <int G, typename V>
__device__ int compact_table(int *keys, V *vals, int H, int *red)
{
    const int t = Group<G>::tid();
    int base = 0;
    for (int s0 = 0; s0 < H; s0 += G) {
        const int s = s0 + t;
        const int k = s < H ? keys[s] : kEmpty;
        V v{};
        if (vals && s < H) v = vals[s];
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
            if (vals) vals[base + pos] = v;
        }
        base += tot;
        Group<G>::sync();
    }
    return base;
}

// Rank of every compacted key among the first n keys (keys are distinct).
__device__ __forceinline__ int key_rank(const int *keys, int n, int k)
{
    int r = 0;
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        r += (keys[j] < k) + (keys[j + 1] < k) + (keys[j + 2] < k) + (keys[j + 3] < k);
    }
    for (; j < n; j++) r += keys[j] < k;
    return r;
}

// Sorts keys[0..N) ascending, moving vals along, by a bitonic network in LDS:
// N is a power of two, the keys are distinct column indices padded with
// INT32_MAX, the G lanes of a group share the compare-exchanges of a stage and
// synchronise between stages.  (log2 N)(log2 N + 1)/2 stages of N/2
// compare-exchanges, against n^2 comparisons for ranking every key.
template <int G, typename V>
__device__ void bitonic_sort(int *keys, V *vals, int N, int t)
{
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = t; i < N / 2; i += G) {
                const int a = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                const int b = a + stride;
                const bool up = (a & size) == 0;
                const int ka = keys[a], kb = keys[b];
                if ((ka > kb) == up) {
                    keys[a] = kb;
                    keys[b] = ka;
                    const V va = vals[a];
                    vals[a] = vals[b];
                    vals[b] = va;
                }
            }
            Group<G>::sync();
        }
    }
}

// exclusive scan of one value per thread over the workgroup; `sh` holds one
// long long per wavefront; *total receives the workgroup sum
__device__ __forceinline__ long long block_exclusive_scan(long long v, long long *sh, long long &total)
{
    // wave inclusive scan
    const int l = lane();
    long long x = v;
    for (int o = 1; o < 64; o <<= 1) {
        long long y = __shfl_up(x, o);
        if (l >= o) x += y;
    }
    const int w = threadIdx.x / 64;
    if (l == 63) sh[w] = x;
    __syncthreads();
    long long before = 0;
    total = 0;
    for (int i = 0; i < (int)(blockDim.x / 64); i++) {
        before += (i < w) ? sh[i] : 0;
        total += sh[i];
    }
    __syncthreads();
    return before + x - v;
}

// Walk of a list of n rows by the groups of a grid (GPB groups per
// workgroup): group g of workgroup b starts at list index i and steps by
// `stride` while i < end.  nxcd == 1: workgroup b takes rows b*GPB.. and the
// grid strides over the whole list.  nxcd > 1 (needs gridDim.x >= nxcd): the
// list is cut into nxcd contiguous segments and the workgroups of XCD
// x = b % nxcd stride through segment x only, so the rows one L2 serves at a
// time are neighbours in the list (and in a mesh-ordered matrix share most of
// the rows of the operands they read).  Every index in [0, n) is visited once.
struct RowWalk {
    int i, stride, end;
};
template <int GPB>
__device__ __forceinline__ RowWalk row_walk(int n, int g, int nxcd)
{
    RowWalk w;
    if (nxcd > 1) {
        const int x = blockIdx.x % nxcd, j = blockIdx.x / nxcd;
        const int nper = ((int)gridDim.x - x + nxcd - 1) / nxcd;
        const int seg = (n + nxcd - 1) / nxcd;
        const int s0 = min(n, x * seg);
        w.end = min(n, s0 + seg);
        w.i = s0 + j * GPB + g;
        w.stride = nper * GPB;
    } else {
        w.i = blockIdx.x * GPB + g;
        w.stride = gridDim.x * GPB;
        w.end = n;
    }
    return w;
}

}  // namespace sp3mm
