// expand.hpp -- expansion of one sparse row product into an LDS hash table,
// shared by the SpGEMM, one-pass direct and mixed-precision kernels.
#pragma once
#include "device_common.hpp"

namespace sp3mm {

// Expands left entries [e0, e1) against the right operand B: for every left
// entry e with column k = loadk(e) and every entry q of row k of B, the
// product of loadl(e) and loadb(q) lands in the slot of key bcol[q].
//
// Lanes form sub-groups of S = 2^slog lanes; a sub-group owns one left entry
// at a time and its lanes walk that entry's right row with stride S (S is
// chosen near the typical right-row length, so one step covers most rows).
// Every lane holds U left entries at once (entries g, g + ngrp, ... of the
// pass), loads all their row bounds, then loads one right entry per left entry
// and inserts the U products: a pass over U*ngrp left entries costs three
// dependent memory round trips.  The first probe of every insert is one
// compare-and-swap issued back to back with the others; only an insert whose
// first slot holds another key enters the probing loop, so the common path has
// no divergent loops.  A probe sequence is bounded by the table size: when the
// table is full the product is dropped and `full` is set (the caller redoes
// the row elsewhere).
//
// acc(hit, slot, left, right, col, fresh) is called by every lane of the group
// for each of its U products of a step (hit false: no product, or not inserted
// yet) and again from the probing loop; with hit set it accumulates the
// product into the occupied slot, fresh marking the insert that claimed it.
template <int G, int U, typename LoadK, typename LoadL, typename LoadB, typename Acc>
__device__ __forceinline__ void expand_gen(int64_t e0, int64_t e1, LoadK loadk, LoadL loadl,
                                           const int32_t *__restrict__ brp, const int32_t *__restrict__ bcol,
                                           LoadB loadb, int slog, int *keys, unsigned mask, int shift, bool &full,
                                           Acc acc)
{
    using LT = decltype(loadl(int64_t{}));
    using BT = decltype(loadb(int{}));
    constexpr int kLogG = G >= 256 ? 8 : G >= 128 ? 7 : G >= 64 ? 6 : G >= 32 ? 5 : 4;
    if (slog > kLogG) slog = kLogG;  // a sub-group never spans more than the group
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
            const int k = ok ? loadk(e) : 0;
            lv[u] = ok ? loadl(e) : LT{};
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
                w[u] = ok ? loadb(q[u]) : BT{};
                q[u] += S;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                sl[u] = hslot(c[u], shift) & mask;
                old[u] = c[u] != kEmpty ? atomicCAS(keys + sl[u], kEmpty, c[u]) : kEmpty;
            }
            __builtin_amdgcn_sched_barrier(0);  // the U swaps stay in flight together
            bool pend = false;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool hit = c[u] != kEmpty && (old[u] == kEmpty || old[u] == c[u]);
                pend |= c[u] != kEmpty && !hit;
                acc(hit, sl[u], lv[u], w[u], c[u], hit && old[u] == kEmpty);
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
                            acc(true, s2, lv[u], w[u], c[u], o == kEmpty);
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

// Branch-free form of the FP64 expansion for the one-pass direct kernel.  Same
// lane layout as expand_gen; on the common path every lane executes the same
// instructions: loads read a valid index and their result is selected, and a
// product that is absent (or not yet inserted) compare-and-swaps and adds into
// the lane's own dummy slot just past the table (keys[H + lane] stays empty:
// the swap writes kEmpty over kEmpty), so no exec-mask branches are needed.
// Only inserts whose first slot holds another key take the probing branch.
// `keys`/`vals` hold H + G slots; `dslot` = H + the lane's index in the group.
template <int G, typename LV, typename BV, int U>
__device__ __forceinline__ void expand_flat(const int32_t *lcol, const LV *lval, int64_t e0, int64_t e1,
                                            const int32_t *__restrict__ brp, const int32_t *__restrict__ bcol,
                                            const BV *__restrict__ bval, int slog, int *keys, double *vals,
                                            unsigned mask, int shift, unsigned dslot, bool &full, int *fullflag,
                                            unsigned hmul = kHashMul)
{
    constexpr int kLogG = G >= 256 ? 8 : G >= 128 ? 7 : G >= 64 ? 6 : G >= 32 ? 5 : 4;
    if (slog > kLogG) slog = kLogG;
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
            const int64_t es = ok ? e : eb;  // eb < e1: a valid entry
            const int k = lcol[es];
            const double x = (double)lval[es];
            const int b0 = brp[k], b1 = brp[k + 1];
            lv[u] = ok ? x : 0.0;
            q[u] = ok ? b0 + sub : 0;
            qe[u] = ok ? b1 : 0;
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
                // some valid entry of B (the loop runs only while B has entries)
                const unsigned qs = (unsigned)max(min(q[u], qe[u] - 1), 0);
                const int cc = bcol[qs];
                const double ww = (double)bval[qs];
                c[u] = ok ? cc : kEmpty;
                w[u] = ww;
                q[u] += S;
            }
#pragma unroll
            for (int u = 0; u < U; u++) sl[u] = c[u] != kEmpty ? (hslot(c[u], shift, hmul) & mask) : dslot;
#pragma unroll
            for (int u = 0; u < U; u++) old[u] = atomicCAS(keys + sl[u], kEmpty, c[u]);
            // keep the U compare-and-swaps in flight together: nothing that
            // consumes their results is scheduled above this point
            __builtin_amdgcn_sched_barrier(0);
            // An absent product swapped its dummy slot (old == kEmpty), so it
            // counts as a hit on sl == dslot and needs no separate test.
            bool pend = false;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool hit = old[u] == kEmpty || old[u] == c[u];
                pend |= !hit;
                atomicAdd(vals + (hit ? sl[u] : dslot), lv[u] * w[u]);
            }
            if (pend) {
                // Linear probing of the inserts whose first slot held another
                // key: every pending insert of the lane advances one slot per
                // iteration (all of them in one loop, so the group iterates as
                // often as the longest probe sequence, not once per insert);
                // inserts no longer pending swap and add on the lane's dummy
                // slot, so an iteration has no per-insert branch.  A probe
                // sequence is bounded by the table size, and once any lane of
                // the group found the table full the row is redone elsewhere.
                bool pd[U];
                unsigned s2[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    pd[u] = !(old[u] == kEmpty || old[u] == c[u]);
                    s2[u] = sl[u];
                }
                for (unsigned probe = 1;; probe++) {
                    bool anyp = false;
#pragma unroll
                    for (int u = 0; u < U; u++) anyp |= pd[u];
                    if (!anyp) break;
                    if (probe > mask ||
                        __hip_atomic_load(fullflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                        full = true;
                        __hip_atomic_store(fullflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        break;
                    }
                    int o[U];
                    unsigned a[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        s2[u] = (s2[u] + 1) & mask;
                        a[u] = pd[u] ? s2[u] : dslot;
                        o[u] = atomicCAS(keys + a[u], kEmpty, pd[u] ? c[u] : kEmpty);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const bool hit = pd[u] && (o[u] == kEmpty || o[u] == c[u]);
                        atomicAdd(vals + (hit ? a[u] : dslot), lv[u] * w[u]);
                        pd[u] = pd[u] && !hit;
                    }
                }
            }
        }
    }
}

// probes per insert that expand_flat_crowded takes for all of a lane's
// pending inserts together
constexpr unsigned kMergedProbes = 4;

// expand_flat for the large tables of groups of several wavefronts, where
// long probe sequences are common: the pending inserts of a lane advance
// together for kMergedProbes slots only, then one at a time (the settled ones
// no longer cost swaps).  The group counts the keys it claims in *claimed
// (LDS, zeroed by the caller); once the count passes `limit` the row is given
// up (`full`), and the caller, after a group barrier, also gives the row up
// when the final count passes `limit`.  The count of distinct keys does not
// depend on the order the lanes insert in, so whether a row is given up
// depends on its pattern only (a probe sequence cannot exceed the table
// while at most limit < H keys are in it).  (A separate function: the
// one-wavefront kernels keep expand_flat's register allocation.)
template <int G, typename LV, typename BV, int U>
__device__ __forceinline__ void expand_flat_crowded(const int32_t *lcol, const LV *lval, int64_t e0, int64_t e1,
                                                    const int32_t *__restrict__ brp,
                                                    const int32_t *__restrict__ bcol, const BV *__restrict__ bval,
                                                    int slog, int *keys, double *vals, unsigned mask, int shift,
                                                    unsigned dslot, bool &full, int *fullflag, int *claimed,
                                                    int limit, unsigned hmul = kHashMul)
{
    static_assert(G > kWave, "the claim count is kept per wavefront of a multi-wavefront group");
    constexpr int kLogG = G >= 256 ? 8 : G >= 128 ? 7 : G >= 64 ? 6 : G >= 32 ? 5 : 4;
    if (slog > kLogG) slog = kLogG;
    const int t = Group<G>::tid();
    const int S = 1 << slog;
    const int sub = t & (S - 1);
    const int ngrp = G >> slog;
    auto given_up = [&]() {
        return __hip_atomic_load(fullflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
               __hip_atomic_load(claimed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > limit;
    };
    for (int64_t eb = e0 + (t >> slog); eb < e1; eb += (int64_t)ngrp * U) {
        int q[U], qe[U];
        double lv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t e = eb + (int64_t)u * ngrp;
            const bool ok = e < e1;
            const int64_t es = ok ? e : eb;  // eb < e1: a valid entry
            const int k = lcol[es];
            const double x = (double)lval[es];
            const int b0 = brp[k], b1 = brp[k + 1];
            lv[u] = ok ? x : 0.0;
            q[u] = ok ? b0 + sub : 0;
            qe[u] = ok ? b1 : 0;
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
                // some valid entry of B (the loop runs only while B has entries)
                const unsigned qs = (unsigned)max(min(q[u], qe[u] - 1), 0);
                const int cc = bcol[qs];
                const double ww = (double)bval[qs];
                c[u] = ok ? cc : kEmpty;
                w[u] = ww;
                q[u] += S;
            }
#pragma unroll
            for (int u = 0; u < U; u++) sl[u] = c[u] != kEmpty ? (hslot(c[u], shift, hmul) & mask) : dslot;
#pragma unroll
            for (int u = 0; u < U; u++) old[u] = atomicCAS(keys + sl[u], kEmpty, c[u]);
            // keep the U compare-and-swaps in flight together: nothing that
            // consumes their results is scheduled above this point
            __builtin_amdgcn_sched_barrier(0);
            // An absent product swapped its dummy slot (old == kEmpty), so it
            // counts as a hit on sl == dslot and needs no separate test.
            bool pend = false;
            int fresh = 0;  // keys this wavefront claimed in this step
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool hit = old[u] == kEmpty || old[u] == c[u];
                pend |= !hit;
                atomicAdd(vals + (hit ? sl[u] : dslot), lv[u] * w[u]);
                fresh += __popcll(ballot(c[u] != kEmpty && old[u] == kEmpty));
            }
            if (fresh && lane() == __ffsll((unsigned long long)ballot(true)) - 1) atomicAdd(claimed, fresh);
            if (pend) {
                // Linear probing of the inserts whose first slot held another
                // key.  For kMergedProbes slots every pending insert of the
                // lane advances one slot per iteration, settled ones swapping
                // and adding on the lane's dummy slot (no per-insert branch);
                // the rest go on one at a time.  Once the row is given up
                // (here or by another lane) the pending inserts stop.
                bool pd[U];
                unsigned s2[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    pd[u] = !(old[u] == kEmpty || old[u] == c[u]);
                    s2[u] = sl[u];
                }
                unsigned probe = 1;
                for (; probe <= kMergedProbes; probe++) {
                    bool anyp = false;
#pragma unroll
                    for (int u = 0; u < U; u++) anyp |= pd[u];
                    if (!anyp) break;
                    if (probe > mask || given_up()) {
                        full = true;
                        __hip_atomic_store(fullflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
                        for (int u = 0; u < U; u++) pd[u] = false;
                        break;
                    }
                    int o[U];
                    unsigned a[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        s2[u] = (s2[u] + 1) & mask;
                        a[u] = pd[u] ? s2[u] : dslot;
                        o[u] = atomicCAS(keys + a[u], kEmpty, pd[u] ? c[u] : kEmpty);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    int nf = 0;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const bool hit = pd[u] && (o[u] == kEmpty || o[u] == c[u]);
                        atomicAdd(vals + (hit ? a[u] : dslot), lv[u] * w[u]);
                        nf += (pd[u] && o[u] == kEmpty) ? 1 : 0;
                        pd[u] = pd[u] && !hit;
                    }
                    if (nf) atomicAdd(claimed, nf);
                }
                // longer sequences (crowded tables) go on one insert at a
                // time, so the settled inserts stop costing swaps
                for (int u = 0; u < U; u++) {
                    if (!pd[u]) continue;
                    if (given_up()) {
                        full = true;
                        __hip_atomic_store(fullflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        continue;
                    }
                    bool done = false;
                    for (unsigned q = probe; q <= mask; q++) {
                        s2[u] = (s2[u] + 1) & mask;
                        const int o = atomicCAS(keys + s2[u], kEmpty, c[u]);
                        if (o == kEmpty || o == c[u]) {
                            atomicAdd(vals + s2[u], lv[u] * w[u]);
                            if (o == kEmpty) atomicAdd(claimed, 1);
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

// Appends, for every lane of the group with `claim` set, `slot` to the group's
// list of claimed slots (list[n++] = slot): one LDS atomic per wavefront.
template <int G>
__device__ __forceinline__ void list_append(bool claim, unsigned slot, int *list, int *count)
{
    uint64_t wm = ballot(claim);
    if constexpr (G < kWave) wm &= ((1ull << G) - 1) << Group<G>::lane0();
    if (wm) {
        const int leader = __ffsll((unsigned long long)wm) - 1;
        int base = 0;
        if (lane() == leader) base = atomicAdd(count, __popcll(wm));
        base = __shfl(base, leader);
        if (claim) list[base + lanes_below(wm)] = (int)slot;
    }
}

}  // namespace sp3mm
