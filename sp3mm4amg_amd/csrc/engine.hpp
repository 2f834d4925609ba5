// engine.hpp -- device context, HBM pool and product pipelines (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sp3mm_amd.h"
#include "kernels.hpp"

namespace sp3mm {

void set_error(const char *fmt, ...);
void clear_error();

#define SP3MM_HIP(call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::sp3mm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                               __FILE__, __LINE__);                                       \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// Caching allocator: blocks are rounded up to a size class (at most 1/8 waste)
// and kept on per-class free lists for reuse; everything is released when a
// fresh hipMalloc fails.
class Pool {
  public:
    void *alloc(size_t bytes);
    void release(void *p);
    void trim();
    size_t cached_bytes() const { return cached_; }

  private:
    static size_t size_class(size_t bytes);
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t cached_ = 0;
    std::mutex mu_;
};

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;      // the stream every call is queued on
    hipStream_t own_stream = nullptr;  // the library's own (non-blocking) stream
    int num_cu = 0;
    size_t max_lds = 65536;
    size_t lds_device = 65536;
    Pool pool;
    // scheduling of the one-pass direct kernels (sp3mm_set_tuning / sp3mm_autotune)
    Tuning tune;
    // device address of the fixed scratch array of the allocation-free entry
    // point on this device (null if it could not be resolved)
    unsigned char *noalloc_base = nullptr;
    // coherent pinned host block for counters the host polls while the
    // stream goes on: the control block, then a flag at byte kPollFlagOff
    // that the binning's scatter kernel sets to the sequence number of the export
    unsigned char *poll_host = nullptr;
    unsigned char *poll_dev = nullptr;
    unsigned long long poll_seq = 0;
    // HIP events kept for reuse by the timers of later calls
    std::vector<hipEvent_t> ev_pool;
    // small pinned host mirror for counters read back between phases
    long long *host_small = nullptr;
    long long *host_small_dev = nullptr;  // its device-side address (kernels write it directly)
    // state of the level consumers (SpMV, Jacobi, dot): a dot result, a sticky
    // error word set by Jacobi sweeps, and the dot's partial sums
    void *cons = nullptr;
    std::recursive_mutex mu;
};

// Device state of the level consumers: the last dot result, a sticky error
// word that Jacobi sweeps set on a zero diagonal entry, then the dot's
// partial sums.  Allocated on first use.
struct ConsState {
    double dot;
    int err;
    int pad;
};
ConsState *cons_state(Context *c);

// Waits for the library stream and reports (and clears) an error that a
// stream-ordered consumer call left in the device error word.
int consumer_wait(Context *c, double *dot_out = nullptr);
// Waits for a stream (an event) by polling for up to ~2 ms, then blocks.  A
// blocking wait returns some 20 us after the GPU is done; a product's final
// wait is on its critical path.
hipError_t spin_stream(hipStream_t s);
// the hash multipliers the autotuner scores, and their raw scores on given
// rows of R (see k_hash_score)
const std::vector<unsigned> &hash_candidate_list();
int hash_scores(Context *c, const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int H1, int H2,
                const int32_t *rows, int nrows, std::vector<unsigned long long> &h, std::vector<int> *nkeys);
hipError_t spin_event(hipEvent_t e);

// Queues `sweeps` damped Jacobi sweeps (x in place, work as the other buffer).
void queue_jacobi(const sp3mm_dmat &A, const double *b, double *x, double *work, double omega, int sweeps, int *err,
                  hipStream_t s);

Context *ctx();  // initialised on first use (device 0 unless sp3mm_init chose one)

// Bump allocator over a fixed device region: while one is installed on the
// calling thread (ArenaScope), every Buf of that thread is carved from it
// instead of the pool, 256-byte aligned, and never returned.
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0, used = 0;
    size_t refused = 0;  // size of the first request that did not fit (0: none)
    void *take(size_t bytes);
};
struct ArenaScope {
    explicit ArenaScope(Arena *a);
    ~ArenaScope();
    Arena *prev;
};

// scratch buffer owned by one call
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    bool arena = false;  // carved from an Arena (not returned to the pool)
    Buf() = default;
    explicit Buf(size_t b);
    ~Buf();
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    Buf(Buf &&o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena) { o.p = nullptr; o.bytes = 0; }
    Buf &operator=(Buf &&o) noexcept;
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// product pipelines (return 0 on success)
// symbolic_only: stop after the row offsets (C.rp, C.nnz) are known
int spgemm(const sp3mm_dmat &A, const sp3mm_dmat &B, sp3mm_dmat &C, sp3mm_stats *st,
           bool symbolic_only = false);
int rap_direct(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C,
               sp3mm_stats *st, bool timing, bool symbolic_only = false);
int rap_direct_mixed(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C, sp3mm_stats *st,
                     int prec, double guard, double *rowsum, bool timing = false);
// R*A*P (FP64, direct path) into caller-owned device arrays, every scratch
// buffer carved from the fixed device scratch array: no device allocation
int rap_direct_noalloc(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int64_t *crp, int32_t *ccol,
                       double *cval, int64_t cap, int64_t &nnz, sp3mm_stats *st);
void dmat_release(sp3mm_dmat &m);
void *batch_create(int count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int prec,
                   double guard, bool graph);
int batch_run(void *b, bool use_graph, double *ms);
int batch_repack(void *b, int count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P);
// times candidate schedules on the batch's packed product, installs the
// fastest and rebuilds the batch (plan and graph) with it
int batch_autotune(void *b, int reps, Tuning &best, double &ms_best, double &ms_default);
int tuning_valid(const Tuning &t);
int transpose(const sp3mm_dmat &A, sp3mm_dmat &T);
// smoothed-aggregation hierarchy of A as one replayable plan (see engine.hip)
void *hier_create(const sp3mm_dmat &A, int nlev, const int64_t *const *aggs, const int64_t *naggs, double omega,
                  int mode, bool graph);
int hier_update(void *h, bool use_graph, double *ms);
int hier_level(void *h, int l, sp3mm_dmat &R, sp3mm_dmat &P, sp3mm_dmat &Ac);
void hier_info(void *h, int32_t *nlev, double *ms_build, int32_t *has_graph);
void hier_free(void *h);
int sa_prolongator(const sp3mm_dmat &A, const int64_t *agg, int64_t nagg, double omega, int mode, sp3mm_dmat &P);
// Queue later calls on `s` (nullptr: the library's own stream), after work
// already queued on the current stream has finished.
int set_stream(hipStream_t s);
int spmv(const sp3mm_dmat &A, const double *x, const double *z, double *y, double alpha, double beta);
int axpby(int64_t n, double alpha, const double *x, double beta, double *y);
int dot(int64_t n, const double *x, const double *y, double &out);
int jacobi(const sp3mm_dmat &A, const double *b, double *x, double *work, double omega, int sweeps);
void *vcycle_create(int nlev, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, const sp3mm_dmat &Ac,
                    int nu1, int nu2, double omega, int coarse_sweeps, bool graph);
int vcycle_solve(void *v, const double *b, double *x, double rtol, int maxit, int *cycles, double *hist,
                 double *ms);
void vcycle_free(void *v);
int autotune(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int prec, double guard, int reps,
             Tuning &best, double &ms_best, double &ms_default);
int batch_result(void *b, int i, sp3mm_dmat &out);
void batch_info(void *b, double *ms_pack, double *ms_build, int64_t *nnz, int32_t *has_graph);
void batch_free(void *b);
int ub_colparts(const sp3mm_dmat &A, const int64_t *host_offsets, int64_t boff_len, int grid_cols,
                int64_t *host_out);
int rap_pair(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C,
             sp3mm_stats *st, bool timing);
int row_ub(const sp3mm_dmat &A, const sp3mm_dmat &B, int64_t *host_ub);

SpOperand operand(const sp3mm_dmat &m);
int dmat_alloc(sp3mm_dmat &m, int64_t M, int64_t N, int64_t nnz, int vtype);

}  // namespace sp3mm
