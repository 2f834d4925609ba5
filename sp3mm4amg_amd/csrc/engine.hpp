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
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t max_lds = 65536;
    size_t lds_device = 65536;
    Pool pool;
    // scheduling of the one-pass direct kernels (sp3mm_set_tuning / sp3mm_autotune)
    Tuning tune;
    // small pinned host mirror for counters read back between phases
    long long *host_small = nullptr;
    long long *host_small_dev = nullptr;  // its device-side address (kernels write it directly)
    std::recursive_mutex mu;
};

Context *ctx();  // initialised on first use (device 0 unless sp3mm_init chose one)

// scratch buffer owned by one call
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    explicit Buf(size_t b);
    ~Buf();
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    Buf(Buf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
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
                     int prec, double guard, double *rowsum);
void dmat_release(sp3mm_dmat &m);
void *batch_create(int count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int prec,
                   double guard, bool graph);
int batch_run(void *b, bool use_graph, double *ms);
int tuning_valid(const Tuning &t);
int transpose(const sp3mm_dmat &A, sp3mm_dmat &T);
int sa_prolongator(const sp3mm_dmat &A, const int64_t *agg, int64_t nagg, double omega, int mode, sp3mm_dmat &P);
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
