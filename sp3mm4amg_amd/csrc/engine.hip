// engine.hip -- device context, HBM pool and the product pipelines.
//
// A product runs as: upper bounds + binning -> per-bin symbolic kernels ->
// scan + re-binning by exact sizes -> per-bin numeric kernels, all on one HIP
// stream.  The host reads small counters back twice per product (bin
// populations, then output size) to size grids and the output arrays.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>

#include "device_common.hpp"
#include "engine.hpp"

namespace sp3mm {

static thread_local char g_err[1024];

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }
const char *last_error_text() { return g_err; }

// --------------------------------------------------------------------------
// Pool
// --------------------------------------------------------------------------
size_t Pool::size_class(size_t b)
{
    if (b <= 4096) return 4096;
    const int lg = 63 - __builtin_clzll((unsigned long long)b);
    const size_t g = (size_t)1 << (lg - 3);
    return (b + g - 1) / g * g;
}

void *Pool::alloc(size_t bytes)
{
    const size_t c = size_class(bytes ? bytes : 1);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.find(c);
    if (it != free_.end()) {
        void *p = it->second;
        free_.erase(it);
        cached_ -= c;
        live_[p] = c;
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        for (auto &kv : free_) (void)hipFree(kv.second);
        free_.clear();
        cached_ = 0;
        e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            size_t live = 0;
            for (auto &kv : live_) live += kv.second;
            set_error("hipMalloc(%zu) failed: %s (library holds %zu bytes in %zu live blocks)", c,
                      hipGetErrorString(e), live, live_.size());
            return nullptr;
        }
    }
    live_[p] = c;
    return p;
}

void Pool::release(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, p);
    cached_ += it->second;
    live_.erase(it);
}

void Pool::trim()
{
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

// --------------------------------------------------------------------------
// Context
// --------------------------------------------------------------------------
static Context *g_ctx = nullptr;
static std::mutex g_ctx_mu;

// Fixed device scratch of the allocation-free entry point (sp3mm_rap_noalloc).
// Every GPU the library runs on holds its own copy until the process ends.  In
// host code `scratch` is only a host-side stand-in for the symbol: its device
// address is resolved once per device at init (hipGetSymbolAddress, with the
// device current) into Context::noalloc_base, and only that address is handed
// to kernels, fills and copies.
static __device__ __align__(256) unsigned char scratch[SP3MM_NOALLOC_SCRATCH_BYTES];

int tuning_valid(const Tuning &t)
{
    const bool pow2h = t.subwave_h == 0 || (t.subwave_h >= 64 && t.subwave_h <= 4096 &&
                                            (t.subwave_h & (t.subwave_h - 1)) == 0);
    return (t.u1 == 2 || t.u1 == 4) && (t.u2 >= 2 && t.u2 <= 4) && pow2h && t.slog1_cap >= 0 &&
           t.slog1_cap <= 8 && t.slog2_cap >= 0 && t.slog2_cap <= 8 && t.h1_shift >= -1 && t.h1_shift <= 3 &&
           (t.h2_div == 1 || t.h2_div == 2 || t.h2_div == 4 || t.h2_div == 8) && (t.xcd_map == 0 || t.xcd_map == 1) &&
           (t.hash1 & 1u) && (t.hash2 & 1u);
}

// initial scheduling: defaults, overridden by SP3MM_FUSED_U ("ab", a and b in
// {2,4}), SP3MM_SUBWAVE_H, SP3MM_SLOG_MAX (both stages), SP3MM_H1_SHIFT,
// SP3MM_H2_DIV, SP3MM_XCD_MAP
static void tuning_from_env(Tuning &t)
{
    Tuning e = t;
    if (const char *v = getenv("SP3MM_FUSED_U")) {
        if (strlen(v) >= 2) {
            e.u1 = v[0] - '0';
            e.u2 = v[1] - '0';
        }
    }
    if (const char *v = getenv("SP3MM_SUBWAVE_H")) e.subwave_h = atoi(v);
    if (const char *v = getenv("SP3MM_SLOG_MAX")) e.slog1_cap = e.slog2_cap = atoi(v);
    if (const char *v = getenv("SP3MM_H1_SHIFT")) e.h1_shift = atoi(v);
    if (const char *v = getenv("SP3MM_H2_DIV")) e.h2_div = atoi(v);
    if (const char *v = getenv("SP3MM_XCD_MAP")) e.xcd_map = atoi(v);
    if (tuning_valid(e)) t = e;
}

int init_device(int dev)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_ctx && g_ctx->device == dev) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return -1;
    }
    if (dev < 0 || dev >= n) {
        set_error("device %d out of range (%d devices)", dev, n);
        return -1;
    }
    Context *c = new Context();
    c->device = dev;
    SP3MM_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    SP3MM_HIP(hipGetDeviceProperties(&prop, dev));
    c->num_cu = prop.multiProcessorCount;
    // host waits spin instead of sleeping: the pipeline reads small counters
    // back between phases, and a blocking wait adds wake-up latency to each
    if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
    // dynamic LDS per workgroup the bin tables may use (SP3MM_LDS_CAP overrides)
    c->lds_device = prop.sharedMemPerBlock;
    c->max_lds = prop.sharedMemPerBlock;
    if (const char *cap = getenv("SP3MM_LDS_CAP")) {
        const size_t v = (size_t)strtoull(cap, nullptr, 10);
        if (v >= 16384 && v < c->max_lds) c->max_lds = v;
    }
    tuning_from_env(c->tune);
    SP3MM_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    SP3MM_HIP(hipHostMalloc((void **)&c->host_small, 4096, hipHostMallocDefault));
    // a coherent block the host polls for counters exported mid-stream (null:
    // the plans read their counters back with a stream synchronisation)
    if (hipHostMalloc((void **)&c->poll_host, 4096, hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess &&
        hipHostGetDevicePointer((void **)&c->poll_dev, c->poll_host, 0) == hipSuccess) {
        memset(c->poll_host, 0, 4096);
    } else {
        (void)hipGetLastError();
        if (c->poll_host) (void)hipHostFree(c->poll_host);
        c->poll_host = nullptr;
        c->poll_dev = nullptr;
    }
    // this device's copy of the fixed scratch array (sp3mm_rap_noalloc reports
    // an error if it cannot be resolved; the other entry points do not use it)
    void *sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess && sp)
        c->noalloc_base = static_cast<unsigned char *>(sp);
    else
        (void)hipGetLastError();
    if (hipHostGetDevicePointer((void **)&c->host_small_dev, c->host_small, 0) != hipSuccess) {
        (void)hipGetLastError();
        c->host_small_dev = nullptr;
    }
    g_ctx = c;
    return 0;
}

Context *ctx()
{
    if (!g_ctx) {
        if (init_device(0)) return nullptr;
    }
    return g_ctx;
}

static thread_local Arena *t_arena = nullptr;

void *Arena::take(size_t b)
{
    const size_t a = (b + 255) & ~(size_t)255;
    if (a > size - used) {
        if (!refused) refused = a;
        return nullptr;
    }
    void *p = base + used;
    used += a;
    return p;
}

ArenaScope::ArenaScope(Arena *a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

Buf::Buf(size_t b) : bytes(b)
{
    if (t_arena) {
        p = t_arena->take(b ? b : 1);
        arena = true;
        return;
    }
    Context *c = ctx();
    p = c ? c->pool.alloc(b) : nullptr;
}

Buf::~Buf()
{
    if (p && !arena && g_ctx) g_ctx->pool.release(p);
}

Buf &Buf::operator=(Buf &&o) noexcept
{
    if (this != &o) {
        if (p && !arena && g_ctx) g_ctx->pool.release(p);
        p = o.p;
        bytes = o.bytes;
        arena = o.arena;
        o.p = nullptr;
        o.bytes = 0;
    }
    return *this;
}

SpOperand operand(const sp3mm_dmat &m)
{
    SpOperand o;
    o.rp = m.rp;
    o.rp32 = m.rp32;
    o.col = m.col;
    o.val = m.val;
    o.M = m.M;
    o.N = m.N;
    return o;
}

int dmat_alloc(sp3mm_dmat &m, int64_t M, int64_t N, int64_t nnz, int vtype)
{
    Context *c = ctx();
    if (!c) return -1;
    memset(&m, 0, sizeof m);
    m.M = M;
    m.N = N;
    m.nnz = nnz;
    m.vtype = vtype;
    m.owned = 1;
    m.rp = static_cast<int64_t *>(c->pool.alloc((size_t)(M + 1) * 8));
    if (nnz < INT32_MAX) m.rp32 = static_cast<int32_t *>(c->pool.alloc((size_t)(M + 1) * 4));
    m.col = static_cast<int32_t *>(c->pool.alloc((size_t)(nnz ? nnz : 1) * 4));
    m.val = c->pool.alloc((size_t)(nnz ? nnz : 1) * (vtype == SP3MM_F32 ? 4 : 8));
    if (!m.rp || !m.col || !m.val || (nnz < INT32_MAX && !m.rp32)) {
        char why[600];
        snprintf(why, sizeof why, "%s", last_error_text());
        set_error("device allocation of a %lld x %lld matrix with %lld nnz failed: %s", (long long)M,
                  (long long)N, (long long)nnz, why);
        return -1;
    }
    return 0;
}

void dmat_release(sp3mm_dmat &m)
{
    if (m.owned && g_ctx) {
        g_ctx->pool.release(m.rp);
        g_ctx->pool.release(m.rp32);
        g_ctx->pool.release(m.col);
        g_ctx->pool.release(m.val);
    }
    memset(&m, 0, sizeof m);
}

// --------------------------------------------------------------------------
// bin configurations
// --------------------------------------------------------------------------
struct BinCfg {
    int G, H;
};

// largest power-of-two table (>= 64) whose LDS footprint fits the device
static int largest_h(size_t max_lds, size_t (*lds)(int, int), int G)
{
    int H = 64;
    while (lds(G, H * 2) <= max_lds && H < (1 << 20)) H *= 2;
    return H;
}
static size_t dsymb_lds(int G, int H) { return direct_lds(G, H, H, false); }
static size_t dnum_lds(int G, int H) { return direct_lds(G, H, H, true); }

// pairwise symbolic: measure = row upper bound; a table of H >= ub keys cannot
// fill.  Rows above the last bin are tried with the largest table that fits
// and raise an error if it fills.
static void symb_bins(size_t max_lds, BinThresholds &thr, std::vector<BinCfg> &cfg)
{
    const int hmax = largest_h(max_lds, symb_lds, 256);
    thr.nb = 0;
    cfg.clear();
    for (long long v = 64; v <= hmax && thr.nb < kMaxBins - 1; v *= 2) {
        thr.v[thr.nb++] = v;
        cfg.push_back({v <= 1024 ? 64 : 256, (int)v});
    }
    cfg.push_back({256, hmax});
}

// pairwise numeric: measure = exact row length n, table of 2n slots
static void num_bins(size_t max_lds, BinThresholds &thr, std::vector<BinCfg> &cfg)
{
    const int hmax = largest_h(max_lds, num_lds, 256);
    thr.nb = 0;
    cfg.clear();
    for (long long v = 32; 2 * v <= hmax && thr.nb < kMaxBins - 1; v *= 2) {
        thr.v[thr.nb++] = v;
        const int H = (int)(2 * v);
        cfg.push_back({H <= 512 ? 64 : 256, H});
    }
    cfg.push_back({0, 0});  // longer rows are not handled by the LDS kernels
}

// direct symbolic: measure = upper bound of the R*A row; both tables of H keys.
// Rows whose second table fills are retried with the largest tables.
static void dsymb_bins(size_t max_lds, BinThresholds &thr, std::vector<BinCfg> &cfg)
{
    const int hmax = largest_h(max_lds, dsymb_lds, 256);
    thr.nb = 0;
    cfg.clear();
    for (long long v = 64; v <= hmax && thr.nb < kMaxBins - 1; v *= 2) {
        thr.v[thr.nb++] = v;
        cfg.push_back({v <= 512 ? 64 : 256, (int)v});
    }
    cfg.push_back({256, hmax});
}

// direct numeric: measure = max(nnz of the R*A row, nnz of the RAP row)
static void dnum_bins(size_t max_lds, BinThresholds &thr, std::vector<BinCfg> &cfg)
{
    const int hmax = largest_h(max_lds, dnum_lds, 256);
    thr.nb = 0;
    cfg.clear();
    for (long long v = 32; 2 * v <= hmax && thr.nb < kMaxBins - 1; v *= 2) {
        thr.v[thr.nb++] = v;
        const int H = (int)(2 * v);
        cfg.push_back({H <= 256 ? 64 : 256, H});
    }
    cfg.push_back({0, 0});
}

// lanes per right-hand row: the mean right-row length rounded up to a power of
// two (at most one wavefront), so one step of the sub-group covers a mean row
static int sub_log(const sp3mm_dmat &B, int G)
{
    const double avg = B.M ? (double)B.nnz / (double)B.M : 1.0;
    int l = 0;
    while ((double)(1 << l) < avg && l < 6) l++;
    const int gl = ilog2((unsigned)G);
    return l > gl ? gl : l;
}

// control block layout (device, zeroed per product)
struct Ctl {
    int counts[16];
    int counts2[16];
    int rawc[16];  // bin populations before small bins are merged (direct plans)
    int pred_count;  // population seen by the launch queued before the read-back (0: abandoned)
    int onebin;      // 1: the binning put every row in one bin without measuring it
    int nretry;
    int err;
    int nspill;  // long rows whose column windows exceed the LDS window kernel
    int pad;
    unsigned long long totals[3];  // bound sum, bound maximum, rows of unstructured spread
    int64_t nnz;
    int64_t pad2;
};

static int max_blocks(const Context *c) { return c->num_cu * 16; }

constexpr int kDenseThreads = 1024;
// persistent workgroups (each with its own dense workspace) for the long-row kernels
static int dense_blocks(const Context *c, int rows)
{
    static const int env = getenv("SP3MM_DENSE_BLOCKS") ? atoi(getenv("SP3MM_DENSE_BLOCKS")) : 0;
    const int cap = env > 0 ? env : c->num_cu * 2;
    return rows < cap ? (rows > 0 ? rows : 1) : cap;
}

// dense workspaces of the direct long-row kernels: bitmap + list over A's
// columns, bitmap over P's columns, and (numeric) the value arrays
struct DenseDirectWs {
    Buf bitsA, valsA, listc, listv, bitsP, valsP, cur;
    DenseWs ws{};
    int alloc(int wb, int64_t ncolA, int64_t ncolP, bool num, hipStream_t s)
    {
        const int nwA = dense_words(ncolA), nwP = dense_words(ncolP);
        const size_t slotsA = (size_t)wb * nwA * 32, slotsP = (size_t)wb * nwP * 32;
        bitsA = Buf((size_t)wb * nwA * 4);
        listc = Buf(slotsA * 4);
        bitsP = Buf((size_t)wb * nwP * 4);
        if (num) {
            valsA = Buf(slotsA * 8);
            listv = Buf(slotsA * 8);
            valsP = Buf(slotsP * 8);
            cur = Buf(slotsA * 4);
        }
        if (!bitsA.p || !listc.p || !bitsP.p || (num && (!valsA.p || !listv.p || !valsP.p || !cur.p))) {
            set_error("dense workspace allocation failed");
            return -1;
        }
        SP3MM_HIP(hipMemsetAsync(bitsA.p, 0, bitsA.bytes, s));
        SP3MM_HIP(hipMemsetAsync(bitsP.p, 0, bitsP.bytes, s));
        if (num) {
            SP3MM_HIP(hipMemsetAsync(valsA.p, 0, valsA.bytes, s));
            SP3MM_HIP(hipMemsetAsync(valsP.p, 0, valsP.bytes, s));
        }
        ws.bitsA = bitsA.as<unsigned>();
        ws.valsA = valsA.as<double>();
        ws.listc = listc.as<int32_t>();
        ws.listv = listv.as<double>();
        ws.bitsP = bitsP.as<unsigned>();
        ws.valsP = valsP.as<double>();
        ws.cur = cur.as<int32_t>();
        ws.nwA = nwA;
        ws.nwP = nwP;
        return 0;
    }
};

static const char *err_text(int e)
{
    if (e & kErrHashFull) return "an LDS hash table overflowed (row too long for the LDS kernels)";
    if (e & kErrOverflow) return "numeric row length differs from the symbolic row length";
    if (e & kErrBadInput) return "invalid input matrix";
    return "device error";
}

struct Timer {
    hipStream_t s;
    bool on;
    std::vector<hipEvent_t> ev;
    explicit Timer(hipStream_t st, bool enabled) : s(st), on(enabled) {}
    ~Timer()
    {
        // events go back to the context's pool (creating one costs host time
        // in the middle of a product's launch sequence)
        for (auto e : ev) {
            if (g_ctx && g_ctx->ev_pool.size() < 64) g_ctx->ev_pool.push_back(e);
            else (void)hipEventDestroy(e);
        }
    }
    int mark()
    {
        if (!on) return -1;  // disabled: no events are created
        hipEvent_t e;
        if (g_ctx && !g_ctx->ev_pool.empty()) {
            e = g_ctx->ev_pool.back();
            g_ctx->ev_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            return -1;
        }
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        return (int)ev.size() - 1;
    }
    double ms(int a, int b)
    {
        float t = 0.f;
        if (a < 0 || b < 0) return 0.0;
        (void)hipEventElapsedTime(&t, ev[a], ev[b]);
        return (double)t;
    }
};

constexpr size_t kPollFlagOff = 4088;

template <typename Query, typename Block>
static hipError_t spin_wait(Query q, Block b)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned k = 0;; k++) {
        const hipError_t e = q();
        if (e != hipErrorNotReady) return e;
        if ((k & 63) == 63 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) return b();
        __builtin_ia32_pause();
    }
}
hipError_t spin_stream(hipStream_t s)
{
    return spin_wait([s] { return hipStreamQuery(s); }, [s] { return hipStreamSynchronize(s); });
}
hipError_t spin_event(hipEvent_t e)
{
    return spin_wait([e] { return hipEventQuery(e); }, [e] { return hipEventSynchronize(e); });
}

// Reads the control block back with a stream synchronisation.  With a
// timer, the event *mark is recorded behind the export, before the
// synchronisation (a product's end event then costs no second round trip).
static int read_ctl(Context *c, const Ctl *d, Ctl &h, Timer *tm = nullptr, int *mark = nullptr);

// Waits for the export `seq` of the control block into the coherent host
// block (the binning's scatter kernel writes it and then raises the flag at
// kPollFlagOff to seq; work queued behind it keeps the GPU busy meanwhile) and
// copies the block out.  A stream that ends without the flag (a failed launch
// or a fault) is reported.  seq 0 (no coherent block) reads back with a stream
// synchronisation instead.
static int wait_ctl(Context *c, const Ctl *d, unsigned long long seq, Ctl &h)
{
    if (seq == 0) return read_ctl(c, d, h);
    SP3MM_HIP(hipGetLastError());
    auto *flag_host = reinterpret_cast<unsigned long long *>(c->poll_host + kPollFlagOff);
    for (unsigned spins = 0;; spins++) {
        if (__atomic_load_n(flag_host, __ATOMIC_ACQUIRE) == seq) break;
        if ((spins & 1023) == 1023) {
            const hipError_t q = hipStreamQuery(c->stream);
            if (q == hipSuccess) {
                if (__atomic_load_n(flag_host, __ATOMIC_ACQUIRE) == seq) break;
                set_error("control block export did not arrive (sequence %llu)", seq);
                return -1;
            }
            if (q != hipErrorNotReady) {
                set_error("stream failed while waiting for the control block: %s", hipGetErrorString(q));
                return -1;
            }
        }
        __builtin_ia32_pause();
    }
    memcpy(&h, c->poll_host, sizeof(Ctl));
    return 0;
}

static int read_ctl(Context *c, const Ctl *d, Ctl &h, Timer *tm, int *mark)
{
    static_assert(sizeof(Ctl) % 4 == 0 && sizeof(Ctl) <= 4096, "control block fits the host mirror");
    SP3MM_HIP(hipGetLastError());  // a kernel launch that was rejected
    if (c->host_small_dev) {
        launch_export_words(d, c->host_small_dev, (int)(sizeof(Ctl) / 4), c->stream);
        SP3MM_HIP(hipGetLastError());
    } else {
        SP3MM_HIP(hipMemcpyAsync(c->host_small, d, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    }
    if (tm && mark) *mark = tm->mark();
    SP3MM_HIP(spin_stream(c->stream));
    memcpy(&h, c->host_small, sizeof(Ctl));
    return 0;
}

// --------------------------------------------------------------------------
// C = A*B
// --------------------------------------------------------------------------
static int check_pair(const sp3mm_dmat &A, const sp3mm_dmat &B)
{
    if (A.N != B.M) {
        set_error("inner dimensions differ: A is %lld x %lld, B is %lld x %lld", (long long)A.M,
                  (long long)A.N, (long long)B.M, (long long)B.N);
        return -1;
    }
    if (!B.rp32) {
        set_error("right operand has >= 2^31 nonzeros");
        return -1;
    }
    if (A.vtype != SP3MM_F64 || B.vtype != SP3MM_F64) {
        set_error("FP64 product needs FP64 operands");
        return -1;
    }
    return 0;
}

int spgemm_impl(const sp3mm_dmat &A, const sp3mm_dmat &B, sp3mm_dmat &C, sp3mm_stats *st, Timer *tm,
                double *ms_num, bool symbolic_only = false)
{
    Context *c = ctx();
    if (!c) return -1;
    if (check_pair(A, B)) return -1;
    hipStream_t s = c->stream;
    const int64_t M = A.M;
    const SpOperand a = operand(A), b = operand(B);
    BinThresholds thr{}, thr2{};
    std::vector<BinCfg> cfg, cfg2;
    symb_bins(c->max_lds, thr, cfg);
    num_bins(c->max_lds, thr2, cfg2);

    const int nlists = (thr.nb > thr2.nb ? thr.nb : thr2.nb) + 1;
    Buf ctl(sizeof(Ctl)), ub((size_t)(M + 1) * 8), rown((size_t)(M + 1) * 4),
        lists((size_t)nlists * (M ? M : 1) * 4), tiles((size_t)scan_tiles(M) * 8),
        binscr((size_t)bin_scratch_ints(M) * 4);
    if (!ctl.p || !ub.p || !rown.p || !lists.p || !tiles.p || !binscr.p) {
        set_error("scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), s));
    launch_ub_bin(a, b, ub.as<int64_t>(), rown.as<int32_t>(), lists.as<int32_t>(), d->counts, thr,
                  d->totals, binscr.as<int>(), s);
    if (read_ctl(c, d, h)) return -1;
    const long long mac = (long long)h.totals[0];

    for (int i = 0; i < thr.nb; i++) {
        if (h.counts[i] == 0) continue;
        const int G = cfg[i].G, H = cfg[i].H;
        launch_symb(G, H, sub_log(B, G), a, b, lists.as<int32_t>() + (int64_t)i * M, d->counts + i,
                    h.counts[i], rown.as<int32_t>(), nullptr, nullptr, &d->err, max_blocks(c), s);
    }
    const int nwB = dense_words(B.N);
    if (h.counts[thr.nb] > 0) {  // long rows: dense bitmap in HBM
        const int wb = dense_blocks(c, h.counts[thr.nb]);
        Buf bits((size_t)wb * nwB * 4);
        if (!bits.p) {
            set_error("dense workspace allocation failed");
            return -1;
        }
        SP3MM_HIP(hipMemsetAsync(bits.p, 0, (size_t)wb * nwB * 4, s));
        launch_dense_spgemm(false, wb, a, b, lists.as<int32_t>() + (int64_t)thr.nb * M, d->counts + thr.nb,
                            bits.as<unsigned>(), nullptr, nwB, sub_log(B, kDenseThreads), rown.as<int32_t>(),
                            nullptr, nullptr, nullptr, &d->err, s, c->max_lds);
    }
    if (dmat_alloc(C, M, B.N, 0, SP3MM_F64)) return -1;
    // C.rp32 was sized for nnz 0; it is re-checked once the total is known
    launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d->nnz, s);
    launch_rebin(rown.as<int32_t>(), nullptr, M, lists.as<int32_t>(), d->counts2, thr2, binscr.as<int>(), s);
    if (read_ctl(c, d, h)) return -1;
    if (h.err) {
        set_error("symbolic phase: %s", err_text(h.err));
        return -1;
    }
    const int64_t nnz = h.nnz;
    if (symbolic_only) {
        C.nnz = nnz;
        if (st) st->mac1 += mac;
        return 0;
    }
    C.nnz = nnz;
    c->pool.release(C.col);
    c->pool.release(C.val);
    C.col = static_cast<int32_t *>(c->pool.alloc((size_t)(nnz ? nnz : 1) * 4));
    C.val = c->pool.alloc((size_t)(nnz ? nnz : 1) * 8);
    if (!C.col || !C.val) {
        set_error("output allocation (%lld nnz) failed", (long long)nnz);
        return -1;
    }
    if (nnz >= INT32_MAX) {
        c->pool.release(C.rp32);
        C.rp32 = nullptr;
    }
    const int t0 = tm ? tm->mark() : -1;
    for (int i = 0; i < thr2.nb; i++) {
        if (h.counts2[i] == 0) continue;
        const int G = cfg2[i].G, H = cfg2[i].H;
        launch_num(G, H, sub_log(B, G), a, b, lists.as<int32_t>() + (int64_t)i * M, d->counts2 + i,
                   h.counts2[i], C.rp, C.col, static_cast<double *>(C.val), &d->err, max_blocks(c), s);
    }
    Buf dbits, dvals;
    if (h.counts2[thr2.nb] > 0) {  // long rows: dense bitmap + values in HBM
        const int wb = dense_blocks(c, h.counts2[thr2.nb]);
        dbits = Buf((size_t)wb * nwB * 4);
        dvals = Buf((size_t)wb * nwB * 32 * 8);
        if (!dbits.p || !dvals.p) {
            set_error("dense workspace allocation failed");
            return -1;
        }
        SP3MM_HIP(hipMemsetAsync(dbits.p, 0, (size_t)wb * nwB * 4, s));
        SP3MM_HIP(hipMemsetAsync(dvals.p, 0, (size_t)wb * nwB * 32 * 8, s));
        launch_dense_spgemm(true, wb, a, b, lists.as<int32_t>() + (int64_t)thr2.nb * M, d->counts2 + thr2.nb,
                            dbits.as<unsigned>(), dvals.as<double>(), nwB, sub_log(B, kDenseThreads), nullptr,
                            C.rp, C.col, static_cast<double *>(C.val), &d->err, s, c->max_lds);
    }
    const int t1 = tm ? tm->mark() : -1;
    if (read_ctl(c, d, h)) return -1;
    if (h.err) {
        set_error("numeric phase: %s", err_text(h.err));
        return -1;
    }
    if (tm && ms_num) *ms_num += tm->ms(t0, t1);
    if (st) {
        st->mac1 += mac;
        st->nnz_out = nnz;
    }
    return 0;
}

// one-pass SpGEMM bins: table of H >= ub slots; 32-lane groups up to 256
// slots, wavefronts up to 1024, workgroups beyond
static void spg_bins(size_t max_lds, BinThresholds &thr, std::vector<BinCfg> &cfg)
{
    thr.nb = 0;
    cfg.clear();
    for (long long v = 64; thr.nb < kMaxBins - 1; v *= 2) {
        const int H = (int)v;
        const int G = H <= 256 ? 32 : H <= 1024 ? 64 : 256;
        if (spgemm_fused_lds(G, H) > max_lds) break;
        thr.v[thr.nb++] = v;
        cfg.push_back({G, H});
    }
    cfg.push_back({0, 0});
}

// C = A*B in one pass over the operands: rows land in scratch slots, then a
// scan and a gather build the output.  Rows beyond the largest LDS table run
// on the dense long-row kernels (count, then numeric).
static int spgemm_onepass(const sp3mm_dmat &A, const sp3mm_dmat &B, sp3mm_dmat &C, sp3mm_stats *st, Timer *tm,
                          double *ms_num, unsigned hmul)
{
    Context *c = ctx();
    if (!c) return -1;
    if (check_pair(A, B)) return -1;
    hipStream_t s = c->stream;
    const int64_t M = A.M;
    const SpOperand a = operand(A), b = operand(B);
    BinThresholds thr{};
    std::vector<BinCfg> cfg;
    spg_bins(c->max_lds, thr, cfg);
    Buf ctl(sizeof(Ctl)), ub((size_t)(M + 1) * 8), rown((size_t)(M + 1) * 4),
        lists((size_t)(thr.nb + 1) * (M ? M : 1) * 4), slotn((size_t)(M ? M : 1) * 4),
        tiles((size_t)scan_tiles(M) * 8), binscr((size_t)bin_scratch_ints(M) * 4);
    if (!ctl.p || !ub.p || !rown.p || !lists.p || !slotn.p || !tiles.p || !binscr.p) {
        set_error("scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), s));
    launch_ub_bin(a, b, ub.as<int64_t>(), rown.as<int32_t>(), lists.as<int32_t>(), d->counts, thr, d->totals,
                  binscr.as<int>(), s, 0, nullptr, nullptr, nullptr, 0, nullptr, 0, 0, nullptr, -1, 0, false, nullptr,
                  M > 0 && A.nnz >= (int64_t)kWaveUbRow * M);
    if (read_ctl(c, d, h)) return -1;
    const long long mac = (long long)h.totals[0];
    std::vector<int64_t> soff(thr.nb + 1, 0), nbase(thr.nb + 1, 0);
    int64_t stot = 0, ntot = 0;
    for (int i = 0; i < thr.nb; i++) {
        soff[i] = stot;
        nbase[i] = ntot;
        stot += (int64_t)h.counts[i] * cfg[i].H;
        ntot += h.counts[i];
    }
    Buf scol((size_t)(stot ? stot : 1) * 4), sval((size_t)(stot ? stot : 1) * 8);
    if (!scol.p || !sval.p) {
        set_error("scratch allocation (%lld entries) failed", (long long)stot);
        return -1;
    }
    const int t0 = tm ? tm->mark() : -1;
    for (int i = 0; i < thr.nb; i++) {
        if (h.counts[i] == 0) continue;
        launch_spgemm_fused(cfg[i].G, cfg[i].H, sub_log(B, cfg[i].G), a, b, lists.as<int32_t>() + (int64_t)i * M,
                            d->counts + i, h.counts[i], rown.as<int32_t>(), slotn.as<int32_t>() + nbase[i],
                            scol.as<int32_t>() + soff[i], sval.as<double>() + soff[i], &d->err, max_blocks(c), s,
                            c->tune.xcd_map, hmul);
    }
    const int t1 = tm ? tm->mark() : -1;
    const int nlong = h.counts[thr.nb];
    const int nwB = dense_words(B.N);
    Buf dbits, dvals;
    const int wb = dense_blocks(c, nlong);
    if (nlong > 0) {
        // long rows: dense bitmap (+ values) in LDS when B's columns fit, else
        // in an HBM workspace; the numeric pass of rows too wide for LDS
        // values runs in LDS column windows when one fits (no value workspace)
        const bool bits_lds = spgemm_dense_in_lds(false, nwB, c->max_lds);
        const bool vals_ws = !spgemm_dense_in_lds(true, nwB, c->max_lds) && spgemm_window_cols(c->max_lds) == 0;
        if (!bits_lds) dbits = Buf((size_t)wb * nwB * 4);
        if (vals_ws) dvals = Buf((size_t)wb * nwB * 32 * 8);
        if ((!bits_lds && !dbits.p) || (vals_ws && !dvals.p)) {
            set_error("dense workspace allocation failed");
            return -1;
        }
        if (dbits.p) SP3MM_HIP(hipMemsetAsync(dbits.p, 0, dbits.bytes, s));
        if (dvals.p) SP3MM_HIP(hipMemsetAsync(dvals.p, 0, dvals.bytes, s));
        launch_dense_spgemm(false, wb, a, b, lists.as<int32_t>() + (int64_t)thr.nb * M, d->counts + thr.nb,
                            dbits.as<unsigned>(), nullptr, nwB, sub_log(B, kDenseThreads), rown.as<int32_t>(),
                            nullptr, nullptr, nullptr, &d->err, s, c->max_lds);
    }
    if (dmat_alloc(C, M, B.N, 0, SP3MM_F64)) return -1;
    launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d->nnz, s);
    // Without long rows every row fits its slot of H >= its upper bound, so the
    // slot total bounds nnz: allocate it and gather without reading the count.
    const bool bounded = nlong == 0 && stot < INT32_MAX;
    int64_t cap = stot;
    if (!bounded) {
        if (read_ctl(c, d, h)) return -1;
        if (h.err) {
            set_error("SpGEMM: %s", err_text(h.err));
            return -1;
        }
        cap = h.nnz;
    }
    c->pool.release(C.col);
    c->pool.release(C.val);
    C.col = static_cast<int32_t *>(c->pool.alloc((size_t)(cap ? cap : 1) * 4));
    C.val = c->pool.alloc((size_t)(cap ? cap : 1) * 8);
    if (!C.col || !C.val) {
        set_error("output allocation (%lld entries) failed", (long long)cap);
        return -1;
    }
    if (!bounded && cap >= INT32_MAX) {
        c->pool.release(C.rp32);
        C.rp32 = nullptr;
    }
    for (int i = 0; i < thr.nb; i++) {
        if (h.counts[i] == 0) continue;
        launch_gather_rows(lists.as<int32_t>() + (int64_t)i * M, d->counts + i, h.counts[i],
                           slotn.as<int32_t>() + nbase[i], cfg[i].H, scol.as<int32_t>() + soff[i],
                           sval.as<double>() + soff[i], C.rp, C.col, static_cast<double *>(C.val), cap, &d->err,
                           max_blocks(c), s);
    }
    if (nlong > 0)
        launch_dense_spgemm(true, wb, a, b, lists.as<int32_t>() + (int64_t)thr.nb * M, d->counts + thr.nb,
                            dbits.as<unsigned>(), dvals.as<double>(), nwB, sub_log(B, kDenseThreads), nullptr, C.rp,
                            C.col, static_cast<double *>(C.val), &d->err, s, c->max_lds);
    if (read_ctl(c, d, h)) return -1;
    if (h.err) {
        set_error("SpGEMM: %s", err_text(h.err));
        return -1;
    }
    C.nnz = h.nnz;
    if (tm && ms_num) *ms_num += tm->ms(t0, t1);
    if (st) {
        st->mac1 += mac;
        st->nnz_out = h.nnz;
    }
    return 0;
}

// the one-pass SpGEMM unless SP3MM_SPGEMM_TWOPASS is set (symbolic + numeric);
// hmul: the multiplier of the one-pass kernel's hash tables
static int spgemm_any(const sp3mm_dmat &A, const sp3mm_dmat &B, sp3mm_dmat &C, sp3mm_stats *st, Timer *tm,
                      double *ms_num, unsigned hmul = kHashMul)
{
    static const bool twopass = getenv("SP3MM_SPGEMM_TWOPASS") != nullptr;
    if (twopass) return spgemm_impl(A, B, C, st, tm, ms_num);
    return spgemm_onepass(A, B, C, st, tm, ms_num, hmul);
}

int spgemm(const sp3mm_dmat &A, const sp3mm_dmat &B, sp3mm_dmat &C, sp3mm_stats *st, bool symbolic_only)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Timer tm(c->stream, true);
    const int e0 = tm.mark();
    double ms_num = 0;
    sp3mm_stats local{};
    const int rc = symbolic_only ? spgemm_impl(A, B, C, &local, &tm, &ms_num, true)
                                 : spgemm_any(A, B, C, &local, &tm, &ms_num);
    if (rc) {
        dmat_release(C);
        return rc;
    }
    const int e1 = tm.mark();
    (void)hipEventSynchronize(tm.ev[e1]);
    if (st) {
        *st = local;
        st->ms_total = tm.ms(e0, e1);
        st->ms_numeric = ms_num;
        st->ms_main = ms_num;
    }
    return 0;
}

// --------------------------------------------------------------------------
// R*A*P, pair form (R*A)*P with the intermediate in HBM
// --------------------------------------------------------------------------
int rap_pair(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C,
             sp3mm_stats *st, bool timing)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Timer tm(c->stream, true);
    const int e0 = tm.mark();
    sp3mm_stats s1{}, s2{};
    double ms1 = 0, ms2 = 0;
    sp3mm_dmat RA;
    memset(&RA, 0, sizeof RA);
    // R*A's keys are A's columns, (R*A)*P's are P's: the installed
    // schedule's multipliers for the direct kernel's two tables (scored for
    // those key sets when the autotuner ran on these operands)
    if (spgemm_any(R, A, RA, &s1, &tm, &ms1, c->tune.hash1)) {
        dmat_release(RA);
        return -1;
    }
    const int rc = spgemm_any(RA, P, C, &s2, &tm, &ms2, c->tune.hash2);
    const int64_t mid = RA.nnz;
    dmat_release(RA);
    if (rc) {
        dmat_release(C);
        return rc;
    }
    const int e1 = tm.mark();
    (void)hipEventSynchronize(tm.ev[e1]);
    if (st) {
        memset(st, 0, sizeof *st);
        st->ms_total = tm.ms(e0, e1);
        st->ms_numeric = ms1 + ms2;
        st->ms_main = ms2;
        st->mac1 = s1.mac1;
        st->mac2 = s2.mac1;
        st->nnz_mid = mid;
        st->nnz_out = C.nnz;
        st->main_kernel_launches = 1;
    }
    (void)timing;
    return 0;
}

// --------------------------------------------------------------------------
// R*A*P, direct form: rows of R*A never leave the CU
// --------------------------------------------------------------------------
static int rap_direct_twopass(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C,
                              sp3mm_stats *st, bool timing, bool symbolic_only)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (check_pair(R, A) || check_pair(A, P)) return -1;
    if (R.vtype != SP3MM_F64) {
        set_error("FP64 product needs FP64 operands");
        return -1;
    }
    hipStream_t s = c->stream;
    Timer tm(s, true);
    const int e0 = tm.mark();
    const int64_t M = R.M;
    const SpOperand r = operand(R), a = operand(A), p = operand(P);
    BinThresholds thr{}, thr2{};
    std::vector<BinCfg> cfg, cfg2;
    dsymb_bins(c->max_lds, thr, cfg);
    dnum_bins(c->max_lds, thr2, cfg2);

    const int nlists = (thr.nb > thr2.nb ? thr.nb : thr2.nb) + 1;
    Buf ctl(sizeof(Ctl)), ub((size_t)(M + 1) * 8), rown((size_t)(M + 1) * 4),
        n1((size_t)(M + 1) * 4), lists((size_t)nlists * (M ? M : 1) * 4),
        retry((size_t)(M ? M : 1) * 4), tiles((size_t)scan_tiles(M) * 8), binscr((size_t)bin_scratch_ints(M) * 4);
    if (!ctl.p || !ub.p || !rown.p || !n1.p || !lists.p || !retry.p || !tiles.p || !binscr.p) {
        set_error("scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), s));
    launch_fill_i32(n1.as<int32_t>(), M, 0, s);
    launch_ub_bin(r, a, ub.as<int64_t>(), rown.as<int32_t>(), lists.as<int32_t>(), d->counts, thr,
                  d->totals, binscr.as<int>(), s);
    const int eu = tm.mark();
    if (read_ctl(c, d, h)) return -1;
    const long long mac1 = (long long)h.totals[0];
    const int slog1 = sub_log(A, 64), slog2 = sub_log(P, 64);
    const int es0 = tm.mark();
    for (int i = 0; i < thr.nb; i++) {
        if (h.counts[i] == 0) continue;
        const int G = cfg[i].G, H = cfg[i].H;
        launch_direct(G, false, H, H, slog1 > ilog2(G) ? ilog2(G) : slog1, slog2, r, a, p,
                      lists.as<int32_t>() + (int64_t)i * M, d->counts + i, h.counts[i], n1.as<int32_t>(),
                      rown.as<int32_t>(), nullptr, nullptr, nullptr, retry.as<int32_t>(), &d->nretry,
                      &d->err, max_blocks(c), s);
    }
    const int dlog1 = sub_log(A, kDenseThreads), dlog2 = sub_log(P, kDenseThreads);
    if (h.counts[thr.nb] > 0) {  // long rows: dense accumulators in HBM
        DenseDirectWs dw;
        if (dw.alloc(dense_blocks(c, h.counts[thr.nb]), A.N, P.N, false, s)) return -1;
        launch_dense_direct(false, dense_blocks(c, h.counts[thr.nb]), r, a, p,
                            lists.as<int32_t>() + (int64_t)thr.nb * M, d->counts + thr.nb, dw.ws, dlog1, dlog2,
                            n1.as<int32_t>(), rown.as<int32_t>(), nullptr, nullptr, nullptr, &d->err, s, false,
                            nullptr, c->max_lds);
    }
    const int es1 = tm.mark();
    if (dmat_alloc(C, M, P.N, 0, SP3MM_F64)) return -1;
    launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d->nnz, s);
    launch_rebin(n1.as<int32_t>(), rown.as<int32_t>(), M, lists.as<int32_t>(), d->counts2, thr2, binscr.as<int>(), s);
    if (read_ctl(c, d, h)) {
        dmat_release(C);
        return -1;
    }
    if (h.nretry > 0 && !h.err) {
        // rows whose RAP table filled: redo them densely, then rescan
        DenseDirectWs dw;
        if (dw.alloc(dense_blocks(c, h.nretry), A.N, P.N, false, s)) {
            dmat_release(C);
            return -1;
        }
        launch_dense_direct(false, dense_blocks(c, h.nretry), r, a, p, retry.as<int32_t>(), &d->nretry, dw.ws,
                            dlog1, dlog2, n1.as<int32_t>(), rown.as<int32_t>(), nullptr, nullptr, nullptr,
                            &d->err, s, false, nullptr, c->max_lds);
        SP3MM_HIP(hipMemsetAsync(d->counts2, 0, sizeof(d->counts2), s));
        launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d->nnz, s);
        launch_rebin(n1.as<int32_t>(), rown.as<int32_t>(), M, lists.as<int32_t>(), d->counts2, thr2, binscr.as<int>(), s);
        if (read_ctl(c, d, h)) {
            dmat_release(C);
            return -1;
        }
    }
    const int es2 = tm.mark();
    if (h.err) {
        set_error("direct symbolic phase: %s", err_text(h.err));
        dmat_release(C);
        return -1;
    }
    const int64_t nnz = h.nnz;
    C.nnz = nnz;
    if (symbolic_only) return 0;
    c->pool.release(C.col);
    c->pool.release(C.val);
    C.col = static_cast<int32_t *>(c->pool.alloc((size_t)(nnz ? nnz : 1) * 4));
    C.val = c->pool.alloc((size_t)(nnz ? nnz : 1) * 8);
    if (!C.col || !C.val) {
        set_error("output allocation (%lld nnz) failed", (long long)nnz);
        dmat_release(C);
        return -1;
    }
    if (nnz >= INT32_MAX) {
        c->pool.release(C.rp32);
        C.rp32 = nullptr;
    }
    const int en0 = tm.mark();
    int launches = 0;
    for (int i = 0; i < thr2.nb; i++) {
        if (h.counts2[i] == 0) continue;
        const int G = cfg2[i].G, H = cfg2[i].H;
        launch_direct(G, true, H, H, slog1 > ilog2(G) ? ilog2(G) : slog1, slog2, r, a, p,
                      lists.as<int32_t>() + (int64_t)i * M, d->counts2 + i, h.counts2[i], nullptr,
                      nullptr, C.rp, C.col, static_cast<double *>(C.val), nullptr, nullptr, &d->err,
                      max_blocks(c), s);
        launches++;
    }
    DenseDirectWs dwn;
    if (h.counts2[thr2.nb] > 0) {
        const int wb = dense_blocks(c, h.counts2[thr2.nb]);
        if (dwn.alloc(wb, A.N, P.N, true, s)) {
            dmat_release(C);
            return -1;
        }
        launch_dense_direct(true, wb, r, a, p, lists.as<int32_t>() + (int64_t)thr2.nb * M, d->counts2 + thr2.nb,
                            dwn.ws, dlog1, dlog2, nullptr, nullptr, C.rp, C.col, static_cast<double *>(C.val),
                            &d->err, s, false, nullptr, c->max_lds);
        launches++;
    }
    const int en1 = tm.mark();
    if (read_ctl(c, d, h)) {
        dmat_release(C);
        return -1;
    }
    if (h.err) {
        set_error("direct numeric phase: %s", err_text(h.err));
        dmat_release(C);
        return -1;
    }
    const int e1 = tm.mark();
    (void)hipEventSynchronize(tm.ev[e1]);
    if (st) {
        memset(st, 0, sizeof *st);
        st->ms_total = tm.ms(e0, e1);
        st->ms_ub = tm.ms(e0, eu);
        st->ms_symbolic = tm.ms(es0, es1);
        st->ms_scan = tm.ms(es1, es2);
        st->ms_numeric = tm.ms(en0, en1);
        st->ms_main = st->ms_numeric;
        st->mac1 = mac1;
        st->nnz_out = nnz;
        st->main_kernel_launches = launches;
    }
    (void)timing;
    return 0;
}

// --------------------------------------------------------------------------
// R*A*P, direct form in one pass: rows computed once into scratch slots, then
// scanned and gathered into the output
// --------------------------------------------------------------------------
struct FusedCfg {
    int G, H1, H2;
};

// bins by the upper bound of the R*A row: bin b takes rows with ub <= v_b and
// runs them with a stage-1 table of H1 = v_b >> h1_shift slots (with shift 0
// the table cannot fill; a full table sends the row to the retry list) and a
// stage-2 table of max(64, G, H1 / h2_div) slots; 32-lane groups up to
// H1 = subwave_h, wavefronts up to 512 slots, workgroups beyond.  Rows above
// the largest bin whose tables fit in LDS go to the long-row kernels.
static void fused_bins(size_t max_lds, bool mixed, const Tuning &tu, BinThresholds &thr,
                       std::vector<FusedCfg> &cfg)
{
    thr.nb = 0;
    cfg.clear();
    for (long long v = 64; thr.nb < kMaxBins - 1; v *= 2) {
        const int H1 = (int)std::max<long long>(64, tu.h1_shift >= 0 ? v >> tu.h1_shift : v << -tu.h1_shift);
        const int G = H1 <= tu.subwave_h ? 32 : H1 <= 512 ? 64 : 256;
        // both tables hold a whole number of group-widths (the compaction and
        // clearing loops step through them G slots at a time)
        const int H2 = std::max(std::max(64, G), H1 / tu.h2_div);
        if ((mixed ? mixed_lds(G, H1, H2) : fused_lds(G, H1, H2)) > max_lds) break;
        thr.v[thr.nb++] = v;
        cfg.push_back({G, H1, H2});
    }
    cfg.push_back({0, 0, 0});
}

// The bin the mean row bound nnz(R)/M * nnz(A)/A.M falls in, -1 beyond the
// bins: in a mesh operator most rows fall in it.
static int likely_bin(const BinThresholds &thr, const sp3mm_dmat &R, const sp3mm_dmat &A)
{
    if (R.M == 0 || A.M == 0) return -1;
    const double est = (double)R.nnz / (double)R.M * ((double)A.nnz / (double)A.M);
    int b = 0;
    while (b < thr.nb && (double)thr.v[b] < est) b++;
    return b >= thr.nb ? -1 : b;
}

// FusedPlan: the one-pass direct pipeline.  build() runs it eagerly, reading
// counters back to size bins, scratch slots and the output; record() enqueues
// the same kernels for the sizes found by build() with no host
// synchronisation, so it can be captured into a HIP graph and replayed.
struct FusedPlan {
    sp3mm_dmat R{}, A{}, P{};
    bool mixed = false;  // FP32-pair accumulation kernel
    bool f32 = false;    // FP32 operands (mixed, or FP64 accumulation of FP32 inputs)
    double guard = 1e-2;
    double *rowsum = nullptr;
    Tuning tu{};
    BinThresholds thr{};
    std::vector<FusedCfg> cfg;
    std::vector<int> cnt, fcnt;
    std::vector<int64_t> soff, nbase;
    int nlong = 0, nretry = 0, nspill = 0;
    bool built = false;  // build() has read the retry and spill counts
    bool spec = false;   // tail launched before the retry count is known
    // the speculative tail leaves rows whose tables filled for later: their
    // window passes (count, then numeric) run only if some row was retried
    bool defer_retry = false;
    int64_t ccap = 0;    // entries allocated for C.col / C.val
    // Output arrays supplied by the caller (C.rp, C.col, C.val set before
    // build(), ext_cap entries in C.col / C.val); the plan then allocates no
    // output and keeps no long-row slots (scratch comes from an Arena).
    bool ext_out = false;
    int64_t ext_cap = 0;
    // per-phase device times (ms_ub, ms_num, ms_scan): HIP events between the
    // phases, which add gaps to the stream; off unless asked for
    bool detail = false;
    int mk(Timer &tm) const { return detail ? tm.mark() : -1; }
    // One-off FP64 products may be reassociated: build() stops after the
    // binning and sets `reassoc` when most rows are beyond the one-wavefront
    // tables and the output is wider than an LDS window (the caller then runs
    // R*(A*P) as two SpGEMMs; see wants_reassoc)
    bool reassoc_ok = false, reassoc = false;

    // (re)allocates the output arrays for `cap` entries
    int alloc_out(Context *c, int64_t cap)
    {
        if (ext_out) {
            if (cap > ext_cap) {
                set_error("output capacity of %lld entries is below the product's %lld nonzeros",
                          (long long)ext_cap, (long long)cap);
                return -1;
            }
            ccap = ext_cap;
            return 0;
        }
        c->pool.release(C.col);
        c->pool.release(C.val);
        C.col = static_cast<int32_t *>(c->pool.alloc((size_t)(cap ? cap : 1) * 4));
        C.val = c->pool.alloc((size_t)(cap ? cap : 1) * 8);
        ccap = cap;
        if (!C.col || !C.val) {
            set_error("output allocation (%lld entries) failed", (long long)cap);
            return -1;
        }
        return 0;
    }
    int64_t nnz = 0, stot = 0;
    long long mac1 = 0;
    int e_end = -1;  // event behind the final read-back of build() (timer of build())
    int slog1 = 0, slog2 = 0, dlog1 = 0, dlog2 = 0;
    Buf ctl, ub, rown, lists, retry, slotn, fix, tiles, scol, sval, binscr, spill;
    // long rows in one pass (k_window_direct slot mode): a scratch slot of
    // lcap entries per row of the long bin, gathered after the scan
    Buf lslotn, lcol, lval;
    int lcap = 0;
    bool lslots = false;
    DenseDirectWs sws;  // global dense workspace of the spilled rows
    sp3mm_dmat C{};
    double ms_ub = 0, ms_num = 0, ms_scan = 0;

    Ctl *d() const { return ctl.as<Ctl>(); }
    SpOperand r() const { return operand(R); }
    SpOperand a() const { return operand(A); }
    SpOperand p() const { return operand(P); }

    // A bin launched before its population is known (`pred`, the bin the mean
    // row bound falls in) owns scratch slots for all M rows; the others share
    // scol/sval (offsets soff) and slotn (offsets nbase), sized by population.
    int pred = -1;
    Buf pcol, pval, pslotn;
    Buf mlen;             // per-workgroup longest rows of R and A (one-bin test)
    bool onebin = false;  // the binning put every row in bin pred
    int32_t *slot_col(int i) const { return i == pred ? pcol.as<int32_t>() : scol.as<int32_t>() + soff[i]; }
    double *slot_val(int i) const { return i == pred ? pval.as<double>() : sval.as<double>() + soff[i]; }
    int32_t *slot_n(int i) const { return i == pred ? pslotn.as<int32_t>() : slotn.as<int32_t>() + nbase[i]; }
    int slot_rows(int i) const { return i == pred ? (int)R.M : cnt[i]; }

    // The bin whose launch can go ahead of the read-back of the populations:
    // the one the mean row bound nnz(R)/M * nnz(A)/A.M falls in, for FP64
    // plans with their own output.  Its launch reads its population from
    // Ctl::pred_count, which the binning sets to 0 when the product is to be
    // reassociated, so that launch then ends at once.
    int predict_bin(Context *c) const
    {
        if (mixed || ext_out || R.M >= INT32_MAX) return -1;
        (void)c;
        return likely_bin(thr, R, A);
    }

    void launch_main(Context *c, hipStream_t s, int i, bool early = false)
    {
        const FusedCfg &f = cfg[i];
        const int64_t M = R.M;
        const int *count = early ? &d()->pred_count : d()->counts + i;
        if (mixed)
            launch_direct_mixed(f.G, f.H1, f.H2, slog1, slog2, (float)guard, r(), a(), p(),
                                lists.as<int32_t>() + (int64_t)i * M, d()->counts + i, cnt[i], rown.as<int32_t>(),
                                slotn.as<int32_t>() + nbase[i], scol.as<int32_t>() + soff[i],
                                sval.as<double>() + soff[i], rowsum, retry.as<int32_t>(), &d()->nretry,
                                fix.as<int32_t>() + nbase[i], d()->counts2 + i, &d()->err, max_blocks(c), s, tu);
        else
            launch_direct_fused(f.G, f.H1, f.H2, slog1, slog2, r(), a(), p(), lists.as<int32_t>() + (int64_t)i * M,
                                count, slot_rows(i), rown.as<int32_t>(), slot_n(i), slot_col(i), slot_val(i),
                                retry.as<int32_t>(), &d()->nretry, &d()->err, max_blocks(c), s, f32, tu);
    }

    // export_ctl: the first gather also copies the control block to the host
    // mirror as it starts (every count in it is final by then, and a gather
    // can only add an overflow bit, which nnz <= ccap rules out); returns
    // whether it does
    bool launch_tail(Context *c, hipStream_t s, bool export_ctl = false)
    {
        const int64_t M = R.M;
        bool exported = false;
        for (int i = 0; i < thr.nb; i++) {
            if (cnt[i] == 0) continue;
            const bool x = export_ctl && !exported && c->host_small_dev;
            launch_gather_rows(lists.as<int32_t>() + (int64_t)i * M, d()->counts + i, cnt[i], slot_n(i), cfg[i].H2,
                               slot_col(i), slot_val(i), C.rp, C.col, static_cast<double *>(C.val), ccap,
                               &d()->err, max_blocks(c), s, x ? d() : nullptr, x ? c->host_small_dev : nullptr,
                               (int)(sizeof(Ctl) / 4));
            exported = exported || x;
        }
        launch_long(c, s, true);
        return exported;
    }

    // rows beyond the hash tables (long rows of the last bin, and rows whose
    // tables filled): the LDS window kernel, and the global dense kernel for
    // the rows whose windows do not fit in LDS.  num = false: count pass.
    void launch_long(Context *c, hipStream_t s, bool num)
    {
        const int64_t M = R.M;
        const size_t wl = c->max_lds - 1024;
        double *cv = num ? static_cast<double *>(C.val) : nullptr;
        int32_t *rn = num ? nullptr : rown.as<int32_t>();
        double *rs = num && mixed ? rowsum : nullptr;
        const int64_t cap = num ? ccap : 0;
        int32_t *llist = lists.as<int32_t>() + (int64_t)thr.nb * M;
        if (nlong > 0 && lslots) {
            // one pass into the slots at count time, a gather in the tail
            if (!num)
                launch_window_direct(true, std::min(nlong, c->num_cu), wl, r(), a(), p(), llist,
                                     d()->counts + thr.nb, dlog1, spill.as<int32_t>(), &d()->nspill,
                                     rown.as<int32_t>(), nullptr, nullptr, nullptr, mixed ? rowsum : nullptr, 0,
                                     &d()->err, s, f32, lslotn.as<int32_t>(), lcol.as<int32_t>(), lval.as<double>(),
                                     nlong);
            else
                launch_gather_rows(llist, d()->counts + thr.nb, nlong, lslotn.as<int32_t>(), lcap,
                                   lcol.as<int32_t>(), lval.as<double>(), C.rp, C.col, static_cast<double *>(C.val),
                                   ccap, &d()->err, max_blocks(c), s);
        } else if (nlong > 0) {
            launch_window_direct(num, std::min(nlong, c->num_cu), wl, r(), a(), p(), llist, d()->counts + thr.nb,
                                 dlog1, spill.as<int32_t>(), &d()->nspill, rn, num ? C.rp : nullptr,
                                 num ? C.col : nullptr, cv, rs, cap, &d()->err, s, f32);
        }
        if (!defer_retry && (nretry > 0 || (!num && !built) || (num && spec)))
            launch_window_direct(num, c->num_cu, wl, r(), a(), p(), retry.as<int32_t>(), &d()->nretry, dlog1,
                                 spill.as<int32_t>(), &d()->nspill, rn, num ? C.rp : nullptr, num ? C.col : nullptr,
                                 cv, rs, cap, &d()->err, s, f32);
        if (nspill > 0)
            launch_dense_direct(num, dense_blocks(c, nspill), r(), a(), p(), spill.as<int32_t>(), &d()->nspill,
                                sws.ws, dlog1, dlog2, nullptr, rn, num ? C.rp : nullptr, num ? C.col : nullptr, cv,
                                &d()->err, s, f32, rs, c->max_lds);
    }

    void launch_fixups(Context *c, hipStream_t s, bool all_bins)
    {
        if (!mixed) return;
        for (int i = 0; i < thr.nb; i++) {
            if (cnt[i] == 0 || (!all_bins && fcnt[i] == 0)) continue;
            const FusedCfg &f = cfg[i];
            launch_direct(f.G, true, f.H1, f.H2, slog1 > ilog2(f.G) ? ilog2(f.G) : slog1, slog2, r(), a(), p(),
                          fix.as<int32_t>() + nbase[i], d()->counts2 + i, all_bins ? cnt[i] : fcnt[i], nullptr,
                          nullptr, C.rp, C.col, static_cast<double *>(C.val), nullptr, nullptr, &d()->err,
                          max_blocks(c), s, true);
        }
    }

    // Bins of at most kMergeRows rows are folded into the next larger non-empty
    // bin while the binning writes its lists (its tables hold their upper
    // bounds too): a launch of a handful of rows costs a whole row's latency,
    // while the appended rows fill groups of the larger launch.  d()->counts
    // holds the merged populations, d()->rawc the populations before.
    static constexpr int kMergeRows = 1024;

    // The binning clears the control block first; with seq > 0 its scatter
    // kernel also exports the block to the coherent host block with that
    // sequence number (see wait_ctl).
    void launch_binning(Context *c, hipStream_t s, unsigned long long seq = 0)
    {
        void *host = seq ? c->poll_dev : nullptr;
        auto *flag = seq ? reinterpret_cast<unsigned long long *>(c->poll_dev + kPollFlagOff) : nullptr;
        // rows of R whose columns spread over more than a quarter of A's rows
        // count as unstructured (see wants_reassoc)
        const int span_lim = reassoc_ok ? (int)std::max<int64_t>(1, R.N / 4) : 0;
        int big_from = thr.nb;
        for (int i = 0; i < thr.nb; i++)
            if (cfg[i].G > kWave) {
                big_from = i;
                break;
            }
        // one-bin mode: when the longest row of R times the longest row of A
        // fits the likely bin's bound, every row goes to that bin unmeasured
        OneBin ob{};
        const bool one = pred >= 0 && cfg[pred].G <= kWave && mlen.p;
        if (one) {
            ob.part = mlen.as<int>();
            ob.nblk = kMaxLenBlocks;
            ob.bound = thr.v[pred];
            ob.u = thr.v[pred];
            ob.word = (int)(offsetof(Ctl, onebin) / 4);
        }
        launch_ub_bin(r(), a(), ub.as<int64_t>(), rown.as<int32_t>(), lists.as<int32_t>(), d()->rawc, thr,
                      d()->totals, binscr.as<int>(), s, kMergeRows, d()->counts, d(), host,
                      (int)(sizeof(Ctl) / 4), flag, seq, span_lim, &d()->pred_count, pred, big_from,
                      reassoc_possible(c), one ? &ob : nullptr, R.M > 0 && R.nnz >= (int64_t)kWaveUbRow * R.M);
    }

    // R*(A*P) instead of the direct forwarding: when at least half of R's
    // rows are long rows or rows of the multi-wavefront bins, at least half
    // of R's rows are unstructured (their columns spread over more than a
    // quarter of A's rows: random aggregates rather than mesh neighbourhoods)
    // and P's columns do not fit one LDS accumulator of the long-row window
    // kernel.  Such rows forward thousands of R*A entries into short rows of
    // P that land all over the output, and past one LDS window every window
    // re-walks all of them; in R*(A*P) a row walks nnz(R_r) rows of A*P
    // instead, each a long contiguous run per window (k_spgemm_window), and
    // the rows of A*P fit the hash tables.  Mesh-structured products (whose
    // output rows stay narrow) keep the direct forwarding.  Same pattern
    // (structural products are associative), FP64 sums in another order.
    unsigned long long wide = 0;  // unstructured rows of R (binning totals[2])
    bool reassoc_possible(Context *c) const
    {
        return reassoc_ok && !mixed && !f32 && P.N > (int64_t)window_cap(c->max_lds - 1024);
    }
    bool wants_reassoc(Context *c) const
    {
        if (mixed || f32) return false;
        int64_t big = nlong;
        for (int i = 0; i < thr.nb; i++)
            if (cfg[i].G > kWave) big += cnt[i];
        return 2 * big >= R.M && 2 * (int64_t)wide >= R.M && R.M > 0 &&
               P.N > (int64_t)window_cap(c->max_lds - 1024);
    }

    int build(Context *c, Timer &tm)
    {
        hipStream_t s = c->stream;
        const int64_t M = R.M;
        fused_bins(c->max_lds, mixed, tu, thr, cfg);
        // the binning's buffers first: its kernels start while the rest of
        // the scratch is allocated
        ctl = Buf(sizeof(Ctl));
        ub = Buf((size_t)(M + 1) * 8);
        rown = Buf((size_t)(M + 1) * 4);
        lists = Buf((size_t)(thr.nb + 1) * (M ? M : 1) * 4);
        binscr = Buf((size_t)bin_scratch_ints(M) * 4);
        if (!ctl.p || !ub.p || !rown.p || !lists.p || !binscr.p) {
            set_error("scratch allocation failed");
            return -1;
        }
        Ctl h;
        slog1 = sub_log(A, 64);
        slog2 = sub_log(P, 64);
        dlog1 = sub_log(A, kDenseThreads);
        dlog2 = sub_log(P, kDenseThreads);
        // the likely bin starts before the populations are read back, so the
        // GPU does not idle through the read-back
        pred = predict_bin(c);
        if (pred >= 0) {
            mlen = Buf((size_t)kMaxLenBlocks * 2 * 4);
            if (!mlen.p) (void)hipGetLastError();  // binning measures every row then
        }
        const int e0 = mk(tm);
        // the binning clears the control block, and its scatter kernel sends
        // the populations to the host while the likely bin already runs
        SP3MM_HIP(hipGetLastError());
        const unsigned long long seq = c->poll_host ? ++c->poll_seq : 0;
        launch_binning(c, s, seq);
        const int eu = mk(tm);
        retry = Buf((size_t)(M ? M : 1) * 4);
        slotn = Buf((size_t)(M ? M : 1) * 4);
        fix = Buf((size_t)(M ? M : 1) * 4);
        tiles = Buf((size_t)scan_tiles(M) * 8);
        spill = Buf((size_t)(M ? M : 1) * 4);
        if (!retry.p || !slotn.p || !fix.p || !tiles.p || !spill.p) {
            set_error("scratch allocation failed");
            return -1;
        }
        // the likely bin's slots; without them it runs with the other bins
        // after the read-back (its rows are listed like any bin's)
        if (pred >= 0) {
            const size_t cap = (size_t)M * cfg[pred].H2;
            pcol = Buf(cap * 4);
            pval = Buf(cap * 8);
            pslotn = Buf((size_t)M * 4);
            if (!pcol.p || !pval.p || !pslotn.p) {
                (void)hipGetLastError();
                pcol = Buf();
                pval = Buf();
                pslotn = Buf();
                pred = -1;
            }
        }
        int en0 = -1;
        if (pred >= 0) {
            en0 = mk(tm);
            launch_main(c, s, pred, true);
        }
        if (wait_ctl(c, d(), seq, h)) return -1;
        onebin = h.onebin != 0;
        mac1 = onebin ? -1 : (long long)h.totals[0];  // one-bin mode: the bounds were not summed
        wide = h.totals[2];
        cnt.assign(h.counts, h.counts + thr.nb + 1);
        nlong = cnt[thr.nb];
        if (reassoc_ok && wants_reassoc(c)) {
            reassoc = true;
            return 0;
        }
        soff.assign(thr.nb + 1, 0);
        nbase.assign(thr.nb + 1, 0);
        int64_t ntot = 0, sown = 0;
        stot = 0;
        for (int i = 0; i < thr.nb; i++) {
            stot += (int64_t)cnt[i] * cfg[i].H2;  // slot entries of every row (bounds nnz)
            if (i == pred) continue;
            soff[i] = sown;
            nbase[i] = ntot;
            sown += (int64_t)cnt[i] * cfg[i].H2;
            ntot += cnt[i];
        }
        scol = Buf((size_t)(sown ? sown : 1) * 4);
        sval = Buf((size_t)(sown ? sown : 1) * 8);
        if (!scol.p || !sval.p) {
            set_error("scratch allocation (%lld entries) failed", (long long)stot);
            return -1;
        }
        // long rows in one pass when their slots stay within 1 GiB
        lcap = window_cap(c->max_lds - 1024);
        lslots = false;
        if (nlong > 0 && !ext_out && (size_t)nlong * lcap * 12 <= ((size_t)1 << 30)) {
            lslotn = Buf((size_t)nlong * 4);
            lcol = Buf((size_t)nlong * lcap * 4);
            lval = Buf((size_t)nlong * lcap * 8);
            lslots = lslotn.p && lcol.p && lval.p;
        }
        if (en0 < 0) en0 = mk(tm);
        for (int i = 0; i < thr.nb; i++)
            if (cnt[i] && i != pred) launch_main(c, s, i);
        const int en1 = mk(tm);
        // long and retried rows: count pass of the window kernel (it lists the
        // rows whose windows exceed LDS), then the global dense count for those
        // Without long rows every row that stayed on the hash tables fits its
        // scratch slot, so the slot total is a likely output size: allocate it
        // and run the tail without reading the counters back.  Every writer
        // checks the capacity; if a row was retried, a retried row outgrew the
        // output, or a row spilled to the global dense kernel, the exact path
        // below redoes the tail.
        const bool speculate = nlong == 0 && !mixed && stot > 0 && stot < INT32_MAX;
        defer_retry = speculate;
        launch_long(c, s, false);
        if (ext_out) {
            C.M = M;
            C.N = P.N;
            C.vtype = SP3MM_F64;
            C.owned = 0;
            C.rp32 = nullptr;
        } else if (dmat_alloc(C, M, P.N, 0, SP3MM_F64)) {
            return -1;
        }
        launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d()->nnz, s);
        if (speculate) {
            if (ext_out)
                ccap = ext_cap;  // the caller's arrays; a product that outgrows them is caught below
            else if (alloc_out(c, stot))
                return -1;
            spec = true;
            const bool early = launch_tail(c, s, true);
            spec = false;
            defer_retry = false;
            if (early) {
                // the control block is in the host mirror once the stream ends
                SP3MM_HIP(hipGetLastError());
                e_end = tm.mark();
                SP3MM_HIP(spin_stream(s));
                memcpy(&h, c->host_small, sizeof(Ctl));
            } else if (read_ctl(c, d(), h, &tm, &e_end)) {
                return -1;
            }
            if (!h.err && h.nspill == 0 && h.nretry == 0 && h.nnz <= ccap) {
                nretry = h.nretry;
                nspill = 0;
                built = true;
                fcnt.assign(h.counts2, h.counts2 + thr.nb + 1);
                nnz = h.nnz;
                C.nnz = nnz;
                const int es = mk(tm);
                ms_ub = tm.ms(e0, eu);
                ms_num = tm.ms(en0, en1);
                ms_scan = tm.ms(en1, es);
                return 0;
            }
            if (h.err & ~kErrOverflow) {
                set_error("direct product: %s", err_text(h.err));
                return -1;
            }
            SP3MM_HIP(hipMemsetAsync(&d()->err, 0, sizeof(int), s));
            if (h.nretry > 0) {
                // the retried rows' lengths (window count pass), then the scan again
                nretry = h.nretry;
                launch_long(c, s, false);
                launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d()->nnz, s);
            }
        }
        if (read_ctl(c, d(), h)) return -1;
        nretry = h.nretry;
        nspill = h.nspill;
        built = true;
        if (nspill > 0 && !h.err) {
            if (sws.alloc(dense_blocks(c, nspill), A.N, P.N, true, s)) return -1;
            launch_dense_direct(false, dense_blocks(c, nspill), r(), a(), p(), spill.as<int32_t>(), &d()->nspill,
                                sws.ws, dlog1, dlog2, nullptr, rown.as<int32_t>(), nullptr, nullptr, nullptr,
                                &d()->err, s, f32, nullptr, c->max_lds);
            launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d()->nnz, s);
            if (read_ctl(c, d(), h)) return -1;
        }
        const int es = mk(tm);
        if (h.err) {
            set_error("direct product: %s", err_text(h.err));
            return -1;
        }
        fcnt.assign(h.counts2, h.counts2 + thr.nb + 1);
        nnz = h.nnz;
        C.nnz = nnz;
        if (alloc_out(c, nnz)) return -1;
        if (nnz >= INT32_MAX) {
            c->pool.release(C.rp32);
            C.rp32 = nullptr;
        }
        launch_tail(c, s);
        launch_fixups(c, s, false);
        if (read_ctl(c, d(), h, &tm, &e_end)) return -1;
        if (h.err) {
            set_error("direct product: %s", err_text(h.err));
            return -1;
        }
        ms_ub = tm.ms(e0, eu);
        ms_num = tm.ms(en0, en1);
        ms_scan = tm.ms(en1, es);
        return 0;
    }

    // the whole pipeline for the sizes found by build(), no host synchronisation
    void record(Context *c, hipStream_t s)
    {
        const int64_t M = R.M;
        launch_binning(c, s);  // clears the control block first
        for (int i = 0; i < thr.nb; i++)
            if (cnt[i]) launch_main(c, s, i);
        launch_long(c, s, false);
        launch_scan(rown.as<int32_t>(), M, C.rp, C.rp32, tiles.as<long long>(), &d()->nnz, s);
        launch_tail(c, s);
        launch_fixups(c, s, true);
        BinCounts want{};
        want.n = thr.nb + 1;
        for (int i = 0; i <= thr.nb; i++) want.c[i] = cnt[i];
        launch_check_nnz(&d()->nnz, nnz, &d()->nretry, nretry, &d()->nspill, nspill, &d()->err, s, d()->counts,
                         &want);
    }

    // after a replay: did the device see the sizes this plan was built for?
    int check(Context *c)
    {
        Ctl h;
        if (read_ctl(c, d(), h)) return -1;
        if (h.err) {
            set_error("replayed product: %s", err_text(h.err));
            return -1;
        }
        return 0;
    }

    void release_all()
    {
        dmat_release(C);
    }
};

static int rap_direct_fused(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C,
                            sp3mm_stats *st, int prec, double guard, double *rowsum, bool timing)
{
    Context *c = ctx();
    Timer tm(c->stream, true);
    const int e0 = tm.mark();
    FusedPlan plan;
    plan.R = R;
    plan.A = A;
    plan.P = P;
    plan.mixed = prec == SP3MM_PREC_MIXED;
    plan.f32 = prec != SP3MM_PREC_FP64;
    plan.tu = c->tune;
    plan.guard = guard;
    plan.rowsum = rowsum;
    plan.reassoc_ok = prec == SP3MM_PREC_FP64 && !rowsum;
    plan.detail = timing;
    if (plan.build(c, tm)) {
        plan.release_all();
        return -1;
    }
    if (plan.reassoc) {
        // R*(A*P): the one-pass SpGEMM twice, A*P kept in HBM
        sp3mm_stats s1{}, s2{};
        sp3mm_dmat AP;
        memset(&AP, 0, sizeof AP);
        const int en0 = tm.mark();
        int rc = spgemm_any(A, P, AP, &s1, nullptr, nullptr);
        if (!rc) rc = spgemm_any(R, AP, C, &s2, nullptr, nullptr);
        const int64_t mid = AP.nnz;
        dmat_release(AP);
        if (rc) {
            dmat_release(C);
            return -1;
        }
        const int e1 = tm.mark();
        (void)hipEventSynchronize(tm.ev[e1]);
        if (st) {
            memset(st, 0, sizeof *st);
            st->ms_total = tm.ms(e0, e1);
            st->ms_ub = tm.ms(e0, en0);
            st->ms_numeric = tm.ms(en0, e1);
            st->ms_main = st->ms_numeric;
            st->mac1 = plan.mac1;
            st->nnz_mid = mid;
            st->nnz_out = C.nnz;
            st->main_kernel_launches = 2;
            st->guarded_entries = plan.nlong;
            for (int i = 0; i < plan.thr.nb; i++)
                if (plan.cfg[i].G > kWave) st->guarded_entries += plan.cnt[i];
            st->reassociated = 1;
        }
        return 0;
    }
    const int e1 = plan.e_end >= 0 ? plan.e_end : tm.mark();
    (void)spin_event(tm.ev[e1]);
    C = plan.C;
    memset(&plan.C, 0, sizeof plan.C);
    if (st) {
        memset(st, 0, sizeof *st);
        st->ms_total = tm.ms(e0, e1);
        st->ms_ub = plan.ms_ub;
        st->ms_numeric = plan.ms_num;
        st->ms_scan = plan.ms_scan;
        st->ms_main = plan.ms_num;
        st->mac1 = plan.mac1;
        st->nnz_out = plan.nnz;
        long long nf = 0;
        int launches = 0;
        for (int i = 0; i < plan.thr.nb; i++) {
            nf += plan.fcnt[i];
            launches += plan.cnt[i] ? 1 : 0;
        }
        st->main_kernel_launches = launches;
        st->guarded_rows = nf;
        st->guarded_entries = plan.nretry + plan.nlong;  // rows off the hash-table path
        st->spilled_rows = plan.nspill;
    }
    return 0;
}

int rap_direct(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C, sp3mm_stats *st,
               bool timing, bool symbolic_only)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (check_pair(R, A) || check_pair(A, P)) return -1;
    static const bool twopass = getenv("SP3MM_DIRECT_TWOPASS") != nullptr;
    if (symbolic_only || twopass) return rap_direct_twopass(R, A, P, C, st, timing, symbolic_only);
    return rap_direct_fused(R, A, P, C, st, SP3MM_PREC_FP64, 0.0, nullptr, timing);
}

int rap_direct_mixed(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, sp3mm_dmat &C, sp3mm_stats *st,
                     int prec, double guard, double *rowsum, bool timing)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (R.N != A.M || A.N != P.M) {
        set_error("inner dimensions differ");
        return -1;
    }
    if (!A.rp32 || !P.rp32) {
        set_error("right operand has >= 2^31 nonzeros");
        return -1;
    }
    if (R.vtype != SP3MM_F32 || A.vtype != SP3MM_F32 || P.vtype != SP3MM_F32) {
        set_error("the mixed-precision product needs FP32 operands (sp3mm_dmat_convert)");
        return -1;
    }
    if (prec == SP3MM_PREC_MIXED64 && rowsum) {
        set_error("row sums are produced by the FP32-pair form only");
        return -1;
    }
    return rap_direct_fused(R, A, P, C, st, prec, guard > 0 ? guard : 1e-2, rowsum, timing);
}

// The direct FP64 product of rap_direct without a single device allocation:
// the output goes to the caller's arrays (crp: M+1 offsets, ccol/cval: cap
// entries) and every scratch buffer of the plan (control block, bounds, bin
// lists, scratch slots, dense workspaces) is carved from this device's fixed
// scratch array.  The array keeps the previous call's contents; the plan
// writes every buffer it reads, and the control block is cleared first.  No
// HIP events are created.  Calls are serialised by the context mutex and
// each returns after its stream work has finished.
int rap_direct_noalloc(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int64_t *crp, int32_t *ccol,
                       double *cval, int64_t cap, int64_t &nnz, sp3mm_stats *st)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    nnz = -1;
    if (check_pair(R, A) || check_pair(A, P)) return -1;
    if (R.vtype != SP3MM_F64 || A.vtype != SP3MM_F64 || P.vtype != SP3MM_F64) {
        set_error("rap_noalloc: FP64 operands only");
        return -1;
    }
    if (!crp || (cap > 0 && (!ccol || !cval)) || cap < 0) {
        set_error("rap_noalloc: null output array or negative capacity");
        return -1;
    }
    if (!c->noalloc_base) {
        set_error("rap_noalloc: the fixed device scratch of device %d could not be resolved", c->device);
        return -1;
    }
    Arena arena;
    arena.base = c->noalloc_base;
    arena.size = SP3MM_NOALLOC_SCRATCH_BYTES;
    ArenaScope scope(&arena);
    Timer tm(c->stream, false);
    FusedPlan plan;
    plan.R = R;
    plan.A = A;
    plan.P = P;
    plan.tu = c->tune;
    plan.ext_out = true;
    plan.ext_cap = cap;
    plan.C.rp = crp;
    plan.C.col = ccol;
    plan.C.val = cval;
    const int rc = plan.build(c, tm);
    if (rc) {
        if (arena.refused) {
            char why[600];
            snprintf(why, sizeof why, "%s", last_error_text());
            set_error("rap_noalloc: the product needs more than the %lld bytes of fixed device scratch (%lld in use "
                      "when a further %lld were asked for): %s",
                      (long long)SP3MM_NOALLOC_SCRATCH_BYTES, (long long)arena.used, (long long)arena.refused, why);
        }
        if (plan.built && plan.nnz > cap) nnz = plan.nnz;  // the capacity a retry needs
        (void)hipStreamSynchronize(c->stream);
        return -1;
    }
    // build() ends with a read-back of the control block: the stream is idle
    nnz = plan.nnz;
    if (st) {
        memset(st, 0, sizeof *st);
        st->mac1 = plan.mac1;
        st->nnz_out = plan.nnz;
        int launches = 0;
        for (int i = 0; i < plan.thr.nb; i++) launches += plan.cnt[i] ? 1 : 0;
        st->main_kernel_launches = launches;
        st->guarded_entries = plan.nretry + plan.nlong;
        st->spilled_rows = plan.nspill;
    }
    return 0;
}

// --------------------------------------------------------------------------
// setup operations: transpose, smoothed-aggregation prolongator
// --------------------------------------------------------------------------
// AT = A^T.  build() forms it (column counts, scan, scatter, per-row sort);
// with `replayable` it also keeps the permutation that maps AT's entries to
// A's, so that record() recomputes AT's values for new values of A (same
// pattern) with one gather, and can be enqueued without host synchronisation
// (e.g. into a HIP graph).
struct TransposePlan {
    sp3mm_dmat A{}, T{};
    Buf perm;

    int build(Context *c, const sp3mm_dmat &a, bool replayable)
    {
        A = a;
        const bool f32 = A.vtype == SP3MM_F32;
        hipStream_t s = c->stream;
        if (dmat_alloc(T, A.N, A.M, A.nnz, A.vtype)) return -1;
        const size_t nz = (size_t)(A.nnz ? A.nnz : 1);
        Buf cnt((size_t)(A.N + 1) * 4), cursor((size_t)(A.N + 1) * 4), tiles((size_t)scan_tiles(A.N) * 8),
            total(8), tcol(nz * 4), tval(nz * (f32 ? 4 : 8)), tsrc(replayable ? nz * 4 : 0);
        if (replayable) perm = Buf(nz * 4);
        if (!cnt.p || !cursor.p || !tiles.p || !total.p || !tcol.p || !tval.p ||
            (replayable && (!tsrc.p || !perm.p))) {
            set_error("transpose: scratch allocation failed");
            dmat_release(T);
            return -1;
        }
        launch_fill_i32(cnt.as<int32_t>(), A.N + 1, 0, s);
        launch_fill_i32(cursor.as<int32_t>(), A.N + 1, 0, s);
        launch_col_count(A.col, A.nnz, cnt.as<int32_t>(), s);
        launch_scan(cnt.as<int32_t>(), A.N, T.rp, T.rp32, tiles.as<long long>(), total.as<int64_t>(), s);
        launch_transpose_scatter(A.rp, A.col, A.val, f32, A.M, A.nnz, T.rp, cursor.as<int32_t>(),
                                 tcol.as<int32_t>(), tval.p, replayable ? tsrc.as<int32_t>() : nullptr, s);
        launch_sort_rows(T.rp, T.M, tcol.as<int32_t>(), tval.p, replayable ? tsrc.as<int32_t>() : nullptr, f32,
                         T.col, T.val, replayable ? perm.as<int32_t>() : nullptr, s);
        SP3MM_HIP(hipGetLastError());
        // the scratch buffers are released at return: wait for the kernels
        SP3MM_HIP(hipStreamSynchronize(s));
        return 0;
    }

    // AT's values from A's current values (build() with `replayable` first)
    void record(Context *, hipStream_t s)
    {
        launch_gather_vals(A.val, perm.as<int32_t>(), A.nnz, A.vtype == SP3MM_F32, T.val, s);
    }
};

int transpose(const sp3mm_dmat &A, sp3mm_dmat &T)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    TransposePlan tp;
    if (tp.build(c, A, false)) return -1;
    T = tp.T;
    return 0;
}

// S = I - omega D^-1 A on A's pattern plus any missing diagonal.  build()
// sizes it (count pass, scan) and fills it; record() refills S's columns and
// values for new values of A (same pattern: the offsets stay), without host
// synchronisation.  The zero-diagonal flag is in the plan's control block.
struct SaPlan {
    sp3mm_dmat A{}, S{};
    double omega = 0;
    int mode = 0;
    Buf ctl;
    Ctl *d() const { return ctl.as<Ctl>(); }

    int build(Context *c, const sp3mm_dmat &a, double om, int md)
    {
        A = a;
        omega = om;
        mode = md;
        const int64_t M = A.M;
        hipStream_t s = c->stream;
        ctl = Buf(sizeof(Ctl));
        Buf len((size_t)(M + 1) * 4), tiles((size_t)scan_tiles(M) * 8);
        if (!ctl.p || !len.p || !tiles.p || A.nnz + M >= INT32_MAX) {
            set_error("prolongator: scratch allocation failed (or A too large)");
            return -1;
        }
        if (dmat_alloc(S, M, M, A.nnz + M, SP3MM_F64)) return -1;
        launch_fill_i32(reinterpret_cast<int32_t *>(d()), (int64_t)(sizeof(Ctl) / 4), 0, s);
        launch_sa_count(A.rp, A.col, M, A.nnz, len.as<int32_t>(), s);
        launch_scan(len.as<int32_t>(), M, S.rp, S.rp32, tiles.as<long long>(), &d()->nnz, s);
        record(c, s);
        Ctl h;
        if (read_ctl(c, d(), h)) return -1;
        if (h.err) {
            set_error("prolongator: %s (zero diagonal)", err_text(h.err));
            return -1;
        }
        S.nnz = h.nnz;
        return 0;
    }

    void record(Context *, hipStream_t s)
    {
        launch_fill_i32(&d()->err, 1, 0, s);
        launch_sa_fill(A.rp, A.col, static_cast<const double *>(A.val), A.M, A.nnz, omega, mode, S.rp, S.col,
                       static_cast<double *>(S.val), &d()->err, s);
    }
};

int sa_prolongator(const sp3mm_dmat &A, const int64_t *agg, int64_t nagg, double omega, int mode, sp3mm_dmat &P)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (A.M != A.N) {
        set_error("prolongator: A must be square (%lld x %lld)", (long long)A.M, (long long)A.N);
        return -1;
    }
    if (A.vtype != SP3MM_F64) {
        set_error("prolongator: A must hold FP64 values");
        return -1;
    }
    if (nagg <= 0 || nagg >= INT32_MAX || (mode != 0 && mode != 1)) {
        set_error("prolongator: bad aggregate count or diagonal mode");
        return -1;
    }
    hipStream_t s = c->stream;
    const int64_t M = A.M;
    Buf ctl(sizeof(Ctl)), len((size_t)(M + 1) * 4), tiles((size_t)scan_tiles(M) * 8), dagg((size_t)(M ? M : 1) * 8);
    if (!ctl.p || !len.p || !tiles.p || !dagg.p) {
        set_error("prolongator: scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), s));
    if (M) SP3MM_HIP(hipMemcpyAsync(dagg.p, agg, (size_t)M * 8, hipMemcpyHostToDevice, s));
    // S = I - omega D^-1 A (pattern of A plus any missing diagonal)
    sp3mm_dmat S{}, P0{};
    launch_sa_count(A.rp, A.col, M, A.nnz, len.as<int32_t>(), s);
    if (dmat_alloc(S, M, M, 0, SP3MM_F64)) return -1;
    launch_scan(len.as<int32_t>(), M, S.rp, S.rp32, tiles.as<long long>(), &d->nnz, s);
    if (read_ctl(c, d, h)) {
        dmat_release(S);
        return -1;
    }
    S.nnz = h.nnz;
    c->pool.release(S.col);
    c->pool.release(S.val);
    S.col = static_cast<int32_t *>(c->pool.alloc((size_t)(S.nnz ? S.nnz : 1) * 4));
    S.val = c->pool.alloc((size_t)(S.nnz ? S.nnz : 1) * 8);
    if (!S.col || !S.val || S.nnz >= INT32_MAX) {
        set_error("prolongator: allocation of S (%lld nnz) failed", (long long)S.nnz);
        dmat_release(S);
        return -1;
    }
    launch_sa_fill(A.rp, A.col, static_cast<const double *>(A.val), M, A.nnz, omega, mode, S.rp, S.col,
                   static_cast<double *>(S.val), &d->err, s);
    // P0: one 1.0 per row at its aggregate
    if (dmat_alloc(P0, M, nagg, M, SP3MM_F64)) {
        dmat_release(S);
        return -1;
    }
    launch_p0_fill(dagg.as<int64_t>(), M, nagg, P0.rp, P0.rp32, P0.col, static_cast<double *>(P0.val), &d->err,
                   s);
    if (read_ctl(c, d, h) || h.err) {
        if (h.err) set_error("prolongator: %s (zero diagonal, or an aggregate out of range)", err_text(h.err));
        dmat_release(S);
        dmat_release(P0);
        return -1;
    }
    const int rc = spgemm(S, P0, P, nullptr, false);
    dmat_release(S);
    dmat_release(P0);
    return rc;
}

// --------------------------------------------------------------------------
// level consumers: SpMV and damped Jacobi on FP64 operators (the operations a
// V-cycle applies to the operators and transfer matrices built above)
// --------------------------------------------------------------------------
// Calls are stream-ordered on the library stream and return without waiting
// for the device; results are visible to the host after a call that waits
// (dot, copy to host, sync).  A Jacobi sweep that meets a zero diagonal entry
// sets a sticky device error word, reported by the next waiting call.
int set_stream(hipStream_t s)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    c->stream = s ? s : c->own_stream;
    return 0;
}

ConsState *cons_state(Context *c)
{
    if (!c->cons) {
        const size_t bytes = sizeof(ConsState) + (size_t)dot_parts() * 8;
        if (hipMalloc(&c->cons, bytes) != hipSuccess) {
            (void)hipGetLastError();
            c->cons = nullptr;
            set_error("consumer state allocation failed");
            return nullptr;
        }
        if (hipMemsetAsync(c->cons, 0, bytes, c->stream) != hipSuccess) {
            set_error("consumer state initialisation failed");
            return nullptr;
        }
    }
    return static_cast<ConsState *>(c->cons);
}

int consumer_wait(Context *c, double *dot_out)
{
    if (!c->cons) {
        SP3MM_HIP(hipStreamSynchronize(c->stream));
        return 0;
    }
    ConsState h{};
    SP3MM_HIP(hipMemcpyAsync(&h, c->cons, sizeof h, hipMemcpyDeviceToHost, c->stream));
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    if (dot_out) *dot_out = h.dot;
    if (h.err) {
        ConsState *d = static_cast<ConsState *>(c->cons);
        SP3MM_HIP(hipMemsetAsync(&d->err, 0, sizeof(int), c->stream));
        SP3MM_HIP(hipStreamSynchronize(c->stream));
        set_error("jacobi: zero diagonal entry (sweep launched before this call)");
        return -1;
    }
    return 0;
}

int spmv(const sp3mm_dmat &A, const double *x, const double *z, double *y, double alpha, double beta)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (A.vtype != SP3MM_F64) {
        set_error("spmv: A must hold FP64 values");
        return -1;
    }
    if (A.M && (!x || !y || (beta != 0.0 && !z))) {
        set_error("spmv: null vector");
        return -1;
    }
    launch_spmv(A.rp, A.col, static_cast<const double *>(A.val), A.M, A.nnz, x, z, y, alpha, beta, c->stream);
    SP3MM_HIP(hipGetLastError());
    return 0;
}

int axpby(int64_t n, double alpha, const double *x, double beta, double *y)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (n < 0 || (n && (!y || (alpha != 0.0 && !x)))) {
        set_error("axpby: bad length or null vector");
        return -1;
    }
    launch_axpby(n, alpha, x, beta, y, c->stream);
    SP3MM_HIP(hipGetLastError());
    return 0;
}

int dot(int64_t n, const double *x, const double *y, double &out)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (n < 0 || (n && (!x || !y))) {
        set_error("dot: bad length or null vector");
        return -1;
    }
    ConsState *st = cons_state(c);
    if (!st) return -1;
    launch_dot(n, x, y, reinterpret_cast<double *>(st + 1), &st->dot, c->stream);
    SP3MM_HIP(hipGetLastError());
    return consumer_wait(c, &out);
}

int jacobi(const sp3mm_dmat &A, const double *b, double *x, double *work, double omega, int sweeps)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (A.vtype != SP3MM_F64 || A.M != A.N) {
        set_error("jacobi: A must be square with FP64 values");
        return -1;
    }
    if (sweeps < 0 || (A.M && sweeps > 0 && (!b || !x || !work))) {
        set_error("jacobi: bad sweep count or null vector");
        return -1;
    }
    if (A.M == 0 || sweeps == 0) return 0;
    ConsState *st = cons_state(c);
    if (!st) return -1;
    queue_jacobi(A, b, x, work, omega, sweeps, &st->err, c->stream);
    SP3MM_HIP(hipGetLastError());
    return 0;
}

void queue_jacobi(const sp3mm_dmat &A, const double *b, double *x, double *work, double omega, int sweeps, int *err,
                  hipStream_t s)
{
    const double *val = static_cast<const double *>(A.val);
    if (sweeps > 1 && launch_jacobi_small(A.rp, A.col, val, A.M, A.nnz, b, x, omega, sweeps, err,
                                          ctx()->lds_device, s))
        return;
    double *cur = x, *nxt = work;
    for (int k = 0; k < sweeps; k++) {
        launch_jacobi(A.rp, A.col, val, A.M, A.nnz, b, cur, nxt, omega, err, s);
        std::swap(cur, nxt);
    }
    // an odd sweep count ends in the work vector: copy back with a kernel
    // (no copy or fill nodes when the sequence is captured in a graph)
    if (cur != x) launch_axpby(A.M, 1.0, cur, 0.0, x, s);
}

// --------------------------------------------------------------------------
// hash multipliers of the direct kernel's tables
// --------------------------------------------------------------------------
// Candidates: kHashMul, then odd constants from a fixed splitmix64 sequence
// (the same list in every process).
static const std::vector<unsigned> &hash_candidates()
{
    static const std::vector<unsigned> cand = [] {
        std::vector<unsigned> v{kHashMul};
        uint64_t x = 0x243F6A8885A308D3ull;
        while (v.size() < 256) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            v.push_back((unsigned)(z >> 32) | 1u);
        }
        return v;
    }();
    return cand;
}

// Tables of the likely bin under schedule t (false: no such bin, or tables
// beyond what k_hash_score takes)
static bool likely_tables(Context *c, const Tuning &t, const sp3mm_dmat &R, const sp3mm_dmat &A, int &H1, int &H2,
                          bool mixed = false)
{
    BinThresholds thr{};
    std::vector<FusedCfg> cfg;
    fused_bins(c->max_lds, mixed, t, thr, cfg);
    const int b = likely_bin(thr, R, A);
    if (b < 0) return false;
    H1 = cfg[b].H1;
    H2 = cfg[b].H2;
    return H1 <= kScoreTable && H2 <= kScoreTable;
}

// Scores of the candidates for tables of H1 and H2 slots on the given rows of
// R (k_hash_score): h[c] / h[ncand + c] the summed home-slot distances at
// stage 1 / 2, nkeys (optional) the rows' distinct key counts.
int hash_scores(Context *c, const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int H1, int H2,
                const int32_t *rows, int nrows, std::vector<unsigned long long> &h, std::vector<int> *nkeys)
{
    const std::vector<unsigned> &cand = hash_candidates();
    const int nc = (int)cand.size();
    if (nrows <= 0 || !A.rp32 || !P.rp32 || H1 < 64 || H2 < 64 || (H1 & (H1 - 1)) || (H2 & (H2 - 1)) ||
        H1 > kScoreTable || H2 > kScoreTable || hash_score_lds(H1, H2) > c->lds_device) {
        set_error("hash scoring: tables of %d and %d slots (powers of two in [64, %d]), %d rows", H1, H2,
                  kScoreTable, nrows);
        return -1;
    }
    for (int i = 0; i < nrows; i++)
        if (rows[i] < 0 || rows[i] >= R.M) {
            set_error("hash scoring: row %d out of range", rows[i]);
            return -1;
        }
    Buf drows((size_t)nrows * 4), dcand((size_t)nc * 4), dscore((size_t)nc * 2 * 8), dkeys((size_t)nrows * 2 * 4);
    if (!drows.p || !dcand.p || !dscore.p || !dkeys.p) {
        set_error("hash scoring: allocation failed");
        return -1;
    }
    hipStream_t s = c->stream;
    SP3MM_HIP(hipMemcpyAsync(drows.p, rows, (size_t)nrows * 4, hipMemcpyHostToDevice, s));
    SP3MM_HIP(hipMemcpyAsync(dcand.p, cand.data(), (size_t)nc * 4, hipMemcpyHostToDevice, s));
    auto *sc = dscore.as<unsigned long long>();
    launch_hash_score(operand(R), operand(A), operand(P), drows.as<int32_t>(), nrows, dcand.as<unsigned>(), nc, H1,
                      H2, sc, sc + nc, dkeys.as<int>(), s);
    SP3MM_HIP(hipGetLastError());
    h.assign((size_t)nc * 2, 0);
    SP3MM_HIP(hipMemcpyAsync(h.data(), sc, (size_t)nc * 2 * 8, hipMemcpyDeviceToHost, s));
    if (nkeys) {
        nkeys->assign((size_t)nrows * 2, 0);
        SP3MM_HIP(hipMemcpyAsync(nkeys->data(), dkeys.p, (size_t)nrows * 2 * 4, hipMemcpyDeviceToHost, s));
    }
    SP3MM_HIP(hipStreamSynchronize(s));
    return 0;
}

const std::vector<unsigned> &hash_candidate_list() { return hash_candidates(); }

// For tables of H1 and H2 slots: the candidates ranked by the summed
// home-slot distance of the keys of 512 sample rows of R (evenly spaced), at
// stage 1 (rank1) and stage 2 (rank2); equal scores keep the list order, so
// kHashMul leads a tie.
static int score_hashes(Context *c, const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int H1, int H2,
                        std::vector<unsigned> &rank1, std::vector<unsigned> &rank2,
                        std::vector<int> *nkeys = nullptr)
{
    const std::vector<unsigned> &cand = hash_candidates();
    const int nc = (int)cand.size();
    const int nrows = (int)std::min<int64_t>(R.M, 512);
    rank1.assign(1, kHashMul);
    rank2.assign(1, kHashMul);
    if (nrows <= 0 || !A.rp32 || !P.rp32 || H1 > kScoreTable || H2 > kScoreTable ||
        hash_score_lds(H1, H2) > c->lds_device)
        return 0;
    std::vector<int32_t> rows((size_t)nrows);
    for (int i = 0; i < nrows; i++) rows[(size_t)i] = (int32_t)(((2 * (int64_t)i + 1) * R.M) / (2 * (int64_t)nrows));
    std::vector<unsigned long long> h;
    if (hash_scores(c, R, A, P, H1, H2, rows.data(), nrows, h, nkeys)) return -1;
    for (int st = 0; st < 2; st++) {
        std::vector<int> idx((size_t)nc);
        for (int i = 0; i < nc; i++) idx[(size_t)i] = i;
        const unsigned long long *v = h.data() + (size_t)st * nc;
        std::stable_sort(idx.begin(), idx.end(), [v](int x, int y) { return v[x] < v[y]; });
        std::vector<unsigned> &out = st ? rank2 : rank1;
        out.clear();
        for (int i : idx) out.push_back(cand[(size_t)i]);
    }
    return 0;
}

static bool same_tuning(const Tuning &x, const Tuning &y)
{
    return x.u1 == y.u1 && x.u2 == y.u2 && x.subwave_h == y.subwave_h && x.slog1_cap == y.slog1_cap &&
           x.slog2_cap == y.slog2_cap && x.h1_shift == y.h1_shift && x.h2_div == y.h2_div && x.xcd_map == y.xcd_map &&
           x.hash1 == y.hash1 && x.hash2 == y.hash2;
}

// --------------------------------------------------------------------------
// autotuner: times the direct product of the caller's operands under candidate
// schedules and installs the fastest in the context.  Coordinate descent from
// the current setting, one parameter at a time, over all parameters twice
// when the first pass changed anything; `reps` timed products per candidate
// after one warm-up (device time by HIP events, median).  Every candidate
// that looks faster at all is timed again interleaved with the current
// schedule, and taken only when that shows it at least 1% faster.  Each pass
// first tries the hash multipliers ranked on the operands (score_hashes).
// The descent runs from the installed schedule and again from two presets
// (mesh operators; tables at an eighth of the bound); the fastest end point
// wins.
// --------------------------------------------------------------------------
int autotune(const sp3mm_dmat &R, const sp3mm_dmat &A, const sp3mm_dmat &P, int prec, double guard, int reps,
             Tuning &best, double &ms_best, double &ms_default)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    reps = std::min(std::max(reps, 1), 50);
    const Tuning saved = c->tune;
    bool reassociated = false;  // the product runs as R*(A*P): the direct kernels' fields do not apply
    // cutoff > 0: a timed product slower than cutoff ends the candidate at
    // once with that time (a schedule whose tables fill and send every row
    // to the long-row kernels costs tens of products' time per product)
    auto measure = [&](const Tuning &t, double &ms, double cutoff = 0.0) -> int {
        c->tune = t;
        std::vector<double> v;
        for (int k = 0; k <= reps; k++) {
            sp3mm_dmat C{};
            sp3mm_stats st{};
            const int rc = rap_direct_fused(R, A, P, C, &st, prec, guard > 0 ? guard : 1e-2, nullptr, false);
            dmat_release(C);
            if (rc) return -1;
            reassociated = st.reassociated != 0;
            if (k > 0) v.push_back(st.ms_total);
            // the warm-up product may allocate: it is cut at four times the bound
            if (cutoff > 0.0 && st.ms_total > (k > 0 ? cutoff : 4.0 * cutoff)) {
                ms = st.ms_total;
                return 0;
            }
        }
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        return 0;
    };
    // a candidate that looked faster is timed again against the current
    // schedule, interleaved (2*reps runs each, medians): one noisy median of a
    // few runs does not decide the step
    auto confirm = [&](const Tuning &t, const Tuning &u, double &ms_t, double &ms_u) -> int {
        std::vector<double> vt, vu;
        for (int k = 0; k < 2 * reps + 1; k++) {
            for (int w = 0; w < 2; w++) {
                c->tune = w ? u : t;
                sp3mm_dmat C{};
                sp3mm_stats st{};
                const int rc = rap_direct_fused(R, A, P, C, &st, prec, guard > 0 ? guard : 1e-2, nullptr, false);
                dmat_release(C);
                if (rc) return -1;
                (w ? vu : vt).push_back(st.ms_total);
            }
        }
        std::sort(vt.begin(), vt.end());
        std::sort(vu.begin(), vu.end());
        ms_t = vt[vt.size() / 2];
        ms_u = vu[vu.size() / 2];
        return 0;
    };
    Tuning cur = saved;
    double cur_ms = 0;
    if (measure(cur, cur_ms)) {
        c->tune = saved;
        return -1;
    }
    ms_default = cur_ms;
    struct Axis {
        int Tuning::*field;
        std::vector<int> values;
    };
    const Axis axes[] = {
        {&Tuning::subwave_h, {0, 128, 256, 512}}, {&Tuning::u1, {2, 4}},       {&Tuning::u2, {2, 3, 4}},
        {&Tuning::slog1_cap, {1, 2, 3}},            {&Tuning::slog2_cap, {1, 2}},
        {&Tuning::h1_shift, {-1, 0, 1, 2, 3}},      {&Tuning::h2_div, {2, 4, 8}},
        {&Tuning::xcd_map, {1, 0}},
    };
    // a candidate is taken when its median looks faster and the interleaved
    // re-timing confirms it at least 1% faster (a median of a few products
    // is noisy)
    auto step = [&](const Tuning &t, bool &took) -> int {
        took = false;
        double ms = 0;
        if (measure(t, ms, 2.0 * cur_ms)) return -1;
        if (ms < cur_ms) {
            double ms_t = 0, ms_cur = 0;
            if (confirm(t, cur, ms_t, ms_cur)) return -1;
            if (ms_t < 0.99 * ms_cur) {
                cur = t;
                cur_ms = ms_t;
                took = true;
            }
        }
        return 0;
    };
    // Hash multipliers ranked for a schedule's likely-bin tables
    // (score_hashes, cached per pair of table sizes).  Once the best-ranked
    // multipliers are taken (`hashed`), every candidate schedule gets the
    // best-ranked for its own tables, and each pass times the next kHashTries
    // of each stage's ranking against the current ones: multipliers that leave
    // every key on its home slot can still differ in LDS bank conflicts.
    constexpr int kHashTries = 3;
    std::map<std::pair<int, int>, std::pair<std::vector<unsigned>, std::vector<unsigned>>> hcache;
    auto ranked = [&](const Tuning &t, const std::vector<unsigned> *&r1, const std::vector<unsigned> *&r2) -> int {
        r1 = r2 = nullptr;
        int H1 = 0, H2 = 0;
        if (!likely_tables(c, t, R, A, H1, H2, prec == SP3MM_PREC_MIXED)) return 0;
        auto it = hcache.find({H1, H2});
        if (it == hcache.end()) {
            std::vector<unsigned> a, b;
            if (score_hashes(c, R, A, P, H1, H2, a, b)) return -1;
            it = hcache.emplace(std::make_pair(H1, H2), std::make_pair(std::move(a), std::move(b))).first;
        }
        r1 = &it->second.first;
        r2 = &it->second.second;
        return 0;
    };
    auto rehash = [&](Tuning &t) -> int {
        const std::vector<unsigned> *r1, *r2;
        if (ranked(t, r1, r2)) return -1;
        if (r1) {
            t.hash1 = (*r1)[0];
            t.hash2 = (*r2)[0];
        }
        return 0;
    };
    // Distinct key counts of the scoring's sample rows: a candidate whose
    // likely-bin tables would not hold them for more than a tenth of the
    // sampled rows is not timed (such rows go to the long-row kernels, at up
    // to hundreds of products' time each).
    std::vector<int> sample_keys;
    {
        int H1 = 0, H2 = 0;
        std::vector<unsigned> r1, r2;
        if (likely_tables(c, cur, R, A, H1, H2, prec == SP3MM_PREC_MIXED) &&
            score_hashes(c, R, A, P, H1, H2, r1, r2, &sample_keys)) {
            c->tune = saved;
            return -1;
        }
    }
    auto would_fill = [&](const Tuning &t) -> bool {
        if (sample_keys.empty()) return false;
        BinThresholds thr{};
        std::vector<FusedCfg> cfg;
        fused_bins(c->max_lds, prec == SP3MM_PREC_MIXED, t, thr, cfg);
        const int b = likely_bin(thr, R, A);
        if (b < 0) return false;
        const FusedCfg &f = cfg[(size_t)b];
        const int cap1 = f.G > kWave ? f.H1 - f.H1 / 8 : f.H1, cap2 = f.G > kWave ? f.H2 - f.H2 / 8 : f.H2;
        int n = 0, over = 0;
        for (size_t i = 0; i + 1 < sample_keys.size(); i += 2) {
            if (sample_keys[i] == INT32_MAX || sample_keys[i + 1] == INT32_MAX) continue;
            n++;
            over += sample_keys[i] > cap1 || sample_keys[i + 1] > cap2;
        }
        return n > 0 && 10 * over > n;
    };
    bool hashed = false;
    // coordinate descent from `cur` (measured: cur_ms)
    auto descend = [&]() -> int {
        hashed = false;
        for (int pass = 0; pass < 2; pass++) {
            bool changed = false, took = false;
            if (reassociated) hashed = true;  // no direct kernel: no tables to hash
            if (!hashed) {
                Tuning t = cur;
                if (rehash(t)) return -1;
                if (t.hash1 == cur.hash1 && t.hash2 == cur.hash2) {
                    hashed = true;
                } else {
                    if (step(t, took)) {
                        return -1;
                    }
                    hashed = changed = took;
                }
            }
            if (hashed && !reassociated) {
                const std::vector<unsigned> *r1, *r2;
                if (ranked(cur, r1, r2)) {
                    return -1;
                }
                for (int st = 0; r1 && st < 2; st++) {
                    const std::vector<unsigned> &r = st ? *r2 : *r1;
                    for (int k = 0; k <= kHashTries && k < (int)r.size(); k++) {
                        Tuning t = cur;
                        (st ? t.hash2 : t.hash1) = r[(size_t)k];
                        if (t.hash1 == cur.hash1 && t.hash2 == cur.hash2) continue;
                        if (step(t, took)) {
                            return -1;
                        }
                        changed = changed || took;
                    }
                }
            }
            for (const Axis &ax : axes) {
                // a reassociated product only walks rows (xcd_map)
                if (reassociated && ax.field != &Tuning::xcd_map) continue;
                for (int v : ax.values) {
                    if (cur.*ax.field == v) continue;
                    Tuning t = cur;
                    t.*ax.field = v;
                    if (hashed && !reassociated && rehash(t)) return -1;
                    if (!tuning_valid(t) || would_fill(t)) continue;
                    if (step(t, took)) {
                        return -1;
                    }
                    changed = changed || took;
                }
            }
            if (!changed) break;
        }
        return 0;
    };
    if (descend()) {
        c->tune = saved;
        return -1;
    }
    // Further descents from presets, each end point kept only when an
    // interleaved re-timing shows it faster than the best so far:
    // * a schedule that suits mesh operators (three-lane stage-2 batches,
    //   four-lane sub-groups, the stage-1 table at half the bound): on the
    //   128^3 7-point product one descent from the defaults settled once at
    //   0.63 ms where others reach 0.55 ms (u2 = 3 and slog2_cap = 2 pay only
    //   together);
    // * tables at an eighth of the bound (stage 1) and of the stage-2 bound
    //   over 8: on the 27-point 100^3 product with FP32 operands a descent
    //   from the other starts settled at 0.566 ms with quarter-bound tables,
    //   where this one reaches 0.485 ms (h1_shift, h2_div and the stage-2
    //   multiplier pay only together).
    Tuning presets[2]{};
    presets[0].u1 = 4;
    presets[0].u2 = 3;
    presets[0].slog2_cap = 2;
    presets[0].h1_shift = 1;
    presets[0].h2_div = 4;
    presets[1].u1 = 4;
    presets[1].u2 = 2;
    presets[1].slog2_cap = 2;
    presets[1].h1_shift = 3;
    presets[1].h2_div = 8;
    Tuning endb = cur;
    double endb_ms = cur_ms;
    for (const Tuning &pre : presets) {
        if (reassociated || same_tuning(pre, saved) || would_fill(pre)) continue;
        // a start more than 3x behind the best end point is not descended
        // from (tables that fill send rows to the long-row kernels, and a
        // descent among such schedules costs many slow products; a start
        // with unscored multipliers can be 1.5x behind and still descend to
        // the best end point)
        cur = pre;
        if (measure(cur, cur_ms, 3.0 * endb_ms)) {
            c->tune = saved;
            return -1;
        }
        if (cur_ms > 3.0 * endb_ms) continue;
        if (descend()) {
            c->tune = saved;
            return -1;
        }
        if (same_tuning(cur, endb)) continue;
        double ms2 = 0, ms1 = 0;
        if (confirm(cur, endb, ms2, ms1)) {
            c->tune = saved;
            return -1;
        }
        if (ms2 < ms1) {
            endb = cur;
            endb_ms = ms2;
        } else {
            endb_ms = ms1;
        }
    }
    cur = endb;
    cur_ms = endb_ms;
    c->tune = cur;
    best = cur;
    ms_best = cur_ms;
    return 0;
}

// --------------------------------------------------------------------------
// batches: many independent products as one block-diagonal product
// --------------------------------------------------------------------------
struct Batch {
    int count = 0;
    bool mixed = false;
    int prec = SP3MM_PREC_FP64;
    double guard = 1e-2;
    std::vector<int64_t> roff, coff, M;
    sp3mm_dmat Rb{}, Ab{}, Pb{};
    FusedPlan plan;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double ms_pack = 0, ms_build = 0;
};

// builds the batch's plan under the context's current schedule and, with
// `graph`, captures its replay into a HIP graph
static int batch_build(Context *c, Batch *b, bool graph)
{
    Timer tm(c->stream, true);
    const int e1 = tm.mark();
    b->plan.R = b->Rb;
    b->plan.A = b->Ab;
    b->plan.P = b->Pb;
    b->plan.mixed = b->prec == SP3MM_PREC_MIXED;
    b->plan.f32 = b->mixed;
    b->plan.tu = c->tune;
    b->plan.guard = b->guard;
    if (b->plan.build(c, tm)) return -1;
    const int e2 = tm.mark();
    (void)hipEventSynchronize(tm.ev[e2]);
    b->ms_build = tm.ms(e1, e2);
    if (graph) {
        hipStream_t s = c->stream;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return 0;  // usable without a graph
        }
        b->plan.record(c, s);
        if (hipStreamEndCapture(s, &b->graph) != hipSuccess ||
            hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
            b->exec = nullptr;
        }
    }
    return 0;
}

// Packs `count` matrices block-diagonally into `out` (row and entry offsets
// added on the GPU, column offsets coloff).  reuse: `out` already holds a
// packing of matrices of the same shapes and nonzero counts; its arrays are
// overwritten in place.
static int pack(Context *c, int count, const sp3mm_dmat *const *ms, int vtype, sp3mm_dmat &out,
                const std::vector<int64_t> &coloff, bool reuse = false)
{
    int64_t M = 0, N = 0, nnz = 0;
    std::vector<PackItem> items((size_t)count);
    for (int i = 0; i < count; i++) {
        const sp3mm_dmat &m = *ms[i];
        items[i] = PackItem{m.rp, m.col, m.val, m.M, m.nnz, M, nnz, (int32_t)coloff[i], 0};
        M += m.M;
        N += m.N;
        nnz += m.nnz;
    }
    if (N >= INT32_MAX || M >= INT32_MAX || nnz >= INT32_MAX) {
        set_error("batch: packed operand too large (%lld x %lld, %lld nnz)", (long long)M, (long long)N,
                  (long long)nnz);
        return -1;
    }
    if (reuse) {
        if (out.M != M || out.N != N || out.nnz != nnz) {
            set_error("batch: operands differ in shape or nonzero count from the batch's (%lld x %lld, %lld nnz "
                      "packed; %lld x %lld, %lld nnz given)",
                      (long long)out.M, (long long)out.N, (long long)out.nnz, (long long)M, (long long)N,
                      (long long)nnz);
            return -1;
        }
    } else if (dmat_alloc(out, M, N, nnz, vtype)) {
        return -1;
    }
    Buf it(items.size() * sizeof(PackItem));
    if (!it.p) return -1;
    SP3MM_HIP(hipMemcpyAsync(it.p, items.data(), items.size() * sizeof(PackItem), hipMemcpyHostToDevice, c->stream));
    launch_pack(it.as<PackItem>(), count, vtype, out.rp, out.col, out.val, c->stream);
    // last row offset, then the int32 copy of the offsets (right operands)
    SP3MM_HIP(hipMemcpyAsync(out.rp + M, &nnz, 8, hipMemcpyHostToDevice, c->stream));
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    launch_rp_convert(out.rp, M, nnz, 0, out.rp, out.rp32, nullptr, c->stream);
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

void *batch_create(int count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int prec,
                   double guard, bool graph)
{
    Context *c = ctx();
    if (!c) return nullptr;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (count <= 0) {
        set_error("batch: empty batch");
        return nullptr;
    }
    const bool mixed = prec != SP3MM_PREC_FP64;
    const int vt = mixed ? SP3MM_F32 : SP3MM_F64;
    for (int i = 0; i < count; i++) {
        if (R[i].N != A[i].M || A[i].N != P[i].M) {
            set_error("batch: problem %d has mismatched shapes", i);
            return nullptr;
        }
        if (R[i].vtype != vt || A[i].vtype != vt || P[i].vtype != vt) {
            set_error("batch: problem %d operands must be %s", i, mixed ? "FP32" : "FP64");
            return nullptr;
        }
    }
    Batch *b = new Batch();
    b->count = count;
    b->mixed = mixed;
    std::vector<const sp3mm_dmat *> rs, as, ps;
    std::vector<int64_t> offR, offA, offP;
    int64_t kr = 0, ka = 0, kp = 0, mr = 0;
    for (int i = 0; i < count; i++) {
        rs.push_back(R + i);
        as.push_back(A + i);
        ps.push_back(P + i);
        offR.push_back(kr);  // R's columns are A's rows
        offA.push_back(ka);  // A's columns are P's rows
        offP.push_back(kp);  // P's columns are the product's columns
        b->roff.push_back(mr);
        b->coff.push_back(kp);
        b->M.push_back(R[i].M);
        kr += R[i].N;
        ka += A[i].N;
        kp += P[i].N;
        mr += R[i].M;
    }
    Timer tm(c->stream, true);
    const int e0 = tm.mark();
    if (pack(c, count, rs.data(), vt, b->Rb, offR) || pack(c, count, as.data(), vt, b->Ab, offA) ||
        pack(c, count, ps.data(), vt, b->Pb, offP)) {
        dmat_release(b->Rb);
        dmat_release(b->Ab);
        dmat_release(b->Pb);
        delete b;
        return nullptr;
    }
    const int e1 = tm.mark();
    (void)hipEventSynchronize(tm.ev[e1]);
    b->ms_pack = tm.ms(e0, e1);
    b->prec = prec;
    b->guard = guard > 0 ? guard : 1e-2;
    if (batch_build(c, b, graph)) {
        b->plan.release_all();
        dmat_release(b->Rb);
        dmat_release(b->Ab);
        dmat_release(b->Pb);
        delete b;
        return nullptr;
    }
    return b;
}

int batch_autotune(void *h, int reps, Tuning &best, double &ms_best, double &ms_default)
{
    Context *c = ctx();
    if (!c || !h) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Batch *b = static_cast<Batch *>(h);
    if (autotune(b->Rb, b->Ab, b->Pb, b->prec, b->guard, reps, best, ms_best, ms_default)) return -1;
    // rebuild the plan (and the graph, if there was one) for the schedule found
    const bool graph = b->exec != nullptr;
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->plan.release_all();
    b->plan = FusedPlan();
    return batch_build(c, b, graph);
}

// Copies the current entries of the batch's operands (the same count, shapes
// and nonzero counts as at creation) into its packed operands; the next run()
// recomputes every product from them.  A pattern change that moves a product
// off the sizes the plan was built for is reported by run().
int batch_repack(void *h, int count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P)
{
    Context *c = ctx();
    if (!c || !h) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Batch *b = static_cast<Batch *>(h);
    if (count != b->count) {
        set_error("batch: %d problems given, the batch holds %d", count, b->count);
        return -1;
    }
    const int vt = b->mixed ? SP3MM_F32 : SP3MM_F64;
    std::vector<const sp3mm_dmat *> rs, as, ps;
    std::vector<int64_t> offR, offA, offP;
    int64_t kr = 0, ka = 0, kp = 0;
    for (int i = 0; i < count; i++) {
        if (R[i].M != b->M[i] || R[i].N != A[i].M || A[i].N != P[i].M) {
            set_error("batch: problem %d has shapes that differ from the batch's", i);
            return -1;
        }
        if (R[i].vtype != vt || A[i].vtype != vt || P[i].vtype != vt) {
            set_error("batch: problem %d operands must be %s", i, b->mixed ? "FP32" : "FP64");
            return -1;
        }
        rs.push_back(R + i);
        as.push_back(A + i);
        ps.push_back(P + i);
        offR.push_back(kr);
        offA.push_back(ka);
        offP.push_back(kp);
        kr += R[i].N;
        ka += A[i].N;
        kp += P[i].N;
    }
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    if (pack(c, count, rs.data(), vt, b->Rb, offR, true) || pack(c, count, as.data(), vt, b->Ab, offA, true) ||
        pack(c, count, ps.data(), vt, b->Pb, offP, true))
        return -1;
    return 0;
}

int batch_run(void *h, bool use_graph, double *ms)
{
    Context *c = ctx();
    if (!c || !h) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Batch *b = static_cast<Batch *>(h);
    if (use_graph && !b->exec) {
        set_error("batch: no graph was instantiated");
        return -1;
    }
    hipEvent_t e0, e1;
    SP3MM_HIP(hipEventCreate(&e0));
    SP3MM_HIP(hipEventCreate(&e1));
    SP3MM_HIP(hipEventRecord(e0, c->stream));
    if (use_graph) SP3MM_HIP(hipGraphLaunch(b->exec, c->stream));
    else b->plan.record(c, c->stream);
    SP3MM_HIP(hipEventRecord(e1, c->stream));
    const int rc = b->plan.check(c);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return rc;
}

int batch_result(void *h, int i, sp3mm_dmat &out)
{
    Context *c = ctx();
    if (!c || !h) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Batch *b = static_cast<Batch *>(h);
    if (i < 0 || i >= b->count) {
        set_error("batch: index %d out of range", i);
        return -1;
    }
    const sp3mm_dmat &C = b->plan.C;
    int64_t lo = 0, hi = 0;
    SP3MM_HIP(hipMemcpy(&lo, C.rp + b->roff[i], 8, hipMemcpyDeviceToHost));
    SP3MM_HIP(hipMemcpy(&hi, C.rp + b->roff[i] + b->M[i], 8, hipMemcpyDeviceToHost));
    const int64_t N = (i + 1 < b->count ? b->coff[i + 1] : b->plan.P.N) - b->coff[i];
    if (dmat_alloc(out, b->M[i], N, hi - lo, SP3MM_F64)) return -1;
    launch_extract(C.rp, C.col, static_cast<const double *>(C.val), b->roff[i], b->M[i], (int32_t)b->coff[i], out.rp,
                   out.rp32, out.col, static_cast<double *>(out.val), c->stream);
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

void batch_info(void *h, double *ms_pack, double *ms_build, int64_t *nnz, int32_t *has_graph)
{
    Batch *b = static_cast<Batch *>(h);
    if (!b) return;
    if (ms_pack) *ms_pack = b->ms_pack;
    if (ms_build) *ms_build = b->ms_build;
    if (nnz) *nnz = b->plan.nnz;
    if (has_graph) *has_graph = b->exec ? 1 : 0;
}

void batch_free(void *h)
{
    Context *c = ctx();
    Batch *b = static_cast<Batch *>(h);
    if (!b || !c) return;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipStreamSynchronize(c->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->plan.release_all();
    dmat_release(b->Rb);
    dmat_release(b->Ab);
    dmat_release(b->Pb);
    delete b;
}

// --------------------------------------------------------------------------
// a smoothed-aggregation hierarchy as one replayable plan
// --------------------------------------------------------------------------
// Level l, from its operator A_l: S = I - omega D^-1 A_l (SaPlan), P = I*S*P0
// (the triple product with the identity on the left: P0 sums S's columns by
// aggregate), R = P^T (TransposePlan), A_{l+1} = R*A_l*P (FusedPlan).  The
// identity and P0 are fixed; every other step is re-enqueued by record(), so
// that the whole chain is captured as one HIP graph and replayed when A_0's
// values change (its pattern must not: a fingerprint of it is checked first).
struct HierLevel {
    SaPlan sa;
    sp3mm_dmat I{}, P0{};
    FusedPlan pplan;    // P = I * S * P0
    TransposePlan tr;   // R = P^T
    FusedPlan rap;      // A_{l+1} = R * A_l * P

    void record(Context *c, hipStream_t s)
    {
        sa.record(c, s);
        pplan.record(c, s);
        tr.record(c, s);
        rap.record(c, s);
    }
    void release_all()
    {
        dmat_release(sa.S);
        dmat_release(I);
        dmat_release(P0);
        pplan.release_all();
        dmat_release(tr.T);
        rap.release_all();
    }
};

struct Hier {
    sp3mm_dmat A0{};
    std::vector<std::unique_ptr<HierLevel>> lv;
    unsigned long long fp = 0;
    Buf fpbuf, errs, errp;  // fingerprint, OR of the plans' error words, their addresses
    int nerr = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double ms_build = 0;
};

static int pattern_fingerprint(Context *c, const sp3mm_dmat &A, Buf &buf, unsigned long long &out)
{
    launch_fingerprint(A.rp, A.M, A.col, A.nnz, buf.as<unsigned long long>(), c->stream);
    SP3MM_HIP(hipMemcpyAsync(c->host_small, buf.p, 8, hipMemcpyDeviceToHost, c->stream));
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    memcpy(&out, c->host_small, 8);
    return 0;
}

// the fixed 0/1 matrix with one entry per row, in column agg[row] (device ids)
static int one_per_row(Context *c, const int64_t *dagg, int64_t M, int64_t ncol, sp3mm_dmat &m, Buf &ctl)
{
    if (dmat_alloc(m, M, ncol, M, SP3MM_F64)) return -1;
    Ctl *d = ctl.as<Ctl>();
    launch_fill_i32(reinterpret_cast<int32_t *>(d), (int64_t)(sizeof(Ctl) / 4), 0, c->stream);
    launch_p0_fill(dagg, M, ncol, m.rp, m.rp32, m.col, static_cast<double *>(m.val), &d->err, c->stream);
    Ctl h;
    if (read_ctl(c, d, h)) return -1;
    if (h.err) {
        set_error("hierarchy: an aggregate id is out of range");
        return -1;
    }
    return 0;
}

// records a failed HIP call as the library error; true if it failed
static bool hip_failed(hipError_t e)
{
    if (e == hipSuccess) return false;
    set_error("HIP call failed: %s", hipGetErrorString(e));
    return true;
}

static void hier_release(Hier *h)
{
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    for (auto &l : h->lv) l->release_all();
    delete h;
}

void *hier_create(const sp3mm_dmat &A, int nlev, const int64_t *const *aggs, const int64_t *naggs, double omega,
                  int mode, bool graph)
{
    Context *c = ctx();
    if (!c) return nullptr;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (A.M != A.N || A.vtype != SP3MM_F64 || nlev <= 0 || (mode != 0 && mode != 1) || !A.rp32) {
        set_error("hierarchy: A must be square FP64 (nnz < 2^31), with at least one level and diag mode 0 or 1");
        return nullptr;
    }
    Hier *h = new Hier();
    h->A0 = A;
    h->A0.owned = 0;
    h->fpbuf = Buf(8);
    Buf ctl(sizeof(Ctl));
    if (!h->fpbuf.p || !ctl.p) {
        set_error("hierarchy: allocation failed");
        delete h;
        return nullptr;
    }
    Timer tm(c->stream, true);
    const int e0 = tm.mark();
    if (pattern_fingerprint(c, A, h->fpbuf, h->fp)) {
        delete h;
        return nullptr;
    }
    sp3mm_dmat cur = A;
    for (int l = 0; l < nlev; l++) {
        const int64_t M = cur.M, nagg = naggs[l];
        if (nagg <= 0 || nagg > M) {
            set_error("hierarchy: level %d has %lld aggregates for %lld rows", l, (long long)nagg, (long long)M);
            hier_release(h);
            return nullptr;
        }
        h->lv.emplace_back(new HierLevel());
        HierLevel &L = *h->lv.back();
        std::vector<int64_t> iota((size_t)M);
        for (int64_t i = 0; i < M; i++) iota[(size_t)i] = i;
        Buf dagg((size_t)(M ? M : 1) * 8), diota((size_t)(M ? M : 1) * 8);
        if (!dagg.p || !diota.p) {
            set_error("hierarchy: allocation failed");
            hier_release(h);
            return nullptr;
        }
        if (hip_failed(hipMemcpyAsync(dagg.p, aggs[l], (size_t)M * 8, hipMemcpyHostToDevice, c->stream)) ||
            hip_failed(hipMemcpyAsync(diota.p, iota.data(), (size_t)M * 8, hipMemcpyHostToDevice, c->stream)) ||
            L.sa.build(c, cur, omega, mode) || one_per_row(c, dagg.as<int64_t>(), M, nagg, L.P0, ctl) ||
            one_per_row(c, diota.as<int64_t>(), M, M, L.I, ctl)) {
            hier_release(h);
            return nullptr;
        }
        Timer t2(c->stream, false);
        L.pplan.R = L.I;
        L.pplan.A = L.sa.S;
        L.pplan.P = L.P0;
        // with the identity on the left, a row's bound is its exact distinct
        // count: a table of twice the bound keeps probe sequences short
        L.pplan.tu = c->tune;
        L.pplan.tu.h1_shift = -1;
        if (L.pplan.build(c, t2) || L.tr.build(c, L.pplan.C, true)) {
            hier_release(h);
            return nullptr;
        }
        L.rap.R = L.tr.T;
        L.rap.A = cur;
        L.rap.P = L.pplan.C;
        L.rap.tu = c->tune;
        if (L.rap.build(c, t2)) {
            hier_release(h);
            return nullptr;
        }
        cur = L.rap.C;
    }
    // the error words every replay must leave clear
    std::vector<const int *> words;
    for (auto &l : h->lv) {
        words.push_back(&l->sa.d()->err);
        words.push_back(&l->pplan.d()->err);
        words.push_back(&l->rap.d()->err);
    }
    h->nerr = (int)words.size();
    h->errp = Buf(words.size() * sizeof(int *));
    h->errs = Buf(8);
    if (!h->errp.p || !h->errs.p) {
        set_error("hierarchy: allocation failed");
        hier_release(h);
        return nullptr;
    }
    const int e1 = tm.mark();
    if (hip_failed(hipMemcpyAsync(h->errp.p, words.data(), words.size() * sizeof(int *), hipMemcpyHostToDevice,
                                  c->stream)) ||
        hip_failed(hipStreamSynchronize(c->stream))) {
        hier_release(h);
        return nullptr;
    }
    h->ms_build = tm.ms(e0, e1);
    if (graph) {
        hipStream_t s = c->stream;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return h;  // usable without a graph
        }
        for (auto &l : h->lv) l->record(c, s);
        if (hipStreamEndCapture(s, &h->graph) != hipSuccess ||
            hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (h->graph) (void)hipGraphDestroy(h->graph);
            h->graph = nullptr;
            h->exec = nullptr;
        }
    }
    return h;
}

int hier_update(void *hp, bool use_graph, double *ms)
{
    Context *c = ctx();
    if (!c || !hp) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Hier *h = static_cast<Hier *>(hp);
    if (use_graph && !h->exec) {
        set_error("hierarchy: no graph was instantiated");
        return -1;
    }
    unsigned long long fp = 0;
    if (pattern_fingerprint(c, h->A0, h->fpbuf, fp)) return -1;
    if (fp != h->fp) {
        set_error("hierarchy: the pattern of A changed since the hierarchy was built (only values may change)");
        return -1;
    }
    hipEvent_t e0, e1;
    SP3MM_HIP(hipEventCreate(&e0));
    SP3MM_HIP(hipEventCreate(&e1));
    SP3MM_HIP(hipEventRecord(e0, c->stream));
    if (use_graph) SP3MM_HIP(hipGraphLaunch(h->exec, c->stream));
    else
        for (auto &l : h->lv) l->record(c, c->stream);
    SP3MM_HIP(hipEventRecord(e1, c->stream));
    launch_or_words(h->errp.as<const int *>(), h->nerr, h->errs.as<int>(), c->stream);
    SP3MM_HIP(hipMemcpyAsync(c->host_small, h->errs.p, 4, hipMemcpyDeviceToHost, c->stream));
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    int err = 0;
    memcpy(&err, c->host_small, 4);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    if (err) {
        set_error("hierarchy update: %s", err_text(err));
        return -1;
    }
    return 0;
}

int hier_level(void *hp, int l, sp3mm_dmat &R, sp3mm_dmat &P, sp3mm_dmat &Ac)
{
    Hier *h = static_cast<Hier *>(hp);
    if (!h || l < 0 || l >= (int)h->lv.size()) {
        set_error("hierarchy: level %d out of range", l);
        return -1;
    }
    HierLevel &L = *h->lv[(size_t)l];
    R = L.tr.T;
    P = L.pplan.C;
    Ac = L.rap.C;
    R.owned = P.owned = Ac.owned = 0;
    return 0;
}

void hier_info(void *hp, int32_t *nlev, double *ms_build, int32_t *has_graph)
{
    Hier *h = static_cast<Hier *>(hp);
    if (!h) return;
    if (nlev) *nlev = (int32_t)h->lv.size();
    if (ms_build) *ms_build = h->ms_build;
    if (has_graph) *has_graph = h->exec ? 1 : 0;
}

void hier_free(void *hp)
{
    Context *c = ctx();
    Hier *h = static_cast<Hier *>(hp);
    if (!h || !c) return;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipStreamSynchronize(c->stream);
    hier_release(h);
}

// --------------------------------------------------------------------------
// sizing helpers exposed through the ABI
// --------------------------------------------------------------------------
int row_ub(const sp3mm_dmat &A, const sp3mm_dmat &B, int64_t *host_ub)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (A.N != B.M || !B.rp32) {
        set_error("upper bound: shapes differ or right operand too large");
        return -1;
    }
    const int64_t M = A.M;
    BinThresholds thr{};
    thr.nb = 0;
    Buf ctl(sizeof(Ctl)), ub((size_t)(M + 1) * 8), lists((size_t)(M ? M : 1) * 4),
        binscr((size_t)bin_scratch_ints(M) * 4);
    if (!ctl.p || !ub.p || !lists.p || !binscr.p) {
        set_error("scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), c->stream));
    launch_ub_bin(operand(A), operand(B), ub.as<int64_t>(), nullptr, lists.as<int32_t>(), d->counts, thr,
                  d->totals, binscr.as<int>(), c->stream);
    if (M) SP3MM_HIP(hipMemcpyAsync(host_ub, ub.p, (size_t)M * 8, hipMemcpyDeviceToHost, c->stream));
    if (read_ctl(c, d, h)) return -1;
    host_ub[M] = (int64_t)h.totals[0];
    return 0;
}

int ub_colparts(const sp3mm_dmat &A, const int64_t *host_offsets, int64_t boff_len, int grid_cols,
                int64_t *host_out)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int64_t M = A.M;
    Buf ctl(sizeof(Ctl)), boff((size_t)boff_len * 8), out((size_t)(M * grid_cols + 1) * 8);
    if (!ctl.p || !boff.p || !out.p) {
        set_error("scratch allocation failed");
        return -1;
    }
    Ctl *d = ctl.as<Ctl>();
    Ctl h;
    SP3MM_HIP(hipMemsetAsync(d, 0, sizeof(Ctl), c->stream));
    SP3MM_HIP(hipMemcpyAsync(boff.p, host_offsets, (size_t)boff_len * 8, hipMemcpyHostToDevice, c->stream));
    launch_ub_colparts(A.rp, A.col, M, boff.as<int64_t>(), grid_cols, out.as<int64_t>(), d->totals, c->stream);
    if (M)
        SP3MM_HIP(hipMemcpyAsync(host_out, out.p, (size_t)M * grid_cols * 8, hipMemcpyDeviceToHost,
                                 c->stream));
    if (read_ctl(c, d, h)) return -1;
    host_out[M * grid_cols] = (int64_t)h.totals[0];
    return 0;
}

}  // namespace sp3mm
