// api.hip -- the C ABI declared in include/sp3mm_amd.h.
//
// Device-resident functions (sp3mm_*) and the Sp3MM4AMG-compatible drop-in
// entry points (_0 / _1).  A drop-in call uploads the borrowed host spmat
// operands (64-bit indices, narrowed to int32 on the GPU and shifted to
// 0-based there), runs the product on the GPU and downloads the result into
// malloc'd host arrays, widening indices back to 64 bits (and to 1-based for
// the _1 variants) on the GPU before the copy.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>

#include "engine.hpp"

namespace sp3mm {
int init_device(int dev);
const char *last_error_text();
}  // namespace sp3mm

using namespace sp3mm;

extern "C" {

double Start = 0, End = 0, Elapsed = 0, ElapsedInternal = 0;

// ---------------------------------------------------------------------------
// device-resident interface
// ---------------------------------------------------------------------------
int sp3mm_init(int device) { return init_device(device); }

const char *sp3mm_last_error(void) { return last_error_text(); }

int sp3mm_set_stream(void *stream)
{
    clear_error();
    return set_stream(static_cast<hipStream_t>(stream));
}

void *sp3mm_get_stream(void)
{
    Context *c = ctx();
    return c ? static_cast<void *>(c->stream) : nullptr;
}

int sp3mm_sync(void)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    return consumer_wait(c);
}

int sp3mm_dmat_upload(sp3mm_dmat *out, int64_t M, int64_t N, int64_t nnz, const int64_t *irp,
                      const int64_t *ja, const void *val, int32_t vtype, int32_t base)
{
    clear_error();
    if (!out || !irp || (nnz && (!ja || !val))) {
        set_error("upload: null argument");
        return -1;
    }
    if (M < 0 || N < 0 || nnz < 0 || N >= INT32_MAX || M >= INT32_MAX) {
        set_error("upload: unsupported shape %lld x %lld (%lld nnz)", (long long)M, (long long)N,
                  (long long)nnz);
        return -1;
    }
    if ((vtype != SP3MM_F64 && vtype != SP3MM_F32) || (base != 0 && base != 1)) {
        set_error("upload: bad value type or index base");
        return -1;
    }
    // row offsets are checked on the host before any device work: monotone,
    // from base to nnz+base (the column kernel trusts them for its bounds)
    if (irp[0] != base || irp[M] != nnz + base) {
        set_error("upload: row offsets must run from %d to nnz+%d (got %lld .. %lld, nnz %lld)", base, base,
                  (long long)irp[0], (long long)irp[M], (long long)nnz);
        return -1;
    }
    for (int64_t r = 0; r < M; r++) {
        if (irp[r + 1] < irp[r]) {
            set_error("upload: row offsets decrease at row %lld", (long long)r);
            return -1;
        }
    }
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipStream_t s = c->stream;
    sp3mm_dmat m;
    if (dmat_alloc(m, M, N, nnz, vtype)) return -1;
    Buf trp((size_t)(M + 1) * 8), tja((size_t)(nnz ? nnz : 1) * 8), err(64);
    if (!trp.p || !tja.p || !err.p) {
        dmat_release(m);
        set_error("upload: staging allocation failed");
        return -1;
    }
    int rc = 0;
    do {
        if (hipMemsetAsync(err.p, 0, 64, s) != hipSuccess) { rc = -1; break; }
        if (hipMemcpyAsync(trp.p, irp, (size_t)(M + 1) * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = -1;
            break;
        }
        if (nnz && hipMemcpyAsync(tja.p, ja, (size_t)nnz * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = -1;
            break;
        }
        if (nnz && hipMemcpyAsync(m.val, val, (size_t)nnz * (vtype == SP3MM_F32 ? 4 : 8),
                                  hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = -1;
            break;
        }
        launch_rp_convert(trp.as<int64_t>(), M, nnz, base, m.rp, m.rp32, err.as<int>(), s);
        launch_col_convert(tja.as<int64_t>(), m.rp, M, N, base, m.col, err.as<int>(), s);
        int herr = 0;
        if (hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess) { rc = -1; break; }
        if (hipStreamSynchronize(s) != hipSuccess) { rc = -1; break; }
        if (herr) {
            set_error("upload: matrix is not valid sorted CSR (offsets must be monotone from %d to "
                      "nnz+%d, columns in [%d, N+%d) and strictly increasing inside each row)",
                      base, base, base, base);
            rc = -2;
        }
    } while (0);
    if (rc) {
        if (rc == -1) set_error("upload: HIP error %s", hipGetErrorString(hipGetLastError()));
        dmat_release(m);
        return -1;
    }
    *out = m;
    return 0;
}

int sp3mm_dmat_import(sp3mm_dmat *out, int64_t M, int64_t N, int64_t nnz, const int64_t *d_irp,
                      const int64_t *d_ja, const void *d_val, int32_t vtype, int32_t base)
{
    clear_error();
    if (!out || !d_irp || (nnz && (!d_ja || !d_val))) {
        set_error("import: null argument");
        return -1;
    }
    if (M < 0 || N < 0 || nnz < 0 || N >= INT32_MAX || M >= INT32_MAX) {
        set_error("import: unsupported shape %lld x %lld (%lld nnz)", (long long)M, (long long)N, (long long)nnz);
        return -1;
    }
    if ((vtype != SP3MM_F64 && vtype != SP3MM_F32) || (base != 0 && base != 1)) {
        set_error("import: bad value type or index base");
        return -1;
    }
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipStream_t s = c->stream;
    sp3mm_dmat m;
    if (dmat_alloc(m, M, N, nnz, vtype)) return -1;
    Buf err(64);
    if (!err.p) {
        dmat_release(m);
        set_error("import: allocation failed");
        return -1;
    }
    // The row offsets are validated on the device before anything walks the
    // columns (the column kernel trusts them for its bounds): one wait here.
    int herr = 0;
    int rc = 0;
    do {
        if (hipMemsetAsync(err.p, 0, 64, s) != hipSuccess) { rc = -1; break; }
        launch_rp_convert(d_irp, M, nnz, base, m.rp, m.rp32, err.as<int>(), s);
        if (hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = -1; break; }
        if (herr) {
            set_error("import: row offsets must be monotone and run from %d to nnz+%d", base, base);
            rc = -2;
            break;
        }
        if (nnz) {
            launch_col_convert(d_ja, m.rp, M, N, base, m.col, err.as<int>(), s);
            if (hipMemcpyAsync(m.val, d_val, (size_t)nnz * (vtype == SP3MM_F32 ? 4 : 8), hipMemcpyDeviceToDevice,
                               s) != hipSuccess) { rc = -1; break; }
        }
        if (hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = -1; break; }
        if (herr) {
            set_error("import: columns must lie in [%d, N+%d) and increase strictly inside each row", base, base);
            rc = -2;
        }
    } while (0);
    if (rc) {
        if (rc == -1) set_error("import: HIP error %s", hipGetErrorString(hipGetLastError()));
        dmat_release(m);
        return -1;
    }
    *out = m;
    return 0;
}

int sp3mm_dmat_export(const sp3mm_dmat *m, int64_t *d_irp, int64_t *d_ja, void *d_val, int32_t base)
{
    clear_error();
    if (!m || !d_irp || (base != 0 && base != 1)) {
        set_error("export: null argument or bad index base");
        return -1;
    }
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipStream_t s = c->stream;
    launch_rp_shift(m->rp, m->M + 1, base, d_irp, s);
    if (m->nnz) {
        if (d_ja) launch_col_widen(m->col, m->nnz, base, d_ja, s);
        if (d_val)
            SP3MM_HIP(hipMemcpyAsync(d_val, m->val, (size_t)m->nnz * (m->vtype == SP3MM_F32 ? 4 : 8),
                                     hipMemcpyDeviceToDevice, s));
    }
    SP3MM_HIP(hipGetLastError());
    SP3MM_HIP(hipStreamSynchronize(s));
    return 0;
}

int sp3mm_dmat_download(const sp3mm_dmat *m, int64_t *irp, int64_t *ja, void *val, int32_t base)
{
    clear_error();
    if (!m || !irp) {
        set_error("download: null argument");
        return -1;
    }
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipStream_t s = c->stream;
    const int64_t M = m->M, nnz = m->nnz;
    Buf trp((size_t)(M + 1) * 8), tja((size_t)(nnz ? nnz : 1) * 8);
    if (!trp.p || !tja.p) {
        set_error("download: staging allocation failed");
        return -1;
    }
    launch_rp_shift(m->rp, M + 1, base, trp.as<int64_t>(), s);
    SP3MM_HIP(hipMemcpyAsync(irp, trp.p, (size_t)(M + 1) * 8, hipMemcpyDeviceToHost, s));
    if (nnz) {
        launch_col_widen(m->col, nnz, base, tja.as<int64_t>(), s);
        if (ja) SP3MM_HIP(hipMemcpyAsync(ja, tja.p, (size_t)nnz * 8, hipMemcpyDeviceToHost, s));
        if (val)
            SP3MM_HIP(hipMemcpyAsync(val, m->val, (size_t)nnz * (m->vtype == SP3MM_F32 ? 4 : 8),
                                     hipMemcpyDeviceToHost, s));
    }
    SP3MM_HIP(hipStreamSynchronize(s));
    return 0;
}

int sp3mm_dmat_convert(sp3mm_dmat *m, int32_t vtype)
{
    clear_error();
    if (!m) return -1;
    if (m->vtype == vtype) return 0;
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int64_t n = m->nnz;
    void *nv = c->pool.alloc((size_t)(n ? n : 1) * (vtype == SP3MM_F32 ? 4 : 8));
    if (!nv) {
        set_error("convert: allocation failed");
        return -1;
    }
    if (vtype == SP3MM_F32)
        launch_f64_to_f32(static_cast<const double *>(m->val), n, static_cast<float *>(nv), c->stream);
    else
        launch_f32_to_f64(static_cast<const float *>(m->val), n, static_cast<double *>(nv), c->stream);
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    if (m->owned) c->pool.release(m->val);
    m->val = nv;
    m->vtype = vtype;
    return 0;
}

void sp3mm_dmat_free(sp3mm_dmat *m)
{
    if (!m) return;
    Context *c = ctx();
    if (c) (void)spin_stream(c->stream);
    dmat_release(*m);
}

int sp3mm_rap(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, sp3mm_dmat *out,
              const sp3mm_opts *opts, sp3mm_stats *stats)
{
    clear_error();
    if (!R || !A || !P || !out) {
        set_error("rap: null argument");
        return -1;
    }
    sp3mm_opts o{};
    if (opts) o = *opts;
    memset(out, 0, sizeof *out);
    if (o.precision == SP3MM_PREC_MIXED || o.precision == SP3MM_PREC_MIXED64) {
        if (o.path != SP3MM_PATH_DIRECT) {
            set_error("rap: the mixed-precision product runs on the direct path only");
            return -1;
        }
        return rap_direct_mixed(*R, *A, *P, *out, stats, o.precision, o.guard, o.rowsum, o.timing != 0);
    }
    if (o.precision != SP3MM_PREC_FP64) {
        set_error("rap: unknown precision %d", o.precision);
        return -1;
    }
    if (o.path == SP3MM_PATH_DIRECT) return rap_direct(*R, *A, *P, *out, stats, o.timing != 0);
    return rap_pair(*R, *A, *P, *out, stats, o.timing != 0);
}

int sp3mm_rap_noalloc(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int64_t *out_rp,
                      int32_t *out_col, double *out_val, int64_t cap, int64_t *nnz, sp3mm_stats *stats)
{
    clear_error();
    if (!R || !A || !P || !nnz) {
        set_error("rap_noalloc: null argument");
        return -1;
    }
    return rap_direct_noalloc(*R, *A, *P, out_rp, out_col, out_val, cap, *nnz, stats);
}

int sp3mm_spgemm(const sp3mm_dmat *A, const sp3mm_dmat *B, sp3mm_dmat *out, sp3mm_stats *stats)
{
    clear_error();
    if (!A || !B || !out) {
        set_error("spgemm: null argument");
        return -1;
    }
    memset(out, 0, sizeof *out);
    return spgemm(*A, *B, *out, stats);
}

int sp3mm_ub(const sp3mm_dmat *A, const sp3mm_dmat *B, int64_t *ub)
{
    clear_error();
    if (!A || !B || !ub) return -1;
    return row_ub(*A, *B, ub);
}

int sp3mm_symbolic(const sp3mm_dmat *A, const sp3mm_dmat *B, const sp3mm_dmat *C, int64_t *rowlens)
{
    clear_error();
    if (!A || !B || !rowlens) return -1;
    sp3mm_dmat out;
    memset(&out, 0, sizeof out);
    int rc = C ? rap_direct(*A, *B, *C, out, nullptr, false, true) : spgemm(*A, *B, out, nullptr, true);
    if (rc) return rc;
    std::vector<int64_t> rp((size_t)out.M + 1);
    Context *c = ctx();
    if (hipMemcpy(rp.data(), out.rp, (size_t)(out.M + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("symbolic: copy failed");
        dmat_release(out);
        return -1;
    }
    (void)c;
    for (int64_t r = 0; r < out.M; r++) rowlens[r] = rp[r + 1] - rp[r];
    rowlens[out.M] = rp[out.M];
    dmat_release(out);
    return 0;
}

int sp3mm_device_info(int32_t *num_cu, int64_t *lds_per_block, int64_t *lds_used)
{
    Context *c = ctx();
    if (!c) return -1;
    if (num_cu) *num_cu = c->num_cu;
    if (lds_per_block) *lds_per_block = (int64_t)c->lds_device;
    if (lds_used) *lds_used = (int64_t)c->max_lds;
    return 0;
}

int sp3mm_batch_create(int32_t count, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                       const sp3mm_opts *opts, int32_t use_graph, sp3mm_batch **out)
{
    clear_error();
    if (!R || !A || !P || !out) {
        set_error("batch: null argument");
        return -1;
    }
    sp3mm_opts o{};
    if (opts) o = *opts;
    if (o.path != SP3MM_PATH_DIRECT) {
        set_error("batch: batches run on the direct path");
        return -1;
    }
    if (o.precision != SP3MM_PREC_FP64 && o.precision != SP3MM_PREC_MIXED && o.precision != SP3MM_PREC_MIXED64) {
        set_error("batch: unknown precision %d", o.precision);
        return -1;
    }
    void *b = batch_create(count, R, A, P, o.precision, o.guard, use_graph != 0);
    *out = static_cast<sp3mm_batch *>(b);
    return b ? 0 : -1;
}

int sp3mm_batch_run(sp3mm_batch *b, int32_t use_graph, double *ms)
{
    clear_error();
    return batch_run(b, use_graph != 0, ms);
}

int sp3mm_batch_repack(sp3mm_batch *b, int32_t count, const sp3mm_dmat *R, const sp3mm_dmat *A,
                       const sp3mm_dmat *P)
{
    clear_error();
    if (!b || !R || !A || !P) {
        set_error("batch: null argument");
        return -1;
    }
    return batch_repack(b, count, R, A, P);
}

int sp3mm_batch_result(sp3mm_batch *b, int32_t i, sp3mm_dmat *out)
{
    clear_error();
    if (!out) return -1;
    memset(out, 0, sizeof *out);
    return batch_result(b, i, *out);
}

int sp3mm_batch_info(sp3mm_batch *b, double *ms_pack, double *ms_build, int64_t *nnz, int32_t *has_graph)
{
    if (!b) return -1;
    batch_info(b, ms_pack, ms_build, nnz, has_graph);
    return 0;
}

void sp3mm_batch_free(sp3mm_batch *b) { batch_free(b); }

static void to_c(const Tuning &t, sp3mm_tuning &o);

int sp3mm_batch_autotune(sp3mm_batch *b, int32_t reps, sp3mm_tuning *best, double *ms_best, double *ms_default)
{
    clear_error();
    if (!b) {
        set_error("batch: null handle");
        return -1;
    }
    Tuning t{};
    double mb = 0, md = 0;
    if (batch_autotune(b, reps, t, mb, md)) return -1;
    if (best) to_c(t, *best);
    if (ms_best) *ms_best = mb;
    if (ms_default) *ms_default = md;
    return 0;
}

int sp3mm_transpose(const sp3mm_dmat *A, sp3mm_dmat *AT)
{
    clear_error();
    if (!A || !AT) {
        set_error("transpose: null argument");
        return -1;
    }
    memset(AT, 0, sizeof *AT);
    return transpose(*A, *AT);
}

int sp3mm_sa_prolongator(const sp3mm_dmat *A, const int64_t *agg, int64_t nagg, double omega, int32_t diag_mode,
                         sp3mm_dmat *P)
{
    clear_error();
    if (!A || !P || (!agg && A->M > 0)) {
        set_error("prolongator: null argument");
        return -1;
    }
    memset(P, 0, sizeof *P);
    return sa_prolongator(*A, agg, nagg, omega, diag_mode, *P);
}

int sp3mm_hier_create(const sp3mm_dmat *A, int32_t nlevels, const int64_t *const *aggs, const int64_t *naggs,
                      double omega, int32_t diag_mode, int32_t use_graph, sp3mm_hier **out)
{
    clear_error();
    if (!A || !aggs || !naggs || !out || nlevels <= 0) {
        set_error("hierarchy: null argument or no levels");
        return -1;
    }
    for (int32_t l = 0; l < nlevels; l++)
        if (!aggs[l]) {
            set_error("hierarchy: level %d has no aggregate ids", l);
            return -1;
        }
    *out = static_cast<sp3mm_hier *>(hier_create(*A, nlevels, aggs, naggs, omega, diag_mode, use_graph != 0));
    return *out ? 0 : -1;
}

int sp3mm_hier_update(sp3mm_hier *h, int32_t use_graph, double *ms)
{
    clear_error();
    if (!h) {
        set_error("hierarchy: null handle");
        return -1;
    }
    return hier_update(h, use_graph != 0, ms);
}

int sp3mm_hier_level(sp3mm_hier *h, int32_t level, sp3mm_dmat *R, sp3mm_dmat *P, sp3mm_dmat *Ac)
{
    clear_error();
    if (!h || !R || !P || !Ac) {
        set_error("hierarchy: null argument");
        return -1;
    }
    return hier_level(h, level, *R, *P, *Ac);
}

int sp3mm_hier_info(sp3mm_hier *h, int32_t *nlevels, double *ms_build, int32_t *has_graph)
{
    clear_error();
    if (!h) {
        set_error("hierarchy: null handle");
        return -1;
    }
    hier_info(h, nlevels, ms_build, has_graph);
    return 0;
}

void sp3mm_hier_free(sp3mm_hier *h) { hier_free(h); }

int sp3mm_spmv(const sp3mm_dmat *A, const double *x, double *y, double alpha, double beta)
{
    clear_error();
    if (!A) {
        set_error("spmv: null argument");
        return -1;
    }
    return spmv(*A, x, y, y, alpha, beta);
}

int sp3mm_residual(const sp3mm_dmat *A, const double *b, const double *x, double *r)
{
    clear_error();
    if (!A) {
        set_error("residual: null argument");
        return -1;
    }
    if (A->M != A->N) {
        set_error("residual: A must be square");
        return -1;
    }
    return spmv(*A, x, b, r, -1.0, 1.0);
}

int sp3mm_axpby(int64_t n, double alpha, const double *x, double beta, double *y)
{
    clear_error();
    return axpby(n, alpha, x, beta, y);
}

int sp3mm_dot(int64_t n, const double *x, const double *y, double *out)
{
    clear_error();
    if (!out) {
        set_error("dot: null result pointer");
        return -1;
    }
    return dot(n, x, y, *out);
}

int sp3mm_jacobi(const sp3mm_dmat *A, const double *b, double *x, double *work, double omega, int32_t sweeps)
{
    clear_error();
    if (!A) {
        set_error("jacobi: null argument");
        return -1;
    }
    return jacobi(*A, b, x, work, omega, sweeps);
}

int sp3mm_copy_to_device(void *dst, const void *src, int64_t bytes)
{
    clear_error();
    if (bytes < 0 || (bytes && (!dst || !src))) {
        set_error("copy: bad arguments");
        return -1;
    }
    if (!bytes) return 0;
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    // stream-ordered after earlier calls that may still read dst
    SP3MM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    SP3MM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int sp3mm_vcycle_create(int32_t nlevels, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P,
                        const sp3mm_dmat *coarse, int32_t nu1, int32_t nu2, double omega, int32_t coarse_sweeps,
                        int32_t use_graph, sp3mm_vcycle **out)
{
    clear_error();
    if (!out || !coarse || nlevels < 1 || !R || !A || !P) {
        set_error("vcycle: null argument or no level");
        return -1;
    }
    *out = nullptr;
    void *v = vcycle_create(nlevels, R, A, P, *coarse, nu1, nu2, omega, coarse_sweeps, use_graph != 0);
    if (!v) return -1;
    *out = static_cast<sp3mm_vcycle *>(v);
    return 0;
}

int sp3mm_vcycle_solve(sp3mm_vcycle *v, const double *b, double *x, double rtol, int32_t maxit, int32_t *cycles,
                       double *hist, double *ms)
{
    clear_error();
    if (!v) {
        set_error("vcycle: null handle");
        return -1;
    }
    int k = 0;
    const int rc = vcycle_solve(v, b, x, rtol, maxit, &k, hist, ms);
    if (cycles) *cycles = k;
    return rc;
}

void sp3mm_vcycle_free(sp3mm_vcycle *v) { vcycle_free(v); }

static void to_c(const Tuning &t, sp3mm_tuning &o)
{
    o.u1 = t.u1;
    o.u2 = t.u2;
    o.subwave_h = t.subwave_h;
    o.slog1_cap = t.slog1_cap;
    o.slog2_cap = t.slog2_cap;
    o.h1_shift = t.h1_shift;
    o.h2_div = t.h2_div;
    o.xcd_map = t.xcd_map;
    o.hash1 = t.hash1;
    o.hash2 = t.hash2;
}

int sp3mm_get_tuning(sp3mm_tuning *t)
{
    clear_error();
    Context *c = ctx();
    if (!c || !t) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    to_c(c->tune, *t);
    return 0;
}

int sp3mm_set_tuning(const sp3mm_tuning *t)
{
    clear_error();
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Tuning n{};
    if (t) {
        n.u1 = t->u1;
        n.u2 = t->u2;
        n.subwave_h = t->subwave_h;
        n.slog1_cap = t->slog1_cap;
        n.slog2_cap = t->slog2_cap;
        n.h1_shift = t->h1_shift;
        n.h2_div = t->h2_div;
        n.xcd_map = t->xcd_map;
        n.hash1 = t->hash1 ? t->hash1 : kHashMul;
        n.hash2 = t->hash2 ? t->hash2 : kHashMul;
        if (!tuning_valid(n)) {
            set_error("set_tuning: parameter out of range");
            return -1;
        }
    }
    c->tune = n;
    return 0;
}

int sp3mm_autotune(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, const sp3mm_opts *opts,
                   int32_t reps, sp3mm_tuning *best, double *ms_best, double *ms_default)
{
    clear_error();
    if (!R || !A || !P) {
        set_error("autotune: null argument");
        return -1;
    }
    sp3mm_opts o{};
    if (opts) o = *opts;
    o.path = SP3MM_PATH_DIRECT;
    o.rowsum = nullptr;
    o.timing = 0;
    // one checked product first: validates the operands and warms the pool
    sp3mm_dmat C{};
    if (sp3mm_rap(R, A, P, &C, &o, nullptr)) return -1;
    sp3mm_dmat_free(&C);
    Tuning t{};
    double mb = 0, md = 0;
    if (autotune(*R, *A, *P, o.precision, o.guard, reps, t, mb, md)) return -1;
    if (best) to_c(t, *best);
    if (ms_best) *ms_best = mb;
    if (ms_default) *ms_default = md;
    return 0;
}

int32_t sp3mm_hash_scores(const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, int32_t H1, int32_t H2,
                          const int32_t *rows, int32_t nrows, int32_t cap, uint32_t *cand, uint64_t *score1,
                          uint64_t *score2, int32_t *nkeys)
{
    clear_error();
    if (!R || !A || !P || !rows || nrows <= 0 || !cand || !score1 || !score2 || !nkeys) {
        set_error("hash_scores: null argument or no rows");
        return -1;
    }
    if (R->N != A->M || A->N != P->M) {
        set_error("hash_scores: inner dimensions differ");
        return -1;
    }
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const std::vector<unsigned> &list = hash_candidate_list();
    if (cap < (int32_t)list.size()) {
        set_error("hash_scores: room for %d candidates, %zu needed", cap, list.size());
        return -1;
    }
    std::vector<unsigned long long> h;
    std::vector<int> keys;
    if (hash_scores(c, *R, *A, *P, H1, H2, rows, nrows, h, &keys)) return -1;
    const size_t nc = list.size();
    for (size_t i = 0; i < nc; i++) {
        cand[i] = list[i];
        score1[i] = h[i];
        score2[i] = h[nc + i];
    }
    for (size_t i = 0; i < keys.size(); i++) nkeys[i] = keys[i];
    return (int32_t)nc;
}

void *sp3mm_device_alloc(int64_t bytes)
{
    Context *c = ctx();
    if (!c || bytes < 0) return nullptr;
    return c->pool.alloc((size_t)(bytes ? bytes : 1));
}

void sp3mm_device_free(void *p)
{
    Context *c = ctx();
    if (!c || !p) return;
    (void)hipStreamSynchronize(c->stream);
    c->pool.release(p);
}

int sp3mm_copy_to_host(void *dst, const void *src, int64_t bytes)
{
    Context *c = ctx();
    if (!c) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    SP3MM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    return consumer_wait(c);
}

double sp3mm_copy_probe(int64_t bytes, int32_t reps)
{
    Context *c = ctx();
    if (!c || bytes <= 0 || reps <= 0) return -1.0;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    Buf a((size_t)bytes), b((size_t)bytes);
    if (!a.p || !b.p) return -1.0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch_copy_probe(a.p, b.p, bytes, c->stream);
    (void)hipEventRecord(e0, c->stream);
    for (int i = 0; i < reps; i++) launch_copy_probe(a.p, b.p, bytes, c->stream);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

}  // extern "C"
