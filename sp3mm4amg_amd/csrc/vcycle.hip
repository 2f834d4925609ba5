// vcycle.hip -- a V-cycle over a device hierarchy, captured as one HIP graph.
//
// The consumer of the coarse operators: damped Jacobi smoothing on every
// level, restriction with R, Jacobi sweeps from zero on the coarsest level,
// prolongation with P and correction.  One cycle is a fixed sequence of
// kernel launches on the library stream (SpMV, Jacobi, axpby, dot), so it is
// recorded once into a graph and replayed per cycle; the host waits once per
// cycle, for the residual norm that ends the graph.
#include <cmath>

#include "engine.hpp"

namespace sp3mm {

namespace {

struct VCycle {
    int nlev = 0;                      // levels above the coarsest
    std::vector<sp3mm_dmat> R, A, P;   // the caller's matrices (not owned)
    sp3mm_dmat Ac{};
    int nu1 = 2, nu2 = 2, csweeps = 50;
    double omega = 2.0 / 3.0;
    // per level 0..nlev: solution, right-hand side, residual, Jacobi buffer
    std::vector<double *> x, b, r, w;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int64_t rows_of(const VCycle &v, int l) { return l < v.nlev ? v.A[l].M : v.Ac.M; }

// One cycle for b[0] on x[0], then r[0] = b[0] - A[0] x[0] and its squared
// norm into the consumer state's dot word.  Same operations, in the same
// order, as a cycle driven call by call through the ABI.
void record_cycle(Context *c, VCycle &v, hipStream_t s)
{
    ConsState *st = static_cast<ConsState *>(c->cons);
    int *err = &st->err;
    auto spmv_q = [&](const sp3mm_dmat &M, const double *x, const double *z, double *y, double al, double be) {
        launch_spmv(M.rp, M.col, static_cast<const double *>(M.val), M.M, M.nnz, x, z, y, al, be, s);
    };
    for (int l = 0; l < v.nlev; l++) {
        if (l > 0) launch_axpby(v.A[l].M, 0.0, nullptr, 0.0, v.x[l], s);  // a correction starts at zero
        queue_jacobi(v.A[l], v.b[l], v.x[l], v.w[l], v.omega, v.nu1, err, s);
        spmv_q(v.A[l], v.x[l], v.b[l], v.r[l], -1.0, 1.0);             // r = b - A x
        spmv_q(v.R[l], v.r[l], nullptr, v.b[l + 1], 1.0, 0.0);          // restrict
    }
    const int n = v.nlev;
    launch_axpby(v.Ac.M, 0.0, nullptr, 0.0, v.x[n], s);
    queue_jacobi(v.Ac, v.b[n], v.x[n], v.w[n], v.omega, v.csweeps, err, s);
    for (int l = n - 1; l >= 0; l--) {
        spmv_q(v.P[l], v.x[l + 1], v.x[l], v.x[l], 1.0, 1.0);          // prolongate and correct
        queue_jacobi(v.A[l], v.b[l], v.x[l], v.w[l], v.omega, v.nu2, err, s);
    }
    spmv_q(v.A[0], v.x[0], v.b[0], v.r[0], -1.0, 1.0);
    launch_dot(v.A[0].M, v.r[0], v.r[0], reinterpret_cast<double *>(st + 1), &st->dot, s);
}

void release(Context *c, VCycle *v)
{
    if (v->exec) (void)hipGraphExecDestroy(v->exec);
    if (v->graph) (void)hipGraphDestroy(v->graph);
    for (auto *vec : {&v->x, &v->b, &v->r, &v->w})
        for (double *p : *vec)
            if (p) c->pool.release(p);
    delete v;
}

}  // namespace

void *vcycle_create(int nlev, const sp3mm_dmat *R, const sp3mm_dmat *A, const sp3mm_dmat *P, const sp3mm_dmat &Ac,
                    int nu1, int nu2, double omega, int coarse_sweeps, bool graph)
{
    Context *c = ctx();
    if (!c) return nullptr;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    if (nlev < 1 || nu1 < 0 || nu2 < 0 || coarse_sweeps < 1 || nu1 > 100 || nu2 > 100 || coarse_sweeps > 10000) {
        set_error("vcycle: bad level count or sweep counts");
        return nullptr;
    }
    for (int l = 0; l < nlev; l++) {
        const sp3mm_dmat &a = A[l], &r = R[l], &p = P[l];
        const sp3mm_dmat &nxt = l + 1 < nlev ? A[l + 1] : Ac;
        if (a.M != a.N || p.M != a.M || r.N != a.M || r.M != p.N || nxt.M != r.M || nxt.N != r.M) {
            set_error("vcycle: level %d: operator and transfer shapes do not chain", l);
            return nullptr;
        }
        if (a.vtype != SP3MM_F64 || r.vtype != SP3MM_F64 || p.vtype != SP3MM_F64) {
            set_error("vcycle: level %d: FP64 matrices are required", l);
            return nullptr;
        }
    }
    if (Ac.vtype != SP3MM_F64 || Ac.M < 1) {
        set_error("vcycle: the coarsest operator must be a non-empty FP64 matrix");
        return nullptr;
    }
    if (!cons_state(c)) return nullptr;
    VCycle *v = new VCycle;
    v->nlev = nlev;
    v->R.assign(R, R + nlev);
    v->A.assign(A, A + nlev);
    v->P.assign(P, P + nlev);
    v->Ac = Ac;
    v->nu1 = nu1;
    v->nu2 = nu2;
    v->csweeps = coarse_sweeps;
    v->omega = omega;
    for (int l = 0; l <= nlev; l++) {
        const size_t bytes = (size_t)std::max<int64_t>(rows_of(*v, l), 1) * 8;
        for (auto *vec : {&v->x, &v->b, &v->r, &v->w}) {
            vec->push_back(static_cast<double *>(c->pool.alloc(bytes)));
            if (!vec->back()) {
                set_error("vcycle: vector allocation failed");
                release(c, v);
                return nullptr;
            }
        }
    }
    if (graph) {
        hipStream_t s = c->stream;
        if (hipStreamSynchronize(s) != hipSuccess || hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            set_error("vcycle: stream capture failed to start");
            release(c, v);
            return nullptr;
        }
        record_cycle(c, *v, s);
        if (hipStreamEndCapture(s, &v->graph) != hipSuccess ||
            hipGraphInstantiate(&v->exec, v->graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            set_error("vcycle: graph capture failed");
            release(c, v);
            return nullptr;
        }
    }
    return v;
}

int vcycle_solve(void *h, const double *b, double *x, double rtol, int maxit, int *cycles, double *hist, double *ms)
{
    Context *c = ctx();
    if (!c || !h) return -1;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    VCycle &v = *static_cast<VCycle *>(h);
    if (!b || !x || maxit < 0) {
        set_error("vcycle: null vector or negative cycle limit");
        return -1;
    }
    hipStream_t s = c->stream;
    ConsState *st = static_cast<ConsState *>(c->cons);
    const int64_t n0 = v.A[0].M;
    launch_axpby(n0, 1.0, b, 0.0, v.b[0], s);
    launch_axpby(n0, 1.0, x, 0.0, v.x[0], s);
    double bb = 0.0, rr = 0.0;
    launch_dot(n0, v.b[0], v.b[0], reinterpret_cast<double *>(st + 1), &st->dot, s);
    SP3MM_HIP(hipGetLastError());
    if (consumer_wait(c, &bb)) return -1;
    const double bn = bb > 0.0 ? std::sqrt(bb) : 1.0;
    launch_spmv(v.A[0].rp, v.A[0].col, static_cast<const double *>(v.A[0].val), n0, v.A[0].nnz, v.x[0], v.b[0],
                v.r[0], -1.0, 1.0, s);
    launch_dot(n0, v.r[0], v.r[0], reinterpret_cast<double *>(st + 1), &st->dot, s);
    SP3MM_HIP(hipGetLastError());
    if (consumer_wait(c, &rr)) return -1;
    double rel = std::sqrt(rr) / bn;
    if (hist) hist[0] = rel;
    hipEvent_t e0, e1;
    SP3MM_HIP(hipEventCreate(&e0));
    SP3MM_HIP(hipEventCreate(&e1));
    SP3MM_HIP(hipEventRecord(e0, s));
    int k = 0;
    int rc = 0;
    while (k < maxit && rel > rtol) {
        if (v.exec) {
            if (hipGraphLaunch(v.exec, s) != hipSuccess) {
                set_error("vcycle: graph launch failed");
                rc = -1;
                break;
            }
        } else {
            record_cycle(c, v, s);
        }
        if (consumer_wait(c, &rr)) {
            rc = -1;
            break;
        }
        rel = std::sqrt(rr) / bn;
        k++;
        if (hist) hist[k] = rel;
    }
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    if (cycles) *cycles = k;
    if (rc) return rc;
    launch_axpby(n0, 1.0, v.x[0], 0.0, x, s);
    SP3MM_HIP(hipGetLastError());
    return consumer_wait(c);
}

void vcycle_free(void *h)
{
    Context *c = ctx();
    if (!c || !h) return;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    (void)hipStreamSynchronize(c->stream);
    release(c, static_cast<VCycle *>(h));
}

}  // namespace sp3mm
