// setup.hip -- the step before the Galerkin product in an aggregation AMG
// setup, on the GPU: the transpose R = P^T and the smoothed-aggregation
// prolongator P = (I - omega D^-1 A) P0 (P0 = the 0/1 aggregation matrix).
// The reference receives P and R from its caller; these kernels let a whole
// hierarchy be built level after level without leaving HBM.
#include <cstdlib>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sp3mm {

// lanes per row for row-parallel kernels: a power of two near the average row
// length / SP3MM_SPMV_ROW_DIV (default 4), at most a wavefront
static int lanes_for(int64_t M, int64_t nnz)
{
    static const double div = [] {
        const char *e = getenv("SP3MM_SPMV_ROW_DIV");
        const double v = e ? atof(e) : 4.0;
        return v >= 1.0 && v <= 64.0 ? v : 4.0;
    }();
    const double avg = M ? (double)nnz / (double)M / div : 1.0;
    int L = 1;
    while (L < 64 && L < avg) L *= 2;
    return L;
}

#define SP3MM_BY_LANES(K, ...)                                                                   \
    switch (L) {                                                                                 \
    case 1: hipLaunchKernelGGL(K<1>, __VA_ARGS__); break;                                        \
    case 2: hipLaunchKernelGGL(K<2>, __VA_ARGS__); break;                                        \
    case 4: hipLaunchKernelGGL(K<4>, __VA_ARGS__); break;                                        \
    case 8: hipLaunchKernelGGL(K<8>, __VA_ARGS__); break;                                        \
    case 16: hipLaunchKernelGGL(K<16>, __VA_ARGS__); break;                                      \
    case 32: hipLaunchKernelGGL(K<32>, __VA_ARGS__); break;                                      \
    default: hipLaunchKernelGGL(K<64>, __VA_ARGS__); break;                                      \
    }

// ---------------------------------------------------------------------------
// transpose: per-column counts, scan (the caller), scatter, per-row sort
// ---------------------------------------------------------------------------
__global__ void k_col_count(const int32_t *__restrict__ col, int64_t nnz, int32_t *__restrict__ cnt)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) atomicAdd(cnt + col[e], 1);
}

// L lanes per row of A: its entries go to the rows of A^T named by their
// columns, at the next free position of that row (order fixed afterwards);
// with tsrc, each entry's index in A goes along (the transpose's permutation)
template <typename VT, int L>
__global__ void __launch_bounds__(256) k_transpose_scatter(const int64_t *__restrict__ rp,
                                                            const int32_t *__restrict__ col,
                                                            const VT *__restrict__ val, int64_t M,
                                                            const int64_t *__restrict__ trp,
                                                            int32_t *__restrict__ cursor, int32_t *__restrict__ tcol,
                                                            VT *__restrict__ tval, int32_t *__restrict__ tsrc)
{
    const int sub = threadIdx.x & (L - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; r < M; r += stride) {
        for (int64_t e = rp[r] + sub; e < rp[r + 1]; e += L) {
            const int c = col[e];
            const int64_t pos = trp[c] + atomicAdd(cursor + c, 1);
            tcol[pos] = (int32_t)r;
            tval[pos] = val[e];
            if (tsrc) tsrc[pos] = (int32_t)e;
        }
    }
}

// one wavefront per row of A^T: rank sort of its (distinct) column indices;
// keys of rows up to kSortLds entries are staged in LDS, longer rows are
// ranked from global memory; isrc/osrc (optional) move with the values
constexpr int kSortLds = 1024;

template <typename VT>
__global__ void __launch_bounds__(256) k_sort_rows(const int64_t *__restrict__ rp, int64_t M,
                                                    const int32_t *__restrict__ icol, const VT *__restrict__ ival,
                                                    const int32_t *__restrict__ isrc, int32_t *__restrict__ ocol,
                                                    VT *__restrict__ oval, int32_t *__restrict__ osrc)
{
    __shared__ int keys[4][kSortLds];
    const int w = threadIdx.x / kWave;
    const int l = lane();
    int *kw = keys[w];
    for (int64_t r = (int64_t)blockIdx.x * 4 + w; r < M; r += (int64_t)gridDim.x * 4) {
        const int64_t b = rp[r];
        const int n = (int)(rp[r + 1] - b);
        const bool staged = n <= kSortLds;
        if (staged)
            for (int j = l; j < n; j += kWave) kw[j] = icol[b + j];
        wave_sync();
        const int *src = staged ? kw : icol + b;
        for (int j = l; j < n; j += kWave) {
            const int k = src[j];
            int rank = 0;
            for (int i = 0; i < n; i++) rank += src[i] < k;
            ocol[b + rank] = k;
            oval[b + rank] = ival[b + j];
            if (osrc) osrc[b + rank] = isrc[b + j];
        }
        wave_sync();
    }
}

// out[j] = src[perm[j]]: the values of a transpose whose pattern is known
template <typename VT>
__global__ void __launch_bounds__(256) k_gather_vals(const VT *__restrict__ src, const int32_t *__restrict__ perm,
                                                      int64_t n, VT *__restrict__ out)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        out[j] = src[perm[j]];
}

// ---------------------------------------------------------------------------
// S = I - omega D^-1 A on the pattern of A (a missing diagonal is inserted)
// ---------------------------------------------------------------------------
// d_i = a_ii (mode 0) or |a_ii| + 1 (mode 1).  L lanes per row; row r's
// diagonal value, whether it is stored and how many of its columns are
// smaller, summed over the lanes.
template <int L>
__device__ __forceinline__ void sa_row_scan(const int32_t *__restrict__ col, const double *__restrict__ val,
                                            int64_t b, int64_t e1, int64_t r, int sub, double &d, int &nless,
                                            int &has)
{
    d = 0.0;
    nless = 0;
    has = 0;
    for (int64_t e = b + sub; e < e1; e += L) {
        const int c = col[e];
        if (c == r) {
            if (val) d = val[e];
            has = 1;
        }
        nless += c < r;
    }
    for (int o = L / 2; o > 0; o >>= 1) {
        d += __shfl_xor(d, o, L);
        nless += __shfl_xor(nless, o, L);
        has |= __shfl_xor(has, o, L);
    }
}

// count pass: row lengths of S
template <int L>
__global__ void __launch_bounds__(256) k_sa_count(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                   int64_t M, int32_t *__restrict__ len)
{
    const int sub = threadIdx.x & (L - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; r < M; r += stride) {
        double d;
        int nless, has;
        sa_row_scan<L>(col, nullptr, rp[r], rp[r + 1], r, sub, d, nless, has);
        if (sub == 0) len[r] = (int32_t)(rp[r + 1] - rp[r]) + (has ? 0 : 1);
    }
}

// fill pass: S's columns and values at the offsets srp of the count pass (the
// missing diagonal goes before the first larger column)
template <int L>
__global__ void __launch_bounds__(256) k_sa_fill(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const double *__restrict__ val, int64_t M, double omega, int mode,
                                                  const int64_t *__restrict__ srp, int32_t *__restrict__ scol,
                                                  double *__restrict__ sval, int *__restrict__ err)
{
    const int sub = threadIdx.x & (L - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; r < M; r += stride) {
        const int64_t b = rp[r], e1 = rp[r + 1];
        double d;
        int nless, has;
        sa_row_scan<L>(col, val, b, e1, r, sub, d, nless, has);
        if (mode == 1) d = fabs(d) + 1.0;
        if (d == 0.0) {
            if (sub == 0) atomicOr(err, kErrBadInput);
            continue;
        }
        const double w = omega / d;
        const int64_t o0 = srp[r];
        // S's row r was sized when the plan was built; a row of A that has
        // grown since (its pattern changed) is reported, not written past
        if (o0 + (e1 - b) + (has ? 0 : 1) != srp[r + 1]) {
            if (sub == 0) atomicOr(err, kErrOverflow);
            continue;
        }
        for (int64_t e = b + sub; e < e1; e += L) {
            const int c = col[e];
            const int64_t pos = o0 + (e - b) + (!has && c > r ? 1 : 0);
            scol[pos] = c;
            sval[pos] = -w * val[e] + (c == r ? 1.0 : 0.0);
        }
        if (!has && sub == 0) {
            scol[o0 + nless] = (int32_t)r;
            sval[o0 + nless] = 1.0;
        }
    }
}

// P0: one entry 1.0 per row at column agg[r]; rejects aggregates out of range
__global__ void k_p0_fill(const int64_t *__restrict__ agg, int64_t M, int64_t nagg, int64_t *__restrict__ rp,
                          int32_t *__restrict__ rp32, int32_t *__restrict__ col, double *__restrict__ val,
                          int *__restrict__ err)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= M; r += stride) {
        rp[r] = r;
        if (rp32) rp32[r] = (int32_t)r;
        if (r < M) {
            const int64_t a = agg[r];
            if (a < 0 || a >= nagg) atomicOr(err, kErrBadInput);
            col[r] = (int32_t)(a < 0 || a >= nagg ? 0 : a);
            val[r] = 1.0;
        }
    }
}

static unsigned grid1(int64_t n, int threads = 256)
{
    int64_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

void launch_col_count(const int32_t *col, int64_t nnz, int32_t *cnt, hipStream_t st)
{
    if (nnz > 0) hipLaunchKernelGGL(k_col_count, dim3(grid1(nnz)), dim3(256), 0, st, col, nnz, cnt);
}

void launch_transpose_scatter(const int64_t *rp, const int32_t *col, const void *val, bool f32, int64_t M,
                              int64_t nnz, const int64_t *trp, int32_t *cursor, int32_t *tcol, void *tval,
                              int32_t *tsrc, hipStream_t st)
{
    if (M == 0) return;
    const int L = lanes_for(M, nnz);
    const unsigned g = grid1(M, 256 / L);
    if (f32) {
        const float *v = static_cast<const float *>(val);
        float *tv = static_cast<float *>(tval);
        switch (L) {
        case 1: hipLaunchKernelGGL((k_transpose_scatter<float, 1>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 2: hipLaunchKernelGGL((k_transpose_scatter<float, 2>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 4: hipLaunchKernelGGL((k_transpose_scatter<float, 4>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 8: hipLaunchKernelGGL((k_transpose_scatter<float, 8>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 16: hipLaunchKernelGGL((k_transpose_scatter<float, 16>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 32: hipLaunchKernelGGL((k_transpose_scatter<float, 32>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        default: hipLaunchKernelGGL((k_transpose_scatter<float, 64>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        }
    } else {
        const double *v = static_cast<const double *>(val);
        double *tv = static_cast<double *>(tval);
        switch (L) {
        case 1: hipLaunchKernelGGL((k_transpose_scatter<double, 1>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 2: hipLaunchKernelGGL((k_transpose_scatter<double, 2>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 4: hipLaunchKernelGGL((k_transpose_scatter<double, 4>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 8: hipLaunchKernelGGL((k_transpose_scatter<double, 8>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 16: hipLaunchKernelGGL((k_transpose_scatter<double, 16>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        case 32: hipLaunchKernelGGL((k_transpose_scatter<double, 32>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        default: hipLaunchKernelGGL((k_transpose_scatter<double, 64>), dim3(g), dim3(256), 0, st, rp, col, v, M, trp, cursor, tcol, tv, tsrc); break;
        }
    }
}

void launch_sort_rows(const int64_t *rp, int64_t M, const int32_t *icol, const void *ival, const int32_t *isrc,
                      bool f32, int32_t *ocol, void *oval, int32_t *osrc, hipStream_t st)
{
    if (M == 0) return;
    const unsigned g = grid1(M, 4);
    if (f32)
        hipLaunchKernelGGL(k_sort_rows<float>, dim3(g), dim3(256), 0, st, rp, M, icol,
                           static_cast<const float *>(ival), isrc, ocol, static_cast<float *>(oval), osrc);
    else
        hipLaunchKernelGGL(k_sort_rows<double>, dim3(g), dim3(256), 0, st, rp, M, icol,
                           static_cast<const double *>(ival), isrc, ocol, static_cast<double *>(oval), osrc);
}

void launch_gather_vals(const void *src, const int32_t *perm, int64_t n, bool f32, void *out, hipStream_t st)
{
    if (n == 0) return;
    if (f32)
        hipLaunchKernelGGL(k_gather_vals<float>, dim3(grid1(n)), dim3(256), 0, st, static_cast<const float *>(src),
                           perm, n, static_cast<float *>(out));
    else
        hipLaunchKernelGGL(k_gather_vals<double>, dim3(grid1(n)), dim3(256), 0, st,
                           static_cast<const double *>(src), perm, n, static_cast<double *>(out));
}

void launch_sa_count(const int64_t *rp, const int32_t *col, int64_t M, int64_t nnz, int32_t *len, hipStream_t st)
{
    if (M == 0) return;
    const int L = lanes_for(M, nnz);
    SP3MM_BY_LANES(k_sa_count, dim3(grid1(M, 256 / L)), dim3(256), 0, st, rp, col, M, len)
}

void launch_sa_fill(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, double omega,
                    int mode, const int64_t *srp, int32_t *scol, double *sval, int *err, hipStream_t st)
{
    if (M == 0) return;
    const int L = lanes_for(M, nnz);
    SP3MM_BY_LANES(k_sa_fill, dim3(grid1(M, 256 / L)), dim3(256), 0, st, rp, col, val, M, omega, mode, srp, scol,
                   sval, err)
}

void launch_p0_fill(const int64_t *agg, int64_t M, int64_t nagg, int64_t *rp, int32_t *rp32, int32_t *col,
                    double *val, int *err, hipStream_t st)
{
    hipLaunchKernelGGL(k_p0_fill, dim3(grid1(M + 1)), dim3(256), 0, st, agg, M, nagg, rp, rp32, col, val, err);
}

}  // namespace sp3mm

namespace sp3mm {

// ---------------------------------------------------------------------------
// consumers of a level: y = alpha A x + beta y, and damped Jacobi sweeps
// ---------------------------------------------------------------------------
// y = alpha A x + beta z (z may be y; with beta == 0 z is not read).
// L lanes per row (L a power of two up to 64), partial sums reduced by shuffles
template <int L>
__global__ void __launch_bounds__(256) k_spmv(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                               const double *__restrict__ val, int64_t M,
                                               const double *__restrict__ x, const double *z, double *y,
                                               double alpha, double beta)
{
    const int sub = threadIdx.x & (L - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; r < M; r += stride) {
        double s = 0.0;
        for (int64_t e = rp[r] + sub; e < rp[r + 1]; e += L) s += val[e] * x[col[e]];
        for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (sub == 0) y[r] = alpha * s + (beta != 0.0 ? beta * z[r] : 0.0);
    }
}

// y = alpha x + beta y; a zero coefficient skips the read of its vector
__global__ void __launch_bounds__(256) k_axpby(int64_t n, double alpha, const double *x, double beta, double *y)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (alpha != 0.0 ? alpha * x[i] : 0.0) + (beta != 0.0 ? beta * y[i] : 0.0);
}

// Dot product in two launches with a fixed summation order (the result does
// not depend on scheduling): kDotBlocks partial sums, then one block adds them.
constexpr int kDotBlocks = 1024;
__global__ void __launch_bounds__(256) k_dot_partial(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ y, double *__restrict__ part)
{
    __shared__ double red[256 / kWave];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += x[i] * y[i];
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane() == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 256 / kWave; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(256) k_dot_final(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double red[256 / kWave];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane() == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 256 / kWave; w++) t += red[w];
        *out = t;
    }
}

// xo = xi + omega (b - A xi) / a_ii
template <int L>
__global__ void __launch_bounds__(256) k_jacobi(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ val, int64_t M,
                                                 const double *__restrict__ b, const double *__restrict__ xi,
                                                 double *__restrict__ xo, double omega, int *__restrict__ err)
{
    const int sub = threadIdx.x & (L - 1);
    const int64_t stride = (int64_t)gridDim.x * (blockDim.x / L);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / L) + threadIdx.x / L; r < M; r += stride) {
        double s = 0.0, d = 0.0;
        for (int64_t e = rp[r] + sub; e < rp[r + 1]; e += L) {
            const int c = col[e];
            s += val[e] * xi[c];
            if (c == r) d = val[e];
        }
        for (int o = L / 2; o > 0; o >>= 1) {
            s += __shfl_xor(s, o);
            d += __shfl_xor(d, o);
        }
        if (sub == 0) {
            if (d == 0.0) atomicOr(err, kErrBadInput);
            else xo[r] = xi[r] + omega * (b[r] - s) / d;
        }
    }
}

// All `sweeps` sweeps of a small system in one workgroup, with the matrix,
// b and both x vectors staged in LDS once (a coarsest-level solve is a few
// kilobytes; a launch per sweep would cost more than the sweep).  Rows are
// split over L-lane groups as in k_jacobi, with the same per-row summation
// order, so the result equals `sweeps` launches of k_jacobi.
constexpr int kSmallRows = 2048;
template <int L>
__global__ void __launch_bounds__(1024) k_jacobi_small(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                        const double *__restrict__ val, int M, int nnz,
                                                        const double *__restrict__ b, double *__restrict__ x,
                                                        double omega, int sweeps, int *__restrict__ err)
{
    extern __shared__ double lds[];
    double *sv = lds;                   // nnz values
    double *sb = sv + nnz;              // M
    double *xs0 = sb + M, *xs1 = xs0 + M;
    int *sc = reinterpret_cast<int *>(xs1 + M);  // nnz columns
    int *sr = sc + nnz;                 // M + 1 offsets from rp[0]
    const int64_t base = rp[0];
    for (int i = threadIdx.x; i < nnz; i += blockDim.x) {
        sv[i] = val[base + i];
        sc[i] = col[base + i];
    }
    for (int i = threadIdx.x; i <= M; i += blockDim.x) sr[i] = (int)(rp[i] - base);
    for (int i = threadIdx.x; i < M; i += blockDim.x) {
        sb[i] = b[i];
        xs0[i] = x[i];
    }
    __syncthreads();
    const int sub = threadIdx.x & (L - 1);
    double *xi = xs0, *xo = xs1;
    for (int k = 0; k < sweeps; k++) {
        for (int r = threadIdx.x / L; r < M; r += blockDim.x / L) {
            double s = 0.0, d = 0.0;
            for (int e = sr[r] + sub; e < sr[r + 1]; e += L) {
                const int c = sc[e];
                s += sv[e] * xi[c];
                if (c == r) d = sv[e];
            }
            for (int o = L / 2; o > 0; o >>= 1) {
                s += __shfl_xor(s, o);
                d += __shfl_xor(d, o);
            }
            if (sub == 0) {
                if (d == 0.0) atomicOr(err, kErrBadInput);
                else xo[r] = xi[r] + omega * (sb[r] - s) / d;
            }
        }
        __syncthreads();
        double *t = xi;
        xi = xo;
        xo = t;
    }
    for (int i = threadIdx.x; i < M; i += blockDim.x) x[i] = xi[i];
}

static size_t jacobi_small_lds(int64_t M, int64_t nnz) { return (size_t)nnz * 12 + (size_t)M * 24 + (size_t)(M + 1) * 4; }

// Lanes per row: the smallest power of two (at most 64) not below the mean
// row length divided by SP3MM_SPMV_ROW_DIV (default 4: about four entries
// per lane, measured best on SA hierarchies of 7-point operators).
void launch_spmv(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, const double *x,
                 const double *z, double *y, double alpha, double beta, hipStream_t st)
{
    if (M == 0) return;
    const int L = lanes_for(M, nnz);
    const unsigned g = grid1(M, 256 / L);
    SP3MM_BY_LANES(k_spmv, dim3(g), dim3(256), 0, st, rp, col, val, M, x, z, y, alpha, beta)
}

void launch_axpby(int64_t n, double alpha, const double *x, double beta, double *y, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_axpby, dim3(grid1(n, 256)), dim3(256), 0, st, n, alpha, x, beta, y);
}

void launch_dot(int64_t n, const double *x, const double *y, double *part, double *out, hipStream_t st)
{
    const unsigned nb = n ? (unsigned)std::min<int64_t>((n + 255) / 256, kDotBlocks) : 1;
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, st, n, x, y, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, part, (int)nb, out);
}

int dot_parts() { return kDotBlocks; }

bool launch_jacobi_small(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz,
                         const double *b, double *x, double omega, int sweeps, int *err, size_t max_lds,
                         hipStream_t st)
{
    if (M <= 0 || M > kSmallRows || jacobi_small_lds(M, nnz) > max_lds) return false;
    const int L = lanes_for(M, nnz);
    const size_t lds = jacobi_small_lds(M, nnz);
    SP3MM_BY_LANES(k_jacobi_small, dim3(1), dim3(1024), lds, st, rp, col, val, (int)M, (int)nnz, b, x, omega,
                   sweeps, err)
    return true;
}

void launch_jacobi(const int64_t *rp, const int32_t *col, const double *val, int64_t M, int64_t nnz, const double *b,
                   const double *xi, double *xo, double omega, int *err, hipStream_t st)
{
    if (M == 0) return;
    const int L = lanes_for(M, nnz);
    const unsigned g = grid1(M, 256 / L);
    SP3MM_BY_LANES(k_jacobi, dim3(g), dim3(256), 0, st, rp, col, val, M, b, xi, xo, omega, err)
}
#undef SP3MM_BY_LANES

// ---------------------------------------------------------------------------
// replay guards of a hierarchy plan
// ---------------------------------------------------------------------------
// 64-bit finaliser of the splitmix64 generator (Steele, Lea, Flood 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// out += sum over the M+1 offsets and nnz columns of mix64(mix64(position) +
// entry): an order-sensitive fingerprint of a CSR pattern (wrapping sums, so
// the result does not depend on the schedule)
__global__ void __launch_bounds__(256) k_fingerprint(const int64_t *__restrict__ rp, int64_t M,
                                                     const int32_t *__restrict__ col, int64_t nnz,
                                                     unsigned long long *out)
{
    unsigned long long acc = 0;
    const int64_t n = M + 1 + nnz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t v = i <= M ? (uint64_t)rp[i] : (uint64_t)(uint32_t)col[i - M - 1];
        acc += mix64(mix64((uint64_t)i) + v);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane() == 0) atomicAdd(out, acc);
}

void launch_fingerprint(const int64_t *rp, int64_t M, const int32_t *col, int64_t nnz, unsigned long long *out,
                        hipStream_t st)
{
    launch_fill_i32(reinterpret_cast<int32_t *>(out), 2, 0, st);
    const unsigned g = (unsigned)std::min<int64_t>((M + 1 + nnz + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fingerprint, dim3(g), dim3(256), 0, st, rp, M, col, nnz, out);
}

// *out = bitwise OR of the n words *words[i]
__global__ void k_or_words(const int *const *words, int n, int *out)
{
    int v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v |= *words[i];
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    if (threadIdx.x == 0) *out = v;
}

void launch_or_words(const int *const *words, int n, int *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_or_words, dim3(1), dim3(64), 0, st, words, n, out);
}

}  // namespace sp3mm
