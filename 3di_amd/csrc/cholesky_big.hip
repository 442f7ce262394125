// cholesky_big.hip — FP64 Cholesky solve for normal matrices too large for one workgroup's LDS.
//
// Solver::solve (solver.h:70-105: cusolverDnDpotrf + Dpotrs on the FP64 copy of the FP32 normal
// matrix, copyMatsFloatToDouble / copyVectorDoubleToFloatNegated, solver.cu:11-81) for the
// joint identity-stage system (K up to a few thousand: 199 + 199 + 94 per frame).  Blocked
// right-looking factorisation in place on an FP64 copy in HBM, 32-column panels:
//   k_chol_diag    one wavefront per matrix: the 32x32 diagonal block factored in registers,
//                  lane i holding row i, the column entries passed by lane reads;
//   k_chol_offdiag a grid of 64-row blocks per matrix: the panel below it, L21 = A21 L11^-T
//                  (one row per thread in registers, L11 and its reciprocal diagonal in LDS);
//   k_chol_update  a grid of 32x32 tiles of the trailing lower triangle per matrix:
//                  A22 -= L21 L21^T, both 32-row slices of the panel staged in LDS;
//   k_chol_trsv    one workgroup per matrix: L y = -g, then L^T delta = y, by 32-row panels,
//                  right-looking (the solution and the panel's diagonal block in LDS; each
//                  panel's triangle on one wavefront from registers, then every later entry
//                  updated by all threads with coalesced reads of the panel's rows or columns).
// A non-positive (or NaN) pivot marks the matrix TDI_NOT_PD; its later launches do nothing.
#include "common.h"

namespace tdi {

namespace {

constexpr int CB = 32;   // panel width

__global__ void __launch_bounds__(256) k_f32_to_f64(size_t n, const float* __restrict__ in, double* __restrict__ out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// lane j's value of v (j uniform)
__device__ __forceinline__ double lane_value(double v, int j)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}

// The 32 x 32 diagonal block of the panel at k0, factored on one wavefront and written back:
// lane i holds row i (entries j <= i; rows and columns past the matrix padded with the identity).
__global__ void __launch_bounds__(64) k_chol_diag(int K, int k0, double* __restrict__ A, int* __restrict__ status)
{
    const int mtx = blockIdx.x;
    if (status[mtx] != TDI_OK) return;
    double* M = A + (size_t)mtx * K * K;
    const int nb = min(CB, K - k0), lane = threadIdx.x;
    const bool row = lane < nb;
    const double* Mr = M + (size_t)(k0 + (row ? lane : 0)) * K + k0;
    double r[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) r[j] = (row && j < nb && j <= lane) ? Mr[min(j, nb - 1)] : (lane == j ? 1.0 : 0.0);
    // column k: pivot from lane k, the rows below scaled, then the trailing rows updated
#pragma unroll
    for (int k = 0; k < CB; ++k) {
        const double d = lane_value(r[k], k);
        if (!(d > 0.0)) {
            if (lane == 0) status[mtx] = TDI_NOT_PD;
            return;
        }
        const double dk = sqrt(d), inv = 1.0 / dk;
        r[k] = lane == k ? dk : (lane > k ? r[k] * inv : r[k]);
#pragma unroll
        for (int j = k + 1; j < CB; ++j) {
            const double ljk = lane_value(r[k], j);
            if (lane >= j) r[j] = fma(-r[k], ljk, r[j]);
        }
    }
    if (row) {
        double* Mw = M + (size_t)(k0 + lane) * K + k0;
#pragma unroll
        for (int j = 0; j < CB; ++j)
            if (j < nb && j <= lane) Mw[j] = r[j];
    }
}

// The panel below the diagonal block: row r of A21 solves x L11^T = a (forward substitution over
// the panel's columns), 64 rows per workgroup (a grid over the rows), each row in its thread's
// registers, L11 and the reciprocals of its diagonal in LDS (read as broadcasts).
__global__ void __launch_bounds__(64) k_chol_offdiag(int K, int k0, double* __restrict__ A,
                                                      const int* __restrict__ status)
{
    const int mtx = blockIdx.y;
    if (status[mtx] != TDI_OK) return;
    double* M = A + (size_t)mtx * K * K;
    const int nb = min(CB, K - k0), tid = threadIdx.x;
    const int r0 = k0 + nb + blockIdx.x * 64;
    if (r0 >= K) return;
    __shared__ double D[CB][CB];
    __shared__ double sinv[CB];
    {
        // every load of the block issued before the LDS stores (64 threads, 16 entries each)
        double v[CB * CB / 64];
#pragma unroll
        for (int q = 0; q < CB * CB / 64; ++q) {
            const int e = tid + 64 * q, i = e / CB, j = e - i * CB;
            const bool in = i < nb && j < nb && j <= i;
            const double x = M[(size_t)(k0 + min(i, nb - 1)) * K + k0 + min(j, nb - 1)];
            v[q] = in ? x : (i == j ? 1.0 : 0.0);
        }
#pragma unroll
        for (int q = 0; q < CB * CB / 64; ++q) {
            const int e = tid + 64 * q;
            D[e / CB][e % CB] = v[q];
        }
    }
    __syncthreads();
    if (tid < CB) sinv[tid] = 1.0 / D[tid][tid];
    __syncthreads();
    const int gr = r0 + tid;
    if (gr >= K) return;
    double* Mr = M + (size_t)gr * K + k0;
    double p[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) p[j] = j < nb ? Mr[min(j, nb - 1)] : 0.0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        double sacc = p[j];
#pragma unroll
        for (int q = 0; q < j; ++q) sacc = fma(-p[q], D[j][q], sacc);
        p[j] = sacc * sinv[j];
    }
#pragma unroll
    for (int j = 0; j < CB; ++j)
        if (j < nb) Mr[j] = p[j];
}

// Tile (ti, tj), ti >= tj, of the trailing matrix below/right of the panel at k0.
__global__ void __launch_bounds__(256) k_chol_update(int K, int k0, double* __restrict__ A,
                                                      const int* __restrict__ status)
{
    const int mtx = blockIdx.y;
    if (status[mtx] != TDI_OK) return;
    const int r0 = k0 + CB;                        // first trailing row
    const int T = (K - r0 + CB - 1) / CB;
    int t = blockIdx.x, ti = 0;
    while (ti < T && t > ti) { t -= ti + 1; ++ti; }   // tile t of row ti: tj = t, 0 <= tj <= ti
    if (ti >= T) return;
    const int tj = t;
    double* M = A + (size_t)mtx * K * K;
    __shared__ double Li[CB][CB + 1], Lj[CB][CB + 1];
    const int tid = threadIdx.x;
    const int nb = min(CB, K - k0);
    {
        // every load of both slices issued before the LDS stores (256 threads, 4 entries each)
        double vi[CB * CB / 256], vj[CB * CB / 256];
#pragma unroll
        for (int q = 0; q < CB * CB / 256; ++q) {
            const int e = tid + 256 * q, r = e / CB, c = e - r * CB;
            const int ri = r0 + ti * CB + r, rj = r0 + tj * CB + r;
            const double xi = M[(size_t)min(ri, K - 1) * K + k0 + min(c, nb - 1)];
            const double xj = M[(size_t)min(rj, K - 1) * K + k0 + min(c, nb - 1)];
            vi[q] = (ri < K && c < nb) ? xi : 0.0;
            vj[q] = (rj < K && c < nb) ? xj : 0.0;
        }
#pragma unroll
        for (int q = 0; q < CB * CB / 256; ++q) {
            const int e = tid + 256 * q, r = e / CB, c = e - r * CB;
            Li[r][c] = vi[q];
            Lj[r][c] = vj[q];
        }
    }
    __syncthreads();
    // 256 threads x 4 entries: thread (a, b4) computes row a, columns b4..b4+3 of the tile
    const int a = tid >> 3, b4 = (tid & 7) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int c = 0; c < CB; ++c) {
        const double li = Li[a][c];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = fma(li, Lj[b4 + q][c], acc[q]);
    }
    const int gi = r0 + ti * CB + a;
    if (gi >= K) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int gj = r0 + tj * CB + b4 + q;
        if (gj < K && gj <= gi) M[(size_t)gi * K + gj] -= acc[q];
    }
}

// L y = -g, then L^T delta = y, one workgroup per matrix, by 32-row panels with the solution in
// LDS (K <= CHOL_BIG_MAXK), right-looking: each panel's 32 x 32 diagonal block staged in LDS by all
// threads, its triangle solved on wavefront 0 (lane i holding row i, forward, or column i,
// backward, in registers, the solved entries passed by lane reads), then every entry not yet
// solved updated with the panel's solution by all threads: forward, eight lanes per later row,
// each reading four consecutive entries of the panel's columns (a row's 32 entries in one
// coalesced read) and the eight partial sums combined by lane shuffles; backward, one earlier
// column per thread, the panel's rows read coalesced across the threads.
constexpr int TRSV_THREADS = 1024;

template <typename TG, typename TD>
__global__ void __launch_bounds__(TRSV_THREADS) k_chol_trsv(int K, const double* __restrict__ A,
                                                             const TG* __restrict__ g, TD* __restrict__ delta,
                                                             const int* __restrict__ status)
{
    const int mtx = blockIdx.x;
    if (status[mtx] != TDI_OK) return;
    const double* L = A + (size_t)mtx * K * K;
    __shared__ double sy[CHOL_BIG_MAXK];
    __shared__ double sD[CB][CB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < K; i += blockDim.x) sy[i] = -(double)g[(size_t)mtx * K + i];
    // the diagonal block of the panel at p0 (zero past the matrix), one entry per thread
    auto stage = [&](int p0, int nb) {
        const int i = tid >> 5, j = tid & 31;
        const double x = L[(size_t)(p0 + min(i, nb - 1)) * K + p0 + min(j, nb - 1)];
        sD[i][j] = (i < nb && j < nb) ? x : 0.0;
    };
    // forward: y_p = L_pp^-1 b_p, then b_i -= L_ip y_p for the rows below
    for (int p0 = 0; p0 < K; p0 += CB) {
        const int nb = min(CB, K - p0);
        stage(p0, nb);
        __syncthreads();
        if (wave == 0) {
            const bool in = lane < nb;
            const int r0 = in ? lane : 0;
            double lr[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) lr[j] = (in && j < lane) ? sD[r0][j] : 0.0;
            const double rinv = in ? 1.0 / sD[r0][r0] : 1.0;
            double r = in ? sy[p0 + r0] : 0.0;
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                if (lane == j) r *= rinv;
                const double yj = lane_value(r, j);
                if (lane > j) r = fma(-lr[j], yj, r);
            }
            if (in) sy[p0 + lane] = r;
        }
        __syncthreads();
        {
            const int sub = tid & 7, c0 = 4 * sub;
            double yv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) yv[q] = c0 + q < nb ? sy[p0 + c0 + q] : 0.0;
            for (int i0 = p0 + nb; i0 < K; i0 += TRSV_THREADS / 8) {
                const int i = i0 + (tid >> 3);
                const int ic = min(i, K - 1);
                const double* Lr = L + (size_t)ic * K + p0;
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = fma(Lr[min(c0 + q, nb - 1)], yv[q], acc);
                acc += __shfl_xor(acc, 1, 64);
                acc += __shfl_xor(acc, 2, 64);
                acc += __shfl_xor(acc, 4, 64);
                if (sub == 0 && i < K) sy[i] -= acc;
            }
        }
        __syncthreads();
    }
    // backward: x_p = L_pp^-T y_p, then y_i -= L_pi x_p for the entries before the panel
    const int last = ((K - 1) / CB) * CB;
    for (int p0 = last; p0 >= 0; p0 -= CB) {
        const int nb = min(CB, K - p0);
        stage(p0, nb);
        __syncthreads();
        if (wave == 0) {
            const bool in = lane < nb;
            const int c = in ? lane : 0;
            double lc[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) lc[j] = (in && j > lane) ? sD[j][c] : 0.0;
            const double rinv = in ? 1.0 / sD[c][c] : 1.0;
            double r = in ? sy[p0 + c] : 0.0;
#pragma unroll
            for (int j = CB - 1; j >= 0; --j) {
                if (lane == j) r *= rinv;
                const double xj = lane_value(r, j);
                if (lane < j) r = fma(-lc[j], xj, r);
            }
            if (in) sy[p0 + lane] = r;
        }
        __syncthreads();
        for (int i = tid; i < p0; i += blockDim.x) {
            const double* Lc = L + (size_t)p0 * K + i;
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const int jj = min(j, nb - 1);
                const double l = Lc[(size_t)jj * K];
                acc = fma(j < nb ? l : 0.0, sy[p0 + jj], acc);
            }
            sy[i] -= acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < K; i += blockDim.x) delta[(size_t)mtx * K + i] = (TD)sy[i];
}

} // namespace

// Factorisation and solve on the FP64 matrices A (overwritten by the factor); g FP32 or FP64.
template <typename TG, typename TD>
static int chol_big_run(int n, int K, double* A, const TG* g, TD* delta, int* status, hipStream_t s)
{
    for (int k0 = 0; k0 < K; k0 += CB) {
        hipLaunchKernelGGL(k_chol_diag, dim3(n), dim3(64), 0, s, K, k0, A, status);
        const int below = K - k0 - CB;
        if (below > 0)
            hipLaunchKernelGGL(k_chol_offdiag, dim3((below + 63) / 64, n), dim3(64), 0, s, K, k0, A, status);
        const int T = (K - k0 - CB + CB - 1) / CB;
        if (T > 0) hipLaunchKernelGGL(k_chol_update, dim3(T * (T + 1) / 2, n), dim3(256), 0, s, K, k0, A, status);
        if (hipGetLastError() != hipSuccess) return TDI_ERR_HIP;
    }
    hipLaunchKernelGGL((k_chol_trsv<TG, TD>), dim3(n), dim3(TRSV_THREADS), 0, s, K, A, g, delta, status);
    return hipGetLastError() != hipSuccess ? TDI_ERR_HIP : TDI_OK;
}

int launch_cholesky_big(int n, int K, const float* H, const float* g, float* delta, int* status, hipStream_t s)
{
    if (K > CHOL_BIG_MAXK) {
        set_error("cholesky: K=%d > %d", K, CHOL_BIG_MAXK);
        return TDI_ERR_ARG;
    }
    double* A = nullptr;
    const size_t nkk = (size_t)n * K * K;
    if (hipMalloc((void**)&A, sizeof(double) * nkk) != hipSuccess) {
        set_error("cholesky: %zu bytes of FP64 scratch not available", sizeof(double) * nkk);
        return TDI_ERR_NOMEM;
    }
    int rc = TDI_OK;
    std::vector<int> ok(n, TDI_OK);
    if (hipMemcpyAsync(status, ok.data(), sizeof(int) * n, hipMemcpyHostToDevice, s) != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)std::min<size_t>((nkk + 255) / 256, 65536)), dim3(256), 0, s, nkk,
                           H, A);
        rc = chol_big_run(n, K, A, g, delta, status, s);
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = TDI_ERR_HIP;
    if (rc) set_error("cholesky (large K): %s", hipGetErrorString(e));
    (void)hipFree(A);
    return rc;
}

template <typename TD>
static int cholesky_big_f64(int n, int K, double* A, const double* g, TD* delta, int* status, hipStream_t s)
{
    if (K > CHOL_BIG_MAXK) {
        set_error("cholesky: K=%d > %d", K, CHOL_BIG_MAXK);
        return TDI_ERR_ARG;
    }
    std::vector<int> ok(n, TDI_OK);
    TDI_HIP(hipMemcpyAsync(status, ok.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    const int rc = chol_big_run(n, K, A, g, delta, status, s);
    // the status upload reads the host vector: it must finish before the vector goes
    TDI_HIP(hipStreamSynchronize(s));
    if (rc) set_error("cholesky (large K): launch failed");
    return rc;
}

int launch_cholesky_big_f64(int n, int K, double* A, const double* g, float* delta, int* status, hipStream_t s)
{
    return cholesky_big_f64(n, K, A, g, delta, status, s);
}

int launch_cholesky_big_f64(int n, int K, double* A, const double* g, double* delta, int* status, hipStream_t s)
{
    return cholesky_big_f64(n, K, A, g, delta, status, s);
}

} // namespace tdi
