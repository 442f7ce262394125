// cholesky_big.hip — FP64 Cholesky solve for normal matrices too large for one workgroup's LDS.
//
// Solver::solve (solver.h:70-105: cusolverDnDpotrf + Dpotrs on the FP64 copy of the FP32 normal
// matrix, copyMatsFloatToDouble / copyVectorDoubleToFloatNegated, solver.cu:11-81) for the
// joint identity-stage system (K up to a few thousand: 199 + 199 + 94 per frame).  Blocked
// right-looking factorisation in place on an FP64 copy in HBM, 32-column panels:
//   k_chol_panel   one workgroup per matrix: factor the 32x32 diagonal block in LDS, then the
//                  panel below it (L21 = A21 L11^-T, 128 rows at a time staged in LDS, one row
//                  per thread);
//   k_chol_update  a grid of 32x32 tiles of the trailing lower triangle per matrix:
//                  A22 -= L21 L21^T, both 32-row slices of the panel staged in LDS;
//   k_chol_trsv    one workgroup per matrix: L y = -g, then L^T delta = y (column-oriented,
//                  one barrier pair per row).
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

__global__ void __launch_bounds__(256) k_chol_panel(int K, int k0, double* __restrict__ A, int* __restrict__ status)
{
    const int mtx = blockIdx.x;
    if (status[mtx] != TDI_OK) return;
    double* M = A + (size_t)mtx * K * K;
    __shared__ double D[CB][CB + 1];
    __shared__ int bad;
    const int nb = min(CB, K - k0), tid = threadIdx.x;
    if (tid == 0) bad = 0;
    for (int e = tid; e < nb * nb; e += blockDim.x) {
        const int i = e / nb, j = e - i * nb;
        D[i][j] = M[(size_t)(k0 + i) * K + k0 + j];
    }
    __syncthreads();
    // unblocked Cholesky of the diagonal block (lower), column by column
    for (int k = 0; k < nb; ++k) {
        if (tid == 0) {
            const double d = D[k][k];
            if (!(d > 0.0)) bad = 1;
            else D[k][k] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        const double dk = D[k][k];
        for (int i = k + 1 + tid; i < nb; i += blockDim.x) D[i][k] /= dk;
        __syncthreads();
        const int w = nb - k - 1;
        for (int e = tid; e < w * w; e += blockDim.x) {
            const int i = k + 1 + e / w, j = k + 1 + e % w;
            if (j <= i) D[i][j] -= D[i][k] * D[j][k];
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) status[mtx] = TDI_NOT_PD;
        return;
    }
    for (int e = tid; e < nb * nb; e += blockDim.x) {
        const int i = e / nb, j = e - i * nb;
        if (j <= i) M[(size_t)(k0 + i) * K + k0 + j] = D[i][j];
    }
    // panel: row r of A21 solves x L11^T = a (forward substitution over the panel's columns),
    // 128 rows at a time staged in LDS (coalesced loads and stores, LDS operands)
    __shared__ double Pn[128][CB + 1];
    for (int r0 = k0 + nb; r0 < K; r0 += 128) {
        const int nr = min(128, K - r0);
        for (int e = tid; e < nr * nb; e += blockDim.x) {
            const int i = e / nb, j = e - i * nb;
            Pn[i][j] = M[(size_t)(r0 + i) * K + k0 + j];
        }
        __syncthreads();
        if (tid < nr) {
            for (int j = 0; j < nb; ++j) {
                double s = Pn[tid][j];
                for (int q = 0; q < j; ++q) s -= Pn[tid][q] * D[j][q];
                Pn[tid][j] = s / D[j][j];
            }
        }
        __syncthreads();
        for (int e = tid; e < nr * nb; e += blockDim.x) {
            const int i = e / nb, j = e - i * nb;
            M[(size_t)(r0 + i) * K + k0 + j] = Pn[i][j];
        }
        __syncthreads();
    }
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
    for (int e = tid; e < CB * CB; e += blockDim.x) {
        const int r = e / CB, c = e - r * CB;
        const int ri = r0 + ti * CB + r, rj = r0 + tj * CB + r;
        Li[r][c] = (ri < K && c < nb) ? M[(size_t)ri * K + k0 + c] : 0.0;
        Lj[r][c] = (rj < K && c < nb) ? M[(size_t)rj * K + k0 + c] : 0.0;
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

template <typename TG>
__global__ void __launch_bounds__(256) k_chol_trsv(int K, const double* __restrict__ A, const TG* __restrict__ g,
                                                    double* __restrict__ y, float* __restrict__ delta,
                                                    const int* __restrict__ status)
{
    const int mtx = blockIdx.x;
    if (status[mtx] != TDI_OK) return;
    const double* L = A + (size_t)mtx * K * K;
    double* b = y + (size_t)mtx * K;
    const int tid = threadIdx.x;
    for (int i = tid; i < K; i += blockDim.x) b[i] = -(double)g[(size_t)mtx * K + i];
    __syncthreads();
    // L y = -g: column k finishes y_k, then updates the rows below
    for (int k = 0; k < K; ++k) {
        const double yk = b[k] / L[(size_t)k * K + k];
        __syncthreads();
        if (tid == 0) b[k] = yk;
        for (int i = k + 1 + tid; i < K; i += blockDim.x) b[i] -= L[(size_t)i * K + k] * yk;
        __syncthreads();
    }
    // L^T delta = y: row k of L is column k of L^T
    for (int k = K - 1; k >= 0; --k) {
        const double dk = b[k] / L[(size_t)k * K + k];
        __syncthreads();
        if (tid == 0) b[k] = dk;
        for (int i = tid; i < k; i += blockDim.x) b[i] -= L[(size_t)k * K + i] * dk;
        __syncthreads();
    }
    for (int i = tid; i < K; i += blockDim.x) delta[(size_t)mtx * K + i] = (float)b[i];
}

} // namespace

// Factorisation and solve on the FP64 matrices A (overwritten by the factor); g FP32 or FP64.
template <typename TG>
static int chol_big_run(int n, int K, double* A, const TG* g, double* y, float* delta, int* status, hipStream_t s)
{
    for (int k0 = 0; k0 < K; k0 += CB) {
        hipLaunchKernelGGL(k_chol_panel, dim3(n), dim3(256), 0, s, K, k0, A, status);
        const int T = (K - k0 - CB + CB - 1) / CB;
        if (T > 0) hipLaunchKernelGGL(k_chol_update, dim3(T * (T + 1) / 2, n), dim3(256), 0, s, K, k0, A, status);
        if (hipGetLastError() != hipSuccess) return TDI_ERR_HIP;
    }
    hipLaunchKernelGGL(k_chol_trsv<TG>, dim3(n), dim3(256), 0, s, K, A, g, y, delta, status);
    return hipGetLastError() != hipSuccess ? TDI_ERR_HIP : TDI_OK;
}

int launch_cholesky_big(int n, int K, const float* H, const float* g, float* delta, int* status, hipStream_t s)
{
    if (K > CHOL_BIG_MAXK) {
        set_error("cholesky: K=%d > %d", K, CHOL_BIG_MAXK);
        return TDI_ERR_ARG;
    }
    double* A = nullptr;
    double* y = nullptr;
    const size_t nkk = (size_t)n * K * K;
    if (hipMalloc((void**)&A, sizeof(double) * nkk) != hipSuccess ||
        hipMalloc((void**)&y, sizeof(double) * (size_t)n * K) != hipSuccess) {
        (void)hipFree(A);
        set_error("cholesky: %zu bytes of FP64 scratch not available", sizeof(double) * nkk);
        return TDI_ERR_NOMEM;
    }
    int rc = TDI_OK;
    std::vector<int> ok(n, TDI_OK);
    if (hipMemcpyAsync(status, ok.data(), sizeof(int) * n, hipMemcpyHostToDevice, s) != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)std::min<size_t>((nkk + 255) / 256, 65536)), dim3(256), 0, s, nkk,
                           H, A);
        rc = chol_big_run(n, K, A, g, y, delta, status, s);
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = TDI_ERR_HIP;
    if (rc) set_error("cholesky (large K): %s", hipGetErrorString(e));
    (void)hipFree(A);
    (void)hipFree(y);
    return rc;
}

int launch_cholesky_big_f64(int n, int K, double* A, const double* g, double* y, float* delta, int* status,
                            hipStream_t s)
{
    if (K > CHOL_BIG_MAXK) {
        set_error("cholesky: K=%d > %d", K, CHOL_BIG_MAXK);
        return TDI_ERR_ARG;
    }
    std::vector<int> ok(n, TDI_OK);
    TDI_HIP(hipMemcpyAsync(status, ok.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    const int rc = chol_big_run(n, K, A, g, y, delta, status, s);
    // the status upload reads the host vector: it must finish before the vector goes
    TDI_HIP(hipStreamSynchronize(s));
    if (rc) set_error("cholesky (large K): launch failed");
    return rc;
}

} // namespace tdi
