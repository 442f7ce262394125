// cholesky_big.hip — FP64 Cholesky solve for normal matrices too large for one workgroup's LDS.
//
// Solver::solve (solver.h:70-105: cusolverDnDpotrf + Dpotrs on the FP64 copy of the FP32 normal
// matrix, copyMatsFloatToDouble / copyVectorDoubleToFloatNegated, solver.cu:11-81) for the
// joint identity-stage system (K up to a few thousand: 199 + 199 + 94 per frame).  Blocked
// right-looking factorisation in place on an FP64 copy in HBM, 32-column panels:
//   k_chol_diag    one workgroup per matrix: factor the 32x32 diagonal block in LDS;
//   k_chol_offdiag a grid of 128-row blocks per matrix: the panel below it, L21 = A21 L11^-T
//                  (rows and L11 staged in LDS, one row per thread);
//   k_chol_update  a grid of 32x32 tiles of the trailing lower triangle per matrix:
//                  A22 -= L21 L21^T, both 32-row slices of the panel staged in LDS;
//   k_chol_trsv    one workgroup per matrix: L y = -g, then L^T delta = y, by 32-row panels
//                  (the solution in LDS, the panel triangles on one wavefront).
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

// The 32 x 32 diagonal block of the panel at k0, factored in LDS and written back.
__global__ void __launch_bounds__(256) k_chol_diag(int K, int k0, double* __restrict__ A, int* __restrict__ status)
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
}

// The panel below the diagonal block: row r of A21 solves x L11^T = a (forward substitution over
// the panel's columns), 128 rows per workgroup (a grid over the rows), rows and L11 staged in LDS.
__global__ void __launch_bounds__(128) k_chol_offdiag(int K, int k0, double* __restrict__ A,
                                                       const int* __restrict__ status)
{
    const int mtx = blockIdx.y;
    if (status[mtx] != TDI_OK) return;
    double* M = A + (size_t)mtx * K * K;
    const int nb = min(CB, K - k0), tid = threadIdx.x;
    const int r0 = k0 + nb + blockIdx.x * 128;
    if (r0 >= K) return;
    const int nr = min(128, K - r0);
    __shared__ double D[CB][CB + 1];
    __shared__ double Pn[128][CB + 1];
    for (int e = tid; e < nb * nb; e += blockDim.x) {
        const int i = e / nb, j = e - i * nb;
        D[i][j] = j <= i ? M[(size_t)(k0 + i) * K + k0 + j] : 0.0;
    }
    for (int e = tid; e < nr * nb; e += blockDim.x) {
        const int i = e / nb, j = e - i * nb;
        Pn[i][j] = M[(size_t)(r0 + i) * K + k0 + j];
    }
    __syncthreads();
    if (tid < nr) {
        for (int j = 0; j < nb; ++j) {
            double sacc = Pn[tid][j];
            for (int q = 0; q < j; ++q) sacc -= Pn[tid][q] * D[j][q];
            Pn[tid][j] = sacc / D[j][j];
        }
    }
    __syncthreads();
    for (int e = tid; e < nr * nb; e += blockDim.x) {
        const int i = e / nb, j = e - i * nb;
        M[(size_t)(r0 + i) * K + k0 + j] = Pn[i][j];
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

// L y = -g, then L^T delta = y, one workgroup per matrix, blocked by 32-row panels with the
// solution in LDS (K <= CHOL_BIG_MAXK): per panel, the product of the rows (forward) or columns
// (backward) already solved with the solution so far, by all threads with coalesced reads of L,
// then the panel's triangle by one wavefront from LDS, the running values in lanes and each
// solved entry broadcast by a shuffle.
template <typename TG>
__global__ void __launch_bounds__(256) k_chol_trsv(int K, const double* __restrict__ A, const TG* __restrict__ g,
                                                    float* __restrict__ delta, const int* __restrict__ status)
{
    const int mtx = blockIdx.x;
    if (status[mtx] != TDI_OK) return;
    const double* L = A + (size_t)mtx * K * K;
    __shared__ double sy[CHOL_BIG_MAXK];
    __shared__ double sL[CB][CB + 1];
    __shared__ double sred[8][CB];
    __shared__ double sr[CB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < K; i += blockDim.x) sy[i] = -(double)g[(size_t)mtx * K + i];
    __syncthreads();
    // forward: y_p = L_pp^-1 (b_p - L_p,<p y_<p)
    for (int p0 = 0; p0 < K; p0 += CB) {
        const int nb = min(CB, K - p0);
        {
            const int row = tid >> 3, sub = tid & 7;
            double acc = 0.0;
            if (row < nb)
                for (int c = sub; c < p0; c += 8) acc += L[(size_t)(p0 + row) * K + c] * sy[c];
            for (int o = 4; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (sub == 0 && row < nb) sr[row] = sy[p0 + row] - acc;
        }
        for (int e = tid; e < nb * nb; e += blockDim.x) {
            const int i = e / nb, j = e - i * nb;
            sL[i][j] = L[(size_t)(p0 + i) * K + p0 + j];
        }
        __syncthreads();
        if (wave == 0) {
            double r = lane < nb ? sr[lane] : 0.0;
            for (int j = 0; j < nb; ++j) {
                const double yj = __shfl(r, j, 64) / sL[j][j];
                if (lane == j) r = yj;
                else if (lane > j && lane < nb) r -= sL[lane][j] * yj;
            }
            if (lane < nb) sy[p0 + lane] = r;
        }
        __syncthreads();
    }
    // backward: x_p = L_pp^-T (y_p - L_>p,p^T x_>p)
    const int last = ((K - 1) / CB) * CB;
    for (int p0 = last; p0 >= 0; p0 -= CB) {
        const int nb = min(CB, K - p0);
        {
            const int c = tid & 31, rg = tid >> 5;
            double acc = 0.0;
            if (c < nb)
                for (int r = p0 + nb + rg; r < K; r += 8) acc += L[(size_t)r * K + p0 + c] * sy[r];
            sred[rg][c] = acc;
        }
        for (int e = tid; e < nb * nb; e += blockDim.x) {
            const int i = e / nb, j = e - i * nb;
            sL[i][j] = L[(size_t)(p0 + i) * K + p0 + j];
        }
        __syncthreads();
        if (wave == 0) {
            double r = 0.0;
            if (lane < nb) {
                double sacc = 0.0;
                for (int q = 0; q < 8; ++q) sacc += sred[q][lane];
                r = sy[p0 + lane] - sacc;
            }
            for (int j = nb - 1; j >= 0; --j) {
                const double xj = __shfl(r, j, 64) / sL[j][j];
                if (lane == j) r = xj;
                else if (lane < j) r -= sL[j][lane] * xj;
            }
            if (lane < nb) sy[p0 + lane] = r;
        }
        __syncthreads();
    }
    for (int i = tid; i < K; i += blockDim.x) delta[(size_t)mtx * K + i] = (float)sy[i];
}

} // namespace

// Factorisation and solve on the FP64 matrices A (overwritten by the factor); g FP32 or FP64.
template <typename TG>
static int chol_big_run(int n, int K, double* A, const TG* g, float* delta, int* status, hipStream_t s)
{
    for (int k0 = 0; k0 < K; k0 += CB) {
        hipLaunchKernelGGL(k_chol_diag, dim3(n), dim3(256), 0, s, K, k0, A, status);
        const int below = K - k0 - CB;
        if (below > 0)
            hipLaunchKernelGGL(k_chol_offdiag, dim3((below + 127) / 128, n), dim3(128), 0, s, K, k0, A, status);
        const int T = (K - k0 - CB + CB - 1) / CB;
        if (T > 0) hipLaunchKernelGGL(k_chol_update, dim3(T * (T + 1) / 2, n), dim3(256), 0, s, K, k0, A, status);
        if (hipGetLastError() != hipSuccess) return TDI_ERR_HIP;
    }
    hipLaunchKernelGGL(k_chol_trsv<TG>, dim3(n), dim3(256), 0, s, K, A, g, delta, status);
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

int launch_cholesky_big_f64(int n, int K, double* A, const double* g, float* delta, int* status, hipStream_t s)
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

} // namespace tdi
