// geometry.hip — vertex synthesis, view transform and projection on gfx950.
//
// Reference: Renderer::compute_nonrigid_shape2 (renderer.cu:806-861) does three cuBLAS
// SGEMVs per basis (X0 = X0_mean + IX alpha + EX eps, ...) and then
// view_transform_3d_pts_and_render_2d (funcs.cu:395-453): P = R P0 + tau,
// xp = phix X/Z + cx, yp = phiy Y/Z + cy.
//
// Both forms accumulate the basis product, rotate, translate and project in FP64 and round
// the results to FP32 once: the projected vertices then equal the correctly rounded values
// (what an FP64 computation rounded to FP32 gives), so the FP32 coverage test downstream
// sees the same inputs as an FP64 transcription would.  MI355X runs FP64 FMA and FP64 MFMA
// at the same 78.6 TFLOP/s, so this costs little next to the rest of an iteration.
//
//   k_synth_gemv   — one frame per grid row (the plain path): thread per vertex.
//   k_synth_mfma64 — many frames: [3V x K] x [K x B] on v_mfma_f64_16x16x4f64; a block
//                    covers 16 vertices (48 basis rows) x 64 frames, each basis value is
//                    read once per 64 frames; the epilogue fuses mean, rotation,
//                    translation and projection.
#include "common.h"

namespace tdi {

namespace {

constexpr int MAXK_LDS = 1024;

__device__ __forceinline__ void store_vertex(const FrameGPU& fr, const double* R, const double* tau, const double* p,
                                             size_t o, float* P0, float* P, float4* vproj)
{
    P0[3 * o + 0] = (float)p[0];
    P0[3 * o + 1] = (float)p[1];
    P0[3 * o + 2] = (float)p[2];
    const double X = R[0] * p[0] + R[3] * p[1] + R[6] * p[2] + tau[0];
    const double Y = R[1] * p[0] + R[4] * p[1] + R[7] * p[2] + tau[1];
    const double Z = R[2] * p[0] + R[5] * p[1] + R[8] * p[2] + tau[2];
    if (P) {                               // camera-space vertices only when asked for
        P[3 * o + 0] = (float)X;
        P[3 * o + 1] = (float)Y;
        P[3 * o + 2] = (float)Z;
    }
    vproj[o] = make_float4((float)((double)fr.phix * X / Z + (double)fr.cx), (float)((double)fr.phiy * Y / Z + (double)fr.cy),
                           (float)Z, 0.f);
}

// FP64 rotation and translation of every listed frame, once per launch (the synthesis
// workgroups read them instead of evaluating Rodrigues' formula per workgroup).
__global__ void k_frame_rt(Layout L, const float* __restrict__ theta, const int* __restrict__ act, int n,
                           double* __restrict__ rt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fl_count(act, n)) return;
    const int f = fl_frame(act, i);
    const float* tf = theta + (size_t)f * L.K;
    double R[9];
    rodrigues_f64(tf + L.off_pose + 3, R);
    double* o = rt + (size_t)f * 12;
    for (int k = 0; k < 9; ++k) o[k] = R[k];
    for (int a = 0; a < 3; ++a) o[9 + a] = tf[L.off_pose + a];
}

__global__ void __launch_bounds__(256) k_synth_gemv(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                     const float* __restrict__ theta, const int* __restrict__ act,
                                                     int n, const double* __restrict__ rt, float* __restrict__ P0,
                                                     float* __restrict__ P, float4* __restrict__ vproj)
{
    __shared__ double R[12];
    __shared__ float th[MAXK_LDS];
    const double* tau = R + 9;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const float* tf = theta + (size_t)f * L.K;
        for (int i = threadIdx.x; i < L.K; i += blockDim.x) th[i] = tf[i];
        if (threadIdx.x < 12) R[threadIdx.x] = rt[(size_t)f * 12 + threadIdx.x];
        __syncthreads();
        const int v = blockIdx.x * blockDim.x + threadIdx.x;
        if (v < m.V) {
            double p[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int r = 3 * v + a;
                double s = m.mean[r];
                const float* e = m.exp_basis + (size_t)r * m.k_exp;
                for (int k = 0; k < L.nexp; ++k) s = fma((double)e[k], (double)th[L.off_exp + k], s);
                if (L.nid) {
                    const float* ib = m.id_basis + (size_t)r * m.k_id;
                    for (int k = 0; k < L.nid; ++k) s = fma((double)ib[k], (double)th[L.off_id + k], s);
                }
                p[a] = s;
            }
            store_vertex(frames[f], R, tau, p, (size_t)f * m.V + v, P0, P, vproj);
        }
        __syncthreads();
    }
}

// FP64 MFMA operand layout, v_mfma_f64_16x16x4f64 (one double per lane for A and B):
//   A[i][k]: lane = i + 16*k          B[k][j]: lane = j + 16*k
// Accumulator: lane l, element r (0..3) holds C[l/16 + 4*r][l % 16] (measured on gfx950;
// the FP32/BF16 16x16 forms use row 4*(l/16) + r instead).
constexpr int S64_ROWS = 48;    // 16 vertices
constexpr int S64_FR = 64;      // 4 waves x 16 frames

typedef double v4d __attribute__((ext_vector_type(4)));

// Columns are positions in the frame list: workgroup column group y covers list entries
// 64y .. 64y+63, so a launch over a short list only runs its first column group(s).
__global__ void __launch_bounds__(256) k_synth_mfma64(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                       const float* __restrict__ theta, const int* __restrict__ act,
                                                       int n, const double* __restrict__ rt, float* __restrict__ P0,
                                                       float* __restrict__ P, float4* __restrict__ vproj)
{
    const int row0 = blockIdx.x * S64_ROWS;
    const int c0 = blockIdx.y * S64_FR;
    const int nact = fl_count(act, n);
    if (c0 >= nact) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ double sacc[S64_ROWS][S64_FR + 1];
    __shared__ double sR[S64_FR][12];
    __shared__ int sfr[S64_FR];
    if (tid < S64_FR) {
        const int c = c0 + tid;
        const int f = c < nact ? fl_frame(act, c) : -1;
        sfr[tid] = f;
        if (f >= 0)
            for (int k = 0; k < 12; ++k) sR[tid][k] = rt[(size_t)f * 12 + k];
    }
    __syncthreads();

    const int Ke = L.nexp;
    const int rowsTotal = 3 * m.V;
    {
        v4d acc[3] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        const int i = lane & 15, kk = lane >> 4;
        const int fj = sfr[wave * 16 + i];
        const bool frok = fj >= 0;
        const float* bcol = theta + (size_t)(frok ? fj : 0) * L.K + L.off_exp;
        const float* arow[3];
        bool rowok[3];
#pragma unroll
        for (int rt = 0; rt < 3; ++rt) {
            const int row = row0 + 16 * rt + i;
            rowok[rt] = row < rowsTotal;
            arow[rt] = m.exp_basis + (size_t)(rowok[rt] ? row : 0) * m.k_exp;
        }
        for (int k = 0; k < Ke; k += 4) {
            const int kx = k + kk;
            const double b = (frok && kx < Ke) ? (double)bcol[kx] : 0.0;
#pragma unroll
            for (int rt = 0; rt < 3; ++rt) {
                const double a = (rowok[rt] && kx < Ke) ? (double)arow[rt][kx] : 0.0;
                acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < 3; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) sacc[16 * rt + (lane >> 4) + 4 * r][wave * 16 + (lane & 15)] = acc[rt][r];
    }
    __syncthreads();
    // epilogue: 16 vertices x 64 frames over 256 threads
    for (int e = tid; e < 16 * S64_FR; e += blockDim.x) {
        const int vl = e & 15, j = e >> 4;
        const int v = row0 / 3 + vl;
        const int f = sfr[j];
        if (v >= m.V || f < 0) continue;
        double p[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = (double)m.mean[3 * v + a] + sacc[3 * vl + a][j];
        store_vertex(frames[f], sR[j], sR[j] + 9, p, (size_t)f * m.V + v, P0, P, vproj);
    }
}

} // namespace

static int launch_frame_rt(const Layout& L, FrameList fl, const float* theta, double* rt, hipStream_t s)
{
    hipLaunchKernelGGL(k_frame_rt, dim3((fl.n + 255) / 256 > 0 ? (fl.n + 255) / 256 : 1), dim3(256), 0, s, L, theta,
                       fl.act, fl.n, rt);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_synthesize(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                      double* rt, float* P0, float* P, float4* vproj, hipStream_t s)
{
    if (L.K > MAXK_LDS) {
        set_error("synthesize: K=%d exceeds %d", L.K, MAXK_LDS);
        return TDI_ERR_ARG;
    }
    int rc = launch_frame_rt(L, fl, theta, rt, s);
    if (rc) return rc;
    dim3 grid((m.V + 255) / 256, grid_frames(fl.n, 64));
    hipLaunchKernelGGL(k_synth_gemv, grid, dim3(256), 0, s, m, L, frames, theta, fl.act, fl.n, rt, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_synthesize_gemm(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                           const float* theta, double* rt, float* P0, float* P, float4* vproj, hipStream_t s)
{
    if (L.nid) return launch_synthesize(m, L, frames, fl, theta, rt, P0, P, vproj, s);
    int rc = launch_frame_rt(L, fl, theta, rt, s);
    if (rc) return rc;
    dim3 grid((3 * m.V + S64_ROWS - 1) / S64_ROWS, (fl.n + S64_FR - 1) / S64_FR);
    hipLaunchKernelGGL(k_synth_mfma64, grid, dim3(256), 0, s, m, L, frames, theta, fl.act, fl.n, rt, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
