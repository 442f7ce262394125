// geometry.hip — vertex synthesis, view transform and projection on gfx950.
//
// Reference: Renderer::compute_nonrigid_shape2 (renderer.cu:806-861) does three cuBLAS
// SGEMVs per basis (X0 = X0_mean + IX alpha + EX eps, ...) and then
// view_transform_3d_pts_and_render_2d (funcs.cu:395-453): P = R P0 + tau,
// xp = phix X/Z + cx, yp = phiy Y/Z + cy.
//
// Two forms:
//   k_synth_gemv  — one frame per grid row (the plain path): thread per vertex, basis
//                   rows streamed from HBM/MALL once per frame.
//   k_synth_mfma  — many frames: the basis product is a GEMM [3V x K] x [K x B] on the
//                   FP32 MFMA (v_mfma_f32_32x32x2f32), so a basis tile is read once per
//                   32 frames instead of once per frame; the epilogue fuses the mean,
//                   the rotation, the translation and the projection.
#include "common.h"

namespace tdi {

namespace {

constexpr int MAXK_LDS = 1024;

__global__ void __launch_bounds__(256) k_synth_gemv(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                     const float* __restrict__ theta, const int* __restrict__ mask,
                                                     float* __restrict__ P0, float* __restrict__ P,
                                                     float4* __restrict__ vproj)
{
    const int f = blockIdx.y;
    if (mask && !mask[f]) return;
    __shared__ float R[9];
    __shared__ float th[MAXK_LDS];
    const float* tf = theta + (size_t)f * L.K;
    for (int i = threadIdx.x; i < L.K; i += blockDim.x) th[i] = tf[i];
    __syncthreads();
    if (threadIdx.x == 0) rodrigues_f32(th + L.off_pose + 3, R, nullptr);
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.V) return;
    const FrameGPU fr = frames[f];
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int r = 3 * v + a;
        float s = m.mean[r];
        const float* e = m.exp_basis + (size_t)r * m.k_exp;
        for (int k = 0; k < L.nexp; ++k) s = fmaf(e[k], th[L.off_exp + k], s);
        if (L.nid) {
            const float* ib = m.id_basis + (size_t)r * m.k_id;
            for (int k = 0; k < L.nid; ++k) s = fmaf(ib[k], th[L.off_id + k], s);
        }
        p[a] = s;
    }
    const size_t o = (size_t)f * m.V + v;
    P0[3 * o + 0] = p[0];
    P0[3 * o + 1] = p[1];
    P0[3 * o + 2] = p[2];
    const float X = R[0] * p[0] + R[3] * p[1] + R[6] * p[2] + th[L.off_pose + 0];
    const float Y = R[1] * p[0] + R[4] * p[1] + R[7] * p[2] + th[L.off_pose + 1];
    const float Z = R[2] * p[0] + R[5] * p[1] + R[8] * p[2] + th[L.off_pose + 2];
    P[3 * o + 0] = X;
    P[3 * o + 1] = Y;
    P[3 * o + 2] = Z;
    vproj[o] = make_float4(fr.phix * X / Z + fr.cx, fr.phiy * Y / Z + fr.cy, Z, 0.f);
}

// ---------------------------------------------------------------------------------------
// MFMA synthesis.  Output tile: 32 basis rows (= part of 3V) x 32 frames per wave, K in
// steps of 2 (v_mfma_f32_32x32x2f32: A 32x2, B 2x32, C 32x32, 16 accumulators per lane).
// Operand layout for 32x32x2 f32 (one lane holds one A and one B element):
//   A[i][k]: lane = i + 32*k       B[k][j]: lane = j + 32*k
// Accumulator layout: lane l, accumulator a (0..15) holds C[i][j] with
//   j = l % 32, i = 8*(a/4) + 4*(l/32) + (a%4).
// Each workgroup (4 waves) covers 96 rows (32 vertices x 3 axes) x 32 frames, so that the
// epilogue can rotate whole vertices.  The expression block is the only basis in the
// batched per-frame stage (identity is folded into the mean there).
constexpr int SM_ROWS = 96;    // 32 vertices
constexpr int SM_FR = 32;

__global__ void __launch_bounds__(256) k_synth_mfma(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames, int n,
                                                     const float* __restrict__ theta, const int* __restrict__ mask,
                                                     float* __restrict__ P0, float* __restrict__ P,
                                                     float4* __restrict__ vproj)
{
    const int row0 = blockIdx.x * SM_ROWS;      // first basis row (3*v0)
    const int f0 = blockIdx.y * SM_FR;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ float sacc[SM_ROWS][SM_FR + 1];
    __shared__ int sactive;
    if (tid == 0) {
        int any = 0;
        for (int j = 0; j < SM_FR; ++j) any |= (f0 + j < n) && (!mask || mask[f0 + j]);
        sactive = any;
    }
    __syncthreads();
    if (!sactive) return;

    const int Ke = L.nexp;
    const int rowsTotal = 3 * m.V;
    // waves 0..2 each compute a 32-row slab; wave 3 idles during the MFMA phase
    if (wave < 3) {
        typedef float v16 __attribute__((ext_vector_type(16)));
        v16 acc = {0};
        const int i = lane & 31, kk = lane >> 5;
        const int row = row0 + wave * 32 + i;
        const int fj = f0 + (lane & 31);
        const bool rowok = row < rowsTotal;
        const bool frok = fj < n;
        const float* arow = m.exp_basis + (size_t)(rowok ? row : 0) * m.k_exp;
        const float* bcol = theta + (size_t)(frok ? fj : 0) * L.K + L.off_exp;
        for (int k = 0; k < Ke; k += 2) {
            const int kx = k + kk;
            float a = (rowok && kx < Ke) ? arow[kx] : 0.f;
            float b = (frok && kx < Ke) ? bcol[kx] : 0.f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int a = 0; a < 16; ++a) {
            const int r = 8 * (a >> 2) + 4 * (lane >> 5) + (a & 3);
            sacc[wave * 32 + r][lane & 31] = acc[a];
        }
    }
    __shared__ float sR[SM_FR][12];
    if (tid < SM_FR) {
        const int f = f0 + tid;
        if (f < n) {
            const float* tf = theta + (size_t)f * L.K;
            rodrigues_f32(tf + L.off_pose + 3, sR[tid], nullptr);
            sR[tid][9] = tf[L.off_pose + 0];
            sR[tid][10] = tf[L.off_pose + 1];
            sR[tid][11] = tf[L.off_pose + 2];
        }
    }
    __syncthreads();
    // epilogue: 32 vertices x 32 frames = 1024 (vertex, frame) pairs over 256 threads
    for (int e = tid; e < 32 * SM_FR; e += blockDim.x) {
        const int vl = e % 32, j = e / 32;
        const int v = row0 / 3 + vl;
        const int f = f0 + j;
        if (v >= m.V || f >= n) continue;
        if (mask && !mask[f]) continue;
        const float* R = sR[j];
        float p[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = m.mean[3 * v + a] + sacc[3 * vl + a][j];
        const size_t o = (size_t)f * m.V + v;
        P0[3 * o + 0] = p[0];
        P0[3 * o + 1] = p[1];
        P0[3 * o + 2] = p[2];
        const float X = R[0] * p[0] + R[3] * p[1] + R[6] * p[2] + R[9];
        const float Y = R[1] * p[0] + R[4] * p[1] + R[7] * p[2] + R[10];
        const float Z = R[2] * p[0] + R[5] * p[1] + R[8] * p[2] + R[11];
        P[3 * o + 0] = X;
        P[3 * o + 1] = Y;
        P[3 * o + 2] = Z;
        const FrameGPU fr = frames[f];
        vproj[o] = make_float4(fr.phix * X / Z + fr.cx, fr.phiy * Y / Z + fr.cy, Z, 0.f);
    }
}

} // namespace

int launch_synthesize(const ModelGPU& m, const Layout& L, const FrameGPU* frames, int n, const float* theta,
                      const int* mask, float* P0, float* P, float4* vproj, hipStream_t s)
{
    if (L.K > MAXK_LDS) {
        set_error("synthesize: K=%d exceeds %d", L.K, MAXK_LDS);
        return TDI_ERR_ARG;
    }
    dim3 grid((m.V + 255) / 256, n);
    hipLaunchKernelGGL(k_synth_gemv, grid, dim3(256), 0, s, m, L, frames, theta, mask, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_synthesize_gemm(const ModelGPU& m, const Layout& L, const FrameGPU* frames, int n, const float* theta,
                           const int* mask, float* P0, float* P, float4* vproj, float* /*scratch*/, hipStream_t s)
{
    if (L.nid) return launch_synthesize(m, L, frames, n, theta, mask, P0, P, vproj, s);
    dim3 grid((3 * m.V + SM_ROWS - 1) / SM_ROWS, (n + SM_FR - 1) / SM_FR);
    hipLaunchKernelGGL(k_synth_mfma, grid, dim3(256), 0, s, m, L, frames, n, theta, mask, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
