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
//   k_synth_tile   — many frames: [3V x K] x [K x B] on v_mfma_f64_16x16x4f64; a workgroup
//                    owns 16 vertices, stages their basis rows in LDS once and walks every
//                    frame of the list; the epilogue fuses mean, rotation, translation and
//                    projection from registers.
#include "common.h"

namespace tdi {

namespace {

constexpr int MAXK_LDS = 1024;

__device__ __forceinline__ void store_vertex(const FrameGPU& fr, const double* R, const double* tau, const double* p,
                                             size_t o, double* P0, float* P, float4* vproj)
{
    // model-space vertices stay FP64: the shading normal is the cross product of edges of
    // about a millimetre between coordinates of about 100 mm, which FP32 would round to 1e-5
    P0[3 * o + 0] = p[0];
    P0[3 * o + 1] = p[1];
    P0[3 * o + 2] = p[2];
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
                                                     int n, const double* __restrict__ rt, double* __restrict__ P0,
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
                const float* e = m.exp_basis + (size_t)r * m.exp_ld;
                for (int k = 0; k < L.nexp; ++k) s = fma((double)e[k], (double)th[L.off_exp + k], s);
                if (L.nid) {
                    const float* ib = m.id_basis + (size_t)r * m.id_ld;
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

typedef double v4d __attribute__((ext_vector_type(4)));

// Vertex tiles for v_mfma_f64_16x16x4f64: a workgroup owns SY_VERTS = 16 vertices and keeps one
// accumulator per coordinate, whose tile row i is vertex i (A = that coordinate's basis rows of the
// 16 vertices, B = the coefficients of 16 frames), so lane l, element r of the x, y and z
// accumulators holds vertex l/16 + 4 r for frame l%16: every MFMA row is a vertex coordinate
// (three MFMAs per k-step for 16 vertices).  The basis rows are staged in LDS (as FP64) once; then
// each wavefront walks its own groups of 16 list entries with no further barrier.  A lane loads
// the coefficients of its frame for its four k residues up front (one batch of independent loads),
// and the epilogue (mean, FP64 rotation, translation, projection) runs from registers.
constexpr int SY_VERTS = 16;               // vertices per workgroup
constexpr int SY_KMAX = 80;                // expression coefficients held in LDS (K_exp <= 80)

// Three wavefronts per SIMD (register budget 168) hide the coefficient loads of one behind the
// MFMAs of the others.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_synth_tile(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames, const float* __restrict__ theta,
             const int* __restrict__ act, int n, const double* __restrict__ rt, const double* __restrict__ mean64,
             double* __restrict__ P0, float* __restrict__ P, float4* __restrict__ vproj)
{
    __shared__ double sA[3][SY_KMAX][16];              // [coordinate][k][vertex]
    __shared__ double smean[SY_VERTS * 3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (fl_count(act, n) == 0) return;       // an emptied list: skip the basis staging
    const int v0 = blockIdx.x * SY_VERTS;
    const int Ke = L.nexp, ke = m.exp_ld;
    const int K4 = (Ke + 3) & ~3;
    // basis rows of the workgroup's vertices, zero padded (k >= Ke, v >= V)
    // (a fixed count of loads per thread, issued unconditionally from clamped indices and zeroed
    // by a select, so that they are in flight together)
    constexpr int NST = 3 * SY_KMAX * 16 / 256;
    float av[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int e = tid + 256 * q;
        const int k = e % K4, i = (e / K4) & 15, c = (e / K4) >> 4;
        const int v = v0 + i;
        av[q] = m.exp_basis[(size_t)(3 * min(v, m.V - 1) + min(c, 2)) * ke + min(k, Ke - 1)];
    }
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int e = tid + 256 * q;
        if (e < 3 * K4 * 16) {
            const int k = e % K4, i = (e / K4) & 15, c = (e / K4) >> 4;
            const int v = v0 + i;
            sA[c][k][i] = (k < Ke && v < m.V) ? (double)av[q] : 0.0;
        }
    }
    // the base shape: the model's mean, or a given FP64 shape (the identity stage's shared identity)
    for (int e = tid; e < SY_VERTS * 3; e += blockDim.x) {
        const int v = v0 + e / 3;
        smean[e] = v < m.V ? (mean64 ? mean64[3 * v0 + e] : (double)m.mean[3 * v0 + e]) : 0.0;
    }
    __syncthreads();
    const int nact = fl_count(act, n);
    const int ngroups = Ke > 0 ? (nact + 15) >> 4 : 0;   // Ke == 0 is handled by k_synth_gemv
    const int j = lane & 15, kq = lane >> 4;
    for (int g = wave; g < ngroups; g += 4) {
        const int c = 16 * g + j;
        const bool live = c < nact;
        const int f = fl_frame(act, live ? c : 0);
        // this lane's B entries: coefficient kq + 4 q of its frame, every load issued
        // unconditionally (clamped to a valid column) so the batch has no branches.  Nothing
        // needs zeroing: the A rows past K_exp are zero, and an output column depends on its
        // own frame's B column only, so lanes without a live frame just skip the epilogue.
        const float* tb = theta + (size_t)f * L.K + L.off_exp;
        float bv[SY_KMAX / 4];
#pragma unroll
        for (int q = 0; q < SY_KMAX / 4; ++q) bv[q] = tb[min(kq + 4 * q, Ke - 1)];
        v4d acc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[a] = (v4d){0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < SY_KMAX / 4; ++q) {
            if (4 * q < K4) {
                const double b = (double)bv[q];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(sA[a][4 * q + kq][j], b, acc[a], 0, 0, 0);
            }
        }
        if (live) {
            const FrameGPU fr = frames[f];
            double R[12];
#pragma unroll
            for (int q = 0; q < 12; ++q) R[q] = rt[(size_t)f * 12 + q];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int vl = kq + 4 * r, v = v0 + vl;
                if (v < m.V) {
                    double p[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) p[a] = smean[3 * vl + a] + acc[a][r];
                    store_vertex(fr, R, R + 9, p, (size_t)f * m.V + v, P0, P, vproj);
                }
            }
        }
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
                      double* rt, double* P0, float* P, float4* vproj, hipStream_t s)
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
    if (m.alb && L.ntex) return launch_albedo(m, L, fl, theta, const_cast<float*>(m.alb), s);
    return TDI_OK;
}

namespace {

// alb[f][v] = tex_mean[v] + sum_q tex_basis[v][q] beta_q, summed in q order with fmaf as the
// shading's per-pixel loop does, so that both give the same value.  A workgroup owns 64 vertices:
// their basis rows are staged in LDS 128 columns at a time (coalesced row segments), then its four
// wavefronts run the sums of four frames at once, lane = vertex (conflict-free LDS rows, the
// frame's coefficients broadcast); a sum carried over column chunks goes through alb itself.
constexpr int AL_V = 64, AL_Q = 128;

__global__ void __launch_bounds__(256) k_albedo(ModelGPU m, Layout L, const int* __restrict__ act, int n,
                                                 const float* __restrict__ theta, float* __restrict__ alb)
{
    __shared__ float sB[AL_V][AL_Q + 1];
    __shared__ float sbeta[4][AL_Q];
    const int v0 = blockIdx.x * AL_V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int v = v0 + lane;
    const int nact = fl_count(act, n);
    for (int q0 = 0; q0 < L.ntex; q0 += AL_Q) {
        const int nq = min(AL_Q, L.ntex - q0);
        __syncthreads();   // the previous chunk's staging is no longer read
        for (int e = threadIdx.x; e < AL_V * nq; e += blockDim.x) {
            const int vl = e / nq, q = e - vl * nq;
            sB[vl][q] = v0 + vl < m.V ? m.tex_basis[(size_t)(v0 + vl) * m.tex_ld + q0 + q] : 0.f;
        }
        for (int yi0 = 0; yi0 < nact; yi0 += 4) {
            __syncthreads();
            for (int e = threadIdx.x; e < 4 * nq; e += blockDim.x) {
                const int w = e / nq, q = e - w * nq;
                const int yi = yi0 + w;
                sbeta[w][q] = yi < nact ? theta[(size_t)fl_frame(act, yi) * L.K + L.off_tex + q0 + q] : 0.f;
            }
            __syncthreads();
            const int yi = yi0 + wave;
            if (yi < nact && v < m.V) {
                const int f = fl_frame(act, yi);
                float r = q0 == 0 ? m.tex_mean[v] : alb[(size_t)f * m.V + v];
                for (int q = 0; q < nq; ++q) r = fmaf(sB[lane][q], sbeta[wave][q], r);
                alb[(size_t)f * m.V + v] = r;
            }
        }
    }
}

} // namespace

int launch_albedo(const ModelGPU& m, const Layout& L, FrameList fl, const float* theta, float* alb, hipStream_t s)
{
    if (L.ntex > 512) {
        set_error("albedo: k_tex %d > 512", L.ntex);
        return TDI_ERR_ARG;
    }
    hipLaunchKernelGGL(k_albedo, dim3((m.V + AL_V - 1) / AL_V), dim3(256), 0, s, m, L, fl.act, fl.n, theta, alb);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_synthesize_gemm(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                           const float* theta, double* rt, double* P0, float* P, float4* vproj, hipStream_t s)
{
    if (L.nid) return launch_synthesize(m, L, frames, fl, theta, rt, P0, P, vproj, s);
    int rc = launch_frame_rt(L, fl, theta, rt, s);
    if (rc) return rc;
    if (L.nexp > SY_KMAX || L.nexp == 0) return launch_synthesize(m, L, frames, fl, theta, rt, P0, P, vproj, s);
    const dim3 grid((m.V + SY_VERTS - 1) / SY_VERTS);
    hipLaunchKernelGGL(k_synth_tile, grid, dim3(256), 0, s, m, L, frames, theta, fl.act, fl.n, rt,
                       (const double*)nullptr, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    if (m.alb && L.ntex) return launch_albedo(m, L, fl, theta, const_cast<float*>(m.alb), s);
    return TDI_OK;
}

namespace {

// mean + identity basis x alpha in FP64 for every vertex, alpha of the first listed frame (the
// frames of an identity-stage evaluation share it).
__global__ void __launch_bounds__(256) k_id_shape(ModelGPU m, Layout L, const int* __restrict__ act, int n,
                                                   const float* __restrict__ theta, double* __restrict__ S)
{
    __shared__ float sa[MAXK_LDS];
    if (fl_count(act, n) == 0) return;
    const float* tf = theta + (size_t)fl_frame(act, 0) * L.K + L.off_id;
    for (int k = threadIdx.x; k < L.nid; k += blockDim.x) sa[k] = tf[k];
    __syncthreads();
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.V) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int r = 3 * v + a;
        double sacc = m.mean[r];
        const float* ib = m.id_basis + (size_t)r * m.id_ld;
        for (int k = 0; k < L.nid; ++k) sacc = fma((double)ib[k], (double)sa[k], sacc);
        S[r] = sacc;
    }
}

} // namespace

int launch_synthesize_shared_id(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                                const float* theta, double* rt, double* P0, float* P, float4* vproj, double* idshape,
                                hipStream_t s)
{
    if (L.nexp == 0 || L.nexp > SY_KMAX || L.nid > MAXK_LDS) return launch_synthesize(m, L, frames, fl, theta, rt, P0, P, vproj, s);
    int rc = launch_frame_rt(L, fl, theta, rt, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_id_shape, dim3((m.V + 255) / 256), dim3(256), 0, s, m, L, fl.act, fl.n, theta, idshape);
    const dim3 grid((m.V + SY_VERTS - 1) / SY_VERTS);
    hipLaunchKernelGGL(k_synth_tile, grid, dim3(256), 0, s, m, L, frames, theta, fl.act, fl.n, rt,
                       (const double*)idshape, P0, P, vproj);
    TDI_HIP(hipGetLastError());
    if (m.alb && L.ntex) return launch_albedo(m, L, fl, theta, const_cast<float*>(m.alb), s);
    return TDI_OK;
}

} // namespace tdi
