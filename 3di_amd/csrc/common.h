// common.h — internal types shared by the gfx950 kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/tdi_api.h"

namespace tdi {

constexpr int NLMK = TDI_NLANDMARKS;
constexpr uint64_t KEY_EMPTY = ~0ull;

// Offsets of the unknown blocks, OptimizationVariables order (Optimizer.h:71-91).
struct Layout {
    int K, off_tex, off_id, off_exp, off_pose, off_light;
    int ntex, nid, nexp, nlight;
};

inline Layout make_layout(const tdi_model_desc* m, const tdi_fit_options* o)
{
    Layout L{};
    L.ntex = o->fit_tex ? m->k_tex : 0;
    L.nid = o->fit_id ? m->k_id : 0;
    L.nexp = o->fit_exp ? m->k_exp : 0;
    L.nlight = o->fit_light == TDI_LIGHT_SH2 ? 9 : (o->fit_light ? 4 : 0);
    L.off_tex = 0;
    L.off_id = L.ntex;
    L.off_exp = L.off_id + L.nid;
    L.off_pose = L.off_exp + L.nexp;
    L.off_light = L.off_pose + 6;
    L.K = L.off_light + L.nlight;
    return L;
}

// Model resident in HBM (Renderer's device arrays, renderer.cu:28-72).
struct ModelGPU {
    int V, NT, k_id, k_tex, k_exp;
    int id_ld, exp_ld, tex_ld; // row strides of the bases: the column counts rounded up to 16, so
                               // that a 16-lane row segment is one aligned 64-byte sector
    const float* mean;       // [V][3]
    const float* id_basis;   // [3V][id_ld]
    const float* exp_basis;  // [3V][exp_ld]
    const float* tex_mean;   // [V]
    const float* tex_basis;  // [V][tex_ld]
    const int* tris;         // [NT][3]
    const int* lmk_ids;      // [51]
    const float* lb;         // [K] lower bounds of the unknowns (-inf = unbounded)
    const float* ub;         // [K]
    // per-frame vertex albedos [frames][V] of the current render (tex_mean + tex_basis beta, in
    // the order the shading sums them), when the albedo is fitted: the shading and the Jacobian
    // then read them instead of summing the basis at every pixel (NULL: they sum it)
    const float* alb = nullptr;
};

// One frame of a batch, device-resident.
struct FrameGPU {
    float phix, phiy, cx, cy;
    int x0, y0, W, H;
    float face_size;
    float lam[3];
    float Lint;
    int pad0;
    long long pix_off;       // offset of this frame in the workspace's crop-sized buffers
    long long obs_off;       // offset of this frame in the batch's observation buffers
};

// Per-frame fit state (device) for the batched path.
struct FitState {
    int active;              // 1 while the GN loop runs for this frame
    int searching;           // 1 while the line search of this iteration runs
    int stop;                // tdi_stop flags
    int status;              // tdi_status
    int iters, outer, ls_evals, inner;
    int fresh;               // 1 when the frame's render in the workspace is that of its theta
    float t;                 // current step
    float dg;                // delta . g
    double obj0;             // F(theta) of the current iteration
};

struct Options {
    tdi_fit_options o;
    Layout L;
    float taps[15];
};

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ uint32_t orderable_bits(float z)
{
    uint32_t u = __float_as_uint(z);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// FP32 coverage test, reference operation order (funcs.cu:950-979), no contraction:
// every product and sum is rounded separately so it matches a scalar CPU build.  The
// operators are written out here, under the pragma: HIP's __fmul_rn/__fadd_rn carry the
// header's contraction setting and clang would fuse them into v_fma_f32.
__device__ __forceinline__ bool cover_rn(float bx, float by, float cx, float cy, float ax, float ay,
                                         int x, int y, float* al, float* be)
{
#pragma clang fp contract(off)
    const float v0x = bx - ax, v1x = cx - ax, v2x = (float)x - ax;
    const float v0y = by - ay, v1y = cy - ay, v2y = (float)y - ay;
    const float d00 = v0x * v0x + v0y * v0y;
    const float d01 = v0x * v1x + v0y * v1y;
    const float d11 = v1x * v1x + v1y * v1y;
    const float d20 = v0x * v2x + v0y * v2y;
    const float d21 = v1x * v2x + v1y * v2y;
    const float den = d00 * d11 - d01 * d01;
    const float na = d11 * d20 - d01 * d21;
    const float nb = d00 * d21 - d01 * d20;
    // Sign pre-test: alpha > 0 and beta > 0 need num/den > 0.  For every finite non-zero
    // den the quotients here are normal numbers (|num| >= ~1e-21, |den| <= ~1e11 at pixel
    // coordinates), so the quotient's sign is sign(num) * sign(den) and the test only skips
    // the divisions of samples the full test rejects; den == 0 never covers (inf + inf > 1).
    const bool pa = (na > 0.f && den > 0.f) || (na < 0.f && den < 0.f);
    const bool pb = (nb > 0.f && den > 0.f) || (nb < 0.f && den < 0.f);
    if (!(pa && pb)) {
        *al = 0.f;
        *be = 0.f;
        return false;
    }
    const float a = na / den;
    const float b = nb / den;
    *al = a;
    *be = b;
    return a > 0.f && b > 0.f && (a + b <= 1.f);
}

// The second-order light terms (TDI_LIGHT_SH2) of the camera-space unit normal n:
// Y = (nx ny, ny nz, nx nz, nx^2 - ny^2, 3 nz^2 - 1).  Y_k alone (s = e_k) is the Jacobian
// column of s_k divided by the albedo.
__device__ __forceinline__ void sh2_terms(const float* n, float* Y)
{
    Y[0] = n[0] * n[1];
    Y[1] = n[1] * n[2];
    Y[2] = n[0] * n[2];
    Y[3] = n[0] * n[0] - n[1] * n[1];
    Y[4] = 3.f * n[2] * n[2] - 1.f;
}

// 1 - alpha - beta rounded as the reference does ((1 - alpha) - beta in FP32).
__device__ __forceinline__ float gamma_rn(float al, float be)
{
#pragma clang fp contract(off)
    return (1.f - al) - be;
}

// Compacted neighbour positions of sorted position `pos` (fill_krels, funcs.cu:288-320,
// with cumM0 = inclusive occupancy count): left/right = last rendered pixel at or before
// the horizontal neighbour, above/below = previous/next rendered pixel; the first and last
// positions keep 0.  Returns (kl, kr, ka, kb).
__device__ inline int4 pixel_nbrs(const int* __restrict__ cum, int k, long long base, int pos, int N, int W, int H)
{
    if (pos == 0 || pos == N - 1) return make_int4(0, 0, 0, 0);
    const int x = k / H;
    const int kleft = (x > 0) ? k - H : 0;
    const int kright = (x < W - 1) ? k + H : 0;
    const int cl = cum[base + kleft];
    int kl = cl > 1 ? cl - 1 : 0;
    if (kl >= N - 1) kl = N - 1;
    int kr = cum[base + kright] - 1;
    if (kr < 0) kr = 0;
    return make_int4(kl, kr, pos - 1, pos + 1);
}

// Rodrigues rotation and its derivatives in FP32 (fill_theta_etc, rotation_header.cu:46-149).
// Column-major R[col*3+row].
__device__ inline void rodrigues_f32(const float* u, float* R, float* dR)
{
    float th = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    float n0 = u[0] / th, n1 = u[1] / th, n2 = u[2] / th;
    float N[9] = {0, n2, -n1, -n2, 0, n0, n1, -n0, 0};
    float s = sinf(th), c1 = 1.f - cosf(th);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += N[k * 3 + r] * N[c * 3 + k];
            R[c * 3 + r] = c1 * acc + s * N[c * 3 + r] + (r == c ? 1.f : 0.f);
        }
    if (!dR) return;
    float U[9] = {0, u[2], -u[1], -u[2], 0, u[0], u[1], -u[0], 0};
    float inv = 1.f / (th * th);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float e0 = -R[i * 3 + 0] + (i == 0), e1 = -R[i * 3 + 1] + (i == 1), e2 = -R[i * 3 + 2] + (i == 2);
        float w0 = u[1] * e2 - u[2] * e1, w1 = u[2] * e0 - u[0] * e2, w2 = u[0] * e1 - u[1] * e0;
        float A[9] = {u[i] * U[0], u[i] * U[1] + w2, u[i] * U[2] - w1,
                      u[i] * U[3] - w2, u[i] * U[4], u[i] * U[5] + w0,
                      u[i] * U[6] + w1, u[i] * U[7] - w0, u[i] * U[8]};
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float acc = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc += A[k * 3 + r] * R[c * 3 + k];
                dR[9 * i + c * 3 + r] = acc * inv;
            }
    }
}

// Rodrigues rotation in FP64 (used for the vertex transform so that projected vertices come
// out correctly rounded to FP32), and its derivatives dR/du_i in FP64 (same formula as
// rodrigues_f32) for the barrier's landmark Jacobian.
template <typename TU>
__device__ inline void rodrigues_f64_d(const TU* uf, double* R, double* dR)
{
    const double u[3] = {uf[0], uf[1], uf[2]};
    const double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double N[9] = {0, u[2] / th, -u[1] / th, -u[2] / th, 0, u[0] / th, u[1] / th, -u[0] / th, 0};
    const double s = sin(th), c1 = 1.0 - cos(th);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += N[k * 3 + r] * N[c * 3 + k];
            R[c * 3 + r] = c1 * acc + s * N[c * 3 + r] + (r == c ? 1.0 : 0.0);
        }
    const double U[9] = {0, u[2], -u[1], -u[2], 0, u[0], u[1], -u[0], 0};
    const double inv = 1.0 / (th * th);
    for (int i = 0; i < 3; ++i) {
        const double e0 = -R[i * 3 + 0] + (i == 0), e1 = -R[i * 3 + 1] + (i == 1), e2 = -R[i * 3 + 2] + (i == 2);
        const double w0 = u[1] * e2 - u[2] * e1, w1 = u[2] * e0 - u[0] * e2, w2 = u[0] * e1 - u[1] * e0;
        const double A[9] = {u[i] * U[0], u[i] * U[1] + w2, u[i] * U[2] - w1,
                             u[i] * U[3] - w2, u[i] * U[4], u[i] * U[5] + w0,
                             u[i] * U[6] + w1, u[i] * U[7] - w0, u[i] * U[8]};
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += A[k * 3 + r] * R[c * 3 + k];
                dR[9 * i + c * 3 + r] = acc * inv;
            }
    }
}

template <typename TU>
__device__ inline void rodrigues_f64(const TU* uf, double* R)
{
    const double u[3] = {uf[0], uf[1], uf[2]};
    const double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double n0 = u[0] / th, n1 = u[1] / th, n2 = u[2] / th;
    const double N[9] = {0, n2, -n1, -n2, 0, n0, n1, -n0, 0};
    const double s = sin(th), c1 = 1.0 - cos(th);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += N[k * 3 + r] * N[c * 3 + k];
            R[c * 3 + r] = c1 * acc + s * N[c * 3 + r] + (r == c ? 1.0 : 0.0);
        }
}

// ------------------------------------------------------------------ host helpers
void set_error(const char* fmt, ...);

#define TDI_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::tdi::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return TDI_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

// ------------------------------------------------------------------ kernel launchers
// All launchers take device pointers and a stream; n = frames in the call; mask selects
// which frames take part (NULL = all).

// Landmarks per frame, [n][6][51]: x, y (image pixels), box lx, ux, ly, uy (face-size units).
// The frames a per-frame launch works on: every workspace slot 0..n-1 (act == nullptr), or
// the compacted list act[1 .. act[0]] of slot indices built on the device from a mask
// (launch_compact_frames).  Per-frame kernels walk the list with a grid-stride loop over
// blockIdx.y (or blockIdx.x), so a launch over few frames costs little even though a
// captured graph fixes its grid.
struct FrameList {
    const int* act;
    int n;              // workspace slots: bound of the list and of the grids
};
inline FrameList all_frames(int n) { return FrameList{nullptr, n}; }
__device__ __forceinline__ int fl_count(const int* act, int n) { return act ? act[0] : n; }
__device__ __forceinline__ int fl_frame(const int* act, int i) { return act ? act[1 + i] : i; }
// grid extent over frames: every slot when there are few, else `cap` blocks that loop
inline int grid_frames(int n, int cap) { return n < 1 ? 1 : (n < cap ? n : cap); }

struct LmkArrays {
    const float* base;
};

// geometry.hip
// rt: [slots][12] scratch for the per-frame FP64 rotation (column-major) and translation;
// P (camera-space vertices) may be null: the fit itself only needs P0 and the projections
int launch_synthesize(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                      double* rt, double* P0, float* P, float4* vproj, hipStream_t s);
// Synthesis of frames that share their identity coefficients (the identity stage's evaluations):
// the identity shape once (FP64, idshape [V][3]), then the expression MFMA tiles on it.
int launch_synthesize_shared_id(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                                const float* theta, double* rt, double* P0, float* P, float4* vproj, double* idshape,
                                hipStream_t s);
// The per-frame vertex albedos of the listed frames (ModelGPU::alb) from their theta.
int launch_albedo(const ModelGPU& m, const Layout& L, FrameList fl, const float* theta, float* alb, hipStream_t s);
int launch_synthesize_gemm(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                           const float* theta, double* rt, double* P0, float* P, float4* vproj, hipStream_t s);

// raster.hip
int launch_clear_keys(uint64_t* keys, long long total, hipStream_t s);
// box: [slots] visible bounding box (xmin, xmax, ymin, ymax), reset here, filled by launch_resolve
int launch_clear_keys_frames(const FrameGPU* frames, FrameList fl, uint64_t* keys, int4* box, int max_area,
                             hipStream_t s);
int launch_raster(const ModelGPU& m, const FrameGPU* frames, FrameList fl, const float4* vproj, int ntmp,
                  uint64_t* keys, int4* box, hipStream_t s);
int launch_box_full(const FrameGPU* frames, int n, int4* box, hipStream_t s);
// tcount: [slots][tmax] tile counts, tmax >= resolve_tiles(largest crop)
int resolve_tiles(int max_area);
// pvert/pbary: per visible position (vb, vc, va, crop index k) and (alpha, beta), so that the
// per-pixel kernels after resolve reach the vertices in one load
int launch_resolve(const ModelGPU& m, const FrameGPU* frames, FrameList fl, const float4* vproj,
                   const uint64_t* keys, int* cum, int* vis_list, int* npix, float* bary, int4* pvert,
                   float2* pbary, int max_area, int* tcount, int tmax, int4* box, hipStream_t s);

// visible box (and tile counts) only, for the passes that work from the keys
int launch_visible_box(const FrameGPU* frames, FrameList fl, const uint64_t* keys, int max_area, int* tcount,
                       int tmax, int4* box, hipStream_t s);

// shade.hip
int launch_shade(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                 const double* P0, const int4* pvert, const float2* pbary, const int* npix, const int4* box,
                 float* img, int max_area, hipStream_t s);
// shading and objective straight from the keys over the visible box (line-search evaluations)
int launch_shade_keys(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                      const double* P0, const float4* vproj, const uint64_t* keys, const int4* box, float* img,
                      hipStream_t s);
// objp: [frames][256] per-workgroup partials of the data objective (reduced in order into obj)
int launch_obj_from_barrier(FrameList fl, const double* objb, double* obj, hipStream_t s);
int launch_objective(const FrameGPU* frames, FrameList fl, const uint64_t* keys, const int4* box,
                     const float* blurred, const float2* obs_g, float w, double* objp, double* obj, hipStream_t s);
// box == nullptr: whole crops; else only the region the normalised gradients need
int launch_blur(const FrameGPU* frames, FrameList fl, const float* taps, const int4* box, const float* in, float* tmp,
                float* out, int max_area, hipStream_t s);
int launch_grads(const FrameGPU* frames, FrameList fl, const int* vis_list, const int* npix,
                 const int* cum, const float* blurred, const float2* obs_g, float w, float4* pixg, int4* pnb,
                 float* pcoef, double* objp, double* obj, int max_area, hipStream_t s);
int launch_obs_prepare(const FrameGPU* frames, int n, const float* taps, const float* obs, float* tmp, float* blurred,
                       float2* obs_g, int max_area, hipStream_t s);

// jacobian.hip
// init.hip
int launch_ortho_init(const ModelGPU& m, int n, const float* lmk, const float* cams, float* x_orth, float* pose,
                      float* angles, int* info, hipStream_t s);

int launch_dI(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
              const double* P0, const int4* pvert, const float2* pbary, const int* npix, const float4* pixg,
              float* dI, int mean_area, hipStream_t s);
// the light columns alone ([pixel][4]: lambda, Lintensity) of the four appended light unknowns
int launch_dI_light(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                    const double* P0, const int4* pvert, const float2* pbary, const int* npix, float* dIl, int max_area,
                    hipStream_t s);
// the light sub-fits' 4 x 4 normal equations from those rows: H [n][4][4] FP64, g [n][4]
int launch_jtj_light(const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb, const float* pcoef,
                     const float4* pixg, const float* dIl, float w, double* H, float* g, hipStream_t s);
int jtj_splits(int n, int max_area);
size_t jtj_partials_floats(int n, int S, int K, bool mfma);
// J^T J and J^T r for frames with more columns than one workgroup's LDS holds (the identity
// stage), FP64 MFMA on FP32 rows: U [crop pixels][jtj_big_row_stride(K)] floats and partials
// [jtj_big_partials_doubles] of scratch.
size_t jtj_big_partials_doubles(int n, int S, int K);
int jtj_big_row_stride(int K);
int launch_jtj_big(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                   const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g, float* U,
                   double* partials, int S, int max_area, hipStream_t s);
int launch_jtj_f32(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                   const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g,
                   float* partials, int S, hipStream_t s);
int launch_jtj_mfma(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                    const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g,
                    float* partials, int S, int g_on_mfma, int two_term, hipStream_t s);

// solve.hip
int launch_barrier(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                   const float* theta, const LmkArrays& lm, int use_landmarks, int mode, double* obj, double* objb,
                   double* H, float* g, hipStream_t s);
// The barrier's derivative terms (mode 1) at any K: D is [n][2][NLMK][K] doubles of scratch.
int launch_barrier_big(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                       const LmkArrays& lm, int use_landmarks, double* D, double* H, float* g, hipStream_t s);
// cholesky_big.hip: FP64 blocked Cholesky solve for K above what one workgroup holds in LDS
constexpr int CHOL_BIG_MAXK = 4096;
int launch_cholesky_big(int n, int K, const float* H, const float* g, float* delta, int* status, hipStream_t s);
// The same on FP64 matrices A [n][K][K] factored in place, g FP64 [n][K].
int launch_cholesky_big_f64(int n, int K, double* A, const double* g, float* delta, int* status, hipStream_t s);
int launch_cholesky_big_f64(int n, int K, double* A, const double* g, double* delta, int* status, hipStream_t s);
// identity.hip: the identity stage's joint system from the frames' normal equations.
int launch_assemble_joint(int n, int K, int Ks, const double* H, const float* g, double* HJ, double* gJ, hipStream_t s);
int launch_assemble_joint(int n, int K, int Ks, const double* H, const double* g, double* HJ, double* gJ, hipStream_t s);
// the joint landmark-only fit over a group of frames: per-frame terms (mode 0 objective and least
// constraint value, mode 1 also gradient and Gauss-Newton matrix over the frame's
// [alpha | s | eps | tau | u]) and the terms counted once (identity bounds, tau_b s)
constexpr int LMK_JOINT_MAXN = 320;   // unknowns per frame: k_id + slack + k_exp + 6
size_t lmk_joint_scratch_doubles(int n, int Ka, int Ke, int S1);
// each = 1: every frame a group of its own (z [n][nf], theta0 [n][K]); mask [n] (may be null)
// selects the frames to evaluate
int launch_lmk_joint(const ModelGPU& m, const Layout& L, const FrameGPU* frames, const float* lmk,
                     const float* theta0, int n, int Ka, int Ke, int S1, const double* z, int mode, int each,
                     const int* mask, double* objf, double* gminf, double* H, double* g, double* J, hipStream_t s);
int launch_lmk_joint_shared(const ModelGPU& m, const Layout& L, int Ka, int S1, long long KJ, const double* z,
                            double tau_b, const double* taus, int mode, int groups, const int* mask, double* obj,
                            double* HJ, double* gJ, hipStream_t s);
int launch_shared_bounds(int Ks, long long KJ, const float* x, const float* lb, const float* ub, int mode, double* obj,
                         double* HJ, double* gJ, hipStream_t s);

// adjoint.hip: the render loss and its adjoint (scratch: device buffers of the caller)
struct AdjScratch {
    double* dB;        // [cap] dL/d(blurred render)
    long long cap;
    double* dP0;       // [n][V][3] dL/d(model-space vertices)
    double* gpl;       // [n][15] pose (6) and light (9)
    double* gexp;      // [n][nexp]
    double* gid;       // [n][nid]
    double* loss;      // [n]
};
int launch_add_offsets(int V, const FrameGPU* frames, int n, const double* rt, const float* dverts, double* P0,
                       float4* vproj, hipStream_t s);
int launch_adjoint(const ModelGPU& m, const Layout& L, const FrameGPU* frames, int n, const float* theta,
                   const double* P0, const int4* pvert, const float2* pbary, const int* vis, const int* npix,
                   const float4* pixg, const float* lmk, const float* taps, int max_area, float* img, float* tmp,
                   float* blurred, const AdjScratch& a, const double* gloss, double* loss, float* gtheta,
                   float* gverts, hipStream_t s);

// lmkfit.hip: the landmark-only fit with slack (mode 0; rep: device tdi_lmk_report[n]) or its
// terms at zin (mode 1); frames and lmk [n][6][51] cover the n frames of the launch
int launch_lmk_slack(const ModelGPU& m, const Layout& L, const FrameGPU* frames, const float* lmk, int n,
                     float* theta, int Kl, int mode, const double* zin, double tau, double* H, double* g,
                     double* obj, void* rep, hipStream_t s);
// H: the normal matrices, FP32 (the building-block ABI) or FP64 (the fits' workspace)
int launch_cholesky(int n, int K, const float* H, const float* g, float* delta, int* status, int mode, const int* mask,
                    hipStream_t s);
int launch_cholesky(int n, int K, const double* H, const float* g, float* delta, int* status, int mode, const int* mask,
                    hipStream_t s);
int launch_ls_init(int n, int K, FitState* st, const int* chol_status, const float* delta, const float* g,
                   const double* obj, int* searching, hipStream_t s);
// unknowns [hold0, hold1) keep theta's values in the trials (hold0 == hold1: none)
int launch_ls_propose(int n, int K, int off_tz, float t_min, FitState* st, const float* theta, const float* delta,
                      float* theta_try, int* searching, int* evalmask, int hold0, int hold1, hipStream_t s);
int launch_ls_decide(int n, int K, FitState* st, const int* evalmask, int* searching, const double* obj,
                     const double* objb, const int* face_ok, float* theta, const float* theta_try,
                     const tdi_fit_options& o, hipStream_t s);
int launch_face_check(int V, const int* lmk_ids, const float4* vproj, float ref_face_size, FrameList fl,
                      int* face_ok, hipStream_t s);
int launch_copy_mask(int n, const FitState* st, int* mask, hipStream_t s);
int launch_split_fresh(int n, const FitState* st, int* mask_r, int* mask_u, hipStream_t s);
// act[0] = number of set entries of mask[0..n-1], act[1..] = their indices in increasing order
int launch_compact_frames(int n, const int* mask, int* act, hipStream_t s);

} // namespace tdi
