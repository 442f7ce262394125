// jacobian.hip — per-pixel Jacobian and the normal equations on gfx950.
//
// Reference:
//   fill_optimization_auxiliary_variables_phase1 (newfuncs.cu:173-255),
//   fill_optimization_dI_depsilon_userotated (265-299), ..._dI_dalpha (314-373),
//   ..._dI_dbeta (383-403): dI/dtheta_j = -(gx dWx/dtheta_j + gy dWy/dtheta_j) for the
//   geometric unknowns (W = pinhole projection of v = a P_b + b P_c + g P_a), and
//   T_j (1 + L Id) for albedo; the light columns reuse dI_dlambda = L rho R n and
//   dI_dLintensity = rho Id (newfuncs.cu:151-155); with nine light unknowns (a, L, s) the
//   columns are rho R n, rho Id (lambda = 0) and rho Y_k(R n) (sh2_terms);
//   fill_optimization_auxiliary_variables_phase2_new (newfuncs.cu:417-463):
//   J_x = c dy s, J_y = -c dx s, s = dy (dI[r]-dI[l]) - dx (dI[b]-dI[a]), c = 1/(0.001+h^3),
//   with the compacted-neighbour positions of fill_krels/fill_krels2 (funcs.cu:288-340);
//   cublasSgemm (derivative_computer.cu:159-164): H += w (Jx^T Jx + Jy^T Jy),
//   g -= w (Jx^T g^s_x + Jy^T g^s_y).
//
// k_dI          : four pixels per wavefront, 16 lanes over the K columns of each, so four
//                 independent gather chains (triangle -> vertices -> basis rows) are in flight
//                 per wave; basis rows are read with unit stride across the 16 lanes.
// k_jtj_f32     : FP32 J^T J / J^T r from LDS-staged J rows, 4x4 register tiles (plain path).
// k_jtj_lds     : one row u = c |d| s per pixel (the reference's two rows are parallel), formed
//                 from branch-free gathers of dI, split into BF16 parts (hi/mid/lo) in LDS in
//                 the MFMA operand layout, [u | r]^T [u | r] on v_mfma_f32_16x16x16bf16_1k with
//                 FP32 accumulation (six products, or three for the two-term split).
//                 The residual rides along as column K, so J^T r comes out of the same MFMA
//                 tiles; J^T r is also accumulated in FP32, and the caller picks which one feeds
//                 the gradient (the GN fixed point depends on g alone, so FP32 J^T r keeps
//                 batched trajectories on the FP64 ones).
// k_jtj_reduce  : sums the per-split partial tiles, scales by w, writes a dense symmetric H.
#include "common.h"

namespace tdi {

namespace {

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

struct FrameRot {
    float R[9], dR[27], tau[3], lam[3], L;
    float a[3], sh[5];   // directional and second-order light terms (zero unless nine light unknowns)
};

__device__ inline void load_frame_rot(const Layout& L, const FrameGPU& fr, const float* tf, FrameRot& s)
{
    rodrigues_f32(tf + L.off_pose + 3, s.R, s.dR);
    for (int a = 0; a < 3; ++a) s.tau[a] = tf[L.off_pose + a];
    if (L.nlight == 4) {
        for (int a = 0; a < 3; ++a) s.lam[a] = tf[L.off_light + a];
        s.L = tf[L.off_light + 3];
    } else if (L.nlight == 9) {
        for (int a = 0; a < 3; ++a) s.lam[a] = 0.f;   // Id = -(R n).v_c; a enters as (R n).a
        s.L = tf[L.off_light + 3];
    } else {
        for (int a = 0; a < 3; ++a) s.lam[a] = fr.lam[a];
        s.L = fr.Lint;
    }
    for (int k = 0; k < 3; ++k) s.a[k] = L.nlight == 9 ? tf[L.off_light + k] : 0.f;
    for (int k = 0; k < 5; ++k) s.sh[k] = L.nlight == 9 ? tf[L.off_light + 4 + k] : 0.f;
}

constexpr int DI_PX_PER_WG = 256;   // pixels per workgroup and frame in k_dI
constexpr int DI_PASSES = 4;        // strip passes of k_dI (its grid's z)

// Four wavefronts per SIMD (register budget 128): more pixels' gathers in flight per CU.
// IDTEX: the layout has identity or albedo columns (the identity stage); the per-frame fit's
// instantiation carries none of their code, which would cost it registers.
template <bool IDTEX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_dI(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                             const int* __restrict__ act, int n, const float* __restrict__ theta,
                                             const double* __restrict__ P0, const int4* __restrict__ pvert,
                                             const float2* __restrict__ pbary, const int* __restrict__ npix,
                                             const float4* __restrict__ pixg, float* __restrict__ dI)
{
    __shared__ FrameRot sr;
    __shared__ float stex[512];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
    const int f = fl_frame(act, yi);
    const int N = npix[f];
    // The visible list runs column by column, so a contiguous range of positions is a vertical
    // strip of the face.  Strip c goes to the workgroups with blockIdx.x % 8 == c, which the
    // dispatcher places on one XCD (workgroups are dealt round-robin to the 8 XCDs and gridDim.x
    // is a multiple of 8): each XCD's L2 then holds the basis rows of one strip of the mesh.
    // A strip is shared by as many workgroups as give each about DI_PX_PER_WG pixels; the
    // others leave before the frame set-up.
    // With gridDim.z passes the frame is cut into 8 gridDim.z strips: pass z takes strips
    // 8 z .. 8 z + 7 of every frame, and the dispatcher issues pass 0's workgroups (all frames)
    // before pass 1's, so an XCD's L2 holds the rows of a narrower strip at a time.
    const int nst = 8 * (int)gridDim.z;
    const int strip = (blockIdx.x & 7) + 8 * (int)blockIdx.z, sb = blockIdx.x >> 3;
    const int nsb = max(1, min((int)(gridDim.x >> 3), (N + nst * DI_PX_PER_WG - 1) / (nst * DI_PX_PER_WG)));
    if (sb >= nsb) continue;
    const FrameGPU fr = frames[f];
    const float* tf = theta + (size_t)f * L.K;
    if (threadIdx.x == 0) load_frame_rot(L, fr, tf, sr);
    for (int q = threadIdx.x; q < L.ntex; q += blockDim.x) stex[q] = tf[L.off_tex + q];
    __syncthreads();
    const long long base = fr.pix_off;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int sub = lane >> 4, l16 = lane & 15;      // four pixels per wavefront, 16 lanes each
    const double* P0f = P0 + (size_t)f * m.V * 3;
    const int K = L.K;
    const int q0 = (int)((long long)N * strip / nst), q1 = (int)((long long)N * (strip + 1) / nst);
    for (int pos0 = q0 + (sb * 4 + wave) * 4; pos0 < q1; pos0 += nsb * 16) {
        const int pos = pos0 + sub;
        if (pos >= q1) continue;
        const int4 pv = pvert[base + pos];
        const float2 ab = pbary[base + pos];
        const float al = ab.x, be = ab.y, ga = gamma_rn(al, be);
        const int vb = pv.x, vc = pv.y, va = pv.z;
        float v[3], p0[3];
        double p0d[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            p0d[a] = (double)al * P0f[3 * vb + a] + (double)be * P0f[3 * vc + a] + (double)ga * P0f[3 * va + a];
            p0[a] = (float)p0d[a];
        }
        // camera-space surface point R p0 + tau (R column-major)
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = sr.R[a] * p0[0] + sr.R[3 + a] * p0[1] + sr.R[6 + a] * p0[2] + sr.tau[a];
        const float4 gg = pixg[base + pos];
        const float iz = 1.f / v[2];
        // dI = c . dv with c = (-gx phix / vz, -gy phiy / vz, (gx phix vx + gy phiy vy) / vz^2)
        const float c0 = -gg.x * fr.phix * iz;
        const float c1 = -gg.y * fr.phiy * iz;
        const float c2 = (gg.x * fr.phix * v[0] + gg.y * fr.phiy * v[1]) * iz * iz;
        // q = R^T c, so that dI for a model-space direction e is q . e
        const float q0 = sr.R[0] * c0 + sr.R[1] * c1 + sr.R[2] * c2;
        const float q1 = sr.R[3] * c0 + sr.R[4] * c1 + sr.R[5] * c2;
        const float q2 = sr.R[6] * c0 + sr.R[7] * c1 + sr.R[8] * c2;
        // shading quantities for albedo and light columns
        float rho = 0.f, Id = 0.f, rn[3] = {0.f, 0.f, 0.f}, Y[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, S = 0.f;
        if (L.ntex || L.nlight) {
            // normal and Id in FP64 from the FP64 corners (edges of about a millimetre)
            const double* B0 = P0f + 3 * vb;
            const double* C0 = P0f + 3 * vc;
            const double* A0 = P0f + 3 * va;
            const double e1x = B0[0] - A0[0], e1y = B0[1] - A0[1], e1z = B0[2] - A0[2];
            const double e2x = C0[0] - A0[0], e2y = C0[1] - A0[1], e2z = C0[2] - A0[2];
            double ndx = e1y * e2z - e1z * e2y, ndy = e1z * e2x - e1x * e2z, ndz = e1x * e2y - e1y * e2x;
            const double inv = 1.0 / sqrt(ndx * ndx + ndy * ndy + ndz * ndz);
            ndx *= inv; ndy *= inv; ndz *= inv;
            const double d0 = (double)sr.lam[0] - sr.tau[0], d1 = (double)sr.lam[1] - sr.tau[1],
                         d2 = (double)sr.lam[2] - sr.tau[2];
            const double lx = sr.R[0] * d0 + sr.R[1] * d1 + sr.R[2] * d2;
            const double ly = sr.R[3] * d0 + sr.R[4] * d1 + sr.R[5] * d2;
            const double lz = sr.R[6] * d0 + sr.R[7] * d1 + sr.R[8] * d2;
            Id = (float)(ndx * lx + ndy * ly + ndz * lz - (ndx * p0d[0] + ndy * p0d[1] + ndz * p0d[2]));
            const float nx = (float)ndx, ny = (float)ndy, nz = (float)ndz;
            const float* rsrc = m.alb ? m.alb + (size_t)f * m.V : m.tex_mean;
            float rb = rsrc[vb], rc = rsrc[vc], ra = rsrc[va];
            for (int q = 0; !m.alb && q < L.ntex; ++q) {
                rb = fmaf(m.tex_basis[(size_t)vb * m.tex_ld + q], stex[q], rb);
                rc = fmaf(m.tex_basis[(size_t)vc * m.tex_ld + q], stex[q], rc);
                ra = fmaf(m.tex_basis[(size_t)va * m.tex_ld + q], stex[q], ra);
            }
            rho = al * rb + be * rc + ga * ra;
            rn[0] = sr.R[0] * nx + sr.R[3] * ny + sr.R[6] * nz;
            rn[1] = sr.R[1] * nx + sr.R[4] * ny + sr.R[7] * nz;
            rn[2] = sr.R[2] * nx + sr.R[5] * ny + sr.R[8] * nz;
            if (L.nlight == 9) {
                sh2_terms(rn, Y);
                S = rn[0] * sr.a[0] + rn[1] * sr.a[1] + rn[2] * sr.a[2];
#pragma unroll
                for (int k = 0; k < 5; ++k) S += sr.sh[k] * Y[k];
            }
        }
        float* out = dI + (size_t)(base + pos) * K;
        // expression columns, 80 per group: every basis gather of the group is issued before
        // the arithmetic (clamped columns, masked stores), so the group costs one round trip
        if (L.nexp) {
            const int ke = m.exp_ld;
            const float* Eb = m.exp_basis + (size_t)(3 * vb) * ke;   // rows 3v, 3v+1, 3v+2
            const float* Ec = m.exp_basis + (size_t)(3 * vc) * ke;
            const float* Ea = m.exp_basis + (size_t)(3 * va) * ke;
            for (int c0 = 0; c0 < L.nexp; c0 += 80) {
                float eb[5][3], ec[5][3], ea[5][3];
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = min(c0 + 16 * i + l16, L.nexp - 1);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        eb[i][a] = Eb[(size_t)a * ke + c];
                        ec[i][a] = Ec[(size_t)a * ke + c];
                        ea[i][a] = Ea[(size_t)a * ke + c];
                    }
                }
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = c0 + 16 * i + l16;
                    const float ex = al * eb[i][0] + be * ec[i][0] + ga * ea[i][0];
                    const float ey = al * eb[i][1] + be * ec[i][1] + ga * ea[i][1];
                    const float ez = al * eb[i][2] + be * ec[i][2] + ga * ea[i][2];
                    if (c < L.nexp) out[L.off_exp + c] = q0 * ex + q1 * ey + q2 * ez;
                }
            }
        }
        // identity columns (the identity stage), 80 per group as the expression columns
        if (IDTEX && L.nid) {
            const int ki = m.id_ld;
            const float* Ib = m.id_basis + (size_t)(3 * vb) * ki;
            const float* Ic = m.id_basis + (size_t)(3 * vc) * ki;
            const float* Ia = m.id_basis + (size_t)(3 * va) * ki;
            for (int c0 = 0; c0 < L.nid; c0 += 80) {
                float eb[5][3], ec[5][3], ea[5][3];
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = min(c0 + 16 * i + l16, L.nid - 1);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        eb[i][a] = Ib[(size_t)a * ki + c];
                        ec[i][a] = Ic[(size_t)a * ki + c];
                        ea[i][a] = Ia[(size_t)a * ki + c];
                    }
                }
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = c0 + 16 * i + l16;
                    const float ex = al * eb[i][0] + be * ec[i][0] + ga * ea[i][0];
                    const float ey = al * eb[i][1] + be * ec[i][1] + ga * ea[i][1];
                    const float ez = al * eb[i][2] + be * ec[i][2] + ga * ea[i][2];
                    if (c < L.nid) out[L.off_id + c] = q0 * ex + q1 * ey + q2 * ez;
                }
            }
        }
        // albedo columns (the identity stage), 80 per group
        if (IDTEX && L.ntex) {
            const int kt = m.tex_ld;
            const float* Tb = m.tex_basis + (size_t)vb * kt;
            const float* Tc = m.tex_basis + (size_t)vc * kt;
            const float* Ta = m.tex_basis + (size_t)va * kt;
            for (int c0 = 0; c0 < L.ntex; c0 += 80) {
                float tb[5], tc[5], ta[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = min(c0 + 16 * i + l16, L.ntex - 1);
                    tb[i] = Tb[c];
                    tc[i] = Tc[c];
                    ta[i] = Ta[c];
                }
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    const int c = c0 + 16 * i + l16;
                    const float tt = al * tb[i] + be * tc[i] + ga * ta[i];
                    if (c < L.ntex) out[L.off_tex + c] = tt + sr.L * Id * tt + S * tt;
                }
            }
        }
        // the other columns: pose and light
        for (int j = L.off_pose + l16; j < K; j += 16) {
            float val;
            if (j < L.off_pose + 3) {
                val = (j == L.off_pose) ? c0 : (j == L.off_pose + 1) ? c1 : c2;
            } else if (j < L.off_light) {
                const float* d = sr.dR + 9 * (j - L.off_pose - 3);
                const float dvx = d[0] * p0[0] + d[3] * p0[1] + d[6] * p0[2];
                const float dvy = d[1] * p0[0] + d[4] * p0[1] + d[7] * p0[2];
                const float dvz = d[2] * p0[0] + d[5] * p0[1] + d[8] * p0[2];
                val = c0 * dvx + c1 * dvy + c2 * dvz;
            } else {
                const int c = j - L.off_light;
                val = (c < 3) ? (L.nlight == 9 ? rho : sr.L * rho) * rn[c] : (c == 3) ? rho * Id : rho * Y[min(c - 4, 4)];
            }
            out[j] = val;
        }
    }
    __syncthreads();   // sr and stex are rewritten by the next iteration
    }
}

// The light columns alone (the light sub-fits, fit_3DMM_lambdas / fit_3DMM_Lintensity, which move
// only the light): one thread per visible pixel, rows of nl = L.K - L.off_light values (lambda,
// then Lintensity for the four appended light unknowns), the same shading quantities and FP32
// expressions as k_dI's light columns.
__global__ void __launch_bounds__(256) k_dI_light(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                   const int* __restrict__ act, int n, const float* __restrict__ theta,
                                                   const double* __restrict__ P0, const int4* __restrict__ pvert,
                                                   const float2* __restrict__ pbary, const int* __restrict__ npix,
                                                   float* __restrict__ dIl)
{
    __shared__ FrameRot sr;
    __shared__ float stex[512];
    const int nl = L.K - L.off_light;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const int N = npix[f];
        if ((int)blockIdx.x * 256 >= N) continue;
        const FrameGPU fr = frames[f];
        const float* tf = theta + (size_t)f * L.K;
        if (threadIdx.x == 0) load_frame_rot(L, fr, tf, sr);
        for (int q = threadIdx.x; q < L.ntex; q += blockDim.x) stex[q] = tf[L.off_tex + q];
        __syncthreads();
        const long long base = fr.pix_off;
        const double* P0f = P0 + (size_t)f * m.V * 3;
        for (int pos = blockIdx.x * 256 + threadIdx.x; pos < N; pos += gridDim.x * 256) {
            const int4 pv = pvert[base + pos];
            const float2 ab = pbary[base + pos];
            const float al = ab.x, be = ab.y, ga = gamma_rn(al, be);
            const int vb = pv.x, vc = pv.y, va = pv.z;
            double p0d[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                p0d[a] = (double)al * P0f[3 * vb + a] + (double)be * P0f[3 * vc + a] + (double)ga * P0f[3 * va + a];
            const double* B0 = P0f + 3 * vb;
            const double* C0 = P0f + 3 * vc;
            const double* A0 = P0f + 3 * va;
            const double e1x = B0[0] - A0[0], e1y = B0[1] - A0[1], e1z = B0[2] - A0[2];
            const double e2x = C0[0] - A0[0], e2y = C0[1] - A0[1], e2z = C0[2] - A0[2];
            double ndx = e1y * e2z - e1z * e2y, ndy = e1z * e2x - e1x * e2z, ndz = e1x * e2y - e1y * e2x;
            const double inv = 1.0 / sqrt(ndx * ndx + ndy * ndy + ndz * ndz);
            ndx *= inv; ndy *= inv; ndz *= inv;
            const double d0 = (double)sr.lam[0] - sr.tau[0], d1 = (double)sr.lam[1] - sr.tau[1],
                         d2 = (double)sr.lam[2] - sr.tau[2];
            const double lx = sr.R[0] * d0 + sr.R[1] * d1 + sr.R[2] * d2;
            const double ly = sr.R[3] * d0 + sr.R[4] * d1 + sr.R[5] * d2;
            const double lz = sr.R[6] * d0 + sr.R[7] * d1 + sr.R[8] * d2;
            const float Id = (float)(ndx * lx + ndy * ly + ndz * lz - (ndx * p0d[0] + ndy * p0d[1] + ndz * p0d[2]));
            const float nx = (float)ndx, ny = (float)ndy, nz = (float)ndz;
            const float* rsrc = m.alb ? m.alb + (size_t)f * m.V : m.tex_mean;
            float rb = rsrc[vb], rc = rsrc[vc], ra = rsrc[va];
            for (int q = 0; !m.alb && q < L.ntex; ++q) {
                rb = fmaf(m.tex_basis[(size_t)vb * m.tex_ld + q], stex[q], rb);
                rc = fmaf(m.tex_basis[(size_t)vc * m.tex_ld + q], stex[q], rc);
                ra = fmaf(m.tex_basis[(size_t)va * m.tex_ld + q], stex[q], ra);
            }
            const float rho = al * rb + be * rc + ga * ra;
            float rn[3];
            rn[0] = sr.R[0] * nx + sr.R[3] * ny + sr.R[6] * nz;
            rn[1] = sr.R[1] * nx + sr.R[4] * ny + sr.R[7] * nz;
            rn[2] = sr.R[2] * nx + sr.R[5] * ny + sr.R[8] * nz;
            float* out = dIl + (size_t)(base + pos) * nl;
            for (int c = 0; c < nl; ++c) out[c] = (c < 3) ? sr.L * rho * rn[c] : rho * Id;
        }
        __syncthreads();   // sr and stex are rewritten by the next frame
    }
}

// Per-pixel coefficients for the normalised-gradient Jacobian.
struct PixCoef {
    int4 nb;
    float c, dx, dy, gsx, gsy;
    int valid;
};

// The neighbour positions and c were computed once per pixel by k_grads (pnb, pcoef).
__device__ inline PixCoef pix_coef(const FrameGPU& fr, const int4* pnb, const float* pcoef, const float4* pixg, int pos,
                                   int N)
{
    PixCoef pc;
    pc.valid = pos < N;
    if (!pc.valid) {
        pc.nb = make_int4(0, 0, 0, 0);
        pc.c = pc.dx = pc.dy = pc.gsx = pc.gsy = 0.f;
        return pc;
    }
    const long long base = fr.pix_off;
    pc.nb = pnb[base + pos];
    const float4 gg = pixg[base + pos];
    pc.c = pcoef[base + pos];
    pc.dx = 2.f * gg.x;
    pc.dy = 2.f * gg.y;
    pc.gsx = gg.z;
    pc.gsy = gg.w;
    return pc;
}

__device__ __forceinline__ void jrow(const PixCoef& pc, const float* __restrict__ dIf, int K, int j, float& jx,
                                     float& jy)
{
    const float dh = dIf[(size_t)pc.nb.y * K + j] - dIf[(size_t)pc.nb.x * K + j];
    const float dv = dIf[(size_t)pc.nb.w * K + j] - dIf[(size_t)pc.nb.z * K + j];
    const float s = pc.dy * dh - pc.dx * dv;
    jx = pc.c * pc.dy * s;
    jy = -pc.c * pc.dx * s;
}

// ------------------------------------------------------------------ FP32 J^T J
constexpr int F32_CH = 32;     // pixels per LDS chunk -> 64 rows
constexpr int MAXKP = 128;

__device__ __forceinline__ int tri_index(int I, int J, int T) { return I * T - (I * (I - 1)) / 2 + (J - I); }

__global__ void __launch_bounds__(256) k_jtj_f32(Layout L, const FrameGPU* __restrict__ frames,
                                                  const int* __restrict__ act, int n,
                                                  const int* __restrict__ npix, const int4* __restrict__ pnb,
                                                  const float* __restrict__ pcoef,
                                                  const float4* __restrict__ pixg, const float* __restrict__ dI,
                                                  int S, int KP, float* __restrict__ Hp, float* __restrict__ gp)
{
    const int s = blockIdx.x;
    __shared__ float sJ[2 * F32_CH][MAXKP + 1];
    __shared__ float sgs[2 * F32_CH];
    __shared__ PixCoef spc[F32_CH];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
    const int f = fl_frame(act, yi);
    const FrameGPU fr = frames[f];
    const int N = npix[f];
    const int K = L.K;
    const int T = KP / 4;
    const int ntiles = T * (T + 1) / 2;
    const int p0 = (int)((long long)N * s / S), p1 = (int)((long long)N * (s + 1) / S);
    const float* dIf = dI + (size_t)fr.pix_off * K;
    const int tid = threadIdx.x;
    float acc[2][16];
    int tI[2], tJ[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int tile = tid + u * 256;
        int I = 0, rem = tile;
        while (I < T && rem >= T - I) { rem -= T - I; ++I; }
        tI[u] = I;
        tJ[u] = I + rem;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[u][e] = 0.f;
    }
    float gacc = 0.f;
    for (int c0 = p0; c0 < p1; c0 += F32_CH) {
        if (tid < F32_CH) {
            const int pos = c0 + tid;
            spc[tid] = pix_coef(fr, pnb, pcoef, pixg, pos < p1 ? pos : N, N);
            sgs[2 * tid] = spc[tid].gsx;
            sgs[2 * tid + 1] = spc[tid].gsy;
        }
        __syncthreads();
        for (int e = tid; e < F32_CH * KP; e += 256) {
            const int p = e / KP, j = e - p * KP;
            float jx = 0.f, jy = 0.f;
            if (spc[p].valid && j < K) jrow(spc[p], dIf, K, j, jx, jy);
            sJ[2 * p][j] = jx;
            sJ[2 * p + 1][j] = jy;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (tid + u * 256 >= ntiles) continue;
            const int a0 = 4 * tI[u], b0 = 4 * tJ[u];
            for (int r = 0; r < 2 * F32_CH; ++r) {
                float xa[4], xb[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    xa[q] = sJ[r][a0 + q];
                    xb[q] = sJ[r][b0 + q];
                }
#pragma unroll
                for (int qa = 0; qa < 4; ++qa)
#pragma unroll
                    for (int qb = 0; qb < 4; ++qb) acc[u][qa * 4 + qb] = fmaf(xa[qa], xb[qb], acc[u][qa * 4 + qb]);
            }
        }
        if (tid < KP)
            for (int r = 0; r < 2 * F32_CH; ++r) gacc = fmaf(sJ[r][tid], sgs[r], gacc);
        __syncthreads();
    }
    float* Hout = Hp + ((size_t)f * S + s) * ntiles * 16;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int tile = tid + u * 256;
        if (tile >= ntiles) continue;
#pragma unroll
        for (int e = 0; e < 16; ++e) Hout[(size_t)tile * 16 + e] = acc[u][e];
    }
    if (tid < KP) gp[((size_t)f * S + s) * KP + tid] = gacc;
    __syncthreads();
    }
}

// Sum the S partials in FP64 into the FP64 normal matrix; tiles are 4x4 (f32 kernel) or 16x16
// (MFMA kernel), upper triangle.
__global__ void __launch_bounds__(256) k_jtj_reduce(const int* __restrict__ act, int n, int K, int KP, int TS, int S,
                                                     const float* __restrict__ Hp, const float* __restrict__ gp,
                                                     float w, double* __restrict__ H, float* __restrict__ g,
                                                     int g_from_tiles)
{
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
    const int f = fl_frame(act, yi);
    const int T = KP / TS;
    const int ntiles = T * (T + 1) / 2;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K * K + K; e += gridDim.x * blockDim.x) {
        if (e < K * K) {
            int a = e / K, b = e - (e / K) * K;
            if (a / TS > b / TS) { const int t = a; a = b; b = t; }
            const int tile = tri_index(a / TS, b / TS, T);
            const int off = (a % TS) * TS + (b % TS);
            double sum = 0.0;
            for (int s = 0; s < S; ++s) sum += Hp[(((size_t)f * S + s) * ntiles + tile) * TS * TS + off];
            H[(size_t)f * K * K + e] = (double)w * sum;
        } else {
            const int j = e - K * K;
            double sum = 0.0;
            if (g_from_tiles) {
                // J^T r lives in column K of the augmented product: entry (j, K)
                const int tile = tri_index(j / TS, K / TS, T);
                const int off = (j % TS) * TS + (K % TS);
                for (int s = 0; s < S; ++s) sum += Hp[(((size_t)f * S + s) * ntiles + tile) * TS * TS + off];
            } else {
                for (int s = 0; s < S; ++s) sum += gp[((size_t)f * S + s) * KP + j];
            }
            g[(size_t)f * K + j] = (float)(-w * sum);
        }
    }
    }
}

// The light sub-fits' 4 x 4 normal equations from the light columns' rows ([pixel][4], k_dI_light):
// one workgroup per listed frame, each thread summing its pixels' two rows in FP64, then a
// workgroup reduction; H = w sum J^T J (FP64), g = -w sum J^T (g^s) as k_jtj_reduce writes them.
__global__ void __launch_bounds__(256) k_jtj_light(const FrameGPU* __restrict__ frames, const int* __restrict__ act,
                                                    int n, const int* __restrict__ npix, const int4* __restrict__ pnb,
                                                    const float* __restrict__ pcoef, const float4* __restrict__ pixg,
                                                    const float* __restrict__ dIl, float w, double* __restrict__ H,
                                                    float* __restrict__ g)
{
    constexpr int NL4 = 4, NS = 14;   // 10 entries of the upper triangle, 4 of the gradient
    __shared__ double sred[4][NS];
    const int nact = fl_count(act, n);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int yi = blockIdx.x; yi < nact; yi += gridDim.x) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const int N = npix[f];
        const float* dIf = dIl + (size_t)fr.pix_off * NL4;
        double acc[NS];
#pragma unroll
        for (int e = 0; e < NS; ++e) acc[e] = 0.0;
        for (int pos = tid; pos < N; pos += blockDim.x) {
            const PixCoef pc = pix_coef(fr, pnb, pcoef, pixg, pos, N);
            float jx[NL4], jy[NL4];
#pragma unroll
            for (int j = 0; j < NL4; ++j) jrow(pc, dIf, NL4, j, jx[j], jy[j]);
            int e = 0;
#pragma unroll
            for (int a = 0; a < NL4; ++a)
#pragma unroll
                for (int b = a; b < NL4; ++b) acc[e++] += (double)jx[a] * jx[b] + (double)jy[a] * jy[b];
#pragma unroll
            for (int a = 0; a < NL4; ++a) acc[10 + a] += (double)jx[a] * pc.gsx + (double)jy[a] * pc.gsy;
        }
#pragma unroll
        for (int e = 0; e < NS; ++e) {
            double v = acc[e];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if (lane == 0) sred[wave][e] = v;
        }
        __syncthreads();
        if (tid < NS) {
            const double v = (sred[0][tid] + sred[1][tid]) + (sred[2][tid] + sred[3][tid]);
            if (tid < 10) {
                int a = 0, r = tid;
                while (r >= NL4 - a) { r -= NL4 - a; ++a; }
                const int b = a + r;
                H[(size_t)f * 16 + a * NL4 + b] = (double)w * v;
                H[(size_t)f * 16 + b * NL4 + a] = (double)w * v;
            } else {
                g[(size_t)f * NL4 + (tid - 10)] = (float)(-(double)w * v);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ BF16 split MFMA J^T J
// J is split into BF16 parts: x = hi + mid + lo (three terms, each rounded to nearest even),
// and J^T J is accumulated in FP32 from the six products that matter down to 2^-24:
// hh + hm + mh + hl + lh + mm.  With nsplit == 2 only hi + mid and hh + hm + mh are used.

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// BF16 parts with the packed round-to-nearest-even conversion, two values per instruction.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float lo_f(unsigned p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float hi_f(unsigned p) { return __uint_as_float(p & 0xffff0000u); }

// ------------------------------------------------------------------ LDS-staged MFMA J^T J
// The two rows the reference forms per pixel are parallel: Jx = c dy s, Jy = -c dx s with
// s = dy (dI[r] - dI[l]) - dx (dI[b] - dI[a]) (newfuncs.cu:417-463), so
//   Jx Jx^T + Jy Jy^T = c^2 (dx^2 + dy^2) s s^T   and   Jx g^s_x + Jy g^s_y = c (g^s_x dy - g^s_y dx) s.
// One row per pixel, u = c |d| s (|d| = sqrt(dx^2 + dy^2)), gives the same normal equations
// with half the rows.  Column K of [u | r] carries r = (g^s_x dy - g^s_y dx) / |d|, so J^T r also
// comes out of the MFMA tiles; it is accumulated in FP32 beside them as well, and the caller
// picks which one feeds the gradient.
//
// The gathers of dI are organised so that they never wait on one another.  A 256-thread
// workgroup takes 32 pixels per step:
//   1. 32 threads load each pixel's neighbour positions and row weights into LDS;
//   2. all threads form u for (4-pixel quad, column) items: 16 gathered dI values per item, no
//      branches (columns j >= K read column K-1 and are replaced by selects, a pixel past the
//      range has zero weights), then write the hi/mid/lo parts to LDS in the MFMA operand
//      layout (lane l of a 16-column block holds pixels 4*(l/16)..+3 of column l%16);
//   3. each wavefront runs v_mfma_f32_16x16x16bf16_1k over its rows of upper-triangle tiles.
// Rows of tiles are dealt to the four wavefronts at compile time, largest first, each to the
// least-loaded wavefront (T = 6: 6/5/5/5 tiles).  J^T r is accumulated in FP32 per item.
constexpr int JL_PX = 32;      // pixels per step
constexpr int JL_WAVES = 4;

constexpr int jl_row_owner(int T, int I)
{
    int load[JL_WAVES] = {0, 0, 0, 0};
    for (int r = 0; r < T; ++r) {
        int best = 0;
        for (int w = 1; w < JL_WAVES; ++w)
            if (load[w] < load[best]) best = w;
        if (r == I) return best;
        load[best] += T - r;
    }
    return 0;
}
constexpr int jl_wave_tiles(int T, int W)
{
    int c = 0;
    for (int I = 0; I < T; ++I)
        if (jl_row_owner(T, I) == W) c += T - I;
    return c;
}
constexpr int jl_max_tiles(int T)
{
    int m = 0;
    for (int w = 0; w < JL_WAVES; ++w) m = jl_wave_tiles(T, w) > m ? jl_wave_tiles(T, w) : m;
    return m;
}
constexpr int jl_first_row(int T, int W)
{
    for (int I = 0; I < T; ++I)
        if (jl_row_owner(T, I) == W) return I;
    return T;
}
// slot of tile (I, J) among the tiles of its wavefront
constexpr int jl_slot(int T, int I, int J)
{
    int c = 0;
    const int w = jl_row_owner(T, I);
    for (int r = 0; r < I; ++r)
        if (jl_row_owner(T, r) == w) c += T - r;
    return c + (J - I);
}

// Tiles (I, J >= I) of row I on the wavefront that owns the row; I is a template parameter so
// the ownership test and the accumulator slots are compile-time constants.
template <int T, int W, int I>
__device__ __forceinline__ void jl_rows(const v4s* fh, const v4s* fm, const v4s* fl, v4f* acc, int nsplit)
{
    if constexpr (I < T) {
        if constexpr (jl_row_owner(T, I) == W) {
            constexpr int base = jl_slot(T, I, I);
#pragma unroll
            for (int J = I; J < T; ++J) {
                v4f& a = acc[base + J - I];
                a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fh[I], fh[J], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fh[I], fm[J], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fm[I], fh[J], a, 0, 0, 0);
                if (nsplit == 3) {
                    a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fh[I], fl[J], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fl[I], fh[J], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fm[I], fm[J], a, 0, 0, 0);
                }
            }
        }
        jl_rows<T, W, I + 1>(fh, fm, fl, acc, nsplit);
    }
}

template <int T, int W>
__device__ __forceinline__ void jl_mfma(const uint2* __restrict__ sfrag, v4f* acc, int lane, int nsplit)
{
    constexpr int B0 = jl_first_row(T, W);
    if constexpr (B0 < T) {
#pragma unroll
        for (int ks = 0; ks < JL_PX / 16; ++ks) {
            v4s fh[T], fm[T], fl[T];
#pragma unroll
            for (int b = B0; b < T; ++b) {
                const uint2* q = sfrag + ((ks * T + b) * 3) * 64 + lane;
                fh[b] = __builtin_bit_cast(v4s, q[0]);
                fm[b] = __builtin_bit_cast(v4s, q[64]);
                fl[b] = __builtin_bit_cast(v4s, q[128]);
            }
            jl_rows<T, W, B0>(fh, fm, fl, acc, nsplit);
        }
    }
}

template <int T, int W, int I>
__device__ __forceinline__ void jl_store_rows(const v4f* acc, float* __restrict__ Hout, int grp, int col)
{
    if constexpr (I < T) {
        if constexpr (jl_row_owner(T, I) == W) {
            constexpr int base = jl_slot(T, I, I);
#pragma unroll
            for (int J = I; J < T; ++J) {
                const v4f a = acc[base + J - I];
                float* o = Hout + (size_t)tri_index(I, J, T) * 256;
#pragma unroll
                for (int r = 0; r < 4; ++r) o[(4 * grp + r) * 16 + col] = a[r];
            }
        }
        jl_store_rows<T, W, I + 1>(acc, Hout, grp, col);
    }
}

template <int T, int W>
__device__ __forceinline__ void jl_store(const v4f* acc, float* __restrict__ Hout, int lane)
{
    jl_store_rows<T, W, 0>(acc, Hout, lane >> 4, lane & 15);
}

template <int T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(T >= 7 ? 2 : 4))) k_jtj_lds(Layout L, const FrameGPU* __restrict__ frames,
                                                  const int* __restrict__ act, int n, const int* __restrict__ npix,
                                                  const int4* __restrict__ pnb, const float* __restrict__ pcoef,
                                                  const float4* __restrict__ pixg, const float* __restrict__ dI, int S,
                                                  int nsplit, float* __restrict__ Hp, float* __restrict__ gp)
{
    constexpr int KP = 16 * T;
    constexpr int NT = T * (T + 1) / 2;
    constexpr int NITEM = (JL_PX / 4) * KP;
    constexpr int R = (NITEM + 255) / 256;
    constexpr int MT = jl_max_tiles(T);
    __shared__ uint2 sfrag[(JL_PX / 16) * T * 3 * 64];     // [kstep][block][part][lane]
    __shared__ float4 scf[JL_PX];                          // (c|d| dy, -c|d| dx, q dy, -q dx)
    __shared__ int4 snb[JL_PX];
    __shared__ float srk[JL_PX];
    __shared__ float sg[NITEM];                             // J^T r partial of every item
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int s = blockIdx.x;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const int N = npix[f];
        const int K = L.K;
        const int p0 = (int)((long long)N * s / S), p1 = (int)((long long)N * (s + 1) / S);
        const float* dIf = dI + (size_t)fr.pix_off * K;
        v4f acc[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = (v4f){0.f, 0.f, 0.f, 0.f};
        float gacc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) gacc[r] = 0.f;
        for (int kb = p0; kb < p1; kb += JL_PX) {
            if (tid < JL_PX) {
                const int pos = kb + tid;
                int4 nb = make_int4(0, 0, 0, 0);
                float4 cf = make_float4(0.f, 0.f, 0.f, 0.f);
                float rk = 0.f;
                if (pos < p1) {
                    const long long o = fr.pix_off + pos;
                    nb = pnb[o];
                    const float c = pcoef[o];
                    const float4 gg = pixg[o];
                    const float dx = 2.f * gg.x, dy = 2.f * gg.y;
                    const float dn = sqrtf(dx * dx + dy * dy);
                    const float wgt = c * dn;                       // row scale c |d|
                    const float e = gg.z * dy - gg.w * dx;
                    const float qg = c * e;                         // J^T r weight
                    cf = make_float4(wgt * dy, -wgt * dx, qg * dy, -qg * dx);
                    rk = dn > 0.f ? e / dn : 0.f;
                    // a NaN observation makes e NaN: 0 * NaN would carry it into every column of the
                    // tiles through the multipliers below, so the tiles' residual column takes 0 and
                    // the NaN reaches g through the FP32 J^T r (qg), which stops the frame on a NaN
                    // search direction
                    if (!(rk == rk)) rk = 0.f;
                }
                snb[tid] = nb;
                scf[tid] = cf;
                srk[tid] = rk;
            }
            __syncthreads();
            // every gather of the step first, then the arithmetic
            float ld[R][16];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int it = min(tid + 256 * r, NITEM - 1);
                const int pq = it / KP, j = it - (it / KP) * KP;
                const int jj = min(j, K - 1);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int4 nb = snb[4 * pq + q];
                    ld[r][4 * q + 0] = dIf[(size_t)nb.x * K + jj];
                    ld[r][4 * q + 1] = dIf[(size_t)nb.y * K + jj];
                    ld[r][4 * q + 2] = dIf[(size_t)nb.z * K + jj];
                    ld[r][4 * q + 3] = dIf[(size_t)nb.w * K + jj];
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int it = tid + 256 * r;
                const bool item = NITEM % 256 == 0 || it < NITEM;
                const int itc = min(it, NITEM - 1);
                const int pq = itc / KP, j = itc - (itc / KP) * KP;
                // 0/1 multipliers instead of selects keep this straight-line code (rk is finite, see
                // the pixel set-up)
                const float ma = j < K ? 1.f : 0.f, mk = j == K ? 1.f : 0.f;
                float u[4];
                float gs = 0.f;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 cf = scf[4 * pq + q];
                    const float rk = srk[4 * pq + q];
                    const float dh = ld[r][4 * q + 1] - ld[r][4 * q + 0];
                    const float dv = ld[r][4 * q + 3] - ld[r][4 * q + 2];
                    u[q] = ma * (cf.x * dh + cf.y * dv) + mk * rk;
                    gs += cf.z * dh + cf.w * dv;
                }
                gacc[r] = fmaf(item ? ma : 0.f, gs, gacc[r]);
                // hi/mid/lo parts, two rows per packed conversion
                const unsigned ph0 = pack_bf16(u[0], u[1]), ph1 = pack_bf16(u[2], u[3]);
                const float r0 = u[0] - lo_f(ph0), r1 = u[1] - hi_f(ph0);
                const float r2 = u[2] - lo_f(ph1), r3 = u[3] - hi_f(ph1);
                const unsigned pm0 = pack_bf16(r0, r1), pm1 = pack_bf16(r2, r3);
                const uint2 h = make_uint2(ph0, ph1);
                const uint2 m = make_uint2(pm0, pm1);
                const uint2 l = make_uint2(pack_bf16(r0 - lo_f(pm0), r1 - hi_f(pm0)),
                                           pack_bf16(r2 - lo_f(pm1), r3 - hi_f(pm1)));
                const int ks = pq >> 2, lp = ((pq & 3) << 4) | (j & 15), cb = j >> 4;
                uint2* dst = sfrag + ((ks * T + cb) * 3) * 64 + lp;
                if (item) {
                    dst[0] = h;
                    dst[64] = m;
                    dst[128] = l;
                }
            }
            __syncthreads();
            switch (wave) {
            case 0: jl_mfma<T, 0>(sfrag, acc, lane, nsplit); break;
            case 1: jl_mfma<T, 1>(sfrag, acc, lane, nsplit); break;
            case 2: jl_mfma<T, 2>(sfrag, acc, lane, nsplit); break;
            default: jl_mfma<T, 3>(sfrag, acc, lane, nsplit); break;
            }
            __syncthreads();   // the next step rewrites the LDS staging
        }
        float* Hout = Hp + ((size_t)f * S + s) * NT * 256;
        switch (wave) {
        case 0: jl_store<T, 0>(acc, Hout, lane); break;
        case 1: jl_store<T, 1>(acc, Hout, lane); break;
        case 2: jl_store<T, 2>(acc, Hout, lane); break;
        default: jl_store<T, 3>(acc, Hout, lane); break;
        }
        // J^T r: each item's partial in its own slot, then every column summed over its items
        // in item order (a fixed order, unlike LDS atomics)
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (NITEM % 256 == 0 || tid + 256 * r < NITEM) sg[tid + 256 * r] = gacc[r];
        __syncthreads();
        for (int j = tid; j < KP; j += 256) {
            float v = 0.f;
            for (int it = j; it < NITEM; it += KP) v += sg[it];
            gp[((size_t)f * S + s) * KP + j] = v;
        }
        __syncthreads();   // sg is rewritten for the next frame
    }
}

// ------------------------------------------------------------------ J^T J for hundreds of columns
// The identity stage's frames have K = 199 + 199 + 94 unknowns: more columns than one workgroup's
// LDS holds, and a joint system whose condition number reaches 1e10 (the albedo and identity
// components that barely move the image have eigenvalues of the barrier's size, 0.16, next to
// data eigenvalues of 3e9).  FP32 accumulation leaves errors of 1e-7 of the largest entries in
// J^T J, far above those eigenvalues, and the joint matrix is then not positive definite
// (measured on the identity-stage workload: the factorisation fails at the first iteration).
// So J^T J is formed with exact products and FP64 sums: FP32 rows, FP64 MFMA.  Two passes:
//   k_urows     the one-row form of the normalised-gradient Jacobian, u = c |d| s (see k_jtj_lds),
//               written for the first N positions of each frame to U [pixel][KP] (FP32, row
//               stride KP = K + 1 rounded up to 64); column K holds r = (g^s_x dy - g^s_y dx) / |d|
//               so that J^T r is the last column of the product, columns past K are zero;
//   k_syrk_f64  U^T U over 64 x 64 output tiles (upper triangle of tiles), a pixel split per grid
//               column: 64 pixels per step staged in LDS as FP32 (the next step's slices loaded
//               into registers meanwhile), each wavefront runs
//               v_mfma_f64_16x16x4f64 for its 16-row block against the tile's four 16-column
//               blocks (FP32 values widened exactly to FP64); FP64 partials per split, summed in
//               order by k_syrk_reduce.
constexpr int UB_T = 64;       // output tile
constexpr int UB_PX = 64;      // pixels per step
constexpr int UR_PX = 16;      // pixels per k_urows workgroup row

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_urows(int K, int KP, const FrameGPU* __restrict__ frames,
                                                const int* __restrict__ act, int n, const int* __restrict__ npix,
                                                const int4* __restrict__ pnb, const float* __restrict__ pcoef,
                                                const float4* __restrict__ pixg, const float* __restrict__ dI,
                                                float* __restrict__ U)
{
    const int ncb = KP / 256 + (KP % 256 ? 1 : 0);        // 256-column blocks
    const int cb = blockIdx.x % ncb, pb0 = blockIdx.x / ncb, npb = gridDim.x / ncb;
    const int j = cb * 256 + threadIdx.x;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const int N = npix[f];
        const float* dIf = dI + (size_t)fr.pix_off * K;
        for (int p0 = pb0 * UR_PX; p0 < N; p0 += npb * UR_PX) {
            const int pe = min(N, p0 + UR_PX);
            for (int p = p0; p < pe; ++p) {
                if (j >= KP) continue;
                const long long o = fr.pix_off + p;
                const int4 nb = pnb[o];
                const float c = pcoef[o];
                const float4 gg = pixg[o];
                const float dx = 2.f * gg.x, dy = 2.f * gg.y;
                const float dn = sqrtf(dx * dx + dy * dy);
                float u = 0.f;
                if (j < K) {
                    const float dh = dIf[(size_t)nb.y * K + j] - dIf[(size_t)nb.x * K + j];
                    const float dv = dIf[(size_t)nb.w * K + j] - dIf[(size_t)nb.z * K + j];
                    const float wgt = c * dn;
                    u = wgt * dy * dh - wgt * dx * dv;
                } else if (j == K) {
                    // a NaN observation stays in this column only: it reaches g, not H
                    u = dn > 0.f ? (gg.z * dy - gg.w * dx) / dn : 0.f;
                }
                U[(size_t)o * KP + j] = u;
            }
        }
    }
}

// tile t of the upper triangle of T x T tiles -> (I, J), I <= J (row-major order, as tri_index)
__device__ __forceinline__ void tri_tile(int t, int T, int& I, int& J)
{
    I = 0;
    while (t >= T - I) { t -= T - I; ++I; }
    J = I + t;
}

// FP64 MFMA operand layout, v_mfma_f64_16x16x4f64: A[i][k] at lane i + 16 k, B[k][j] at lane
// j + 16 k; lane l's accumulator element r holds C[l / 16 + 4 r][l % 16].
// LDS row stride of the staged slices: the 64 lanes (kq, li) of an operand read,
// [4 ks + kq][16 w + li], fall in 64 distinct banks (kq * 16 + li).
constexpr int UB_SP = UB_T + 16;

__global__ void __launch_bounds__(256) k_syrk_f64(int KP, const FrameGPU* __restrict__ frames,
                                                   const int* __restrict__ act, int n, const int* __restrict__ npix,
                                                   const float* __restrict__ U, int S, double* __restrict__ Hp)
{
    __shared__ __align__(16) float sA[UB_PX][UB_SP];       // [pixel][column]
    __shared__ __align__(16) float sB[UB_PX][UB_SP];
    const int T = KP / UB_T, NT = T * (T + 1) / 2;
    int I, J;
    tri_tile(blockIdx.x, T, I, J);
    const bool diag = I == J;
    // float4 items per thread and step: 64 pixels x 16 float4 per slice, one slice on the diagonal
    const int nq = diag ? 4 : 8;
    const int s = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, kq = lane >> 4;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.z; yi < nact; yi += gridDim.z) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const int N = npix[f];
        const int p0 = (int)((long long)N * s / S), p1 = (int)((long long)N * (s + 1) / S);
        const float* Uf = U + (size_t)fr.pix_off * KP;
        v4d acc[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = (v4d){0.0, 0.0, 0.0, 0.0};
        // the next step's slices are loaded into registers while the current step's MFMAs run:
        // item it = tid + 256 q is slice it / 1024 (I, then J), pixel (it % 1024) / 16, columns
        // 4 (it % 16) .. + 3 (16 consecutive lanes read a pixel's 64-float row segment)
        float4 pre[8];
        auto fetch = [&](int kb) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (q >= nq) continue;
                const int it = tid + 256 * q;
                const int sl = it >> 10, r = it & 1023;
                const int p = r >> 4, c4 = r & 15;
                const int pix = kb + p;
                pre[q] = pix < p1 ? *(const float4*)(Uf + (size_t)pix * KP + (sl ? J : I) * UB_T + 4 * c4)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        };
        if (p0 < p1) fetch(p0);
        for (int kb = p0; kb < p1; kb += UB_PX) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (q >= nq) continue;
                const int it = tid + 256 * q;
                const int sl = it >> 10, r = it & 1023;
                const int p = r >> 4, c4 = r & 15;
                *(float4*)&(sl ? sB : sA)[p][4 * c4] = pre[q];
            }
            __syncthreads();
            if (kb + UB_PX < p1) fetch(kb + UB_PX);
            const float (*sb)[UB_SP] = diag ? sA : sB;
#pragma unroll 4
            for (int ks = 0; ks < UB_PX / 4; ++ks) {
                const double a = (double)sA[4 * ks + kq][16 * wave + li];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double bv = (double)sb[4 * ks + kq][16 * b + li];
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[b], 0, 0, 0);
                }
            }
            __syncthreads();   // the next step rewrites the staging
        }
        double* Hout = Hp + (((size_t)f * S + s) * NT + blockIdx.x) * (UB_T * UB_T);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) Hout[(size_t)(16 * wave + kq + 4 * r) * UB_T + 16 * b + li] = acc[b][r];
    }
}

// H [f][K][K] = w sum_s of the tiles (both triangles), g [f][K] = -w sum_s of column K.
__global__ void __launch_bounds__(256) k_syrk_reduce(const int* __restrict__ act, int n, int K, int KP, int S,
                                                      const double* __restrict__ Hp, float w, double* __restrict__ H,
                                                      float* __restrict__ g)
{
    const int T = KP / UB_T, NT = T * (T + 1) / 2;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K * K + K; e += gridDim.x * blockDim.x) {
            int a, b;
            if (e < K * K) {
                a = e / K;
                b = e - a * K;
            } else {
                a = e - K * K;
                b = K;
            }
            if (a / UB_T > b / UB_T) { const int t = a; a = b; b = t; }
            const int tile = tri_index(a / UB_T, b / UB_T, T);
            const size_t off = (size_t)(a % UB_T) * UB_T + (b % UB_T);
            double sum = 0.0;
            for (int s = 0; s < S; ++s) sum += Hp[(((size_t)f * S + s) * NT + tile) * (UB_T * UB_T) + off];
            if (e < K * K) H[(size_t)f * K * K + e] = (double)w * sum;
            else g[(size_t)f * K + (e - K * K)] = (float)(-(double)w * sum);
        }
    }
}

} // namespace

size_t jtj_big_partials_doubles(int n, int S, int K)
{
    const int KP = ((K + 1 + UB_T - 1) / UB_T) * UB_T;
    const int T = KP / UB_T;
    return (size_t)n * S * (T * (T + 1) / 2) * UB_T * UB_T;
}

int jtj_big_row_stride(int K) { return ((K + 1 + UB_T - 1) / UB_T) * UB_T; }

int launch_jtj_big(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                   const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g, float* U,
                   double* partials, int S, int max_area, hipStream_t s)
{
    const int K = L.K;
    const int KP = jtj_big_row_stride(K);
    const int T = KP / UB_T;
    const int n = fl.n;
    const int gy = grid_frames(n, 1 << 16);
    const int ncb = KP / 256 + (KP % 256 ? 1 : 0);
    int npb = (max_area / 4 + UR_PX - 1) / UR_PX;     // sized for a quarter-visible crop; larger frames loop
    if (npb > 4096) npb = 4096;
    if (npb < 1) npb = 1;
    hipLaunchKernelGGL(k_urows, dim3(ncb * npb, gy), dim3(256), 0, s, K, KP, frames, fl.act, n, npix, pnb, pcoef, pixg,
                       dI, U);
    hipLaunchKernelGGL(k_syrk_f64, dim3(T * (T + 1) / 2, S, gy), dim3(256), 0, s, KP, frames, fl.act, n, npix, U, S,
                       partials);
    int bx = (K * K + K + 255) / 256;
    if (bx > 4096) bx = 4096;
    hipLaunchKernelGGL(k_syrk_reduce, dim3(bx, gy), dim3(256), 0, s, fl.act, n, K, KP, S, partials, w, H, g);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int jtj_splits(int n, int per_frame_pixels_hint)
{
    // aim for >= 32768 workgroups overall (8 per frame at 4096 frames): frame sizes vary 16x, and
    // with fewer splits the largest frames' workgroups finish last with the GPU nearly idle
    // (measured at 4096 frames: 1, 2, 4, 8, 16 splits give 16.5, 14.5, 14.0, 13.7, 14.6 ms);
    // at least ~2 chunks per workgroup, and at most 64 splits (more split the sums of a frame's
    // normal matrix into partials of a few dozen pixels and cost agreement with FP64: the
    // one-iteration energy-norm median on 32 frames of the 4096-frame workload went from 1.7e-4 at 64 to 2.7e-4 at 256)
    int S = (32768 + n - 1) / n;
    const int maxS = per_frame_pixels_hint / 128 + 1;
    if (S > maxS) S = maxS;
    if (S < 1) S = 1;
    if (S > 64) S = 64;
    return S;
}

int launch_dI(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
              const double* P0, const int4* pvert, const float2* pbary, const int* npix, const float4* pixg,
              float* dI, int mean_area, hipStream_t s)
{
    if (L.ntex > 512) {
        set_error("dI: k_tex %d > 512", L.ntex);
        return TDI_ERR_ARG;
    }
    // 8 DI_PASSES strips per frame, each shared by up to nsb workgroups; one frame per grid row
    // (sized for a quarter-visible mean crop: larger frames loop, and few workgroups of a grid
    // row find no pixels)
    int nsb = (mean_area / 4 + 8 * DI_PASSES * DI_PX_PER_WG - 1) / (8 * DI_PASSES * DI_PX_PER_WG);
    if (nsb > 64) nsb = 64;
    if (nsb < 1) nsb = 1;
    const dim3 grid(8 * nsb, grid_frames(fl.n, 1 << 16), DI_PASSES);
    if (L.nid || L.ntex)
        hipLaunchKernelGGL(k_dI<true>, grid, dim3(256), 0, s, m, L, frames, fl.act, fl.n, theta, P0, pvert, pbary, npix,
                           pixg, dI);
    else
        hipLaunchKernelGGL(k_dI<false>, grid, dim3(256), 0, s, m, L, frames, fl.act, fl.n, theta, P0, pvert, pbary, npix,
                           pixg, dI);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_dI_light(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                    const double* P0, const int4* pvert, const float2* pbary, const int* npix, float* dIl, int max_area,
                    hipStream_t s)
{
    if (L.nlight != 4 || L.K - L.off_light != 4 || L.ntex > 512) {
        set_error("dI (light columns): needs the four appended light unknowns (nlight %d, K %d, off_light %d)", L.nlight,
                  L.K, L.off_light);
        return TDI_ERR_ARG;
    }
    int gx = (max_area + 255) / 256;
    if (gx > 64) gx = 64;
    if (gx < 1) gx = 1;
    hipLaunchKernelGGL(k_dI_light, dim3(gx, grid_frames(fl.n, 1 << 16)), dim3(256), 0, s, m, L, frames, fl.act, fl.n,
                       theta, P0, pvert, pbary, npix, dIl);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_jtj_light(const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb, const float* pcoef,
                     const float4* pixg, const float* dIl, float w, double* H, float* g, hipStream_t s)
{
    hipLaunchKernelGGL(k_jtj_light, dim3(grid_frames(fl.n, 1 << 16)), dim3(256), 0, s, frames, fl.act, fl.n, npix, pnb,
                       pcoef, pixg, dIl, w, H, g);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

size_t jtj_partials_floats(int n, int S, int K, bool mfma)
{
    if (mfma) {
        const int KP = ((K + 1 + 15) / 16) * 16;
        const int T = KP / 16;
        return (size_t)n * S * (T * (T + 1) / 2) * 256 + (size_t)n * S * KP;
    }
    const int KP = ((K + 3) / 4) * 4;
    const int T = KP / 4;
    return (size_t)n * S * (T * (T + 1) / 2) * 16 + (size_t)n * S * KP;
}

int launch_jtj_f32(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                   const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g,
                   float* partials, int S, hipStream_t s)
{
    const int K = L.K;
    const int KP = ((K + 3) / 4) * 4;
    if (KP > MAXKP) {
        set_error("jtj_f32: K=%d > %d", K, MAXKP);
        return TDI_ERR_ARG;
    }
    const int T = KP / 4;
    const int n = fl.n;
    float* Hp = partials;
    float* gp = partials + (size_t)n * S * (T * (T + 1) / 2) * 16;
    const int gy = grid_frames(n, 1 << 16);
    hipLaunchKernelGGL(k_jtj_f32, dim3(S, gy), dim3(256), 0, s, L, frames, fl.act, n, npix, pnb, pcoef, pixg, dI, S,
                       KP, Hp, gp);
    int bx = (K * K + K + 255) / 256;
    hipLaunchKernelGGL(k_jtj_reduce, dim3(bx, gy), dim3(256), 0, s, fl.act, n, K, KP, 4, S, Hp, gp, w, H, g, 0);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_jtj_mfma(const Layout& L, const FrameGPU* frames, FrameList fl, const int* npix, const int4* pnb,
                    const float* pcoef, const float4* pixg, const float* dI, float w, double* H, float* g,
                    float* partials, int S, int g_on_mfma, int two_term, hipStream_t s)
{
    const int K = L.K;
    const int KP = ((K + 1 + 15) / 16) * 16;
    if (KP > 112) {
        set_error("jtj_mfma: K=%d too large (max 111)", K);
        return TDI_ERR_ARG;
    }
    const int nsplit = two_term ? 2 : 3;
    const int T = KP / 16;
    const int n = fl.n;
    float* gp = partials + (size_t)n * S * (T * (T + 1) / 2) * 256;
    const int gy = grid_frames(n, 1 << 16);
    const dim3 grid(S, gy);
    const int* act = fl.act;
    switch (T) {
    case 1: hipLaunchKernelGGL(k_jtj_lds<1>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    case 2: hipLaunchKernelGGL(k_jtj_lds<2>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    case 3: hipLaunchKernelGGL(k_jtj_lds<3>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    case 4: hipLaunchKernelGGL(k_jtj_lds<4>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    case 5: hipLaunchKernelGGL(k_jtj_lds<5>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    case 6: hipLaunchKernelGGL(k_jtj_lds<6>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    default: hipLaunchKernelGGL(k_jtj_lds<7>, grid, dim3(256), 0, s, L, frames, act, n, npix, pnb, pcoef, pixg, dI, S, nsplit, partials, gp); break;
    }
    int bx = (K * K + K + 255) / 256;
    hipLaunchKernelGGL(k_jtj_reduce, dim3(bx, gy), dim3(256), 0, s, act, n, K, KP, 16, S, partials, gp, w, H, g,
                       g_on_mfma);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
