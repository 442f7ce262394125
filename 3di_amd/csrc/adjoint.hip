// adjoint.hip — the render loss and its hand-written adjoint (vector-Jacobian product) on gfx950.
//
// Loss of a frame: over its visible pixels, |g^_r - g^_s|^2, the squared difference of the
// normalised gradients of the blurred render and of the blurred observation (the data term's
// residuals: fill_grads, funcs.cu:842-897; the fit's objective -w sum g^_r.g^_s is this sum up to
// a constant and the factor 2/w while |g^| = 1), plus sum_i |x_i - l_i|^2 over the 51 landmark
// vertices projected to image pixels.  The model-space vertices may carry per-vertex offsets.
//
// The adjoint runs backwards through the forward's stages with the visibility (triangle of each
// pixel) held fixed; the rasterisation is recomputed from theta rather than stored:
//   k_adj_dB     residuals -> dL/dg (through the normalisation g^ = 2 g / h, h = 1e-16 + 2|g|)
//                -> dL/dB scattered to the four neighbours of each visible pixel (FP64 atomics);
//   blur         dL/dI = K^T dL/dB = K dL/dB (the 15 taps are symmetric, zero outside the crop);
//   k_adj_shade  per visible pixel, dL/dI through the shading I = rho (1 + L Id + S) into the
//                triangle's model-space corners (normal, surface point, and the barycentrics of the
//                pixel centre in the projected triangle), the pose (projection and shading) and
//                the light; FP64, corners accumulated with FP64 atomics;
//   k_adj_lmk    the landmark term;
//   k_adj_basis  dL/d(expression, identity) = basis^T dL/dP0, per frame;
//   k_adj_out    gradients in the unknowns' layout (pose, expression, identity, light; albedo 0),
//                scaled by the upstream gradient of each frame's loss.
// The shading derivative matches newfuncs.cu:18-167's forward; it is the exact derivative of the
// rendered value, where the reference's analytic Jacobian (newfuncs.cu:173-463) takes the
// image-gradient (optical-flow) form of the geometric columns.
#include "common.h"

namespace tdi {

namespace {

struct AdjFrame {
    double R[9], dR[27], tau[3], dl[3], a[3], sh[5], L;
    int nsh, point;
};

__device__ void adj_frame_setup(const Layout& L, const FrameGPU& fr, const float* tf, AdjFrame& s)
{
    rodrigues_f64_d(tf + L.off_pose + 3, s.R, s.dR);
    double lam[3] = {fr.lam[0], fr.lam[1], fr.lam[2]};
    double Li = fr.Lint;
    s.point = L.nlight == 4;
    s.nsh = L.nlight == 9 ? 5 : 0;
    if (L.nlight == 4) {
        for (int a = 0; a < 3; ++a) lam[a] = tf[L.off_light + a];
        Li = tf[L.off_light + 3];
    } else if (L.nlight == 9) {
        lam[0] = lam[1] = lam[2] = 0.0;
        Li = tf[L.off_light + 3];
    }
    for (int a = 0; a < 3; ++a) {
        s.tau[a] = tf[L.off_pose + a];
        s.dl[a] = lam[a] - s.tau[a];
        s.a[a] = L.nlight == 9 ? tf[L.off_light + a] : 0.0;
    }
    for (int k = 0; k < 5; ++k) s.sh[k] = L.nlight == 9 ? tf[L.off_light + 4 + k] : 0.0;
    s.L = Li;
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Per-vertex offsets: P0 += dverts, and the projected vertices again (store_vertex's formula).
__global__ void __launch_bounds__(256) k_add_offsets(int V, const FrameGPU* __restrict__ frames, int n,
                                                      const double* __restrict__ rt, const float* __restrict__ dverts,
                                                      double* __restrict__ P0, float4* __restrict__ vproj)
{
    const int f = blockIdx.y;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n || v >= V) return;
    const double* R = rt + (size_t)f * 12;
    const size_t o = (size_t)f * V + v;
    double p[3];
    for (int a = 0; a < 3; ++a) {
        p[a] = P0[3 * o + a] + (double)dverts[3 * o + a];
        P0[3 * o + a] = p[a];
    }
    const FrameGPU fr = frames[f];
    const double X = R[0] * p[0] + R[3] * p[1] + R[6] * p[2] + R[9];
    const double Y = R[1] * p[0] + R[4] * p[1] + R[7] * p[2] + R[10];
    const double Z = R[2] * p[0] + R[5] * p[1] + R[8] * p[2] + R[11];
    vproj[o] = make_float4((float)((double)fr.phix * X / Z + (double)fr.cx), (float)((double)fr.phiy * Y / Z + (double)fr.cy),
                           (float)Z, 0.f);
}

// Residuals of the visible pixels: the data loss (FP64, per frame) and dL/dB at the neighbours.
__global__ void __launch_bounds__(256) k_adj_dB(const FrameGPU* __restrict__ frames, int n, const int* __restrict__ vis,
                                                 const int* __restrict__ npix, const float4* __restrict__ pixg,
                                                 double* __restrict__ dB, double* __restrict__ loss, int want_grad)
{
    const int f = blockIdx.y;
    if (f >= n) return;
    const FrameGPU fr = frames[f];
    const int N = npix[f];
    const long long base = fr.pix_off;
    double acc = 0.0;
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += gridDim.x * blockDim.x) {
        const float4 gg = pixg[base + pos];
        const double gx = gg.x, gy = gg.y;
        const double gn = sqrt(gx * gx + gy * gy);
        const double h = 1e-16 + sqrt(4.0 * gx * gx + 4.0 * gy * gy);
        const double rx = 2.0 * gx / h - (double)gg.z, ry = 2.0 * gy / h - (double)gg.w;
        acc += rx * rx + ry * ry;
        if (!want_grad || !(gn > 0.0)) continue;
        // dL/dg^ = 2 r; dL/dg = 2a/h - 4 (a.g) g / (|g| h^2)
        const double ax = 2.0 * rx, ay = 2.0 * ry;
        const double ag = ax * gx + ay * gy;
        const double dgx = 2.0 * ax / h - 4.0 * ag * gx / (gn * h * h);
        const double dgy = 2.0 * ay / h - 4.0 * ag * gy / (gn * h * h);
        const int k = vis[base + pos];
        const int x = k / fr.H, y = k - x * fr.H;
        const int kr = (x < fr.W - 1) ? k + fr.H : 0, kl = (x > 0) ? k - fr.H : 0;
        const int ka = (y > 0) ? k - 1 : 0, kb = (y < fr.H - 1) ? k + 1 : 0;
        atomicAdd(dB + base + kr, 0.5 * dgx);
        atomicAdd(dB + base + kl, -0.5 * dgx);
        atomicAdd(dB + base + kb, 0.5 * dgy);
        atomicAdd(dB + base + ka, -0.5 * dgy);
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss + f, acc);
}

__global__ void __launch_bounds__(256) k_d2f(const FrameGPU* __restrict__ frames, int n, const double* __restrict__ in,
                                              float* __restrict__ out)
{
    const int f = blockIdx.y;
    if (f >= n) return;
    const FrameGPU fr = frames[f];
    const int A = fr.W * fr.H;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x)
        out[fr.pix_off + k] = (float)in[fr.pix_off + k];
}

constexpr int ADJ_G = 15;   // per-frame accumulators: pose 6, light 9

// dL/dI of the visible pixels through the shading, the barycentrics and the projection.
__global__ void __launch_bounds__(256) k_adj_shade(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames, int n,
                                                    const float* __restrict__ theta, const double* __restrict__ P0,
                                                    const int4* __restrict__ pvert, const float2* __restrict__ pbary,
                                                    const int* __restrict__ npix, const float* __restrict__ gI,
                                                    double* __restrict__ dP0, double* __restrict__ gpl)
{
    __shared__ AdjFrame sf;
    __shared__ float stex[512];
    const int f = blockIdx.y;
    if (f >= n) return;
    const int N = npix[f];
    if (blockIdx.x * blockDim.x >= (unsigned)N) return;
    const FrameGPU fr = frames[f];
    const float* tf = theta + (size_t)f * L.K;
    if (threadIdx.x == 0) adj_frame_setup(L, fr, tf, sf);
    for (int q = threadIdx.x; q < L.ntex; q += blockDim.x) stex[q] = tf[L.off_tex + q];
    __syncthreads();
    const long long base = fr.pix_off;
    const double* P0f = P0 + (size_t)f * m.V * 3;
    double* dPf = dP0 + (size_t)f * m.V * 3;
    double acc[ADJ_G];
#pragma unroll
    for (int j = 0; j < ADJ_G; ++j) acc[j] = 0.0;
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += gridDim.x * blockDim.x) {
        const int4 pv = pvert[base + pos];
        const double g = gI[base + pv.w];
        if (g == 0.0) continue;
        const float2 ab = pbary[base + pos];
        const double al = ab.x, be = ab.y, ga = gamma_rn(ab.x, ab.y);
        const int vv[3] = {pv.x, pv.y, pv.z};   // corners b, c, a
        double C0[3][3], Pc[3][3], x2[3], y2[3], rhov[3];
        for (int c = 0; c < 3; ++c) {
            for (int a = 0; a < 3; ++a) C0[c][a] = P0f[3 * vv[c] + a];
            for (int a = 0; a < 3; ++a)
                Pc[c][a] = sf.R[a] * C0[c][0] + sf.R[3 + a] * C0[c][1] + sf.R[6 + a] * C0[c][2] + sf.tau[a];
            x2[c] = (double)fr.phix * Pc[c][0] / Pc[c][2] + (double)fr.cx;
            y2[c] = (double)fr.phiy * Pc[c][1] / Pc[c][2] + (double)fr.cy;
            float r = m.tex_mean[vv[c]];
            for (int q = 0; q < L.ntex; ++q) r = fmaf(m.tex_basis[(size_t)vv[c] * m.tex_ld + q], stex[q], r);
            rhov[c] = r;
        }
        // forward (shade_eval's formulas, FP64)
        const double e1[3] = {C0[0][0] - C0[2][0], C0[0][1] - C0[2][1], C0[0][2] - C0[2][2]};
        const double e2[3] = {C0[1][0] - C0[2][0], C0[1][1] - C0[2][1], C0[1][2] - C0[2][2]};
        double mm[3];
        cross3(e1, e2, mm);
        const double mn = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
        const double nv[3] = {mm[0] / mn, mm[1] / mn, mm[2] / mn};
        double p0[3], ell[3], rn[3];
        for (int a = 0; a < 3; ++a) p0[a] = al * C0[0][a] + be * C0[1][a] + ga * C0[2][a];
        for (int a = 0; a < 3; ++a) ell[a] = sf.R[3 * a] * sf.dl[0] + sf.R[3 * a + 1] * sf.dl[1] + sf.R[3 * a + 2] * sf.dl[2];
        for (int a = 0; a < 3; ++a) rn[a] = sf.R[a] * nv[0] + sf.R[3 + a] * nv[1] + sf.R[6 + a] * nv[2];
        const double Id = nv[0] * (ell[0] - p0[0]) + nv[1] * (ell[1] - p0[1]) + nv[2] * (ell[2] - p0[2]);
        const double rho = al * rhov[0] + be * rhov[1] + ga * rhov[2];
        double S = 0.0, Y[5] = {0, 0, 0, 0, 0}, dS[3] = {0, 0, 0};   // dS = dS/d(rn)
        if (sf.nsh) {
            Y[0] = rn[0] * rn[1]; Y[1] = rn[1] * rn[2]; Y[2] = rn[0] * rn[2];
            Y[3] = rn[0] * rn[0] - rn[1] * rn[1]; Y[4] = 3.0 * rn[2] * rn[2] - 1.0;
            S = rn[0] * sf.a[0] + rn[1] * sf.a[1] + rn[2] * sf.a[2];
            for (int k = 0; k < 5; ++k) S += sf.sh[k] * Y[k];
            dS[0] = sf.a[0] + sf.sh[0] * rn[1] + sf.sh[2] * rn[2] + 2.0 * sf.sh[3] * rn[0];
            dS[1] = sf.a[1] + sf.sh[0] * rn[0] + sf.sh[1] * rn[2] - 2.0 * sf.sh[3] * rn[1];
            dS[2] = sf.a[2] + sf.sh[1] * rn[1] + sf.sh[2] * rn[0] + 6.0 * sf.sh[4] * rn[2];
        }
        const double Fs = 1.0 + sf.L * Id + S;
        // backward
        const double grho = g * Fs, gF = g * rho;
        const double gId = gF * sf.L;
        // light
        if (sf.point) {
            for (int a = 0; a < 3; ++a) acc[6 + a] += gId * rn[a];   // lambda
            acc[9] += gF * Id;                                        // L
        } else if (sf.nsh) {
            for (int a = 0; a < 3; ++a) acc[6 + a] += gF * rn[a];     // a
            acc[9] += gF * Id;                                        // L
            for (int k = 0; k < 5; ++k) acc[10 + k] += gF * Y[k];     // s
        }
        // pose through the shading: Id = rn.(lambda - tau) - n.p0, S(rn), rn = R n
        double grn[3];
        for (int a = 0; a < 3; ++a) {
            acc[a] += -gId * rn[a];
            grn[a] = gId * sf.dl[a] + gF * dS[a];
        }
        for (int i = 0; i < 3; ++i) {
            const double* d = sf.dR + 9 * i;
            double dn[3];
            for (int a = 0; a < 3; ++a) dn[a] = d[a] * nv[0] + d[3 + a] * nv[1] + d[6 + a] * nv[2];
            acc[3 + i] += grn[0] * dn[0] + grn[1] * dn[1] + grn[2] * dn[2];
        }
        // model-space normal: Id's n.(ell - p0) and S's R n
        double gn[3];
        for (int a = 0; a < 3; ++a)
            gn[a] = gId * (ell[a] - p0[a]) + gF * (sf.R[3 * a] * dS[0] + sf.R[3 * a + 1] * dS[1] + sf.R[3 * a + 2] * dS[2]);
        const double ndg = nv[0] * gn[0] + nv[1] * gn[1] + nv[2] * gn[2];
        double gm[3], ge1[3], ge2[3];
        for (int a = 0; a < 3; ++a) gm[a] = (gn[a] - nv[a] * ndg) / mn;
        cross3(e2, gm, ge1);
        cross3(gm, e1, ge2);
        double gC[3][3];
        for (int a = 0; a < 3; ++a) {
            const double gp0 = -gId * nv[a];   // the surface point p0 in Id
            gC[0][a] = ge1[a] + al * gp0;
            gC[1][a] = ge2[a] + be * gp0;
            gC[2][a] = -ge1[a] - ge2[a] + ga * gp0;
        }
        // barycentrics of the pixel centre in the projected triangle (Id is constant over the
        // triangle: n is normal to its edges, so only the albedo varies with them)
        const double gal = grho * (rhov[0] - rhov[2]), gbe = grho * (rhov[1] - rhov[2]);
        const double ux = x2[0] - x2[2], uy = y2[0] - y2[2], vx = x2[1] - x2[2], vy = y2[1] - y2[2];
        const double D = ux * vy - uy * vx;
        const double wx = -(vy * gal - uy * gbe) / D, wy = -(-vx * gal + ux * gbe) / D;
        const double bw[3] = {al, be, ga};
        for (int c = 0; c < 3; ++c) {
            const double g2x = bw[c] * wx, g2y = bw[c] * wy;
            const double iz = 1.0 / Pc[c][2];
            const double gP[3] = {(double)fr.phix * g2x * iz, (double)fr.phiy * g2y * iz,
                                  -((double)fr.phix * Pc[c][0] * g2x + (double)fr.phiy * Pc[c][1] * g2y) * iz * iz};
            for (int a = 0; a < 3; ++a) acc[a] += gP[a];
            for (int i = 0; i < 3; ++i) {
                const double* d = sf.dR + 9 * i;
                double dv[3];
                for (int a = 0; a < 3; ++a) dv[a] = d[a] * C0[c][0] + d[3 + a] * C0[c][1] + d[6 + a] * C0[c][2];
                acc[3 + i] += gP[0] * dv[0] + gP[1] * dv[1] + gP[2] * dv[2];
            }
            for (int a = 0; a < 3; ++a) gC[c][a] += sf.R[3 * a] * gP[0] + sf.R[3 * a + 1] * gP[1] + sf.R[3 * a + 2] * gP[2];
        }
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) atomicAdd(dPf + 3 * vv[c] + a, gC[c][a]);
    }
    for (int j = 0; j < ADJ_G; ++j) {
        const double s = wave_sum(acc[j]);
        if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(gpl + (size_t)f * ADJ_G + j, s);
    }
}

// The landmark term: one wavefront per frame.
__global__ void __launch_bounds__(64) k_adj_lmk(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames, int n,
                                                 const float* __restrict__ theta, const double* __restrict__ P0,
                                                 const float* __restrict__ lmk, double* __restrict__ dP0,
                                                 double* __restrict__ gpl, double* __restrict__ loss, int want_grad)
{
    __shared__ double sR[9], sdR[27];
    const int f = blockIdx.x, i = threadIdx.x;
    if (f >= n) return;
    const float* tf = theta + (size_t)f * L.K;
    if (i == 0) rodrigues_f64_d(tf + L.off_pose + 3, sR, sdR);
    __syncthreads();
    const FrameGPU fr = frames[f];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < NLMK) {
        const int v = m.lmk_ids[i];
        const double* p = P0 + ((size_t)f * m.V + v) * 3;
        double P[3];
        for (int a = 0; a < 3; ++a) P[a] = sR[a] * p[0] + sR[3 + a] * p[1] + sR[6 + a] * p[2] + (double)tf[L.off_pose + a];
        const float* L6 = lmk + (size_t)f * 6 * NLMK;
        const double dx = (double)fr.phix * P[0] / P[2] + (double)fr.cx - (double)L6[i];
        const double dy = (double)fr.phiy * P[1] / P[2] + (double)fr.cy - (double)L6[NLMK + i];
        acc[6] = dx * dx + dy * dy;
        if (want_grad) {
            const double iz = 1.0 / P[2];
            const double gP[3] = {(double)fr.phix * 2.0 * dx * iz, (double)fr.phiy * 2.0 * dy * iz,
                                  -((double)fr.phix * P[0] * 2.0 * dx + (double)fr.phiy * P[1] * 2.0 * dy) * iz * iz};
            for (int a = 0; a < 3; ++a) acc[a] = gP[a];
            for (int k = 0; k < 3; ++k) {
                const double* d = sdR + 9 * k;
                double dv[3];
                for (int a = 0; a < 3; ++a) dv[a] = d[a] * p[0] + d[3 + a] * p[1] + d[6 + a] * p[2];
                acc[3 + k] = gP[0] * dv[0] + gP[1] * dv[1] + gP[2] * dv[2];
            }
            double* dp = dP0 + ((size_t)f * m.V + v) * 3;
            for (int a = 0; a < 3; ++a)
                atomicAdd(dp + a, sR[3 * a] * gP[0] + sR[3 * a + 1] * gP[1] + sR[3 * a + 2] * gP[2]);
        }
    }
    for (int j = 0; j < 7; ++j) acc[j] = wave_sum(acc[j]);
    if (i == 0) {
        atomicAdd(loss + f, acc[6]);
        if (want_grad)
            for (int j = 0; j < 6; ++j) atomicAdd(gpl + (size_t)f * ADJ_G + j, acc[j]);
    }
}

// out[f][k] += sum_r basis[r][k] dP0[f][r] over the 3V rows, k < nk; rows split over gridDim.x.
constexpr int BAS_COLS = 128;

__global__ void __launch_bounds__(256) k_adj_basis(int V, const float* __restrict__ basis, int ld, int nk, int n,
                                                    const double* __restrict__ dP0, double* __restrict__ out)
{
    __shared__ double red[4][BAS_COLS];
    const int f = blockIdx.y;
    if (f >= n) return;
    const int rows = 3 * V;
    const int r0 = (int)((long long)rows * blockIdx.x / gridDim.x), r1 = (int)((long long)rows * (blockIdx.x + 1) / gridDim.x);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double* d = dP0 + (size_t)f * rows;
    for (int c0 = 0; c0 < nk; c0 += BAS_COLS) {
        double s0 = 0.0, s1 = 0.0;
        const int k0 = c0 + lane, k1 = c0 + 64 + lane;
        for (int r = r0 + wave; r < r1; r += 4) {
            const double dv = d[r];
            if (dv == 0.0) continue;
            const float* b = basis + (size_t)r * ld;
            if (k0 < nk) s0 = fma((double)b[k0], dv, s0);
            if (k1 < nk) s1 = fma((double)b[k1], dv, s1);
        }
        red[wave][lane] = s0;
        red[wave][64 + lane] = s1;
        __syncthreads();
        if (wave == 0) {
            for (int h = 0; h < 2; ++h) {
                const int k = c0 + 64 * h + lane;
                if (k < nk) {
                    const double s = red[0][64 * h + lane] + red[1][64 * h + lane] + red[2][64 * h + lane] +
                                     red[3][64 * h + lane];
                    if (s != 0.0) atomicAdd(out + (size_t)f * nk + k, s);
                }
            }
        }
        __syncthreads();
    }
}

// Gradients in the layout of the unknowns (scaled by the upstream gradient of each frame's loss),
// the per-vertex gradient in FP32 and the loss.
__global__ void __launch_bounds__(256) k_adj_out(Layout L, int V, int n, const double* __restrict__ gloss,
                                                  const double* __restrict__ gpl, const double* __restrict__ gexp,
                                                  const double* __restrict__ gid, const double* __restrict__ dP0,
                                                  float* __restrict__ gtheta, float* __restrict__ gverts)
{
    const int f = blockIdx.y;
    if (f >= n) return;
    const double s = gloss ? gloss[f] : 1.0;
    if (gtheta && blockIdx.x == 0) {
        for (int j = threadIdx.x; j < L.K; j += blockDim.x) {
            double v = 0.0;
            if (j >= L.off_pose && j < L.off_pose + 6) v = gpl[(size_t)f * ADJ_G + (j - L.off_pose)];
            else if (j >= L.off_light && j < L.off_light + L.nlight) v = gpl[(size_t)f * ADJ_G + 6 + (j - L.off_light)];
            else if (j >= L.off_exp && j < L.off_exp + L.nexp) v = gexp[(size_t)f * L.nexp + (j - L.off_exp)];
            else if (j >= L.off_id && j < L.off_id + L.nid) v = gid[(size_t)f * L.nid + (j - L.off_id)];
            gtheta[(size_t)f * L.K + j] = (float)(s * v);
        }
    }
    if (gverts)
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < 3 * V; e += gridDim.x * blockDim.x)
            gverts[(size_t)f * 3 * V + e] = (float)(s * dP0[(size_t)f * 3 * V + e]);
}

} // namespace

int launch_add_offsets(int V, const FrameGPU* frames, int n, const double* rt, const float* dverts, double* P0,
                       float4* vproj, hipStream_t s)
{
    if (n <= 0) return TDI_OK;
    hipLaunchKernelGGL(k_add_offsets, dim3((V + 255) / 256, n), dim3(256), 0, s, V, frames, n, rt, dverts, P0, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_adjoint(const ModelGPU& m, const Layout& L, const FrameGPU* frames, int n, const float* theta,
                   const double* P0, const int4* pvert, const float2* pbary, const int* vis, const int* npix,
                   const float4* pixg, const float* lmk, const float* taps, int max_area, float* img, float* tmp,
                   float* blurred, const AdjScratch& a, const double* gloss, double* loss, float* gtheta,
                   float* gverts, hipStream_t s)
{
    if (n <= 0) return TDI_OK;
    const bool grad = gtheta || gverts;
    const int pb = std::max(1, std::min(64, (max_area + 255) / 256));
    TDI_HIP(hipMemsetAsync(a.loss, 0, sizeof(double) * n, s));
    if (grad) {
        TDI_HIP(hipMemsetAsync(a.dB, 0, sizeof(double) * a.cap, s));
        TDI_HIP(hipMemsetAsync(a.dP0, 0, sizeof(double) * (size_t)n * m.V * 3, s));
        TDI_HIP(hipMemsetAsync(a.gpl, 0, sizeof(double) * (size_t)n * ADJ_G, s));
        if (L.nexp) TDI_HIP(hipMemsetAsync(a.gexp, 0, sizeof(double) * (size_t)n * L.nexp, s));
        if (L.nid) TDI_HIP(hipMemsetAsync(a.gid, 0, sizeof(double) * (size_t)n * L.nid, s));
    }
    hipLaunchKernelGGL(k_adj_dB, dim3(pb, n), dim3(256), 0, s, frames, n, vis, npix, pixg, a.dB, a.loss, (int)grad);
    TDI_HIP(hipGetLastError());
    if (grad) {
        hipLaunchKernelGGL(k_d2f, dim3(pb, n), dim3(256), 0, s, frames, n, a.dB, img);
        TDI_HIP(hipGetLastError());
        int rc = launch_blur(frames, all_frames(n), taps, nullptr, img, tmp, blurred, max_area, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_adj_shade, dim3(pb, n), dim3(256), 0, s, m, L, frames, n, theta, P0, pvert, pbary, npix,
                           blurred, a.dP0, a.gpl);
        TDI_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_adj_lmk, dim3(n), dim3(64), 0, s, m, L, frames, n, theta, P0, lmk, a.dP0, a.gpl, a.loss,
                       (int)grad);
    TDI_HIP(hipGetLastError());
    if (grad) {
        const int rb = 64;
        if (L.nexp) {
            hipLaunchKernelGGL(k_adj_basis, dim3(rb, n), dim3(256), 0, s, m.V, m.exp_basis, m.exp_ld, L.nexp, n, a.dP0,
                               a.gexp);
            TDI_HIP(hipGetLastError());
        }
        if (L.nid) {
            hipLaunchKernelGGL(k_adj_basis, dim3(rb, n), dim3(256), 0, s, m.V, m.id_basis, m.id_ld, L.nid, n, a.dP0,
                               a.gid);
            TDI_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_adj_out, dim3(std::max(1, std::min(256, (3 * m.V + 255) / 256)), n), dim3(256), 0, s, L,
                           m.V, n, gloss, a.gpl, a.gexp, a.gid, a.dP0, gtheta, gverts);
        TDI_HIP(hipGetLastError());
    }
    if (loss) TDI_HIP(hipMemcpyAsync(loss, a.loss, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return TDI_OK;
}

} // namespace tdi
