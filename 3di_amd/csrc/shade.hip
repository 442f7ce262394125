// shade.hip — shading, blur, normalised image gradients and the data objective on gfx950.
//
// Reference:
//   fill_diffuse_component_and_populate_texture_and_shape (newfuncs.cu:18-167):
//       n = normalise((P0_b - P0_a) x (P0_c - P0_a)), p0 = a P0_b + b P0_c + g P0_a,
//       Id = n.(R^T (lambda - tau)) - n.p0, I = rho (1 + L Id), rho barycentric albedo;
//       with nine light unknowns (a, L, s) the factor is 1 + (R n).a + L Id + sum_k s_k Y_k(R n)
//       with lambda = 0 in Id (sh2_terms);
//   fill_tex_im1 (funcs.cu:800-817): scatter I into the crop (zero elsewhere);
//   convolutionRowsGPU/ColumnsGPU (newfuncs.cu:539-715): separable 15-tap Gaussian,
//       zero boundary (here both passes in one LDS-tiled kernel);
//   fill_grads (funcs.cu:842-897): g = (I[x+1]-I[x-1], I[y+1]-I[y-1]) / 2 with crop-border
//       neighbours reading crop index 0, h = 1e-16 + sqrt(4 gx^2 + 4 gy^2), g^ = 2 g / h,
//       the same for the blurred observation;
//   objective SGEMM (renderer.cu:725-729): C = w sum g^_r . g^_s (the fit minimises -C).
#include "common.h"

namespace tdi {

struct Taps {
    float t[15];
};

// Workgroups per frame of the objective kernels: each writes its partial sum into its own
// slot of a [frames][OBJ_PARTS] buffer, reduced in order by k_obj_reduce.
constexpr int OBJ_PARTS = 256;

namespace {

// Region of a crop a pass has to cover, half-open [X0, X1) x [Y0, Y1).  Without a box (or
// when the visible pixels touch the crop border, whose neighbour rule reads crop index 0) it
// is the whole crop.  With the visible box (xl, xh, yl, yh) of the frame: the normalised
// gradients read the blurred image at visible pixels +-1, so the blurred output covers the
// box + 1.  The rendered image is written on the box only (cleared where no triangle is); the
// blur, which reaches 7 pixels along x and y, reads zero outside it.
struct Region {
    int X0, X1, Y0, Y1;
};

__device__ __forceinline__ bool crop_region(const FrameGPU& fr, const int4* box, int f, int mx, int my, Region& r)
{
    r = Region{0, fr.W, 0, fr.H};
    if (!box) return true;
    const int4 b = box[f];
    if (b.y < 0) return false;                  // nothing visible
    if (b.x == 0 || b.y == fr.W - 1 || b.z == 0 || b.w == fr.H - 1) return true;
    r.X0 = max(0, b.x - mx);
    r.X1 = min(fr.W, b.y + mx + 1);
    r.Y0 = max(0, b.z - my);
    r.Y1 = min(fr.H, b.w + my + 1);
    return true;
}

// Column kernels run on a 1-D grid of nb workgroups.  The workgroups per frame follow the
// length of the frame list, read on the device: 8 per frame when the list is long, up to 128
// (one column per wavefront) when it is short, so that the last line-search rounds, whose
// lists hold few frames, spread each frame over more wavefronts.  Slots of fpb consecutive
// workgroups walk the list.
constexpr int COL_MIN = 8, COL_MAX = 128;

struct ColMap {
    int fpb, nslots, slot, sub;
};

__device__ __forceinline__ int col_fpb(int nb, int nact)
{
    const int fpb = nact > 0 ? nb / nact : COL_MAX;
    return fpb < COL_MIN ? COL_MIN : (fpb > COL_MAX ? COL_MAX : fpb);
}

__device__ __forceinline__ ColMap col_map(int nact)
{
    ColMap c;
    c.fpb = col_fpb(gridDim.x, nact);
    c.nslots = gridDim.x / c.fpb;
    c.slot = blockIdx.x / c.fpb;
    c.sub = blockIdx.x - c.slot * c.fpb;
    return c;
}

__global__ void __launch_bounds__(256) k_clear_crops(const FrameGPU* __restrict__ frames, const int* __restrict__ act,
                                                      int n, const int4* __restrict__ box, float* __restrict__ img)
{
    const int nact = fl_count(act, n);
    const ColMap cm = col_map(nact);
    if (cm.slot >= cm.nslots) return;
    for (int yi = cm.slot; yi < nact; yi += cm.nslots) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        Region r;
        if (!crop_region(fr, box, f, 0, 0, r)) continue;
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
        for (int x = r.X0 + cm.sub * nw + wave; x < r.X1; x += cm.fpb * nw)
            for (int y = r.Y0 + lane; y < r.Y1; y += 64) img[fr.pix_off + (long long)x * fr.H + y] = 0.f;
    }
}

// Per-frame shading constants: R (column-major), lambda - tau, L; with nine light unknowns
// (nsh = 5) lambda = 0, la = R^T a (the directional term in model space) and the second-order
// terms s.
struct ShadeFrame {
    float R[9], dl[3], L;
    float la[3], sh[5];
    int nsh;
};

__device__ __forceinline__ void shade_frame_setup(const Layout& L, const FrameGPU& fr, const float* tf, ShadeFrame& sf)
{
    rodrigues_f32(tf + L.off_pose + 3, sf.R, nullptr);
    float lam[3] = {fr.lam[0], fr.lam[1], fr.lam[2]};
    float Li = fr.Lint;
    sf.nsh = L.nlight == 9 ? 5 : 0;
    if (L.nlight == 4) {
        lam[0] = tf[L.off_light]; lam[1] = tf[L.off_light + 1]; lam[2] = tf[L.off_light + 2];
        Li = tf[L.off_light + 3];
    } else if (sf.nsh) {
        lam[0] = lam[1] = lam[2] = 0.f;
        Li = tf[L.off_light + 3];
    }
    for (int a = 0; a < 3; ++a) sf.dl[a] = lam[a] - tf[L.off_pose + a];
    sf.L = Li;
    for (int k = 0; k < 5; ++k) sf.sh[k] = sf.nsh ? tf[L.off_light + 4 + k] : 0.f;
    for (int r = 0; r < 3; ++r)
        sf.la[r] = sf.nsh ? sf.R[3 * r] * tf[L.off_light] + sf.R[3 * r + 1] * tf[L.off_light + 1]
                                + sf.R[3 * r + 2] * tf[L.off_light + 2]
                          : 0.f;
}

// Per-pixel inputs of the shading: the triangle's model-space corners (FP64) and corner albedos.
struct ShadeIn {
    double b[3], c[3], a[3];
    float rb, rc, ra;
};

__device__ __forceinline__ void shade_load(const ModelGPU& m, const double* __restrict__ P0f,
                                           const float* __restrict__ albf, int vb, int vc, int va, ShadeIn& in)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        in.b[i] = P0f[3 * vb + i];
        in.c[i] = P0f[3 * vc + i];
        in.a[i] = P0f[3 * va + i];
    }
    const float* rsrc = albf ? albf : m.tex_mean;
    in.rb = rsrc[vb];
    in.rc = rsrc[vc];
    in.ra = rsrc[va];
}

// I = rho (1 + L Id) at barycentric (al, be) of the triangle (vb, vc, va).  The geometry (normal,
// surface point, Id) and the shading factor are FP64 from the FP64 corners, so the rendered
// value is the FP64 value of the same formula rounded once (the weights al, be, ga are the
// rasteriser's FP32 values, as in the reference).
__device__ __forceinline__ float shade_eval(const ModelGPU& m, const Layout& L, const ShadeFrame& sf,
                                            const float* stex, int vb, int vc, int va, const ShadeIn& in,
                                            float al, float be)
{
    const float gaf = gamma_rn(al, be);
    const double ald = al, bed = be, gad = gaf;
    const double* B0 = in.b;
    const double* C0 = in.c;
    const double* A0 = in.a;
    const double e1x = B0[0] - A0[0], e1y = B0[1] - A0[1], e1z = B0[2] - A0[2];
    const double e2x = C0[0] - A0[0], e2y = C0[1] - A0[1], e2z = C0[2] - A0[2];
    double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= inv; ny *= inv; nz *= inv;
    const double px = ald * B0[0] + bed * C0[0] + gad * A0[0];
    const double py = ald * B0[1] + bed * C0[1] + gad * A0[1];
    const double pz = ald * B0[2] + bed * C0[2] + gad * A0[2];
    // R^T (lambda - tau)
    const double lx = (double)sf.R[0] * sf.dl[0] + (double)sf.R[1] * sf.dl[1] + (double)sf.R[2] * sf.dl[2];
    const double ly = (double)sf.R[3] * sf.dl[0] + (double)sf.R[4] * sf.dl[1] + (double)sf.R[5] * sf.dl[2];
    const double lz = (double)sf.R[6] * sf.dl[0] + (double)sf.R[7] * sf.dl[1] + (double)sf.R[8] * sf.dl[2];
    const double Id = nx * lx + ny * ly + nz * lz - (nx * px + ny * py + nz * pz);
    float rb = in.rb, rc = in.rc, ra = in.ra;
    for (int q = 0; !m.alb && q < L.ntex; ++q) {
        rb = fmaf(m.tex_basis[(size_t)vb * m.tex_ld + q], stex[q], rb);
        rc = fmaf(m.tex_basis[(size_t)vc * m.tex_ld + q], stex[q], rc);
        ra = fmaf(m.tex_basis[(size_t)va * m.tex_ld + q], stex[q], ra);
    }
    const double rho = ald * rb + bed * rc + gad * ra;
    if (!sf.nsh) return (float)(rho + sf.L * rho * Id);
    const double nc[3] = {sf.R[0] * nx + sf.R[3] * ny + sf.R[6] * nz, sf.R[1] * nx + sf.R[4] * ny + sf.R[7] * nz,
                          sf.R[2] * nx + sf.R[5] * ny + sf.R[8] * nz};
    const double Y[5] = {nc[0] * nc[1], nc[1] * nc[2], nc[0] * nc[2], nc[0] * nc[0] - nc[1] * nc[1],
                         3.0 * nc[2] * nc[2] - 1.0};
    double S = nx * sf.la[0] + ny * sf.la[1] + nz * sf.la[2];
#pragma unroll
    for (int k = 0; k < 5; ++k) S += sf.sh[k] * Y[k];
    return (float)(rho + sf.L * rho * Id + rho * S);
}

__device__ __forceinline__ float shade_pixel(const ModelGPU& m, const Layout& L, const double* __restrict__ P0f,
                                             const float* __restrict__ albf, const ShadeFrame& sf, const float* stex,
                                             int vb, int vc, int va, float al, float be)
{
    ShadeIn in;
    shade_load(m, P0f, albf, vb, vc, va, in);
    return shade_eval(m, L, sf, stex, vb, vc, va, in, al, be);
}

// Shading of the visible list (pvert/pbary from the resolve pass), scattered into the crop.
__global__ void __launch_bounds__(256) k_shade(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                const int* __restrict__ act, int n, const float* __restrict__ theta,
                                                const double* __restrict__ P0, const int4* __restrict__ pvert,
                                                const float2* __restrict__ pbary, const int* __restrict__ npix,
                                                float* __restrict__ img)
{
    __shared__ ShadeFrame sf;
    __shared__ float stex[512];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const int N = npix[f];
        if (blockIdx.x * blockDim.x >= (unsigned)N) continue;
        const float* tf = theta + (size_t)f * L.K;
        const FrameGPU fr = frames[f];
        if (threadIdx.x == 0) shade_frame_setup(L, fr, tf, sf);
        for (int k = threadIdx.x; k < L.ntex; k += blockDim.x) stex[k] = tf[L.off_tex + k];
        __syncthreads();
        const long long base = fr.pix_off;
        const double* P0f = P0 + (size_t)f * m.V * 3;
        const float* albf = m.alb ? m.alb + (size_t)f * m.V : nullptr;
        for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += gridDim.x * blockDim.x) {
            const int4 pv = pvert[base + pos];
            const float2 ab = pbary[base + pos];
            img[base + pv.w] = shade_pixel(m, L, P0f, albf, sf, stex, pv.x, pv.y, pv.z, ab.x, ab.y);
        }
        __syncthreads();   // the frame's shared constants are rewritten by the next iteration
    }
}

// Shading straight from the depth|triangle keys over the frame's visible box:
// the barycentrics are recomputed with the coverage test the rasteriser used, so the image is
// the one the visible-list path renders; pixels without a triangle are cleared.  Used by the
// line search, which needs the objective only.
__global__ void __launch_bounds__(256) k_shade_keys(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                     const int* __restrict__ act, int n, const float* __restrict__ theta,
                                                     const double* __restrict__ P0, const float4* __restrict__ vproj,
                                                     const unsigned long long* __restrict__ keys,
                                                     const int4* __restrict__ box, float* __restrict__ img)
{
    __shared__ ShadeFrame sf;
    __shared__ float stex[512];
    const int nact = fl_count(act, n);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const ColMap cm = col_map(nact);
    if (cm.slot >= cm.nslots) return;
    for (int yi = cm.slot; yi < nact; yi += cm.nslots) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        Region r;
        if (!crop_region(fr, box, f, 0, 0, r)) continue;
        if (r.X0 + (int)cm.sub * nw >= r.X1) continue;
        const float* tf = theta + (size_t)f * L.K;
        if (threadIdx.x == 0) shade_frame_setup(L, fr, tf, sf);
        for (int k = threadIdx.x; k < L.ntex; k += blockDim.x) stex[k] = tf[L.off_tex + k];
        __syncthreads();
        const long long base = fr.pix_off;
        const double* P0f = P0 + (size_t)f * m.V * 3;
        const float* albf = m.alb ? m.alb + (size_t)f * m.V : nullptr;
        const float4* vp = vproj + (size_t)f * m.V;
        for (int x = r.X0 + cm.sub * nw + wave; x < r.X1; x += cm.fpb * nw)
            for (int y0 = r.Y0 + lane; y0 < r.Y1; y0 += 128) {
                // two rows per lane: the gathers of both (keys, triangle corners, then vertices
                // and albedo) are issued before the arithmetic, each under its pixel's coverage
                const int y1 = y0 + 64;
                const bool h1 = y1 < r.Y1;
                const int k0 = x * fr.H + y0, k1 = x * fr.H + y1;
                const unsigned long long key0 = keys[base + k0];
                const unsigned long long key1 = h1 ? keys[base + k1] : KEY_EMPTY;
                const bool c0 = key0 != KEY_EMPTY, c1 = key1 != KEY_EMPTY;
                int v0[3] = {0, 0, 0}, v1[3] = {0, 0, 0};
                if (c0) {
                    const int t = (int)(unsigned)(key0 & 0xffffffffull);
                    v0[0] = m.tris[3 * t]; v0[1] = m.tris[3 * t + 1]; v0[2] = m.tris[3 * t + 2];
                }
                if (c1) {
                    const int t = (int)(unsigned)(key1 & 0xffffffffull);
                    v1[0] = m.tris[3 * t]; v1[1] = m.tris[3 * t + 1]; v1[2] = m.tris[3 * t + 2];
                }
                float4 B0, C0, A0, B1, C1, A1;
                ShadeIn in0, in1;
                if (c0) {
                    B0 = vp[v0[0]]; C0 = vp[v0[1]]; A0 = vp[v0[2]];
                    shade_load(m, P0f, albf, v0[0], v0[1], v0[2], in0);
                }
                if (c1) {
                    B1 = vp[v1[0]]; C1 = vp[v1[1]]; A1 = vp[v1[2]];
                    shade_load(m, P0f, albf, v1[0], v1[1], v1[2], in1);
                }
                float s0 = 0.f, s1 = 0.f;
                if (c0) {
                    float al, be;
                    cover_rn(B0.x, B0.y, C0.x, C0.y, A0.x, A0.y, fr.x0 + x, fr.y0 + y0, &al, &be);
                    s0 = shade_eval(m, L, sf, stex, v0[0], v0[1], v0[2], in0, al, be);
                }
                if (c1) {
                    float al, be;
                    cover_rn(B1.x, B1.y, C1.x, C1.y, A1.x, A1.y, fr.x0 + x, fr.y0 + y1, &al, &be);
                    s1 = shade_eval(m, L, sf, stex, v1[0], v1[1], v1[2], in1, al, be);
                }
                img[base + k0] = s0;
                if (h1) img[base + k1] = s1;
            }
        __syncthreads();   // the frame's shared constants are rewritten by the next iteration
    }
}

__device__ __forceinline__ void nbr_index(int k, int W, int H, int& kr, int& kl, int& ka, int& kb)
{
    const int x = k / H, y = k - x * H;
    kr = (x < W - 1) ? k + H : 0;
    kl = (x > 0) ? k - H : 0;
    ka = (y > 0) ? k - 1 : 0;
    kb = (y < H - 1) ? k + 1 : 0;
}

// Both passes of the blur in one kernel, over tiles of the output region (the visible box + 1,
// or the whole crop): a 32-column x 64-row output tile loads its input with a 7-pixel halo
// into LDS (zero outside the crop), blurs along y into an LDS intermediate over the tile's
// 46 columns, then along x.  The taps are applied in the order of the two-pass kernel
// (j = -7..7, FMA) of separate row and column passes.
constexpr int BT_X = 32, BT_Y = 64;
constexpr int BT_IX = BT_X + 14, BT_IY = BT_Y + 14;

__global__ void __launch_bounds__(256) k_blur_tiles(const FrameGPU* __restrict__ frames, const int* __restrict__ act,
                                                     int n, const int4* __restrict__ box, Taps taps,
                                                     const float* __restrict__ in, float* __restrict__ out)
{
    __shared__ float simg[BT_IX][BT_IY + 1];
    __shared__ float stmp[BT_IX][BT_Y + 1];
    const int nact = fl_count(act, n);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const ColMap cm = col_map(nact);
    if (cm.slot >= cm.nslots) return;
    for (int yi = cm.slot; yi < nact; yi += cm.nslots) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        Region rg, rs;
        if (!crop_region(fr, box, f, 1, 1, rg)) continue;
        // the shaded region: the rendered image is written there only and is zero around it
        crop_region(fr, box, f, 0, 0, rs);
        const int W = fr.W, H = fr.H;
        const int ntx = (rg.X1 - rg.X0 + BT_X - 1) / BT_X, nty = (rg.Y1 - rg.Y0 + BT_Y - 1) / BT_Y;
        const float* src = in + fr.pix_off;
        float* dst = out + fr.pix_off;
        for (int t = cm.sub; t < ntx * nty; t += cm.fpb) {
            const int x0 = rg.X0 + (t % ntx) * BT_X, y0 = rg.Y0 + (t / ntx) * BT_Y;
            // input tile: columns x0-7 .. x0+BT_X+6, rows y0-7 .. y0+BT_Y+6, zero outside the crop
            for (int c = wave; c < BT_IX; c += 4) {
                const int x = x0 - 7 + c;
                for (int r = lane; r < BT_IY; r += 64) {
                    const int y = y0 - 7 + r;
                    simg[c][r] = (x >= rs.X0 && x < rs.X1 && y >= rs.Y0 && y < rs.Y1) ? src[(size_t)x * H + y] : 0.f;
                }
            }
            __syncthreads();
            // along y, for the rows of the tile over all its input columns (zero outside the crop)
            for (int c = wave; c < BT_IX; c += 4) {
                const int x = x0 - 7 + c;
                float s = 0.f;
                if (x >= 0 && x < W) {
#pragma unroll
                    for (int j = -7; j <= 7; ++j) s = fmaf(taps.t[7 - j], simg[c][lane + 7 + j], s);
                }
                stmp[c][lane] = s;
            }
            __syncthreads();
            // along x
            for (int c = wave; c < BT_X; c += 4) {
                const int x = x0 + c, y = y0 + lane;
                if (x < rg.X1 && y < rg.Y1) {
                    float s = 0.f;
#pragma unroll
                    for (int j = -7; j <= 7; ++j) s = fmaf(taps.t[7 - j], stmp[c + 7 + j][lane], s);
                    dst[(size_t)x * H + y] = s;
                }
            }
            __syncthreads();   // simg/stmp are rewritten by the next tile
        }
    }
}

// Objective term g^_r . g^_s of one rendered pixel, with the gradient (gx, gy) and h.
__device__ __forceinline__ double pixel_term(const float* __restrict__ I, int kr, int kl, int ka, int kb, float2 gs,
                                             float& gx, float& gy, float& h)
{
    gx = 0.5f * (I[kr] - I[kl]);
    gy = 0.5f * (I[kb] - I[ka]);
    h = 1e-16f + sqrtf(4.f * gx * gx + 4.f * gy * gy);
    return (double)(2.f * gx / h) * gs.x + (double)(2.f * gy / h) * gs.y;
}

// Gradients of the blurred rendered crop at every visible pixel, the objective
// -w sum g^_r.g^_s accumulated in FP64, and the per-pixel records (gx, gy, g^s_x, g^s_y),
// neighbour positions and Jacobian coefficient used by the normal equations.
__global__ void __launch_bounds__(256) k_grads(const FrameGPU* __restrict__ frames, const int* __restrict__ act, int n,
                                                const int* __restrict__ vis, const int* __restrict__ npix,
                                                const int* __restrict__ cum,
                                                const float* __restrict__ blurred, const float2* __restrict__ obs_g,
                                                float w, float4* __restrict__ pixg, int4* __restrict__ pnb,
                                                float* __restrict__ pcoef, double* __restrict__ objp)
{
    __shared__ double sw[4];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
    const int f = fl_frame(act, yi);
    const FrameGPU fr = frames[f];
    const int N = npix[f];
    const long long base = fr.pix_off;
    const float* I = blurred + base;
    double acc = 0.0;
    for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += gridDim.x * blockDim.x) {
        const int k = vis[base + pos];
        int kr, kl, ka, kb;
        nbr_index(k, fr.W, fr.H, kr, kl, ka, kb);
        float gx, gy, h;
        const float2 gs = obs_g[fr.obs_off + k];
        acc += pixel_term(I, kr, kl, ka, kb, gs, gx, gy, h);
        pixg[base + pos] = make_float4(gx, gy, gs.x, gs.y);
        // per-pixel record for the normalised-gradient Jacobian: compacted neighbour
        // positions and c = 1 / (0.001 + h^3) (newfuncs.cu:446)
        pnb[base + pos] = pixel_nbrs(cum, k, base, pos, N, fr.W, fr.H);
        pcoef[base + pos] = 1.f / (0.001f + h * h * h);
    }
    // block reduction (wave shuffles, then LDS)
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sw[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) objp[(size_t)f * OBJ_PARTS + blockIdx.x] = -(double)w * (sw[0] + sw[1] + sw[2] + sw[3]);
    __syncthreads();   // sw is rewritten by the next iteration
    }
}

// Observation: blur once, then normalised gradients at every crop pixel.
// The data objective -w sum g^_r . g^_s alone, over the pixels of the visible box whose key
// holds a triangle (the pixels of the visible list), without the per-pixel records.
__global__ void __launch_bounds__(256) k_objective(const FrameGPU* __restrict__ frames, const int* __restrict__ act,
                                                    int n, const unsigned long long* __restrict__ keys,
                                                    const int4* __restrict__ box, const float* __restrict__ blurred,
                                                    const float2* __restrict__ obs_g, float w,
                                                    double* __restrict__ objp)
{
    __shared__ double sw[4];
    const int nact = fl_count(act, n);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const ColMap cm = col_map(nact);
    if (cm.slot >= cm.nslots) return;
    for (int yi = cm.slot; yi < nact; yi += cm.nslots) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        Region r;
        if (!crop_region(fr, box, f, 0, 0, r) || r.X0 + (int)cm.sub * nw >= r.X1) {
            if (threadIdx.x == 0) objp[(size_t)f * OBJ_PARTS + cm.sub] = 0.0;
            continue;
        }
        const long long base = fr.pix_off;
        const float* I = blurred + base;
        double acc = 0.0;
        for (int x = r.X0 + cm.sub * nw + wave; x < r.X1; x += cm.fpb * nw)
            for (int y = r.Y0 + lane; y < r.Y1; y += 64) {
                const int k = x * fr.H + y;
                if (keys[base + k] == KEY_EMPTY) continue;
                // neighbour indices as nbr_index gives them, from the loop's x and y
                const int kr = (x < fr.W - 1) ? k + fr.H : 0, kl = (x > 0) ? k - fr.H : 0;
                const int ka = (y > 0) ? k - 1 : 0, kb = (y < fr.H - 1) ? k + 1 : 0;
                float gx, gy, h;
                acc += pixel_term(I, kr, kl, ka, kb, obs_g[fr.obs_off + k], gx, gy, h);
            }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (lane == 0) sw[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) objp[(size_t)f * OBJ_PARTS + cm.sub] = -(double)w * (sw[0] + sw[1] + sw[2] + sw[3]);
        __syncthreads();   // sw is rewritten by the next iteration
    }
}

// obj[f] += the workgroup partials of the data objective, summed in workgroup order, so that
// the objective of a frame does not depend on the order in which workgroups finish.
// nparts: partials per frame; col_nb > 0: a column kernel of col_nb workgroups wrote them (as
// many per frame as col_map gave it).
__global__ void k_obj_reduce(const int* __restrict__ act, int n, int nparts, int col_nb,
                             const double* __restrict__ objp, double* __restrict__ obj)
{
    const int nact = fl_count(act, n);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nact) return;
    const int f = fl_frame(act, i);
    const int np = col_nb > 0 ? col_fpb(col_nb, nact) : nparts;
    double s = 0.0;
    for (int b = 0; b < np; ++b) s += objp[(size_t)f * OBJ_PARTS + b];
    obj[f] += s;
}

// obj[f] = objb[f] for the listed frames: the objective restarted from the barrier value that the
// last barrier evaluation left (for evaluations that change only the light).
__global__ void k_obj_from_barrier(const int* __restrict__ act, int n, const double* __restrict__ objb,
                                   double* __restrict__ obj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fl_count(act, n)) return;
    const int f = fl_frame(act, i);
    obj[f] = objb[f];
}

__global__ void __launch_bounds__(256) k_obs_grad(const FrameGPU* __restrict__ frames, int n, const float* __restrict__ blurred,
                                                   float2* __restrict__ obs_g)
{
    for (int f = blockIdx.y; f < n; f += gridDim.y) {
        const FrameGPU fr = frames[f];
        const int A = fr.W * fr.H;
        const float* S = blurred + fr.pix_off;
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
            int kr, kl, ka, kb;
            nbr_index(k, fr.W, fr.H, kr, kl, ka, kb);
            const float sx = 0.5f * (S[kr] - S[kl]);
            const float sy = 0.5f * (S[kb] - S[ka]);
            const float hs = 1e-16f + sqrtf(4.f * sx * sx + 4.f * sy * sy);
            obs_g[fr.obs_off + k] = make_float2(2.f * sx / hs, 2.f * sy / hs);
        }
    }
}

// Column kernels: workgroups of four wavefronts, one column per wavefront at a time; 8
// workgroups per slot of the list capacity for large batches, up to 128 for few frames (the
// per-frame split is then chosen on the device, see col_map).
inline dim3 col_grid(int n)
{
    int bx = n > 0 ? 4096 / n : COL_MAX;
    bx = bx < COL_MIN ? COL_MIN : (bx > COL_MAX ? COL_MAX : bx);
    return dim3(bx * (n > 0 ? n : 1));
}

// Visible-list kernels: workgroups of 256 threads over a frame's visible list (the ones past
// the frame's visible count leave at once), 16 per frame for large batches and up to one per
// 256 pixels of half the largest crop when there are few frames; one frame per grid row.
inline dim3 list_grid(int max_area, int n)
{
    int bx = (max_area / 2 + 255) / 256;
    const int cap = n > 0 ? std::min(OBJ_PARTS, std::max(16, 8192 / n)) : OBJ_PARTS;
    if (bx > cap) bx = cap;
    if (bx < 1) bx = 1;
    return dim3(bx, grid_frames(n, 1 << 16));
}

inline dim3 crop_grid(int max_area, int n)
{
    int bx = (max_area + 255) / 256;
    if (bx > 256) bx = 256;
    if (bx < 1) bx = 1;
    return dim3(bx, grid_frames(n, 64));
}

} // namespace

int launch_shade(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                 const double* P0, const int4* pvert, const float2* pbary, const int* npix, const int4* box,
                 float* img, int max_area, hipStream_t s)
{
    if (L.ntex > 512) {
        set_error("shade: k_tex %d > 512", L.ntex);
        return TDI_ERR_ARG;
    }
    hipLaunchKernelGGL(k_clear_crops, col_grid(fl.n), dim3(256), 0, s, frames, fl.act, fl.n, box, img);
    hipLaunchKernelGGL(k_shade, list_grid(max_area, fl.n), dim3(256), 0, s, m, L, frames, fl.act, fl.n, theta, P0,
                       pvert, pbary, npix, img);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_shade_keys(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                      const double* P0, const float4* vproj, const uint64_t* keys, const int4* box, float* img,
                      hipStream_t s)
{
    if (L.ntex > 512) {
        set_error("shade: k_tex %d > 512", L.ntex);
        return TDI_ERR_ARG;
    }
    hipLaunchKernelGGL(k_shade_keys, col_grid(fl.n), dim3(256), 0, s, m, L, frames, fl.act, fl.n, theta, P0, vproj,
                       (const unsigned long long*)keys, box, img);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_objective(const FrameGPU* frames, FrameList fl, const uint64_t* keys, const int4* box,
                     const float* blurred, const float2* obs_g, float w, double* objp, double* obj, hipStream_t s)
{
    const dim3 grid = col_grid(fl.n);
    hipLaunchKernelGGL(k_objective, grid, dim3(256), 0, s, frames, fl.act, fl.n,
                       (const unsigned long long*)keys, box, blurred, obs_g, w, objp);
    hipLaunchKernelGGL(k_obj_reduce, dim3((fl.n + 255) / 256), dim3(256), 0, s, fl.act, fl.n, 0, (int)grid.x, objp, obj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_obj_from_barrier(FrameList fl, const double* objb, double* obj, hipStream_t s)
{
    hipLaunchKernelGGL(k_obj_from_barrier, dim3((fl.n + 255) / 256), dim3(256), 0, s, fl.act, fl.n, objb, obj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_blur(const FrameGPU* frames, FrameList fl, const float* taps, const int4* box, const float* in, float* tmp,
                float* out, int max_area, hipStream_t s)
{
    Taps t;
    for (int i = 0; i < 15; ++i) t.t[i] = taps[i];
    (void)tmp;   // the fused kernel keeps the intermediate in LDS
    hipLaunchKernelGGL(k_blur_tiles, col_grid(fl.n), dim3(256), 0, s, frames, fl.act, fl.n, box, t, in, out);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_grads(const FrameGPU* frames, FrameList fl, const int* vis_list, const int* npix,
                 const int* cum, const float* blurred, const float2* obs_g, float w, float4* pixg, int4* pnb,
                 float* pcoef, double* objp, double* obj, int max_area, hipStream_t s)
{
    const dim3 grid = list_grid(max_area, fl.n);
    hipLaunchKernelGGL(k_grads, grid, dim3(256), 0, s, frames, fl.act, fl.n, vis_list, npix, cum,
                       blurred, obs_g, w, pixg, pnb, pcoef, objp);
    hipLaunchKernelGGL(k_obj_reduce, dim3((fl.n + 255) / 256), dim3(256), 0, s, fl.act, fl.n, (int)grid.x, 0, objp, obj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_obs_prepare(const FrameGPU* frames, int n, const float* taps, const float* obs, float* tmp, float* blurred,
                       float2* obs_g, int max_area, hipStream_t s)
{
    int rc = launch_blur(frames, all_frames(n), taps, nullptr, obs, tmp, blurred, max_area, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_obs_grad, crop_grid(max_area, n), dim3(256), 0, s, frames, n, blurred, obs_g);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
