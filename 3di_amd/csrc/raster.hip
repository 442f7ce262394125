// raster.hip — z-buffered rasterisation and visible-pixel compaction on gfx950.
//
// Reference: get_pixels_to_render (funcs.cu:915-1010) tests every integer sample point in
// a triangle's bounding box with an FP32 barycentric test, keeps up to NTMP pixels per
// triangle, and does atomicMinFloat on the depth of the triangle's first vertex;
// keep_only_minZ (funcs.cu:1016-1029) and thrust::remove_if (renderer.cu:665) then
// compact the surviving slots, fill_ks_and_M0 (funcs.cu:81-106) builds the occupancy map
// and DerivativeComputer sorts/scans it (derivative_computer.cu:43-72).
//
// Here:
//   k_raster  — thread per triangle, bounding box clipped to the frame's crop, exact
//               reference coverage arithmetic (cover_rn, no FMA contraction), and a single
//               64-bit atomicMin of (orderable depth << 32 | triangle id) per covered pixel.
//               The winner per pixel is therefore unique and deterministic: minimum depth,
//               then lowest triangle id (the reference keeps every exact-depth tie and lets
//               them race in the scatter, funcs.cu:816,1027).
//   k_resolve — one workgroup per frame walks the crop in column-major order, 4 pixels per
//               lane; three wavefront ballots (one per bit of the lane's 0..4 count) and
//               popcounts give each occupied pixel its rank, which is at once the
//               inclusive occupancy prefix sum (the reference's cumM0), the position in the
//               sorted visible-pixel list (the reference's sort_by_key output) and the slot
//               of its barycentric weights.  No sort is needed because the crop is walked in
//               pixel-index order.
#include "common.h"

namespace tdi {

namespace {

__global__ void k_fill_u64(uint64_t* __restrict__ p, long long n, uint64_t v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void __launch_bounds__(256) k_raster(const int* __restrict__ tris, int NT, int V,
                                                 const FrameGPU* __restrict__ frames, const int* __restrict__ mask,
                                                 const float4* __restrict__ vproj, int ntmp,
                                                 unsigned long long* __restrict__ keys)
{
    const int f = blockIdx.y;
    if (mask && !mask[f]) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NT) return;
    const FrameGPU fr = frames[f];
    const float4* vp = vproj + (size_t)f * V;
    const int ib = tris[3 * t], ic = tris[3 * t + 1], ia = tris[3 * t + 2];
    const float4 B = vp[ib], Cc = vp[ic], A = vp[ia];
    const float xmin = floorf(fminf(fminf(B.x, Cc.x), A.x));
    const float xmax = ceilf(fmaxf(fmaxf(B.x, Cc.x), A.x));
    const float ymin = floorf(fminf(fminf(B.y, Cc.y), A.y));
    const float ymax = ceilf(fmaxf(fmaxf(B.y, Cc.y), A.y));
    if (!(xmin == xmin) || !(xmax == xmax) || !(ymin == ymin) || !(ymax == ymax)) return;
    // clip the scan to the crop window [x0, x0+W) x [y0, y0+H)
    const float xlo = fmaxf(xmin, (float)fr.x0), xhi = fminf(xmax, (float)(fr.x0 + fr.W));
    const float ylo = fmaxf(ymin, (float)fr.y0), yhi = fminf(ymax, (float)(fr.y0 + fr.H));
    if (!(xlo < xhi) || !(ylo < yhi)) return;
    const unsigned long long key = ((unsigned long long)orderable_bits(B.z) << 32) | (unsigned)t;
    unsigned long long* kf = keys + fr.pix_off;
    int covered = 0;
    for (int x = (int)xlo; (float)x < xhi; ++x) {
        for (int y = (int)ylo; (float)y < yhi; ++y) {
            float al, be;
            if (!cover_rn(B.x, B.y, Cc.x, Cc.y, A.x, A.y, x, y, &al, &be)) continue;
            if (ntmp > 0 && covered >= ntmp) continue;
            ++covered;
            atomicMin(kf + (size_t)(x - fr.x0) * fr.H + (y - fr.y0), key);
        }
    }
}

constexpr int RES_THREADS = 1024;
constexpr int RES_PER_THREAD = 4;
constexpr int RES_TILE = RES_THREADS * RES_PER_THREAD;

__global__ void __launch_bounds__(RES_THREADS) k_resolve(const int* __restrict__ tris, int V,
                                                          const FrameGPU* __restrict__ frames,
                                                          const int* __restrict__ mask,
                                                          const float4* __restrict__ vproj,
                                                          const unsigned long long* __restrict__ keys,
                                                          int* __restrict__ cum, int* __restrict__ vis,
                                                          int* __restrict__ npix, float2* __restrict__ bary)
{
    const int f = blockIdx.x;
    if (mask && !mask[f]) return;
    const FrameGPU fr = frames[f];
    const float4* vp = vproj + (size_t)f * V;
    const long long base = fr.pix_off;
    const int A = fr.W * fr.H;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = RES_THREADS / 64;
    __shared__ int swave[NW];
    int running = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int tile = 0; tile < A; tile += RES_TILE) {
        const int k0 = tile + RES_PER_THREAD * tid;     // this thread's 4 consecutive pixels
        unsigned long long key[RES_PER_THREAD];
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < RES_PER_THREAD; ++q) {
            key[q] = (k0 + q < A) ? keys[base + k0 + q] : KEY_EMPTY;
            cnt += key[q] != KEY_EMPTY;
        }
        // wave prefix of the per-thread counts (0..4) from three ballots, one per count bit
        const unsigned long long b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4);
        const int before = __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt);
        if (lane == 0) swave[wave] = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
        __syncthreads();
        int woff = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            woff += (w < wave) ? swave[w] : 0;
            total += swave[w];
        }
        int pos = running + woff + before;
#pragma unroll
        for (int q = 0; q < RES_PER_THREAD; ++q) {
            const int k = k0 + q;
            if (k >= A) break;
            const bool occ = key[q] != KEY_EMPTY;
            cum[base + k] = pos + (occ ? 1 : 0);
            if (occ) {
                vis[base + pos] = k;
                const int t = (int)(unsigned)(key[q] & 0xffffffffull);
                const float4 B = vp[tris[3 * t]], Cc = vp[tris[3 * t + 1]], Av = vp[tris[3 * t + 2]];
                float al, be;
                cover_rn(B.x, B.y, Cc.x, Cc.y, Av.x, Av.y, fr.x0 + k / fr.H, fr.y0 + k % fr.H, &al, &be);
                bary[base + k] = make_float2(al, be);
                ++pos;
            }
        }
        running += total;
        __syncthreads();
    }
    if (tid == 0) npix[f] = running;
}

} // namespace

int launch_clear_keys(uint64_t* keys, long long total, hipStream_t s)
{
    int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, keys, total, KEY_EMPTY);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

namespace {
__global__ void __launch_bounds__(256) k_clear_keys_masked(const FrameGPU* __restrict__ frames,
                                                            const int* __restrict__ mask, uint64_t* __restrict__ keys)
{
    const int f = blockIdx.y;
    if (mask && !mask[f]) return;
    const FrameGPU fr = frames[f];
    const int A = fr.W * fr.H;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) keys[fr.pix_off + k] = KEY_EMPTY;
}
} // namespace

int launch_clear_keys_masked(const FrameGPU* frames, int n, const int* mask, uint64_t* keys, int max_area, hipStream_t s)
{
    int bx = (max_area + 255) / 256;
    if (bx > 1024) bx = 1024;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_clear_keys_masked, dim3(bx, n), dim3(256), 0, s, frames, mask, keys);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_raster(const ModelGPU& m, const FrameGPU* frames, int n, const int* mask, const float4* vproj, int ntmp,
                  uint64_t* keys, hipStream_t s)
{
    dim3 grid((m.NT + 255) / 256, n);
    hipLaunchKernelGGL(k_raster, grid, dim3(256), 0, s, m.tris, m.NT, m.V, frames, mask, vproj, ntmp,
                       (unsigned long long*)keys);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_resolve(const ModelGPU& m, const FrameGPU* frames, int n, const int* mask, const float4* vproj,
                   const uint64_t* keys, int* cum, int* vis_list, int* npix, float* bary, int /*max_area*/,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_resolve, dim3(n), dim3(RES_THREADS), 0, s, m.tris, m.V, frames, mask, vproj,
                       (const unsigned long long*)keys, cum, vis_list, npix, (float2*)bary);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
