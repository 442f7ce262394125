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
//   k_raster  — thread per triangle, bounding box clipped to the frame's crop, the reference's
//               coverage decision (cover_rn's FP32 sequence, no FMA contraction; the divisions
//               only for samples next to an edge, raster_frame), and a single 64-bit
//               atomicMin of (orderable depth << 32 | triangle id) per covered pixel.
//               The winner per pixel is therefore unique and deterministic: minimum depth,
//               then lowest triangle id (the reference keeps every exact-depth tie and lets
//               them race in the scatter, funcs.cu:816,1027).
//   k_resolve_count/k_resolve_write — the crop in column-major order, 1024-pixel tiles, 4
//               pixels per lane; three wavefront ballots (one per bit of the lane's 0..4
//               count) and popcounts give each occupied pixel its rank, which is at once the
//               inclusive occupancy prefix sum (the reference's cumM0), the position in the
//               sorted visible-pixel list (the reference's sort_by_key output) and the slot
//               of its barycentric weights.  No sort is needed because pixels are ranked in
//               pixel-index order; tile offsets come from a first counting pass.
#include "common.h"

namespace tdi {

namespace {

__global__ void k_fill_u64(uint64_t* __restrict__ p, long long n, uint64_t v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Coverage of the sample points of one triangle in one frame.  The decision is the one
// cover_rn makes: alpha > 0, beta > 0 (the sign pre-test, exact) and RN(RN(na/den) +
// RN(nb/den)) <= 1.  The sum is first formed from na * rcp(den) and nb * rcp(den): v_rcp_f32
// is within 1 ulp, so that sum is within 1.5 * 2^-22 * (alpha + beta) of the correctly rounded
// one, and only samples whose approximate sum lies within 2^-18 of 1 (or with den outside
// [2^-60, 2^60]) take the two IEEE divisions.  The samples are scanned x-major: sample i of the
// scan box is (xl + i / ny, yl + i % ny).
struct RasterTri {
    float Ax, Ay, v0x, v0y, v1x, v1y, d00, d01, d11, den, q;
    int xl, yl, ny;
    bool den_ok, dpos, dneg;
};

// The scan box and the edge quantities of a triangle; returns the number of samples to test
// (0: none can be covered).
__device__ __forceinline__ int raster_setup(const FrameGPU& fr, float4 B, float4 Cc, float4 A, RasterTri& r)
{
#pragma clang fp contract(off)
    const float xmin = floorf(fminf(fminf(B.x, Cc.x), A.x));
    const float xmax = ceilf(fmaxf(fmaxf(B.x, Cc.x), A.x));
    const float ymin = floorf(fminf(fminf(B.y, Cc.y), A.y));
    const float ymax = ceilf(fmaxf(fmaxf(B.y, Cc.y), A.y));
    if (!(xmin == xmin) || !(xmax == xmax) || !(ymin == ymin) || !(ymax == ymax)) return 0;
    // clip the scan to the crop window [x0, x0+W) x [y0, y0+H); the clipped bounds are integers
    float xlo = fmaxf(xmin, (float)fr.x0), xhi = fminf(xmax, (float)(fr.x0 + fr.W));
    float ylo = fmaxf(ymin, (float)fr.y0), yhi = fminf(ymax, (float)(fr.y0 + fr.H));
    if (!(xlo < xhi) || !(ylo < yhi)) return 0;
    r.Ax = A.x;
    r.Ay = A.y;
    r.v0x = B.x - A.x; r.v1x = Cc.x - A.x;
    r.v0y = B.y - A.y; r.v1y = Cc.y - A.y;
    r.d00 = r.v0x * r.v0x + r.v0y * r.v0y;
    r.d01 = r.v0x * r.v1x + r.v0y * r.v1y;
    r.d11 = r.v1x * r.v1x + r.v1y * r.v1y;
    r.den = r.d00 * r.d11 - r.d01 * r.d01;
    {
        // The [floor, ceil) box holds samples outside the triangle's own extent, which only a
        // near-degenerate triangle can cover through rounding (brute force over millions of
        // triangles of up to 4 px: covered samples outside the vertex box occur for
        // den < 1e-6 d00 d11 only; tests/test_raster_candidates.py keeps that check).  A small,
        // well-shaped triangle therefore scans only the integers within its vertex box widened
        // by 1e-3 px; the decisions for every sample are unchanged.
        const float exn = fmaxf(fmaxf(B.x, Cc.x), A.x) - fminf(fminf(B.x, Cc.x), A.x);
        const float eyn = fmaxf(fmaxf(B.y, Cc.y), A.y) - fminf(fminf(B.y, Cc.y), A.y);
        if (exn <= 4.f && eyn <= 4.f && r.d00 > 0.f && r.d11 > 0.f && r.den >= 1e-4f * (r.d00 * r.d11)) {
            xlo = fmaxf(xlo, ceilf(fminf(fminf(B.x, Cc.x), A.x) - 1e-3f));
            xhi = fminf(xhi, floorf(fmaxf(fmaxf(B.x, Cc.x), A.x) + 1e-3f) + 1.f);
            ylo = fmaxf(ylo, ceilf(fminf(fminf(B.y, Cc.y), A.y) - 1e-3f));
            yhi = fminf(yhi, floorf(fmaxf(fmaxf(B.y, Cc.y), A.y) + 1e-3f) + 1.f);
            if (!(xlo < xhi) || !(ylo < yhi)) return 0;
        }
    }
    r.xl = (int)xlo;
    r.yl = (int)ylo;
    r.ny = (int)yhi - r.yl;
    r.q = __builtin_amdgcn_rcpf(r.den);
    r.den_ok = fabsf(r.den) >= 0x1p-60f && fabsf(r.den) <= 0x1p60f;
    r.dpos = r.den > 0.f;
    r.dneg = r.den < 0.f;
    return ((int)xhi - r.xl) * r.ny;
}

__device__ __forceinline__ bool raster_covers(const RasterTri& r, int x, int y)
{
#pragma clang fp contract(off)
    const float v2x = (float)x - r.Ax, v2y = (float)y - r.Ay;
    const float d20 = r.v0x * v2x + r.v0y * v2y;
    const float d21 = r.v1x * v2x + r.v1y * v2y;
    const float na = r.d11 * d20 - r.d01 * d21;
    const float nb = r.d00 * d21 - r.d01 * d20;
    const bool pa = (na > 0.f && r.dpos) || (na < 0.f && r.dneg);
    const bool pb = (nb > 0.f && r.dpos) || (nb < 0.f && r.dneg);
    const float s1 = na * r.q + nb * r.q;
    bool cov = pa && pb && s1 < 1.f;
    if (pa && pb && !(r.den_ok && fabsf(s1 - 1.f) > 0x1p-18f)) {
        const float a = na / r.den, b = nb / r.den;
        cov = a > 0.f && b > 0.f && (a + b <= 1.f);
    }
    return cov;
}

// Every sample of the scan box in order, at most ntmp covered ones (ntmp <= 0: all).
__device__ __forceinline__ void raster_loop(const FrameGPU& fr, const RasterTri& r, int nxy, unsigned long long key,
                                            int ntmp, unsigned long long* __restrict__ kf)
{
    int covered = 0;
    int x = r.xl, y = r.yl;
    int off = (r.xl - fr.x0) * fr.H + (r.yl - fr.y0);
    for (int i = 0; i < nxy; ++i) {
        if (raster_covers(r, x, y) && (ntmp <= 0 || covered < ntmp)) {
            ++covered;
            atomicMin(kf + off, key);
        }
        ++y;
        ++off;
        if (y - r.yl == r.ny) {
            y = r.yl;
            ++x;
            off += fr.H - r.ny;
        }
    }
}

__device__ __forceinline__ unsigned long long tri_key(float z, int t)
{
    return ((unsigned long long)orderable_bits(z) << 32) | (unsigned)t;
}

__device__ __forceinline__ void raster_frame(const FrameGPU& fr, float4 B, float4 Cc, float4 A, int t, int ntmp,
                                             unsigned long long* __restrict__ keys)
{
    RasterTri r;
    const int nxy = raster_setup(fr, B, Cc, A, r);
    if (nxy > 0) raster_loop(fr, r, nxy, tri_key(B.z, t), ntmp, keys + fr.pix_off);
}

__global__ void __launch_bounds__(256) k_raster(const int* __restrict__ tris, int NT, int V,
                                                 const FrameGPU* __restrict__ frames, const int* __restrict__ act,
                                                 int n, const float4* __restrict__ vproj, int ntmp,
                                                 unsigned long long* __restrict__ keys, int4* __restrict__ box)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NT) return;
    const int ib = tris[3 * t], ic = tris[3 * t + 1], ia = tris[3 * t + 2];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        // empty box for the visible-box or resolve pass that follows (it grows it with atomics)
        if (box && t == 0) box[f] = make_int4(fr.W, -1, fr.H, -1);
        const float4* vp = vproj + (size_t)f * V;
        raster_frame(fr, vp[ib], vp[ic], vp[ia], t, ntmp, keys);
    }
}

// Resolve in two passes over 1024-pixel tiles so that every tile of every listed frame is a
// workgroup of its own (a frame's crop is not walked by one workgroup in sequence):
//   k_resolve_count: occupied pixels per tile -> tcount[f][tile];
//   k_resolve_write: tile offset = sum of the preceding tiles' counts, then the ballot
//                    ranks inside the tile give cum, the sorted visible list and the slot
//                    of the barycentrics; the last tile writes the frame's pixel count.
constexpr int RES_THREADS = 256;
constexpr int RES_PER_THREAD = 4;
constexpr int RES_TILE = RES_THREADS * RES_PER_THREAD;
constexpr int RES_NW = RES_THREADS / 64;

// per-thread count (0..4) of this thread's four pixels and their keys
__device__ __forceinline__ int tile_keys(const unsigned long long* __restrict__ kf, int k0, int A,
                                         unsigned long long* key)
{
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < RES_PER_THREAD; ++q) {
        key[q] = (k0 + q < A) ? kf[k0 + q] : KEY_EMPTY;
        cnt += key[q] != KEY_EMPTY;
    }
    return cnt;
}

__global__ void __launch_bounds__(RES_THREADS) k_resolve_count(const FrameGPU* __restrict__ frames,
                                                                const int* __restrict__ act, int n,
                                                                const unsigned long long* __restrict__ keys,
                                                                int* __restrict__ tcount, int tmax,
                                                                int4* __restrict__ box)
{
    __shared__ int swave[RES_NW];
    __shared__ int4 sbox[RES_NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const int A = fr.W * fr.H;
        const int ntile = (A + RES_TILE - 1) / RES_TILE;
        for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
            unsigned long long key[RES_PER_THREAD];
            const int k0 = tile * RES_TILE + RES_PER_THREAD * tid;
            const int cnt = tile_keys(keys + fr.pix_off, k0, A, key);
            const unsigned long long b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4);
            if (lane == 0) swave[wave] = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
            if (box) {
                // bounding box of the visible pixels (x = k / H, y = k % H), per wavefront
                int xl = fr.W, xh = -1, yl = fr.H, yh = -1;
                if (b0 | b1 | b2) {
#pragma unroll
                    for (int q = 0; q < RES_PER_THREAD; ++q) {
                        if (key[q] == KEY_EMPTY) continue;
                        const int k = k0 + q, x = k / fr.H, y = k - x * fr.H;
                        xl = min(xl, x); xh = max(xh, x); yl = min(yl, y); yh = max(yh, y);
                    }
                    for (int o = 32; o > 0; o >>= 1) {
                        xl = min(xl, __shfl_xor(xl, o, 64)); xh = max(xh, __shfl_xor(xh, o, 64));
                        yl = min(yl, __shfl_xor(yl, o, 64)); yh = max(yh, __shfl_xor(yh, o, 64));
                    }
                }
                if (lane == 0) sbox[wave] = make_int4(xl, xh, yl, yh);
            }
            __syncthreads();
            if (tid == 0) {
                int t = 0;
#pragma unroll
                for (int w = 0; w < RES_NW; ++w) t += swave[w];
                tcount[(size_t)f * tmax + tile] = t;
                if (box && t > 0) {
                    // one workgroup's box; the atomics only where it extends the box read so far
                    // (the box only grows, so a value read before another workgroup's update
                    // can only cause a redundant atomic, never a skipped one)
                    int4 b = sbox[0];
#pragma unroll
                    for (int w = 1; w < RES_NW; ++w) {
                        const int4 c = sbox[w];
                        b.x = min(b.x, c.x); b.y = max(b.y, c.y); b.z = min(b.z, c.z); b.w = max(b.w, c.w);
                    }
                    const int4 cur = box[f];
                    if (b.x < cur.x) atomicMin(&box[f].x, b.x);
                    if (b.y > cur.y) atomicMax(&box[f].y, b.y);
                    if (b.z < cur.z) atomicMin(&box[f].z, b.z);
                    if (b.w > cur.w) atomicMax(&box[f].w, b.w);
                }
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(RES_THREADS) k_resolve_write(const int* __restrict__ tris, int V,
                                                                const FrameGPU* __restrict__ frames,
                                                                const int* __restrict__ act, int n,
                                                                const float4* __restrict__ vproj,
                                                                const unsigned long long* __restrict__ keys,
                                                                const int* __restrict__ tcount, int tmax,
                                                                const int4* __restrict__ box,
                                                                int* __restrict__ cum, int* __restrict__ vis,
                                                                int* __restrict__ npix, float2* __restrict__ bary,
                                                                int4* __restrict__ pvert, float2* __restrict__ pbary)
{
    __shared__ int swave[RES_NW];
    __shared__ int soff[RES_NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        const float4* vp = vproj + (size_t)f * V;
        const long long base = fr.pix_off;
        const int A = fr.W * fr.H;
        const int ntile = (A + RES_TILE - 1) / RES_TILE;
        // With the visible box known, only the tiles over its columns +-1 are ranked: the
        // neighbour lookups read cum there and at crop index 0 (the border rule); the last
        // tile always runs because it writes the pixel count.
        int t_lo = 0, t_hi = ntile - 1;
        if (box) {
            const int4 bb = box[f];
            if (bb.y < 0) {
                t_lo = 1;
                t_hi = 0;
            } else {
                t_lo = (max(bb.x - 1, 0) * fr.H) / RES_TILE;
                t_hi = (min(bb.y + 2, fr.W) * fr.H - 1) / RES_TILE;
            }
        }
        for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
            if (tile != 0 && tile != ntile - 1 && (tile < t_lo || tile > t_hi)) continue;
            // offset of this tile: the preceding tiles' counts
            int part = 0;
            for (int j = tid; j < tile; j += RES_THREADS) part += tcount[(size_t)f * tmax + j];
            for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
            if (lane == 0) soff[wave] = part;
            unsigned long long key[RES_PER_THREAD];
            const int k0 = tile * RES_TILE + RES_PER_THREAD * tid;
            const int cnt = tile_keys(keys + base, k0, A, key);
            const unsigned long long b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4);
            const int before = __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt);
            if (lane == 0) swave[wave] = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
            __syncthreads();
            int off = 0, woff = 0, total = 0;
#pragma unroll
            for (int w = 0; w < RES_NW; ++w) {
                off += soff[w];
                woff += (w < wave) ? swave[w] : 0;
                total += swave[w];
            }
            int pos = off + woff + before;
            // the four pixels' gathers first (triangle corners, then their projected vertices),
            // so that they are in flight together; then the coverage arithmetic and the stores
            int tv[RES_PER_THREAD][3];
            float4 tB[RES_PER_THREAD], tC[RES_PER_THREAD], tA[RES_PER_THREAD];
#pragma unroll
            for (int q = 0; q < RES_PER_THREAD; ++q) {
                tv[q][0] = tv[q][1] = tv[q][2] = 0;
                if (key[q] != KEY_EMPTY) {
                    const int t = (int)(unsigned)(key[q] & 0xffffffffull);
                    tv[q][0] = tris[3 * t];
                    tv[q][1] = tris[3 * t + 1];
                    tv[q][2] = tris[3 * t + 2];
                }
            }
#pragma unroll
            for (int q = 0; q < RES_PER_THREAD; ++q)
                if (key[q] != KEY_EMPTY) {
                    tB[q] = vp[tv[q][0]];
                    tC[q] = vp[tv[q][1]];
                    tA[q] = vp[tv[q][2]];
                }
#pragma unroll
            for (int q = 0; q < RES_PER_THREAD; ++q) {
                const int k = k0 + q;
                if (k >= A) break;
                const bool occ = key[q] != KEY_EMPTY;
                cum[base + k] = pos + (occ ? 1 : 0);
                if (occ) {
                    vis[base + pos] = k;
                    float al, be;
                    cover_rn(tB[q].x, tB[q].y, tC[q].x, tC[q].y, tA[q].x, tA[q].y, fr.x0 + k / fr.H, fr.y0 + k % fr.H,
                             &al, &be);
                    if (bary) bary[base + k] = make_float2(al, be);
                    pvert[base + pos] = make_int4(tv[q][0], tv[q][1], tv[q][2], k);
                    pbary[base + pos] = make_float2(al, be);
                    ++pos;
                }
            }
            if (tile == ntile - 1 && tid == 0) npix[f] = off + total;
            __syncthreads();
        }
    }
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
// Clears the keys a frame's previous raster can have set.  Every raster is followed by the
// visible-box pass (or resolve), so the keys outside box[f] are already empty and only the
// box is cleared; batch chunk loads set box[f] to the whole crop.  Without a box the whole
// crop is cleared.  The box itself is reset by the next k_raster, after this clear has read it.
__global__ void __launch_bounds__(256) k_clear_keys_frames(const FrameGPU* __restrict__ frames,
                                                            const int* __restrict__ act, int n,
                                                            uint64_t* __restrict__ keys, const int4* __restrict__ box)
{
    const int nact = fl_count(act, n);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const FrameGPU fr = frames[f];
        int4 bx = box ? box[f] : make_int4(0, fr.W - 1, 0, fr.H - 1);
        bx.x = max(bx.x, 0); bx.y = min(bx.y, fr.W - 1);
        bx.z = max(bx.z, 0); bx.w = min(bx.w, fr.H - 1);
        uint64_t* kf = keys + fr.pix_off;
        for (int x = bx.x + blockIdx.x * nw + wave; x <= bx.y; x += gridDim.x * nw)
            for (int y = bx.z + lane; y <= bx.w; y += 64) kf[(long long)x * fr.H + y] = KEY_EMPTY;
    }
}

// box[f] = the whole crop, for every slot of a freshly loaded chunk (their keys are stale).
__global__ void k_box_full(const FrameGPU* __restrict__ frames, int n, int4* __restrict__ box)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) box[f] = make_int4(0, frames[f].W - 1, 0, frames[f].H - 1);
}
} // namespace

int launch_box_full(const FrameGPU* frames, int n, int4* box, hipStream_t s)
{
    hipLaunchKernelGGL(k_box_full, dim3((n + 255) / 256), dim3(256), 0, s, frames, n, box);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_clear_keys_frames(const FrameGPU* frames, FrameList fl, uint64_t* keys, int4* box, int max_area,
                             hipStream_t s)
{
    // a workgroup clears four crop columns at a time
    int bx = (int)(sqrtf((float)max_area) + 3) / 4;
    if (bx > 64) bx = 64;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_clear_keys_frames, dim3(bx, grid_frames(fl.n, 64)), dim3(256), 0, s, frames, fl.act, fl.n,
                       keys, box);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_raster(const ModelGPU& m, const FrameGPU* frames, FrameList fl, const float4* vproj, int ntmp,
                  uint64_t* keys, int4* box, hipStream_t s)
{
    dim3 grid((m.NT + 255) / 256, grid_frames(fl.n, 64));
    hipLaunchKernelGGL(k_raster, grid, dim3(256), 0, s, m.tris, m.NT, m.V, frames, fl.act, fl.n, vproj, ntmp,
                       (unsigned long long*)keys, box);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int resolve_tiles(int max_area) { return (max_area + RES_TILE - 1) / RES_TILE; }

int launch_visible_box(const FrameGPU* frames, FrameList fl, const uint64_t* keys, int max_area, int* tcount,
                       int tmax, int4* box, hipStream_t s)
{
    int bx = resolve_tiles(max_area);
    if (bx > tmax) {
        set_error("resolve: crop of %d pixels needs %d tiles > %d", max_area, bx, tmax);
        return TDI_ERR_ARG;
    }
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_resolve_count, dim3(bx, grid_frames(fl.n, 64)), dim3(RES_THREADS), 0, s, frames, fl.act, fl.n,
                       (const unsigned long long*)keys, tcount, tmax, box);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_resolve(const ModelGPU& m, const FrameGPU* frames, FrameList fl, const float4* vproj,
                   const uint64_t* keys, int* cum, int* vis_list, int* npix, float* bary, int4* pvert,
                   float2* pbary, int max_area, int* tcount, int tmax, int4* box, hipStream_t s)
{
    int bx = resolve_tiles(max_area);
    if (bx > tmax) {
        set_error("resolve: crop of %d pixels needs %d tiles > %d", max_area, bx, tmax);
        return TDI_ERR_ARG;
    }
    if (bx < 1) bx = 1;
    const dim3 grid(bx, grid_frames(fl.n, 64));
    hipLaunchKernelGGL(k_resolve_count, grid, dim3(RES_THREADS), 0, s, frames, fl.act, fl.n,
                       (const unsigned long long*)keys, tcount, tmax, box);
    hipLaunchKernelGGL(k_resolve_write, grid, dim3(RES_THREADS), 0, s, m.tris, m.V, frames, fl.act, fl.n, vproj,
                       (const unsigned long long*)keys, tcount, tmax, box, cum, vis_list, npix, (float2*)bary,
                       pvert, pbary);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
