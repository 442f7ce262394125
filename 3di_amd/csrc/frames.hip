// frames.hip — frame preparation: 8-bit BGR video frames -> the grey FP32 crops the fit observes.
//
// One pass per crop pixel replaces the reference's per-frame host chain
// InputData::get_resized_frame (input_data.cu:44-53: cvtColor BGR2GRAY, convertTo CV_32F, /255,
// cv::resize by the camera's resize coefficient, copyMakeBorder with zeros at the bottom and right)
// followed by the crop at (x0_short, y0_short) and its transpose to column-major
// (video_fitter.cu:537-543).  The frames stay resident in HBM; only the crop's footprint in the
// source frame is read.  OpenCV's published semantics are kept operation for operation:
//   - grey = (1868 B + 9617 G + 4899 R + 2^13) >> 14, the BT.601 weights in 14-bit fixed point;
//   - value = float(grey) * float(1/255), the scale of convertTo;
//   - resized size round-half-even(src * coef); when it equals the source size the frame is
//     copied; otherwise INTER_LINEAR with half-pixel centres: f = float((d + 0.5) * (1/coef) - 0.5)
//     evaluated in double, s = floor(f), weight f - s; columns left of 0 or at/after the last
//     take the edge pixel with weight 0, rows are clamped individually and keep their weights;
//     the horizontal pass first, each pass a*(1-w) + b*w with both products and the sum rounded
//     (contraction is off in this file's device code: no fused multiply-add).
#include <climits>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"

namespace tdi {

namespace {

// A workgroup prepares a 64 x 64 tile of one crop in two phases.  Phase 1: each lane owns a crop
// column and walks 16 rows, so the 64 lanes of a wavefront read neighbouring source pixels of the
// same two source rows (a few cache lines per load instead of one per lane); the values go to an
// LDS tile.  Phase 2: each lane owns a crop row and writes 16 columns, so every wavefront store is
// 64 consecutive floats of one column-major crop column.
constexpr int PREP_TILE = 64;
constexpr int PREP_WAVES = 4;

struct PrepFrame {
    long long src;     // byte offset of the frame
    long long dst;     // float offset of the crop
    double scale;      // 1 / coef
    int x0, y0, w, h;  // crop in resized-frame pixels
    int dw, dh;        // resized frame size
    int copy;          // resized size equals the source size: plain copy
    int tiles_y;       // ceil(h / PREP_TILE)
};

__device__ __forceinline__ float grey_at(const uint8_t* __restrict__ f, int sw, int x, int y)
{
#pragma clang fp contract(off)
    const uint8_t* p = f + 3ll * ((long long)y * sw + x);
    const int g = (1868 * (int)p[0] + 9617 * (int)p[1] + 4899 * (int)p[2] + 8192) >> 14;
    return (float)g * (float)(1.0 / 255.0);
}

__device__ __forceinline__ float lerp_rn(float a, float b, float w)
{
#pragma clang fp contract(off)
    return a * (1.f - w) + b * w;
}

__global__ __launch_bounds__(256) void k_prepare_crops(int n, const PrepFrame* __restrict__ fr,
                                                       const int* __restrict__ tile_start,
                                                       const uint8_t* __restrict__ bgr, int sw, int sh,
                                                       float* __restrict__ out)
{
#pragma clang fp contract(off)   // every product and sum rounded, as the restated operation sequence
    __shared__ float tile[PREP_TILE][PREP_TILE + 1];
    // frame of this tile: the last f with tile_start[f] <= blockIdx.x
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const PrepFrame F = fr[lo];
    const int t = b - tile_start[lo];
    const int c0 = (t / F.tiles_y) * PREP_TILE, r0t = (t % F.tiles_y) * PREP_TILE;
    const int lane = (int)threadIdx.x % PREP_TILE, wave = (int)threadIdx.x / PREP_TILE;
    const uint8_t* src = bgr + F.src;

    // phase 1: lane = column
    {
        const int c = c0 + lane;
        const int X = F.x0 + c;
        int sx = 0, sx1 = 0;
        float fx = 0.f;
        if (!F.copy && X < F.dw) {
            fx = (float)(((double)X + 0.5) * F.scale - 0.5);
            sx = (int)floorf(fx);
            fx -= (float)sx;
            if (sx < 0) { sx = 0; fx = 0.f; }
            if (sx >= sw - 1) { sx = sw - 1; fx = 0.f; }
            sx1 = min(sx + 1, sw - 1);
        }
        for (int k = wave; k < PREP_TILE; k += PREP_WAVES) {
            const int r = r0t + k;
            if (c >= F.w || r >= F.h) continue;
            const int Y = F.y0 + r;
            float v = 0.f;
            if (X < F.dw && Y < F.dh) {
                if (F.copy) {
                    v = grey_at(src, sw, X, Y);
                } else {
                    float fy = (float)(((double)Y + 0.5) * F.scale - 0.5);
                    const int sy = (int)floorf(fy);
                    fy -= (float)sy;
                    const int y0 = min(max(sy, 0), sh - 1), y1 = min(max(sy + 1, 0), sh - 1);
                    const float h0 = lerp_rn(grey_at(src, sw, sx, y0), grey_at(src, sw, sx1, y0), fx);
                    const float h1 = lerp_rn(grey_at(src, sw, sx, y1), grey_at(src, sw, sx1, y1), fx);
                    v = lerp_rn(h0, h1, fy);
                }
            }
            tile[lane][k] = v;
        }
    }
    __syncthreads();
    // phase 2: lane = row
    const int r = r0t + lane;
    if (r >= F.h) return;
    for (int k = wave; k < PREP_TILE; k += PREP_WAVES) {
        const int c = c0 + k;
        if (c >= F.w) break;
        out[F.dst + (long long)c * F.h + r] = tile[k][lane];
    }
}

} // namespace

namespace {

// Checks the arguments and builds the per-frame descriptors and tile offsets on the host.
int prepare_setup(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                  const float* coef, const int32_t* crop, float* out, std::vector<PrepFrame>& fr,
                  std::vector<int>& tile_start, long long& tiles)
{
    if (n < 0 || (n > 0 && (!bgr || !coef || !crop || !out)) || src_w < 1 || src_h < 1 ||
        frame_stride < 3ll * src_w * src_h) {
        set_error("tdi_prepare_crops: null argument, negative frame count or bad frame geometry");
        return TDI_ERR_ARG;
    }
    fr.resize(n);
    tile_start.resize(n);
    long long dst = 0;
    tiles = 0;
    for (int i = 0; i < n; ++i) {
        const float k = coef[i];
        const int* q = crop + 4 * i;
        if (!(k > 0.f) || !std::isfinite(k) || q[0] < 0 || q[1] < 0 || q[2] < 1 || q[3] < 1) {
            set_error("tdi_prepare_crops: frame %d: coefficient must be positive and finite, crop origin "
                      "non-negative and crop size positive", i);
            return TDI_ERR_ARG;
        }
        if (q[0] > INT_MAX - q[2] || q[1] > INT_MAX - q[3]) {
            // the kernel forms x0 + column and y0 + row in 32 bits
            set_error("tdi_prepare_crops: frame %d: crop reaches past 2^31 - 1 pixels", i);
            return TDI_ERR_ARG;
        }
        PrepFrame& F = fr[i];
        F.src = (long long)i * frame_stride;
        F.dst = dst;
        F.scale = 1.0 / (double)k;
        F.x0 = q[0]; F.y0 = q[1]; F.w = q[2]; F.h = q[3];
        const double dw = std::nearbyint((double)src_w * (double)k), dh = std::nearbyint((double)src_h * (double)k);
        if (dw < 1.0 || dh < 1.0 || dw > 1e9 || dh > 1e9) {
            set_error("tdi_prepare_crops: frame %d: the resized frame would be empty or too large", i);
            return TDI_ERR_ARG;
        }
        F.dw = (int)dw; F.dh = (int)dh;
        F.copy = (F.dw == src_w && F.dh == src_h);
        F.tiles_y = (F.h + PREP_TILE - 1) / PREP_TILE;
        tile_start[i] = (int)tiles;
        tiles += (long long)F.tiles_y * ((F.w + PREP_TILE - 1) / PREP_TILE);
        dst += (long long)F.w * F.h;
        if (tiles > (1ll << 24) - 1) {
            set_error("tdi_prepare_crops: too many crop pixels for one call");
            return TDI_ERR_ARG;
        }
    }
    return TDI_OK;
}

// Uploads the descriptors to dfr/dts (device memory), runs the kernel and waits for it: the
// host descriptor arrays are the sources of the asynchronous copies.
int prepare_run(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h, float* out,
                const std::vector<PrepFrame>& fr, const std::vector<int>& tile_start, long long tiles,
                PrepFrame* dfr, int* dts, hipStream_t s)
{
    int rc = TDI_OK;
    if (hipMemcpyAsync(dfr, fr.data(), sizeof(PrepFrame) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dts, tile_start.data(), sizeof(int) * n, hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("tdi_prepare_crops: descriptor upload failed");
        rc = TDI_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_prepare_crops, dim3((unsigned)tiles), dim3(PREP_TILE * PREP_WAVES), 0, s, n, dfr, dts, bgr,
                           src_w, src_h, out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error("tdi_prepare_crops: %s", hipGetErrorString(e)); rc = TDI_ERR_HIP; }
    }
    (void)hipStreamSynchronize(s);
    return rc;
}

// The descriptors of tdi_prepare_crops_noalloc: PrepFrame[n] then int[n], n <= TDI_PREP_NOALLOC_MAX_FRAMES,
// one copy per GPU for the life of the process.
constexpr size_t PREP_DESC_BYTES = ((sizeof(PrepFrame) * TDI_PREP_NOALLOC_MAX_FRAMES + 255) / 256) * 256;
constexpr size_t PREP_SCRATCH_BYTES = PREP_DESC_BYTES + sizeof(int) * TDI_PREP_NOALLOC_MAX_FRAMES;
__device__ __align__(256) unsigned char scratch[PREP_SCRATCH_BYTES];

// Device address of `scratch` on each device, looked up once per device with that device current
// (the array's host-side name is not a device address).
constexpr int PREP_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[PREP_MAX_DEVICES] = {};

int scratch_for_current_device(void** ptr)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= PREP_MAX_DEVICES) {
        set_error("tdi_prepare_crops_noalloc: no usable current device");
        return TDI_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("tdi_prepare_crops_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return TDI_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return TDI_OK;
}


// ---------------------------------------------------------------------------------------------
// Inner-mouth masking (PAINT_INNERMOUTH_BLACK; preprocessing.cu:40-103).  The reference strokes
// the inner-lip polygon (landmarks 43-50; 43 and 47 moved 3u inwards, u = face size /
// REF_FACE_SIZE) with 3-pixel lines, traces the outer contour of that ring and blackens every pixel
// of the mouth box (landmarks 31-50 clipped to the frame) whose signed distance to the contour
// exceeds u, when the mouth is open by more than 3 REF_FACE_SIZE units (mean of three lip
// distances).  The traced contour is replaced by its geometry, the polygon moved one pixel outwards:
// a pixel is blackened when it lies inside the polygon by more than u - 1 (or anywhere inside when
// u < 1) or outside by less than 1 - u.  The per-frame quantities are FP32 on the host in the
// reference's order; the per-pixel test is FP64 with squared distances and no contraction.
// One lane per pixel of the box; the boxes of all frames form one flat list of 256-pixel chunks.

constexpr int MOUTH_CHUNK = 256;

struct MouthFrame {
    double px[8], py[8];  // polygon in frame pixels
    double u;             // face size / REF_FACE_SIZE
    long long src;        // byte offset of the frame
    int x0, y0, bw, bh;   // box origin and size (pixels)
    int chunk0;           // first chunk of the frame
};

__device__ __forceinline__ bool mouth_pixel(const MouthFrame& M, double X, double Y)
{
#pragma clang fp contract(off)
    double best = 1e300;
    bool inside = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int k1 = (k + 1) & 7;
        const double ax = M.px[k], ay = M.py[k], bx = M.px[k1], by = M.py[k1];
        const double ex = bx - ax, ey = by - ay;
        const double wx = X - ax, wy = Y - ay;
        const double ee = ex * ex + ey * ey;
        double t = ee > 0.0 ? (wx * ex + wy * ey) / ee : 0.0;
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double dx = wx - t * ex, dy = wy - t * ey;
        const double d2 = dx * dx + dy * dy;
        best = d2 < best ? d2 : best;
        if ((ay > Y) != (by > Y)) {
            const double xc = ax + (Y - ay) / (by - ay) * ex;
            inside = (X < xc) != inside;
        }
    }
    const double m = M.u - 1.0;
    return inside ? (m < 0.0 || best > m * m) : (m < 0.0 && best < m * m);
}

__global__ __launch_bounds__(MOUTH_CHUNK) void k_paint_innermouth(int n, const MouthFrame* __restrict__ fr,
                                                                   uint8_t* __restrict__ bgr, int w)
{
    const int c = blockIdx.x;
    int lo = 0, hi = n - 1;  // last frame with chunk0 <= c
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (fr[mid].chunk0 <= c) lo = mid; else hi = mid - 1;
    }
    const MouthFrame& M = fr[lo];
    const long long p = (long long)(c - M.chunk0) * MOUTH_CHUNK + threadIdx.x;
    if (p >= (long long)M.bw * M.bh) return;
    const int i = M.y0 + (int)(p / M.bw), j = M.x0 + (int)(p % M.bw);
    if (mouth_pixel(M, (double)j, (double)i)) {
        uint8_t* q = bgr + M.src + 3ll * ((long long)i * w + j);
        q[0] = 0; q[1] = 0; q[2] = 0;
    }
}

// Per-frame quantities on the host, FP32 as the reference computes them.  Returns false for a
// frame that is left as it is (mouth closed, empty box, non-finite landmarks).
bool mouth_frame(const float* x, const float* y, int w, int h, float ref, MouthFrame& M)
{
#pragma clang fp contract(off)
    float minx = x[0], maxx = x[0], miny = y[0], maxy = y[0];
    for (int i = 1; i < TDI_NLANDMARKS; ++i) {
        minx = std::min(minx, x[i]); maxx = std::max(maxx, x[i]);
        miny = std::min(miny, y[i]); maxy = std::max(maxy, y[i]);
    }
    const float face = std::max(maxx - minx, maxy - miny);  // compute_face_size, funcs.cu:44-53
    const float unit = face / ref;
    bool finite = std::isfinite(face);
    for (int i = 31; i < TDI_NLANDMARKS; ++i) finite = finite && std::isfinite(x[i]) && std::isfinite(y[i]);
    auto dist = [&](int a, int b) { return std::sqrt((x[a] - x[b]) * (x[a] - x[b]) + (y[a] - y[b]) * (y[a] - y[b])); };
    const float open = (ref / face) * (dist(44, 50) + dist(45, 49) + dist(46, 48)) / 3.0f;
    if (!(finite && face > 0.f && open > 3.0f)) return false;
    for (int k = 0; k < 8; ++k) {
        const float shift = k == 0 ? 3 * unit : (k == 4 ? -(3 * unit) : 0.f);
        M.px[k] = (double)(x[43 + k] + shift);
        M.py[k] = (double)y[43 + k];
    }
    float bx0 = x[31], bx1 = x[31], by0 = y[31], by1 = y[31];
    for (int i = 32; i < TDI_NLANDMARKS; ++i) {
        bx0 = std::min(bx0, x[i]); bx1 = std::max(bx1, x[i]);
        by0 = std::min(by0, y[i]); by1 = std::max(by1, y[i]);
    }
    bx0 = std::min((float)w, std::max(0.f, bx0)); bx1 = std::max(0.f, std::min((float)w, bx1));
    by0 = std::min((float)h, std::max(0.f, by0)); by1 = std::max(0.f, std::min((float)h, by1));
    M.u = unit;
    M.x0 = (int)(unsigned)bx0; M.y0 = (int)(unsigned)by0;
    M.bw = (int)(unsigned)bx1 - M.x0; M.bh = (int)(unsigned)by1 - M.y0;
    return M.bw > 0 && M.bh > 0;
}

} // namespace

} // namespace tdi

using namespace tdi;

extern "C" int tdi_paint_innermouth_black(int32_t n, uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                                          const float* lmk_x, const float* lmk_y, float ref_face_size, void* stream)
{
    if (n < 0 || (n > 0 && (!bgr || !lmk_x || !lmk_y)) || src_w < 1 || src_h < 1 ||
        frame_stride < 3ll * src_w * src_h || !(ref_face_size > 0.f) || !std::isfinite(ref_face_size)) {
        set_error("tdi_paint_innermouth_black: null argument, negative frame count, bad frame geometry or "
                  "reference face size");
        return TDI_ERR_ARG;
    }
    std::vector<MouthFrame> fr;
    fr.reserve(n);
    long long chunks = 0;
    for (int f = 0; f < n; ++f) {
        MouthFrame M{};
        if (!mouth_frame(lmk_x + (size_t)f * TDI_NLANDMARKS, lmk_y + (size_t)f * TDI_NLANDMARKS, src_w, src_h,
                         ref_face_size, M))
            continue;
        M.src = (long long)f * frame_stride;
        M.chunk0 = (int)chunks;
        chunks += ((long long)M.bw * M.bh + MOUTH_CHUNK - 1) / MOUTH_CHUNK;
        if (chunks > INT_MAX) {
            set_error("tdi_paint_innermouth_black: too many mouth pixels for one call");
            return TDI_ERR_ARG;
        }
        fr.push_back(M);
    }
    if (fr.empty()) return TDI_OK;
    hipStream_t s = (hipStream_t)stream;
    MouthFrame* dfr = nullptr;
    if (hipMalloc(&dfr, sizeof(MouthFrame) * fr.size()) != hipSuccess) {
        set_error("tdi_paint_innermouth_black: descriptor allocation failed");
        return TDI_ERR_NOMEM;
    }
    int rc = TDI_OK;
    hipError_t e = hipMemcpyAsync(dfr, fr.data(), sizeof(MouthFrame) * fr.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_paint_innermouth, dim3((unsigned)chunks), dim3(MOUTH_CHUNK), 0, s, (int)fr.size(), dfr,
                           bgr, src_w);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { set_error("tdi_paint_innermouth_black: %s", hipGetErrorString(e)); rc = TDI_ERR_HIP; }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dfr);
    return rc;
}

extern "C" int tdi_prepare_crops(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                                 const float* coef, const int32_t* crop, float* out, void* stream)
{
    std::vector<PrepFrame> fr;
    std::vector<int> tile_start;
    long long tiles = 0;
    int rc = prepare_setup(n, bgr, src_w, src_h, frame_stride, coef, crop, out, fr, tile_start, tiles);
    if (rc || n == 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    PrepFrame* dfr = nullptr;
    int* dts = nullptr;
    if (hipMalloc(&dfr, sizeof(PrepFrame) * n) != hipSuccess || hipMalloc(&dts, sizeof(int) * n) != hipSuccess) {
        (void)hipFree(dfr);
        set_error("tdi_prepare_crops: descriptor allocation failed");
        return TDI_ERR_NOMEM;
    }
    rc = prepare_run(n, bgr, src_w, src_h, out, fr, tile_start, tiles, dfr, dts, s);
    (void)hipFree(dfr);
    (void)hipFree(dts);
    return rc;
}

extern "C" int tdi_prepare_crops_noalloc(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h,
                                         int64_t frame_stride, const float* coef, const int32_t* crop, float* out,
                                         void* stream)
{
    if (n > TDI_PREP_NOALLOC_MAX_FRAMES) {
        set_error("tdi_prepare_crops_noalloc: %d frames, at most %d per call", n, TDI_PREP_NOALLOC_MAX_FRAMES);
        return TDI_ERR_ARG;
    }
    std::vector<PrepFrame> fr;
    std::vector<int> tile_start;
    long long tiles = 0;
    int rc = prepare_setup(n, bgr, src_w, src_h, frame_stride, coef, crop, out, fr, tile_start, tiles);
    if (rc || n == 0) return rc;
    void* base = nullptr;
    if ((rc = scratch_for_current_device(&base))) return rc;
    PrepFrame* dfr = reinterpret_cast<PrepFrame*>(base);
    int* dts = reinterpret_cast<int*>(static_cast<unsigned char*>(base) + PREP_DESC_BYTES);
    return prepare_run(n, bgr, src_w, src_h, out, fr, tile_start, tiles, dfr, dts, (hipStream_t)stream);
}
