// misc.hip — small layout and conversion kernels used by the C ABI.
#include "common.h"

namespace tdi {

namespace {

__global__ void k_pack(int total, const float* __restrict__ xyz, const float* __restrict__ xyp, float4* __restrict__ vp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    vp[i] = make_float4(xyp[2 * i], xyp[2 * i + 1], xyz[3 * i + 2], 0.f);
}

__global__ void k_unpack(int total, const float* __restrict__ P, const float4* __restrict__ vp, float* __restrict__ xyz,
                         float* __restrict__ xyp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    xyz[3 * i] = P[3 * i];
    xyz[3 * i + 1] = P[3 * i + 1];
    xyz[3 * i + 2] = P[3 * i + 2];
    const float4 v = vp[i];
    xyp[2 * i] = v.x;
    xyp[2 * i + 1] = v.y;
}

__global__ void k_fill(float* __restrict__ p, long long n, float v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void k_rotation(int n, const float* __restrict__ u, float* __restrict__ R, float* __restrict__ dR)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r[9], d[27];
    rodrigues_f32(u + 3 * i, r, dR ? d : nullptr);
    for (int k = 0; k < 9; ++k) R[9 * i + k] = r[k];
    if (dR)
        for (int k = 0; k < 27; ++k) dR[27 * i + k] = d[k];
}

// Per slot f < n: the light block of the normal matrix (nd x nd from row/column j0), the light
// gradient, objective, barrier objective and pixel count, packed into 16 doubles.
__global__ void k_light_pack(int n, int K, int j0, int nd, const double* __restrict__ H, const float* __restrict__ g,
                             const double* __restrict__ obj, const double* __restrict__ objb,
                             const int* __restrict__ npix, double* __restrict__ out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double* o = out + (size_t)f * 16;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            o[3 * a + b] = (a < nd && b < nd) ? H[(size_t)f * K * K + (size_t)(j0 + a) * K + j0 + b] : 0.0;
    for (int a = 0; a < 3; ++a) o[9 + a] = a < nd ? (double)g[(size_t)f * K + j0 + a] : 0.0;
    o[12] = obj[f];
    o[13] = objb[f];
    o[14] = (double)npix[f];
    o[15] = 0.0;
}

} // namespace

int launch_misc_pack(int n, int V, const float* xyz, const float* xyp, float4* vproj, hipStream_t s)
{
    const int total = n * V;
    hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, s, total, xyz, xyp, vproj);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_misc_unpack(int n, int V, const float* P, const float4* vproj, float* xyz, float* xyp, hipStream_t s)
{
    const int total = n * V;
    hipLaunchKernelGGL(k_unpack, dim3((total + 255) / 256), dim3(256), 0, s, total, P, vproj, xyz, xyp);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_misc_fill_f32(float* p, long long count, float v, hipStream_t s)
{
    int blocks = (int)std::min<long long>((count + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, count, v);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_misc_f64_to_f32(const double* in, float* out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f64_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_misc_light_pack(int n, int K, int j0, int nd, const double* H, const float* g, const double* obj,
                           const double* objb, const int* npix, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_light_pack, dim3((n + 63) / 64), dim3(64), 0, s, n, K, j0, nd, H, g, obj, objb, npix, out);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_misc_rotation(int n, const float* u, float* R, float* dR, hipStream_t s)
{
    if (n <= 0) return TDI_OK;
    hipLaunchKernelGGL(k_rotation, dim3((n + 63) / 64), dim3(64), 0, s, n, u, R, dR);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
