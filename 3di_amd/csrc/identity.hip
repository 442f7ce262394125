// identity.hip — the joint system of the identity stage (fit_to_multi_images, model_fitter.cu:1252-1479,
// whose shape fit is fit_3DMM_shape_rigid over T frames, model_fitter.cu:121-375).
//
// The frames of a batch share the albedo and identity coefficients (the first Ks = ntex + nid
// unknowns of every frame's layout) and each has its own expression, pose and light block
// (Kf = K - Ks).  The joint unknown vector is
//     x = [shared (Ks) | frame 0's block (Kf) | frame 1's block | ...],  KJ = Ks + T Kf,
// the joint normal matrix the sum of the frames' matrices placed on it: the shared-shared block
// sums every frame's, each frame's block sits on the diagonal with its coupling to the shared
// block beside it, and blocks of different frames are zero (get_obj_hessian_and_gradient_multiframe,
// model_fitter.cu:21-116, accumulates the per-frame products into one Ktotal x Ktotal JTJ).
// The frames' barriers leave out the shared coefficients' bounds; k_shared_bounds adds them once
// (the reference evaluates the coefficient bounds once for all frames, derivative_computer.cu:176-212).
#include "common.h"

namespace tdi {

namespace {

// HJ [KJ][KJ] FP64 and gJ [KJ] FP64 from the frames' H [n][K][K] FP64 and g [n][K] (FP32 or FP64).
template <typename TG>
__global__ void __launch_bounds__(256) k_assemble(int n, int K, int Ks, const double* __restrict__ H,
                                                   const TG* __restrict__ g, double* __restrict__ HJ,
                                                   double* __restrict__ gJ)
{
    const int Kf = K - Ks;
    const long long KJ = Ks + (long long)n * Kf;
    const long long total = KJ * KJ + KJ;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        if (e < KJ * KJ) {
            const long long A = e / KJ, B = e - A * KJ;
            double v = 0.0;
            if (A < Ks && B < Ks) {
                for (int t = 0; t < n; ++t) v += H[(size_t)t * K * K + (size_t)A * K + B];
            } else {
                // frame and local index of each side; a shared index belongs to every frame
                const int ta = A < Ks ? -1 : (int)((A - Ks) / Kf), tb = B < Ks ? -1 : (int)((B - Ks) / Kf);
                const int la = A < Ks ? (int)A : Ks + (int)((A - Ks) % Kf);
                const int lb = B < Ks ? (int)B : Ks + (int)((B - Ks) % Kf);
                const int t = ta >= 0 ? ta : tb;
                if (ta < 0 || tb < 0 || ta == tb) v = H[(size_t)t * K * K + (size_t)la * K + lb];
            }
            HJ[e] = v;
        } else {
            const long long A = e - KJ * KJ;
            double v = 0.0;
            if (A < Ks) {
                for (int t = 0; t < n; ++t) v += g[(size_t)t * K + A];
            } else {
                const int t = (int)((A - Ks) / Kf), l = Ks + (int)((A - Ks) % Kf);
                v = g[(size_t)t * K + l];
            }
            gJ[A] = v;
        }
    }
}

// The shared coefficients' bound terms at x (frame 0's theta): mode 0, the objective into
// *obj; mode 1, gradient and Hessian diagonal added to gJ and HJ.  Only finite bounds take part.
__global__ void __launch_bounds__(256) k_shared_bounds(int Ks, long long KJ, const float* __restrict__ x,
                                                        const float* __restrict__ lb, const float* __restrict__ ub,
                                                        int mode, double* __restrict__ obj, double* __restrict__ HJ,
                                                        double* __restrict__ gJ)
{
    __shared__ double sred[4];
    double acc = 0.0;
    for (int j = threadIdx.x; j < Ks; j += blockDim.x) {
        const float lbf = lb[j], ubf = ub[j];
        if (!(ubf < INFINITY) || !(lbf > -INFINITY)) continue;
        const double xj = x[j], fub = xj - (double)ubf, flb = (double)lbf - xj;
        if (mode == 0) {
            acc += -log(-fub) - log(-flb);
        } else {
            gJ[j] += (double)(float)(-1.0 / fub + 1.0 / flb);
            const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
            HJ[(size_t)j * KJ + j] += a * a + b * b;
        }
    }
    if (mode == 0) {
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) *obj = sred[0] + sred[1] + sred[2] + sred[3];
    }
}

} // namespace

template <typename TG>
static int assemble_joint(int n, int K, int Ks, const double* H, const TG* g, double* HJ, double* gJ, hipStream_t s)
{
    const long long KJ = Ks + (long long)n * (K - Ks);
    const long long total = KJ * KJ + KJ;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_assemble<TG>, dim3(blocks), dim3(256), 0, s, n, K, Ks, H, g, HJ, gJ);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_assemble_joint(int n, int K, int Ks, const double* H, const float* g, double* HJ, double* gJ, hipStream_t s)
{
    return assemble_joint(n, K, Ks, H, g, HJ, gJ, s);
}

int launch_assemble_joint(int n, int K, int Ks, const double* H, const double* g, double* HJ, double* gJ, hipStream_t s)
{
    return assemble_joint(n, K, Ks, H, g, HJ, gJ, s);
}

int launch_shared_bounds(int Ks, long long KJ, const float* x, const float* lb, const float* ub, int mode, double* obj,
                         double* HJ, double* gJ, hipStream_t s)
{
    hipLaunchKernelGGL(k_shared_bounds, dim3(1), dim3(256), 0, s, Ks, KJ, x, lb, ub, mode, obj, HJ, gJ);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
