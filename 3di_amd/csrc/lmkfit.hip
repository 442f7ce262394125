// lmkfit.hip — the landmark-only log-barrier fit with slack, one workgroup per frame, FP64.
//
// Logbarrier_Initializer::set_minimal_slack (logbarrier_initializer.cu:1041-1090) and fit_model
// with use_slack (1095-1267), for T = 1 and no identity unknowns (li_init, video_fitter.cu:30).
// The reference drives each Newton step from the host: ~40 small launches for the terms
// (compute_gradient_and_hessian, 1272-1842), a cuSOLVER factorisation, and one host round trip
// per line-search trial.  Here a workgroup runs the whole fit of its frame — every round,
// Newton step, factorisation and trial — in one launch, with the frame's state in LDS:
//   unknowns z = (eps[0..Kl), tau_x, tau_y, tau_z, u1, u2, u3, s), n = Kl + 7, np = n - 1;
//   J: the projected-landmark derivatives [102][npp] (rows x_i, y_i; npp = np rounded up to 4,
//      padding columns zero), then the same LDS holds the n x n Hessian and its factor;
//   the Gauss-Newton block sum_i wx_i dx_i dx_i^T + wy_i dy_i dy_i^T is formed from J scaled by
//      sqrt(w), each thread owning 4x4 tiles of the upper triangle in registers.
// Objective, gradient and Hessian follow the reference term by term (see tdi_api.h); log, sqrt
// and division are FP64.
#include <cmath>

#include "common.h"

namespace tdi {

namespace {

constexpr int LF_THREADS = 256;
constexpr int LF_NMAX = 96;                     // n = Kl + 7 <= 96
constexpr int LF_NPP = 96;                      // npp <= 96
constexpr int LF_ROWS = 2 * NLMK;               // 102
constexpr int LF_TMAX = LF_NPP / 4;             // 24 tiles a side
constexpr int LF_TILES_PER_THREAD = (LF_TMAX * (LF_TMAX + 1) / 2 + LF_THREADS - 1) / LF_THREADS;   // 2

struct LmkRep {
    int status, iterations, rounds, accepted;
    float slack0, slack;
};

struct LfShared {
    double z[LF_NMAX], d[LF_NMAX], zt[LF_NMAX], g[LF_NMAX], hs[LF_NMAX];
    double base[NLMK][3], p[NLMK][3], c[NLMK][3];
    double R[9], dR[27];
    double cx[NLMK], cy[NLMK], wx[NLMK], wy[NLMK], hx[NLMK], hy[NLMK], qg[NLMK], qh[NLMK], gmin[NLMK];
    double red[LF_THREADS / 64];
    double bc[2];
    int flag;
};

__device__ __forceinline__ double block_sum(double v, LfShared& S)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) S.red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < LF_THREADS / 64; ++w) s += S.red[w];
        S.bc[0] = s;
    }
    __syncthreads();
    const double r = S.bc[0];
    __syncthreads();
    return r;
}

struct LfFrame {
    ModelGPU m;
    Layout L;
    FrameGPU fr;
    const float* lm;   // [6][51]: x, y, box lx, ux, ly, uy
    int Kl, n, np, npp;
};

// Objective at zv (barrier weight tau_b on the slack); with derivs the per-landmark coefficients
// of the gradient and Hessian are left in S (and R, dR, p, c at zv).  Every thread returns the
// objective.
__device__ double lf_eval(const LfFrame& F, const double* zv, bool derivs, double tau_b, LfShared& S)
{
    const ModelGPU& m = F.m;
    const int tid = threadIdx.x, Kl = F.Kl, np = F.np;
    if (tid == 0) {
        if (derivs) rodrigues_f64_d(zv + Kl + 3, S.R, S.dR);
        else rodrigues_f64(zv + Kl + 3, S.R);
    }
    for (int r = tid; r < 3 * NLMK; r += LF_THREADS) {
        const int i = r / 3, a = r - 3 * i;
        const float* e = m.exp_basis + (size_t)(3 * m.lmk_ids[i] + a) * m.exp_ld;
        double x = S.base[i][a];
        for (int k = 0; k < Kl; ++k) x = fma((double)e[k], zv[k], x);
        S.p[i][a] = x;
    }
    __syncthreads();
    double acc = 0.0;
    if (tid < NLMK) {
        const int i = tid;
        const double s = zv[np];
        double c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = S.R[a] * S.p[i][0] + S.R[3 + a] * S.p[i][1] + S.R[6 + a] * S.p[i][2] + zv[Kl + a];
        S.c[i][0] = c[0]; S.c[i][1] = c[1]; S.c[i][2] = c[2];
        const double xp = F.fr.phix * c[0] / c[2] + F.fr.cx, yp = F.fr.phiy * c[1] / c[2] + F.fr.cy;
        const double fs = F.fr.face_size;
        const double lx = F.lm[i], ly = F.lm[NLMK + i];
        const double Gxm = s + xp - (lx + (double)F.lm[2 * NLMK + i] * fs);
        const double Gxp = s + (lx + (double)F.lm[3 * NLMK + i] * fs) - xp;
        const double Gym = s + yp - (ly + (double)F.lm[4 * NLMK + i] * fs);
        const double Gyp = s + (ly + (double)F.lm[5 * NLMK + i] * fs) - yp;
        acc = -log(Gxm) - log(Gym) - log(Gxp) - log(Gyp);
        if (i == 0) acc += tau_b * s;
        S.gmin[i] = fmin(fmin(Gxm, Gym), fmin(Gxp, Gyp));
        if (Gxm != Gxm || Gym != Gym || Gxp != Gxp || Gyp != Gyp) S.gmin[i] = NAN;
        if (derivs) {
            S.cx[i] = -1.0 / Gxm + 1.0 / Gxp;
            S.cy[i] = -1.0 / Gym + 1.0 / Gyp;
            S.wx[i] = 1.0 / (Gxm * Gxm) + 1.0 / (Gxp * Gxp);
            S.wy[i] = 1.0 / (Gym * Gym) + 1.0 / (Gyp * Gyp);
            S.hx[i] = 1.0 / (Gxm * Gxm) - 1.0 / (Gxp * Gxp);
            S.hy[i] = 1.0 / (Gym * Gym) - 1.0 / (Gyp * Gyp);
            const double q0 = 1.0 / (1e-6 + Gxm), q1 = 1.0 / (1e-6 + Gym), q2 = 1.0 / (1e-6 + Gxp),
                         q3 = 1.0 / (1e-6 + Gyp);
            S.qg[i] = q0 + q1 + q2 + q3;
            S.qh[i] = q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3;
        }
    } else if (tid >= 64 && tid < 64 + Kl) {
        const int k = tid - 64;
        const float lbf = m.lb[F.L.off_exp + k], ubf = m.ub[F.L.off_exp + k];
        if (ubf < INFINITY && lbf > -INFINITY) {
            const double x = zv[k], fub = x - (double)ubf, flb = (double)lbf - x;
            acc = -log(-fub) - log(-flb);
        }
    }
    return block_sum(acc, S);
}

// Gradient and Gauss-Newton Hessian at the point of the last lf_eval(derivs = true); H (n x n,
// row-major) is written over J in `dyn`.
__device__ void lf_terms(const LfFrame& F, const double* zv, double tau_b, double* dyn, LfShared& S)
{
    const ModelGPU& m = F.m;
    const int tid = threadIdx.x, Kl = F.Kl, n = F.n, np = F.np, npp = F.npp;
    double* J = dyn;
    for (int e = tid; e < NLMK * npp; e += LF_THREADS) {
        const int i = e / npp, j = e - i * npp;
        double dxj = 0.0, dyj = 0.0;
        if (j < np) {
            double dv[3];
            if (j < Kl) {
                const int v = m.lmk_ids[i];
                double ev[3];
                for (int a = 0; a < 3; ++a) ev[a] = m.exp_basis[(size_t)(3 * v + a) * m.exp_ld + j];
                for (int a = 0; a < 3; ++a) dv[a] = S.R[a] * ev[0] + S.R[3 + a] * ev[1] + S.R[6 + a] * ev[2];
            } else if (j < Kl + 3) {
                dv[0] = dv[1] = dv[2] = 0.0;
                dv[j - Kl] = 1.0;
            } else {
                const double* d = S.dR + 9 * (j - Kl - 3);
                for (int a = 0; a < 3; ++a) dv[a] = d[a] * S.p[i][0] + d[3 + a] * S.p[i][1] + d[6 + a] * S.p[i][2];
            }
            const double cx = S.c[i][0], cy = S.c[i][1], cz = S.c[i][2];
            dxj = F.fr.phix * (dv[0] * cz - dv[2] * cx) / (cz * cz);
            dyj = F.fr.phiy * (dv[1] * cz - dv[2] * cy) / (cz * cz);
        }
        J[(size_t)(2 * i) * npp + j] = dxj;
        J[(size_t)(2 * i + 1) * npp + j] = dyj;
    }
    __syncthreads();
    // gradient and the slack column from the raw derivatives
    for (int j = tid; j < np; j += LF_THREADS) {
        double gs = 0.0, hs = 0.0;
        for (int i = 0; i < NLMK; ++i) {
            const double dx = J[(size_t)(2 * i) * npp + j], dy = J[(size_t)(2 * i + 1) * npp + j];
            gs += S.cx[i] * dx + S.cy[i] * dy;
            hs += S.hx[i] * dx + S.hy[i] * dy;
        }
        if (j < Kl) {
            const float lbf = m.lb[F.L.off_exp + j], ubf = m.ub[F.L.off_exp + j];
            if (ubf < INFINITY && lbf > -INFINITY) {
                const double x = zv[j], fub = x - (double)ubf, flb = (double)lbf - x;
                gs += -1.0 / fub + 1.0 / flb;
            }
        }
        S.g[j] = gs;
        S.hs[j] = hs;
    }
    double hss = 0.0;
    if (tid == 0) {
        double qs = 0.0;
        for (int i = 0; i < NLMK; ++i) { qs += S.qg[i]; hss += S.qh[i]; }
        S.g[np] = tau_b - qs;
    }
    __syncthreads();
    for (int e = tid; e < LF_ROWS * npp; e += LF_THREADS) {
        const int r = e / npp, i = r >> 1;
        J[e] *= sqrt((r & 1) ? S.wy[i] : S.wx[i]);
    }
    __syncthreads();
    // 4x4 tiles (ta <= tb) of the np x np block, accumulated in registers
    const int T = npp / 4, ntiles = T * (T + 1) / 2;
    double acc[LF_TILES_PER_THREAD][16];
    int tta[LF_TILES_PER_THREAD], ttb[LF_TILES_PER_THREAD];
#pragma unroll
    for (int q = 0; q < LF_TILES_PER_THREAD; ++q) {
        const int t = tid + q * LF_THREADS;
        int ta = 0, rem = t;
        while (ta < T && rem >= T - ta) { rem -= T - ta; ++ta; }
        tta[q] = ta;
        ttb[q] = ta + rem;
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[q][u] = 0.0;
        if (t >= ntiles) continue;
        const double* ca = J + 4 * ta;
        const double* cb = J + 4 * (ta + rem);
        for (int r = 0; r < LF_ROWS; ++r) {
            const double2 a01 = *reinterpret_cast<const double2*>(ca + (size_t)r * npp);
            const double2 a23 = *reinterpret_cast<const double2*>(ca + (size_t)r * npp + 2);
            const double2 b01 = *reinterpret_cast<const double2*>(cb + (size_t)r * npp);
            const double2 b23 = *reinterpret_cast<const double2*>(cb + (size_t)r * npp + 2);
            const double av[4] = {a01.x, a01.y, a23.x, a23.y}, bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[q][4 * x + y] = fma(av[x], bv[y], acc[q][4 * x + y]);
        }
    }
    __syncthreads();   // J is read in full: H may overwrite it
    double* H = dyn;
#pragma unroll
    for (int q = 0; q < LF_TILES_PER_THREAD; ++q) {
        const int t = tid + q * LF_THREADS;
        if (t >= ntiles) continue;
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
                const int a = 4 * tta[q] + x, b = 4 * ttb[q] + y;
                if (a >= np || b >= np) continue;
                H[(size_t)a * n + b] = acc[q][4 * x + y];
                H[(size_t)b * n + a] = acc[q][4 * x + y];
            }
    }
    __syncthreads();
    for (int j = tid; j < np; j += LF_THREADS) {
        H[(size_t)j * n + np] = S.hs[j];
        H[(size_t)np * n + j] = S.hs[j];
        if (j < Kl) {
            const float lbf = m.lb[F.L.off_exp + j], ubf = m.ub[F.L.off_exp + j];
            if (ubf < INFINITY && lbf > -INFINITY) {
                const double x = zv[j], fub = x - (double)ubf, flb = (double)lbf - x;
                const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
                H[(size_t)j * n + j] += a * a + b * b;
            }
        }
    }
    if (tid == 0) H[(size_t)np * n + np] = hss;
    __syncthreads();
}

// Cholesky factorisation of H (n x n in LDS, lower triangle) and the solve H d = -g.  Returns
// false (in every thread) on a pivot that is not positive.
__device__ bool lf_solve(int n, double* H, LfShared& S)
{
    const int tid = threadIdx.x;
    for (int k = 0; k < n; ++k) {
        if (tid == 0) {
            const double d = H[(size_t)k * n + k];
            S.flag = d > 0.0;
            if (d > 0.0) H[(size_t)k * n + k] = sqrt(d);
        }
        __syncthreads();
        if (!S.flag) return false;
        const double dk = H[(size_t)k * n + k];
        for (int i = k + 1 + tid; i < n; i += LF_THREADS) H[(size_t)i * n + k] /= dk;
        __syncthreads();
        const int w = n - k - 1;
        for (int e = tid; e < w * w; e += LF_THREADS) {
            const int i = k + 1 + e / w, j = k + 1 + e % w;
            if (j <= i) H[(size_t)i * n + j] -= H[(size_t)i * n + k] * H[(size_t)j * n + k];
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += LF_THREADS) S.zt[i] = -S.g[i];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double yk = S.zt[k] / H[(size_t)k * n + k];
        __syncthreads();
        if (tid == 0) S.zt[k] = yk;
        for (int i = k + 1 + tid; i < n; i += LF_THREADS) S.zt[i] -= H[(size_t)i * n + k] * yk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const double dk = S.zt[k] / H[(size_t)k * n + k];
        __syncthreads();
        if (tid == 0) S.d[k] = dk;
        for (int i = tid; i < k; i += LF_THREADS) S.zt[i] -= H[(size_t)k * n + i] * dk;
        __syncthreads();
    }
    return true;
}

// mode 0: the fit (theta in/out, rep); mode 1: terms at zin (Hout, gout, objout).
__global__ void __launch_bounds__(LF_THREADS) k_lmk_slack(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                          const float* __restrict__ lmk, float* __restrict__ theta,
                                                          int Kl, int mode, const double* __restrict__ zin,
                                                          double tau_in, double* __restrict__ Hout,
                                                          double* __restrict__ gout, double* __restrict__ objout,
                                                          LmkRep* __restrict__ rep)
{
    extern __shared__ __align__(16) double dyn[];
    __shared__ LfShared S;
    const int f = blockIdx.x, tid = threadIdx.x;
    LfFrame F;
    F.m = m;
    F.L = L;
    F.fr = frames[f];
    F.lm = lmk + (size_t)f * 6 * NLMK;
    F.Kl = Kl;
    F.n = Kl + 7;
    F.np = Kl + 6;
    F.npp = (F.np + 3) & ~3;
    const int n = F.n, np = F.np;
    float* th = theta + (size_t)f * L.K;
    // landmark positions without the fitted expression terms: mean + fixed identity terms
    for (int r = tid; r < 3 * NLMK; r += LF_THREADS) {
        const int i = r / 3, a = r - 3 * i;
        const size_t row = (size_t)(3 * m.lmk_ids[i] + a);
        double x = m.mean[row];
        for (int k = 0; k < L.nid; ++k) x += (double)m.id_basis[row * m.id_ld + k] * (double)th[L.off_id + k];
        S.base[i][a] = x;
    }
    if (mode == 1) {
        for (int j = tid; j < n; j += LF_THREADS) S.z[j] = zin[(size_t)f * n + j];
        __syncthreads();
        const double obj = lf_eval(F, S.z, true, tau_in, S);
        lf_terms(F, S.z, tau_in, dyn, S);
        if (objout && tid == 0) objout[f] = obj;
        if (gout)
            for (int j = tid; j < n; j += LF_THREADS) gout[(size_t)f * n + j] = S.g[j];
        if (Hout)
            for (int e = tid; e < n * n; e += LF_THREADS) Hout[(size_t)f * n * n + e] = dyn[e];
        return;
    }
    for (int j = tid; j < n; j += LF_THREADS) {
        double v = 0.0;
        if (j >= Kl && j < Kl + 6) v = th[L.off_pose + j - Kl];
        S.z[j] = v;
    }
    __syncthreads();
    // set_minimal_slack: s = 5 - min G at s = 0
    lf_eval(F, S.z, false, 0.0, S);
    if (tid == 0) {
        double lo = S.gmin[0];
        for (int i = 1; i < NLMK; ++i) lo = (S.gmin[i] != S.gmin[i] || lo != lo) ? NAN : fmin(lo, S.gmin[i]);
        S.z[np] = 5.0 - lo;
    }
    __syncthreads();
    const double slack0 = S.z[np];
    int iterations = 0, rounds = 0, accepted = 0;
    bool success = false;
    double tau_b = 10.0;
    for (int round = 0; round < 5 && !success; ++round) {
        tau_b *= 10.0;
        ++rounds;
        for (int it = 0; it < 100; ++it) {
            if (S.z[np] < -0.01) { success = true; break; }
            const double obj = lf_eval(F, S.z, true, tau_b, S);
            lf_terms(F, S.z, tau_b, dyn, S);
            ++iterations;
            if (!lf_solve(n, dyn, S)) break;
            bool stop = false;
            double t = 1.0;
            for (int inner = 0; inner < 100; ++inner) {
                if (t < 0.01) { stop = true; break; }
                for (int j = tid; j < n; j += LF_THREADS) S.zt[j] = S.z[j] + t * S.d[j];
                __syncthreads();
                const double ot = lf_eval(F, S.zt, false, tau_b, S);
                if (ot < obj) {
                    for (int j = tid; j < n; j += LF_THREADS) S.z[j] = S.zt[j];
                    __syncthreads();
                    ++accepted;
                    break;
                }
                t *= 0.5;
            }
            if (stop) break;
        }
    }
    if (success) {
        for (int k = tid; k < Kl; k += LF_THREADS) th[L.off_exp + k] = (float)S.z[k];
        if (tid < 6) th[L.off_pose + tid] = (float)S.z[Kl + tid];
    }
    if (tid == 0) {
        LmkRep r;
        r.status = success ? TDI_OK : TDI_LMK_INFEASIBLE;
        r.iterations = iterations;
        r.rounds = rounds;
        r.accepted = accepted;
        r.slack0 = (float)slack0;
        r.slack = (float)S.z[np];
        rep[f] = r;
    }
}

} // namespace

int launch_lmk_slack(const ModelGPU& m, const Layout& L, const FrameGPU* frames, const float* lmk, int n,
                     float* theta, int Kl, int mode, const double* zin, double tau, double* H, double* g,
                     double* obj, void* rep, hipStream_t s)
{
    if (Kl < 1 || Kl > L.nexp || Kl + 7 > LF_NMAX) {
        set_error("landmark slack fit: k_exp = %d, expected 1..min(%d, %d)", Kl, L.nexp, LF_NMAX - 7);
        return TDI_ERR_ARG;
    }
    if (n <= 0) return TDI_OK;
    const int npp = (Kl + 6 + 3) & ~3;
    const size_t lds = sizeof(double) * std::max((size_t)LF_ROWS * npp, (size_t)(Kl + 7) * (Kl + 7));
    if (lds > 64 * 1024) {
        static const hipError_t attr = hipFuncSetAttribute(
            reinterpret_cast<const void*>(k_lmk_slack), hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)(sizeof(double) * LF_ROWS * LF_NPP));
        TDI_HIP(attr);
    }
    hipLaunchKernelGGL(k_lmk_slack, dim3(n), dim3(LF_THREADS), lds, s, m, L, frames, lmk, theta, Kl, mode, zin, tau,
                       H, g, obj, reinterpret_cast<LmkRep*>(rep));
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
