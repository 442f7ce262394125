// lmkjoint.hip — the terms of the joint landmark-only fit over a group of frames, FP64.
//
// Logbarrier_Initializer::compute_gradient_and_hessian / evaluate_objective_function over T
// frames with identity unknowns (logbarrier_initializer.cu:1272-1842, 1921-2081), as
// fit_to_multi_images_landmarks_only (model_fitter.cu:1205-1247) and the two landmark fits at the
// start of fit_to_multi_images (1291-1330) run them: the frames share the first Ka identity
// coefficients and the slack s, each frame has its own Ke expression coefficients and pose.
//
// k_lmk_joint: one workgroup per frame.  The frame's unknowns in its own order
//   q = [alpha (Ka) | s (S1 = 0 or 1) | eps (Ke) | tau (3) | u (3)],  nf = Ka + S1 + Ke + 6,
// read from the joint vector z = [alpha | s | frame 0's (eps, tau, u) | frame 1's | ...].  It
// writes the frame's objective (landmark barrier terms and its expression bounds), the least
// constraint value, and (mode 1) its gradient and Gauss-Newton matrix over q, FP64, with the
// landmark derivatives [102][nf] in global scratch (too many columns for LDS at 199 identity
// coefficients).  The terms counted once for the group (the identity bounds, tau_b s and its
// gradient) are k_lmk_joint_shared's; the frames' matrices are placed on the joint one by
// launch_assemble_joint (shared block first).
#include <cmath>

#include "common.h"

namespace tdi {

namespace {

constexpr int LJ_THREADS = 256;
constexpr int LJ_ROWS = 2 * NLMK;

struct LjShared {
    double q[LMK_JOINT_MAXN], hs[LMK_JOINT_MAXN];
    double base[NLMK][3], p[NLMK][3];
    double R[9], dR[27];
    double cx[NLMK], cy[NLMK], wx[NLMK], wy[NLMK], hx[NLMK], hy[NLMK], qg[NLMK], qh[NLMK], gmin[NLMK];
    double cz[NLMK][3];
    double red[LJ_THREADS / 64];
};

// each: every frame is a group of its own, its unknowns z [f][nf] in its own order, and mask [n]
// selects the frames to evaluate (null: all); theta0 is then each frame's own row [n][K].
__global__ void __launch_bounds__(LJ_THREADS) k_lmk_joint(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                          const float* __restrict__ lmk, const float* __restrict__ theta0,
                                                          int n, int Ka, int Ke, int S1, const double* __restrict__ z,
                                                          int mode, int each, const int* __restrict__ mask,
                                                          double* __restrict__ objf,
                                                          double* __restrict__ gminf, double* __restrict__ H,
                                                          double* __restrict__ g, double* __restrict__ J)
{
    const int f = blockIdx.x;
    if (f >= n || (mask && !mask[f])) return;
    __shared__ LjShared S;
    const int tid = threadIdx.x;
    const int Ks = Ka + S1, oe = Ks, op = Ks + Ke, nf = op + 6, nfp = (nf + 3) & ~3;
    const size_t zf = (size_t)Ks + (size_t)f * (Ke + 6);
    for (int j = tid; j < nf; j += LJ_THREADS)
        S.q[j] = each ? z[(size_t)f * nf + j] : (j < Ks ? z[j] : z[zf + (j - Ks)]);
    if (each) theta0 += (size_t)f * L.K;
    const FrameGPU fr = frames[f];
    const float* lm = lmk + (size_t)f * 6 * NLMK;
    // landmark positions without the fitted columns: the mean plus the identity columns past Ka
    // at theta's values
    for (int r = tid; r < 3 * NLMK; r += LJ_THREADS) {
        const int i = r / 3, a = r - 3 * i;
        const size_t row = (size_t)(3 * m.lmk_ids[i] + a);
        double x = m.mean[row];
        for (int k = Ka; k < L.nid; ++k) x += (double)m.id_basis[row * m.id_ld + k] * (double)theta0[L.off_id + k];
        S.base[i][a] = x;
    }
    __syncthreads();
    if (tid == 0) {
        if (mode) rodrigues_f64_d(S.q + op + 3, S.R, S.dR);
        else rodrigues_f64(S.q + op + 3, S.R);
    }
    for (int r = tid; r < 3 * NLMK; r += LJ_THREADS) {
        const int i = r / 3, a = r - 3 * i;
        const size_t row = (size_t)(3 * m.lmk_ids[i] + a);
        double x = S.base[i][a];
        for (int k = 0; k < Ka; ++k) x = fma((double)m.id_basis[row * m.id_ld + k], S.q[k], x);
        for (int k = 0; k < Ke; ++k) x = fma((double)m.exp_basis[row * m.exp_ld + k], S.q[oe + k], x);
        S.p[i][a] = x;
    }
    __syncthreads();
    double acc = 0.0;
    if (tid < NLMK) {
        const int i = tid;
        const double s = S1 ? S.q[Ka] : 0.0;
        double c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = S.R[a] * S.p[i][0] + S.R[3 + a] * S.p[i][1] + S.R[6 + a] * S.p[i][2] + S.q[op + a];
        S.cz[i][0] = c[0]; S.cz[i][1] = c[1]; S.cz[i][2] = c[2];
        const double xp = fr.phix * c[0] / c[2] + fr.cx, yp = fr.phiy * c[1] / c[2] + fr.cy;
        const double fs = fr.face_size;
        const double lx = lm[i], ly = lm[NLMK + i];
        const double Gxm = s + xp - (lx + (double)lm[2 * NLMK + i] * fs);
        const double Gxp = s + (lx + (double)lm[3 * NLMK + i] * fs) - xp;
        const double Gym = s + yp - (ly + (double)lm[4 * NLMK + i] * fs);
        const double Gyp = s + (ly + (double)lm[5 * NLMK + i] * fs) - yp;
        acc = -log(Gxm) - log(Gym) - log(Gxp) - log(Gyp);
        S.gmin[i] = fmin(fmin(Gxm, Gym), fmin(Gxp, Gyp));
        if (Gxm != Gxm || Gym != Gym || Gxp != Gxp || Gyp != Gyp) S.gmin[i] = NAN;
        if (mode) {
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
    } else if (tid >= 64 && tid < 64 + Ke) {
        const int k = tid - 64;
        const float lbf = m.lb[L.off_exp + k], ubf = m.ub[L.off_exp + k];
        if (ubf < INFINITY && lbf > -INFINITY) {
            const double x = S.q[oe + k], fub = x - (double)ubf, flb = (double)lbf - x;
            acc = -log(-fub) - log(-flb);
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((tid & 63) == 0) S.red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0, lo = S.gmin[0];
        for (int w = 0; w < LJ_THREADS / 64; ++w) sum += S.red[w];
        for (int i = 1; i < NLMK; ++i) lo = (S.gmin[i] != S.gmin[i] || lo != lo) ? NAN : fmin(lo, S.gmin[i]);
        objf[f] = sum;
        gminf[f] = lo;
    }
    if (!mode) return;
    // projected-landmark derivatives: rows x_i, y_i over the frame's columns (the slack's zero)
    double* Jf = J + (size_t)f * LJ_ROWS * nfp;
    for (int e = tid; e < NLMK * nfp; e += LJ_THREADS) {
        const int i = e / nfp, j = e - i * nfp;
        double dxj = 0.0, dyj = 0.0;
        if (j < nf && !(S1 && j == Ka)) {
            double dv[3];
            if (j < Ka || (j >= oe && j < op)) {
                const int v = m.lmk_ids[i];
                double ev[3];
                if (j < Ka)
                    for (int a = 0; a < 3; ++a) ev[a] = m.id_basis[(size_t)(3 * v + a) * m.id_ld + j];
                else
                    for (int a = 0; a < 3; ++a) ev[a] = m.exp_basis[(size_t)(3 * v + a) * m.exp_ld + (j - oe)];
                for (int a = 0; a < 3; ++a) dv[a] = S.R[a] * ev[0] + S.R[3 + a] * ev[1] + S.R[6 + a] * ev[2];
            } else if (j < op + 3) {
                dv[0] = dv[1] = dv[2] = 0.0;
                dv[j - op] = 1.0;
            } else {
                const double* d = S.dR + 9 * (j - op - 3);
                for (int a = 0; a < 3; ++a) dv[a] = d[a] * S.p[i][0] + d[3 + a] * S.p[i][1] + d[6 + a] * S.p[i][2];
            }
            const double cx = S.cz[i][0], cy = S.cz[i][1], cz = S.cz[i][2];
            dxj = fr.phix * (dv[0] * cz - dv[2] * cx) / (cz * cz);
            dyj = fr.phiy * (dv[1] * cz - dv[2] * cy) / (cz * cz);
        }
        Jf[(size_t)(2 * i) * nfp + j] = dxj;
        Jf[(size_t)(2 * i + 1) * nfp + j] = dyj;
    }
    __syncthreads();
    double* Hf = H + (size_t)f * nf * nf;
    double* gf = g + (size_t)f * nf;
    // gradient and the slack column from the raw derivatives
    for (int j = tid; j < nf; j += LJ_THREADS) {
        if (S1 && j == Ka) continue;
        double gs = 0.0, hs = 0.0;
        for (int i = 0; i < NLMK; ++i) {
            const double dx = Jf[(size_t)(2 * i) * nfp + j], dy = Jf[(size_t)(2 * i + 1) * nfp + j];
            gs += S.cx[i] * dx + S.cy[i] * dy;
            hs += S.hx[i] * dx + S.hy[i] * dy;
        }
        if (j >= oe && j < op) {
            const float lbf = m.lb[L.off_exp + j - oe], ubf = m.ub[L.off_exp + j - oe];
            if (ubf < INFINITY && lbf > -INFINITY) {
                const double x = S.q[j], fub = x - (double)ubf, flb = (double)lbf - x;
                gs += -1.0 / fub + 1.0 / flb;
            }
        }
        gf[j] = gs;
        S.hs[j] = hs;
    }
    if (S1 && tid == 0) {
        double qs = 0.0, hss = 0.0;
        for (int i = 0; i < NLMK; ++i) { qs += S.qg[i]; hss += S.qh[i]; }
        gf[Ka] = -qs;
        S.hs[Ka] = hss;
    }
    __syncthreads();
    for (int e = tid; e < LJ_ROWS * nfp; e += LJ_THREADS) {
        const int r = e / nfp, i = r >> 1;
        Jf[e] *= sqrt((r & 1) ? S.wy[i] : S.wx[i]);
    }
    __syncthreads();
    // 4 x 4 tiles (ta <= tb) of the nf x nf block
    const int T = nfp / 4, ntiles = T * (T + 1) / 2;
    for (int t = tid; t < ntiles; t += LJ_THREADS) {
        int ta = 0, rem = t;
        while (rem >= T - ta) { rem -= T - ta; ++ta; }
        const int tb = ta + rem;
        double tacc[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) tacc[u] = 0.0;
        const double* ca = Jf + 4 * ta;
        const double* cb = Jf + 4 * tb;
        for (int r = 0; r < LJ_ROWS; ++r) {
            const double2 a01 = *reinterpret_cast<const double2*>(ca + (size_t)r * nfp);
            const double2 a23 = *reinterpret_cast<const double2*>(ca + (size_t)r * nfp + 2);
            const double2 b01 = *reinterpret_cast<const double2*>(cb + (size_t)r * nfp);
            const double2 b23 = *reinterpret_cast<const double2*>(cb + (size_t)r * nfp + 2);
            const double av[4] = {a01.x, a01.y, a23.x, a23.y}, bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) tacc[4 * x + y] = fma(av[x], bv[y], tacc[4 * x + y]);
        }
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
                const int a = 4 * ta + x, b = 4 * tb + y;
                if (a >= nf || b >= nf) continue;
                Hf[(size_t)a * nf + b] = tacc[4 * x + y];
                Hf[(size_t)b * nf + a] = tacc[4 * x + y];
            }
    }
    __syncthreads();
    for (int j = tid; j < nf; j += LJ_THREADS) {
        if (S1) {
            if (j == Ka) {
                Hf[(size_t)Ka * nf + Ka] = S.hs[Ka];
            } else {
                Hf[(size_t)j * nf + Ka] = S.hs[j];
                Hf[(size_t)Ka * nf + j] = S.hs[j];
            }
        }
        if (j >= oe && j < op) {
            const float lbf = m.lb[L.off_exp + j - oe], ubf = m.ub[L.off_exp + j - oe];
            if (ubf < INFINITY && lbf > -INFINITY) {
                const double x = S.q[j], fub = x - (double)ubf, flb = (double)lbf - x;
                const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
                Hf[(size_t)j * nf + j] += a * a + b * b;
            }
        }
    }
}

// The terms counted once for the group, at the joint vector z: the identity bounds (the
// identity coefficients' bounds of the model) and tau_b s.  obj[0] = their objective; mode 1
// also adds their gradient and Hessian diagonal to gJ and HJ (the gradient of tau_b s is tau_b).
// With a grid of n workgroups (every frame a group of its own) workgroup f works on z, obj, HJ
// and gJ offset by f (strides KJ, 1, KJ^2, KJ), skipping frames whose mask is 0.
__global__ void __launch_bounds__(256) k_lmk_joint_shared(ModelGPU m, Layout L, int Ka, int S1, long long KJ,
                                                          const double* __restrict__ z, double tau_b,
                                                          const double* __restrict__ taus, int mode,
                                                          const int* __restrict__ mask,
                                                          double* __restrict__ obj, double* __restrict__ HJ,
                                                          double* __restrict__ gJ)
{
    __shared__ double sred[4];
    const int tid = threadIdx.x, f = blockIdx.x;
    if (mask && !mask[f]) return;
    if (taus) tau_b = taus[f];   // each group's own barrier weight
    z += (size_t)f * KJ;
    obj += f;
    if (mode) {
        HJ += (size_t)f * KJ * KJ;
        gJ += (size_t)f * KJ;
    }
    double acc = 0.0;
    for (int k = tid; k < Ka; k += blockDim.x) {
        const float lbf = m.lb[L.off_id + k], ubf = m.ub[L.off_id + k];
        if (!(ubf < INFINITY) || !(lbf > -INFINITY)) continue;
        const double x = z[k], fub = x - (double)ubf, flb = (double)lbf - x;
        acc += -log(-fub) - log(-flb);
        if (mode) {
            gJ[k] += -1.0 / fub + 1.0 / flb;
            const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
            HJ[(size_t)k * KJ + k] += a * a + b * b;
        }
    }
    if (S1 && tid == 0) {
        acc += tau_b * z[Ka];
        if (mode) gJ[Ka] += tau_b;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((tid & 63) == 0) sred[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) obj[0] = sred[0] + sred[1] + sred[2] + sred[3];
}

} // namespace

size_t lmk_joint_scratch_doubles(int n, int Ka, int Ke, int S1)
{
    const int nf = Ka + S1 + Ke + 6, nfp = (nf + 3) & ~3;
    return (size_t)n * LJ_ROWS * nfp;
}

int launch_lmk_joint(const ModelGPU& m, const Layout& L, const FrameGPU* frames, const float* lmk,
                     const float* theta0, int n, int Ka, int Ke, int S1, const double* z, int mode, int each,
                     const int* mask, double* objf, double* gminf, double* H, double* g, double* J, hipStream_t s)
{
    if (Ka < 0 || Ka > L.nid || Ke < 0 || Ke > L.nexp || Ke > LJ_THREADS - 64 || Ka + S1 + Ke + 6 > LMK_JOINT_MAXN) {
        set_error("joint landmark fit: k_id %d (of %d), k_exp %d (of %d, at most %d), %d unknowns per frame (at most %d)",
                  Ka, L.nid, Ke, L.nexp, LJ_THREADS - 64, Ka + S1 + Ke + 6, LMK_JOINT_MAXN);
        return TDI_ERR_ARG;
    }
    if (n <= 0) return TDI_OK;
    hipLaunchKernelGGL(k_lmk_joint, dim3(n), dim3(LJ_THREADS), 0, s, m, L, frames, lmk, theta0, n, Ka, Ke, S1, z, mode,
                       each, mask, objf, gminf, H, g, J);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_lmk_joint_shared(const ModelGPU& m, const Layout& L, int Ka, int S1, long long KJ, const double* z,
                            double tau_b, const double* taus, int mode, int groups, const int* mask, double* obj,
                            double* HJ, double* gJ, hipStream_t s)
{
    hipLaunchKernelGGL(k_lmk_joint_shared, dim3(groups), dim3(256), 0, s, m, L, Ka, S1, KJ, z, tau_b, taus, mode, mask,
                       obj, HJ, gJ);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
