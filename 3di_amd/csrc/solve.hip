// solve.hip — log barrier, batched Cholesky and the Gauss-Newton/line-search state machine.
//
// Reference:
//   Logbarrier_Initializer::compute_gradient_and_hessian (logbarrier_initializer.cu:1272-1842)
//   with compute_Gs (458-518), fill_optimization_dc_landmark/_dI_depsilon_landmark/
//   _dI_dalpha_landmark (587-910), update_gradient/update_diagonal_of_hessian_wbounds
//   (915-948) and evaluate_objective_function (1921-2081): landmark-box constraints
//   G = box edge distance of the projected landmark, objective -sum log G, gradient
//   sum -dG/G, Gauss-Newton Hessian sum (dG/G)(dG/G)^T; coefficient bounds
//   -log(ub-x) - log(x-lb) with exact gradient and diagonal Hessian.
//   Solver::solve (solver.h:70-105): cusolverDnDpotrf + Dpotrs in FP64, delta = -H^-1 g.
//   fit_3DMM_shape_rigid (model_fitter.cu:121-375): Armijo backtracking with ALPHA 0.4,
//   BETA 0.8, stop when -delta.g < 3e-4 or t < t_min, reject steps with tau_z < 0, shrink
//   on "face too large" or a NaN barrier.
//
// The batched state machine keeps every per-frame decision of the reference but runs it
// for all frames of a batch with masks, so the whole fit is a fixed launch sequence
// (graph-capturable): frames that have accepted a step or stopped sit out later rounds.
#include "common.h"

namespace tdi {

namespace {

constexpr int BMAXK = 112;

// mode 0: objective only -> obj[f] = objb[f] = barrier value
// mode 1: derivatives only -> H[f] += barrier Hessian, g[f] += barrier gradient
// Everything is FP64 (rotation and its derivatives, landmark positions, constraint values,
// weights, projected-landmark derivatives, sums, H); the FP32 g receives one rounding.
// Dynamic LDS: the projected-landmark derivatives dx, dy as double [51][K] each, scaled by the
// square roots of the Gauss-Newton weights, so that the landmark block is sum_i ex ex^T + ey ey^T.
__global__ void __launch_bounds__(256) k_barrier(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                  const int* __restrict__ act, int n, const float* __restrict__ theta,
                                                  LmkArrays lm, int use_lmk, int mode, double* __restrict__ obj,
                                                  double* __restrict__ objb, double* __restrict__ H,
                                                  float* __restrict__ g)
{
    // one workgroup per listed frame (the grid has a workgroup per slot)
    if ((int)blockIdx.x >= fl_count(act, n)) return;
    const int f = fl_frame(act, blockIdx.x);
    const int K = L.K;
    const FrameGPU fr = frames[f];
    const float* tf = theta + (size_t)f * K;
    extern __shared__ double sdyn[];
    double* sdx = sdyn;                    // [NLMK][K], sqrt(wx_i) dx_i/dtheta
    double* sdy = sdyn + (size_t)NLMK * K; // [NLMK][K]
    __shared__ double sR[9], sdR[27];
    __shared__ double sv[NLMK][3], sp[NLMK][3];
    __shared__ double swx[NLMK], swy[NLMK], sgx[NLMK], sgy[NLMK];
    __shared__ double sred[4];
    const int tid = threadIdx.x;
    if (tid == 0) rodrigues_f64_d(tf + L.off_pose + 3, sR, sdR);
    __syncthreads();
    double acc = 0.0;
    // coefficient bounds (only finite bounds take part)
    for (int j = tid; j < K; j += blockDim.x) {
        const float lbf = m.lb[j], ubf = m.ub[j];
        if (!(ubf < INFINITY) || !(lbf > -INFINITY)) continue;
        const double x = tf[j], lb = lbf, ub = ubf;
        const double fub = x - ub, flb = lb - x;
        if (mode == 0) {
            acc += -log(-fub) - log(-flb);
        } else {
            g[(size_t)f * K + j] += (float)(-1.0 / fub + 1.0 / flb);
            const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
            H[(size_t)f * K * K + (size_t)j * K + j] += a * a + b * b;
        }
    }
    if (use_lmk) {
        // the model-space landmark positions, one thread per coordinate
        if (tid < 3 * NLMK) {
            const int i = tid / 3, a = tid - 3 * i, r = 3 * m.lmk_ids[i] + a;
            double s = m.mean[r];
            for (int k = 0; k < L.nexp; ++k) s = fma((double)m.exp_basis[(size_t)r * m.exp_ld + k], (double)tf[L.off_exp + k], s);
            for (int k = 0; k < L.nid; ++k) s = fma((double)m.id_basis[(size_t)r * m.id_ld + k], (double)tf[L.off_id + k], s);
            sp[i][a] = s;
        }
        __syncthreads();
        if (tid < NLMK) {
            const int i = tid;
            const double p[3] = {sp[i][0], sp[i][1], sp[i][2]};
            const double vx = sR[0] * p[0] + sR[3] * p[1] + sR[6] * p[2] + tf[L.off_pose + 0];
            const double vy = sR[1] * p[0] + sR[4] * p[1] + sR[7] * p[2] + tf[L.off_pose + 1];
            const double vz = sR[2] * p[0] + sR[5] * p[1] + sR[8] * p[2] + tf[L.off_pose + 2];
            sv[i][0] = vx; sv[i][1] = vy; sv[i][2] = vz;
            const double xproj = fr.phix * vx / vz + fr.cx;
            const double yproj = fr.phiy * vy / vz + fr.cy;
            const float* L6 = lm.base + (size_t)f * 6 * NLMK;
            const double lx = L6[i], ly = L6[NLMK + i];
            const double blx = L6[2 * NLMK + i], bux = L6[3 * NLMK + i];
            const double bly = L6[4 * NLMK + i], buy = L6[5 * NLMK + i];
            const double fs = fr.face_size;
            const double Gxm = xproj - (lx + blx * fs);
            const double Gym = yproj - (ly + bly * fs);
            const double Gxp = (lx + bux * fs) - xproj;
            const double Gyp = (ly + buy * fs) - yproj;
            if (mode == 0) {
                acc += -log(Gxm) - log(Gym) - log(Gxp) - log(Gyp);
            } else {
                swx[i] = 1.0 / (Gxm * Gxm) + 1.0 / (Gxp * Gxp);
                swy[i] = 1.0 / (Gym * Gym) + 1.0 / (Gyp * Gyp);
                sgx[i] = -1.0 / Gxm + 1.0 / Gxp;
                sgy[i] = -1.0 / Gym + 1.0 / Gyp;
            }
        }
        __syncthreads();
        if (mode == 1) {
            // projected-landmark gradients dx[i][j], dy[i][j]
            for (int e = tid; e < NLMK * K; e += blockDim.x) {
                const int i = e / K, j = e - i * K;
                double dv[3] = {0.0, 0.0, 0.0};
                bool geo = true;
                if (j >= L.off_pose && j < L.off_pose + 3) {
                    dv[j - L.off_pose] = 1.0;
                } else if (j >= L.off_pose + 3 && j < L.off_light) {
                    const double* d = sdR + 9 * (j - L.off_pose - 3);
                    for (int a = 0; a < 3; ++a) dv[a] = d[a] * sp[i][0] + d[3 + a] * sp[i][1] + d[6 + a] * sp[i][2];
                } else if (j >= L.off_exp && j < L.off_pose) {
                    const int v = m.lmk_ids[i], c = j - L.off_exp;
                    double e3[3];
                    for (int a = 0; a < 3; ++a) e3[a] = m.exp_basis[(size_t)(3 * v + a) * m.exp_ld + c];
                    for (int a = 0; a < 3; ++a) dv[a] = sR[a] * e3[0] + sR[3 + a] * e3[1] + sR[6 + a] * e3[2];
                } else if (j >= L.off_id && j < L.off_exp) {
                    const int v = m.lmk_ids[i], c = j - L.off_id;
                    double e3[3];
                    for (int a = 0; a < 3; ++a) e3[a] = m.id_basis[(size_t)(3 * v + a) * m.id_ld + c];
                    for (int a = 0; a < 3; ++a) dv[a] = sR[a] * e3[0] + sR[3 + a] * e3[1] + sR[6 + a] * e3[2];
                } else {
                    geo = false;
                }
                double dxj = 0.0, dyj = 0.0;
                if (geo) {
                    const double vx = sv[i][0], vy = sv[i][1], vz = sv[i][2];
                    dxj = fr.phix * (dv[0] * vz - dv[2] * vx) / (vz * vz);
                    dyj = fr.phiy * (dv[1] * vz - dv[2] * vy) / (vz * vz);
                }
                // gradient: sum_i gx_i dx_i; the Hessian uses sqrt(w) dx
                sdx[(size_t)i * K + j] = dxj;
                sdy[(size_t)i * K + j] = dyj;
            }
            __syncthreads();
            for (int j = tid; j < K; j += blockDim.x) {
                double s = 0.0;
                for (int i = 0; i < NLMK; ++i) s += sgx[i] * sdx[(size_t)i * K + j] + sgy[i] * sdy[(size_t)i * K + j];
                g[(size_t)f * K + j] += (float)s;
            }
            __syncthreads();
            for (int e = tid; e < NLMK * K; e += blockDim.x) {
                const int i = e / K;
                sdx[e] *= sqrt(swx[i]);
                sdy[e] *= sqrt(swy[i]);
            }
            __syncthreads();
            for (int e = tid; e < K * K; e += blockDim.x) {
                const int a = e / K, b = e - a * K;
                if (b < a) continue;
                double s = 0.0;
                for (int i = 0; i < NLMK; ++i)
                    s += sdx[(size_t)i * K + a] * sdx[(size_t)i * K + b] + sdy[(size_t)i * K + a] * sdy[(size_t)i * K + b];
                H[(size_t)f * K * K + e] += s;
                if (b != a) H[(size_t)f * K * K + (size_t)b * K + a] += s;
            }
        }
    }
    if (mode == 0) {
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((tid & 63) == 0) sred[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            const double s = sred[0] + sred[1] + sred[2] + sred[3];
            obj[f] = s;
            objb[f] = s;
        }
    }
}

// ------------------------------------------------------------------ batched Cholesky
__device__ __forceinline__ int pk(int i, int j) { return i * (i + 1) / 2 + j; }   // i >= j

template <typename T>
__device__ bool chol_packed(T* A, int K, T* sflag)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = 0; j < K; ++j) {
        if (tid == 0) {
            const T d = A[pk(j, j)];
            if (!(d > T(0))) sflag[0] = T(1);
            A[pk(j, j)] = sqrt(d);
        }
        __syncthreads();
        if (sflag[0] != T(0)) return false;
        const T djj = A[pk(j, j)];
        for (int i = j + 1 + tid; i < K; i += nt) A[pk(i, j)] /= djj;
        __syncthreads();
        const int m = K - j - 1;
        const int npairs = m * (m + 1) / 2;
        for (int e = tid; e < npairs; e += nt) {
            // e -> (ii, kk) with 0 <= kk <= ii < m
            int ii = (int)((sqrtf(8.f * e + 1.f) - 1.f) * 0.5f);
            while ((ii + 1) * (ii + 2) / 2 <= e) ++ii;
            while (ii * (ii + 1) / 2 > e) --ii;
            const int kk = e - ii * (ii + 1) / 2;
            const int i = j + 1 + ii, k = j + 1 + kk;
            A[pk(i, k)] -= A[pk(i, j)] * A[pk(k, j)];
        }
        __syncthreads();
    }
    return true;
}

// Solve L L^T x = b in place (x overwrites b), b in shared memory.
template <typename T>
__device__ void chol_solve_packed(const T* A, int K, T* b)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = 0; i < K; ++i) {
        if (tid == 0) b[i] /= A[pk(i, i)];
        __syncthreads();
        const T yi = b[i];
        for (int k = i + 1 + tid; k < K; k += nt) b[k] -= A[pk(k, i)] * yi;
        __syncthreads();
    }
    for (int i = K - 1; i >= 0; --i) {
        if (tid == 0) b[i] /= A[pk(i, i)];
        __syncthreads();
        const T xi = b[i];
        for (int k = tid; k < i; k += nt) b[k] -= A[pk(i, k)] * xi;
        __syncthreads();
    }
}

constexpr int CMAXK = 112;

// mode 0: FP64 factorisation and solve.  mode 1: FP32 factorisation of the equilibrated
// matrix and solve, then one FP64 residual-refinement step (r = -g - H x computed in FP64,
// correction solved with the FP32 factor); when the FP32 factorisation breaks down or the
// correction exceeds 1e-3 of the solution (an ill-conditioned frame), the FP64 factorisation.
template <typename TH>
__global__ void __launch_bounds__(256) k_cholesky(int K, const TH* __restrict__ H, const float* __restrict__ g,
                                                   float* __restrict__ delta, int* __restrict__ status,
                                                   const int* __restrict__ mask, int mode)
{
    const int f = blockIdx.x;
    if (mask && !mask[f]) return;
    extern __shared__ double dsm[];
    __shared__ double sflag[1];
    __shared__ double sb[CMAXK];
    __shared__ float sbf[CMAXK];
    __shared__ float sflagf[1];
    const TH* Hf = H + (size_t)f * K * K;
    const float* gf = g + (size_t)f * K;
    const int tid = threadIdx.x;
    const int np = K * (K + 1) / 2;
    if (tid == 0) { sflag[0] = 0.0; sflagf[0] = 0.f; }
    if (mode != 0) {
        // FP32 factorisation of the equilibrated matrix D^-1/2 H D^-1/2 (D = diag H), which
        // removes the scale spread between pose and coefficient columns, then one FP64
        // residual-refinement step on the original system.  A frame whose scaled matrix is
        // not positive definite in FP32 falls through to the FP64 factorisation below.
        float* A = (float*)dsm;
        __shared__ float sdinv[CMAXK];
        for (int i = tid; i < K; i += blockDim.x) {
            const float d = (float)Hf[(size_t)i * K + i];
            sdinv[i] = d > 0.f ? rsqrtf(d) : 1.f;
        }
        __syncthreads();
        for (int e = tid; e < np; e += blockDim.x) {
            int i = (int)((sqrtf(8.f * e + 1.f) - 1.f) * 0.5f);
            while ((i + 1) * (i + 2) / 2 <= e) ++i;
            while (i * (i + 1) / 2 > e) --i;
            const int j = e - i * (i + 1) / 2;
            A[e] = (float)Hf[(size_t)i * K + j] * sdinv[i] * sdinv[j];
        }
        for (int i = tid; i < K; i += blockDim.x) sbf[i] = -gf[i] * sdinv[i];
        __syncthreads();
        if (chol_packed<float>(A, K, sflagf)) {
            chol_solve_packed<float>(A, K, sbf);
            // x0 = D^-1/2 y
            for (int i = tid; i < K; i += blockDim.x) sbf[i] *= sdinv[i];
            __syncthreads();
            // refinement: r = -g - H x0 in FP64, correction solved with the same scaled factor
            for (int i = tid; i < K; i += blockDim.x) {
                double r = -(double)gf[i];
                for (int k = 0; k < K; ++k) r -= (double)Hf[(size_t)i * K + k] * (double)sbf[k];
                sb[i] = r * (double)sdinv[i];
            }
            __syncthreads();
            for (int i = 0; i < K; ++i) {
                if (tid == 0) sb[i] /= (double)A[pk(i, i)];
                __syncthreads();
                const double yi = sb[i];
                for (int k = i + 1 + tid; k < K; k += blockDim.x) sb[k] -= (double)A[pk(k, i)] * yi;
                __syncthreads();
            }
            for (int i = K - 1; i >= 0; --i) {
                if (tid == 0) sb[i] /= (double)A[pk(i, i)];
                __syncthreads();
                const double xi = sb[i];
                for (int k = tid; k < i; k += blockDim.x) sb[k] -= (double)A[pk(i, k)] * xi;
                __syncthreads();
            }
            // the correction measures the FP32 solution's error (about cond * 2^-24 of it); past
            // 1e-3 of the solution one refinement step leaves more than ~1e-6, and the FP64
            // factorisation below takes over
            if (tid == 0) {
                double cn = 0.0, xn = 0.0;
                for (int i = 0; i < K; ++i) {
                    const double c = sb[i] * (double)sdinv[i];
                    cn += c * c;
                    xn += (double)sbf[i] * (double)sbf[i];
                }
                sflag[0] = (cn <= 1e-6 * xn) ? 0.0 : 1.0;
            }
            __syncthreads();
            if (sflag[0] == 0.0) {
                for (int i = tid; i < K; i += blockDim.x)
                    delta[(size_t)f * K + i] = (float)((double)sbf[i] + sb[i] * (double)sdinv[i]);
                if (tid == 0) status[f] = TDI_OK;
                return;
            }
            __syncthreads();
            if (tid == 0) sflag[0] = 0.0;
        }
        __syncthreads();
    }
    // FP64 factorisation and solve (mode 0, and the fallback of mode 1)
    double* A = dsm;
    for (int e = tid; e < np; e += blockDim.x) {
        int i = (int)((sqrtf(8.f * e + 1.f) - 1.f) * 0.5f);
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        while (i * (i + 1) / 2 > e) --i;
        const int j = e - i * (i + 1) / 2;
        A[e] = (double)Hf[(size_t)i * K + j];
    }
    for (int i = tid; i < K; i += blockDim.x) sb[i] = -(double)gf[i];
    __syncthreads();
    if (!chol_packed<double>(A, K, sflag)) {
        if (tid == 0) status[f] = TDI_NOT_PD;
        return;
    }
    chol_solve_packed<double>(A, K, sb);
    for (int i = tid; i < K; i += blockDim.x) delta[(size_t)f * K + i] = (float)sb[i];
    if (tid == 0) status[f] = TDI_OK;
}

// ------------------------------------------------------------------ line-search state machine
__global__ void k_ls_init(int n, int K, FitState* __restrict__ st, const int* __restrict__ chol_status,
                          const float* __restrict__ delta, const float* __restrict__ g, const double* __restrict__ obj,
                          int* __restrict__ searching)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    FitState s = st[f];
    searching[f] = 0;
    if (!s.active) return;
    s.iters++;
    s.fresh = 1;              // the full evaluation just rendered theta
    if (chol_status[f] != TDI_OK) {
        s.active = 0;
        s.stop |= TDI_STOP_NOT_PD;
        s.status = TDI_NOT_PD;
    } else if (isnan(delta[(size_t)f * K])) {
        s.active = 0;
        s.stop |= TDI_STOP_NAN;
        s.status = TDI_NAN;
    } else {
        double dg = 0.0;
        for (int k = 0; k < K; ++k) dg += (double)delta[(size_t)f * K + k] * (double)g[(size_t)f * K + k];
        s.dg = (float)dg;
        s.obj0 = obj[f];
        s.t = 1.f;
        s.inner = 0;
        s.searching = 1;
        searching[f] = 1;
    }
    st[f] = s;
}

// Trials theta + t delta; unknowns [hold0, hold1) keep theta's value (the finetune fit's pose,
// model_fitter.cu:502).
__global__ void k_ls_propose(int n, int K, int off_tz, float t_min, FitState* __restrict__ st,
                             const float* __restrict__ theta, const float* __restrict__ delta,
                             float* __restrict__ theta_try, int* __restrict__ searching, int* __restrict__ evalmask,
                             int hold0, int hold1)
{
    const int f = blockIdx.x;
    if (f >= n) return;
    __shared__ int go;
    if (threadIdx.x == 0) {
        FitState s = st[f];
        int ev = 0;
        if (searching[f]) {
            if (s.t < t_min) {
                s.active = 0;
                s.searching = 0;
                s.stop |= TDI_STOP_STEP;
                searching[f] = 0;
            } else {
                const bool held = off_tz >= hold0 && off_tz < hold1;
                const float tz = theta[(size_t)f * K + off_tz] + (held ? 0.f : s.t * delta[(size_t)f * K + off_tz]);
                if (tz < 0.f) {
                    s.searching = 0;
                    searching[f] = 0;
                } else {
                    ev = 1;
                }
            }
            st[f] = s;
        }
        evalmask[f] = ev;
        go = ev;
    }
    __syncthreads();
    if (!go) return;
    const float t = st[f].t;
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        theta_try[(size_t)f * K + k] = (k >= hold0 && k < hold1) ? theta[(size_t)f * K + k]
                                                                 : theta[(size_t)f * K + k] + t * delta[(size_t)f * K + k];
}

__global__ void k_ls_decide(int n, int K, FitState* __restrict__ st, const int* __restrict__ evalmask,
                            int* __restrict__ searching, const double* __restrict__ obj, const double* __restrict__ objb,
                            const int* __restrict__ face_ok, float* __restrict__ theta,
                            const float* __restrict__ theta_try, tdi_fit_options o)
{
    const int f = blockIdx.x;
    if (f >= n || !evalmask[f]) return;
    __shared__ int accept;
    if (threadIdx.x == 0) {
        FitState s = st[f];
        int acc = 0;
        if (!face_ok[f] || isnan(objb[f])) {
            s.t *= o.armijo_beta;
        } else {
            s.inner++;
            s.ls_evals++;
            if (-s.dg < o.decrement_tol) {
                s.active = 0;
                s.searching = 0;
                s.stop |= TDI_STOP_DECREMENT;
            } else if (obj[f] < s.obj0 + (double)(o.armijo_alpha * s.t * s.dg)) {
                acc = 1;
                s.searching = 0;
            } else {
                s.t *= o.armijo_beta;
            }
            if (s.inner >= o.max_ls) s.searching = 0;
        }
        s.fresh = acc;        // the trial's render is in the workspace; it is theta's if accepted
        searching[f] = s.searching;
        st[f] = s;
        accept = acc;
    }
    __syncthreads();
    if (!accept) return;
    for (int k = threadIdx.x; k < K; k += blockDim.x) theta[(size_t)f * K + k] = theta_try[(size_t)f * K + k];
}

// "Face too large" test of compute_nonrigid_shape2 (renderer.cu:836-857) on projected
// landmark vertices 0, 9 and 40.
__global__ void k_face_check(int V, const int* __restrict__ lmk_ids, const float4* __restrict__ vproj,
                             float ref_face_size, const int* __restrict__ act, int n, int* __restrict__ face_ok)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fl_count(act, n)) return;
    const int f = fl_frame(act, i);
    if (ref_face_size <= 0.f) { face_ok[f] = 1; return; }
    const float4* vp = vproj + (size_t)f * V;
    const float4 a = vp[lmk_ids[0]], b = vp[lmk_ids[9]], c = vp[lmk_ids[40]];
    const float iod = fabsf(b.x - a.x), lem = fabsf(a.y - c.y), rem = fabsf(b.y - c.y);
    const float lim = 2.f * ref_face_size;
    face_ok[f] = !(iod > lim || lem > lim || rem > lim);
}

__global__ void k_copy_mask(int n, const FitState* __restrict__ st, int* __restrict__ mask)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) mask[f] = st[f].active;
}

// Active frames whose workspace render is stale (mask_r) or that of their theta (mask_u).
__global__ void k_split_fresh(int n, const FitState* __restrict__ st, int* __restrict__ mask_r,
                              int* __restrict__ mask_u)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const FitState s = st[f];
    mask_r[f] = s.active && !s.fresh;
    mask_u[f] = s.active && s.fresh;
}

// Stream compaction of a mask into a frame list, one 1024-thread workgroup: per 1024-entry
// tile a wavefront ballot and popcount give each set entry its rank within the wave, the
// wave totals in LDS give the wave offsets, and a running count carries across tiles.
constexpr int CMP_THREADS = 1024;

__global__ void __launch_bounds__(CMP_THREADS) k_compact_frames(int n, const int* __restrict__ mask,
                                                                  int* __restrict__ act)
{
    constexpr int NW = CMP_THREADS / 64;
    __shared__ int swave[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int running = 0;
    for (int t0 = 0; t0 < n; t0 += CMP_THREADS) {
        const int f = t0 + tid;
        const int set = (f < n) && mask[f];
        const unsigned long long b = __ballot(set);
        if (lane == 0) swave[wave] = __popcll(b);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            off += (w < wave) ? swave[w] : 0;
            total += swave[w];
        }
        if (set) act[1 + running + off + __popcll(b & lt)] = f;
        running += total;
        __syncthreads();
    }
    if (tid == 0) act[0] = running;
}

// The barrier's derivative terms (k_barrier mode 1) for frames with more unknowns than the
// landmark-derivative rows fit in LDS (the identity stage): the rows go to global scratch D
// [frame][2][NLMK][K] and the Gauss-Newton sum is formed over 32 x 32 tiles of H by a grid.
// Same formulas, FP64 throughout.
__global__ void __launch_bounds__(256) k_barrier_rows(ModelGPU m, Layout L, const FrameGPU* __restrict__ frames,
                                                       const int* __restrict__ act, int n, const float* __restrict__ theta,
                                                       LmkArrays lm, int use_lmk, double* __restrict__ D,
                                                       double* __restrict__ H, float* __restrict__ g)
{
    if ((int)blockIdx.x >= fl_count(act, n)) return;
    const int f = fl_frame(act, blockIdx.x);
    const int K = L.K;
    const FrameGPU fr = frames[f];
    const float* tf = theta + (size_t)f * K;
    double* sdx = D + (size_t)f * 2 * NLMK * K;
    double* sdy = sdx + (size_t)NLMK * K;
    __shared__ double sR[9], sdR[27];
    __shared__ double sv[NLMK][3], sp[NLMK][3];
    __shared__ double swx[NLMK], swy[NLMK], sgx[NLMK], sgy[NLMK];
    const int tid = threadIdx.x;
    if (tid == 0) rodrigues_f64_d(tf + L.off_pose + 3, sR, sdR);
    __syncthreads();
    for (int j = tid; j < K; j += blockDim.x) {
        const float lbf = m.lb[j], ubf = m.ub[j];
        if (!(ubf < INFINITY) || !(lbf > -INFINITY)) continue;
        const double x = tf[j], lb = lbf, ub = ubf;
        const double fub = x - ub, flb = lb - x;
        g[(size_t)f * K + j] += (float)(-1.0 / fub + 1.0 / flb);
        const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
        H[(size_t)f * K * K + (size_t)j * K + j] += a * a + b * b;
    }
    if (!use_lmk) return;
    if (tid < 3 * NLMK) {
        const int i = tid / 3, a = tid - 3 * i, r = 3 * m.lmk_ids[i] + a;
        double sacc = m.mean[r];
        for (int k = 0; k < L.nexp; ++k) sacc = fma((double)m.exp_basis[(size_t)r * m.exp_ld + k], (double)tf[L.off_exp + k], sacc);
        for (int k = 0; k < L.nid; ++k) sacc = fma((double)m.id_basis[(size_t)r * m.id_ld + k], (double)tf[L.off_id + k], sacc);
        sp[i][a] = sacc;
    }
    __syncthreads();
    if (tid < NLMK) {
        const int i = tid;
        const double p[3] = {sp[i][0], sp[i][1], sp[i][2]};
        const double vx = sR[0] * p[0] + sR[3] * p[1] + sR[6] * p[2] + tf[L.off_pose + 0];
        const double vy = sR[1] * p[0] + sR[4] * p[1] + sR[7] * p[2] + tf[L.off_pose + 1];
        const double vz = sR[2] * p[0] + sR[5] * p[1] + sR[8] * p[2] + tf[L.off_pose + 2];
        sv[i][0] = vx; sv[i][1] = vy; sv[i][2] = vz;
        const double xproj = fr.phix * vx / vz + fr.cx;
        const double yproj = fr.phiy * vy / vz + fr.cy;
        const float* L6 = lm.base + (size_t)f * 6 * NLMK;
        const double lx = L6[i], ly = L6[NLMK + i];
        const double blx = L6[2 * NLMK + i], bux = L6[3 * NLMK + i];
        const double bly = L6[4 * NLMK + i], buy = L6[5 * NLMK + i];
        const double fs = fr.face_size;
        const double Gxm = xproj - (lx + blx * fs);
        const double Gym = yproj - (ly + bly * fs);
        const double Gxp = (lx + bux * fs) - xproj;
        const double Gyp = (ly + buy * fs) - yproj;
        swx[i] = 1.0 / (Gxm * Gxm) + 1.0 / (Gxp * Gxp);
        swy[i] = 1.0 / (Gym * Gym) + 1.0 / (Gyp * Gyp);
        sgx[i] = -1.0 / Gxm + 1.0 / Gxp;
        sgy[i] = -1.0 / Gym + 1.0 / Gyp;
    }
    __syncthreads();
    for (int e = tid; e < NLMK * K; e += blockDim.x) {
        const int i = e / K, j = e - i * K;
        double dv[3] = {0.0, 0.0, 0.0};
        bool geo = true;
        if (j >= L.off_pose && j < L.off_pose + 3) {
            dv[j - L.off_pose] = 1.0;
        } else if (j >= L.off_pose + 3 && j < L.off_light) {
            const double* d = sdR + 9 * (j - L.off_pose - 3);
            for (int a = 0; a < 3; ++a) dv[a] = d[a] * sp[i][0] + d[3 + a] * sp[i][1] + d[6 + a] * sp[i][2];
        } else if (j >= L.off_exp && j < L.off_pose) {
            const int v = m.lmk_ids[i], c = j - L.off_exp;
            double e3[3];
            for (int a = 0; a < 3; ++a) e3[a] = m.exp_basis[(size_t)(3 * v + a) * m.exp_ld + c];
            for (int a = 0; a < 3; ++a) dv[a] = sR[a] * e3[0] + sR[3 + a] * e3[1] + sR[6 + a] * e3[2];
        } else if (j >= L.off_id && j < L.off_exp) {
            const int v = m.lmk_ids[i], c = j - L.off_id;
            double e3[3];
            for (int a = 0; a < 3; ++a) e3[a] = m.id_basis[(size_t)(3 * v + a) * m.id_ld + c];
            for (int a = 0; a < 3; ++a) dv[a] = sR[a] * e3[0] + sR[3 + a] * e3[1] + sR[6 + a] * e3[2];
        } else {
            geo = false;
        }
        double dxj = 0.0, dyj = 0.0;
        if (geo) {
            const double vx = sv[i][0], vy = sv[i][1], vz = sv[i][2];
            dxj = fr.phix * (dv[0] * vz - dv[2] * vx) / (vz * vz);
            dyj = fr.phiy * (dv[1] * vz - dv[2] * vy) / (vz * vz);
        }
        sdx[(size_t)i * K + j] = dxj;
        sdy[(size_t)i * K + j] = dyj;
    }
    __syncthreads();
    for (int j = tid; j < K; j += blockDim.x) {
        double sacc = 0.0;
        for (int i = 0; i < NLMK; ++i) sacc += sgx[i] * sdx[(size_t)i * K + j] + sgy[i] * sdy[(size_t)i * K + j];
        g[(size_t)f * K + j] += (float)sacc;
    }
    __syncthreads();
    for (int e = tid; e < NLMK * K; e += blockDim.x) {
        const int i = e / K;
        sdx[e] *= sqrt(swx[i]);
        sdy[e] *= sqrt(swy[i]);
    }
}

constexpr int BH_T = 32;

// H += sum_i dx_i dx_i^T + dy_i dy_i^T over the tiles (I, J >= I) of 32 x 32; both halves written.
__global__ void __launch_bounds__(256) k_barrier_hess(int K, const int* __restrict__ act, int n,
                                                       const double* __restrict__ D, double* __restrict__ H)
{
    const int T = (K + BH_T - 1) / BH_T;
    int t = blockIdx.x, I = 0;
    while (t >= T - I) { t -= T - I; ++I; }
    const int J = I + t;
    __shared__ double sa[2][NLMK][BH_T], sb[2][NLMK][BH_T];
    const int nact = fl_count(act, n);
    for (int yi = blockIdx.y; yi < nact; yi += gridDim.y) {
        const int f = fl_frame(act, yi);
        const double* Df = D + (size_t)f * 2 * NLMK * K;
        for (int e = threadIdx.x; e < 2 * NLMK * BH_T; e += blockDim.x) {
            const int c = e % BH_T, i = (e / BH_T) % NLMK, q = e / (BH_T * NLMK);
            const int a = I * BH_T + c, b = J * BH_T + c;
            sa[q][i][c] = a < K ? Df[((size_t)q * NLMK + i) * K + a] : 0.0;
            sb[q][i][c] = b < K ? Df[((size_t)q * NLMK + i) * K + b] : 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < BH_T * BH_T; e += blockDim.x) {
            const int r = e / BH_T, c = e - r * BH_T;
            const int a = I * BH_T + r, b = J * BH_T + c;
            if (a >= K || b >= K || (I == J && c < r)) continue;
            double sacc = 0.0;
            for (int i = 0; i < NLMK; ++i) sacc += sa[0][i][r] * sb[0][i][c] + sa[1][i][r] * sb[1][i][c];
            H[(size_t)f * K * K + (size_t)a * K + b] += sacc;
            if (a != b) H[(size_t)f * K * K + (size_t)b * K + a] += sacc;
        }
        __syncthreads();
    }
}

} // namespace

int launch_barrier_big(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl, const float* theta,
                       const LmkArrays& lm, int use_landmarks, double* D, double* H, float* g, hipStream_t s)
{
    hipLaunchKernelGGL(k_barrier_rows, dim3(fl.n > 0 ? fl.n : 1), dim3(256), 0, s, m, L, frames, fl.act, fl.n, theta,
                       lm, use_landmarks, D, H, g);
    if (use_landmarks) {
        const int T = (L.K + BH_T - 1) / BH_T;
        hipLaunchKernelGGL(k_barrier_hess, dim3(T * (T + 1) / 2, grid_frames(fl.n, 1 << 16)), dim3(256), 0, s, L.K,
                           fl.act, fl.n, D, H);
    }
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_barrier(const ModelGPU& m, const Layout& L, const FrameGPU* frames, FrameList fl,
                   const float* theta, const LmkArrays& lm, int use_landmarks, int mode, double* obj, double* objb,
                   double* H, float* g, hipStream_t s)
{
    // the objective needs no LDS rows; the derivative terms above BMAXK go through launch_barrier_big
    if (L.K > BMAXK && mode == 1 && use_landmarks) {
        set_error("barrier: K=%d > %d (derivative terms: launch_barrier_big)", L.K, BMAXK);
        return TDI_ERR_ARG;
    }
    const size_t lds = mode == 1 && use_landmarks ? 2 * sizeof(double) * NLMK * L.K : 0;
    if (lds > 64 * 1024) {
        // above the default dynamic-LDS limit: allow up to what K <= BMAXK needs (91 KB), once per process
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(k_barrier),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           (int)(2 * sizeof(double) * NLMK * BMAXK));
        TDI_HIP(attr);
    }
    hipLaunchKernelGGL(k_barrier, dim3(fl.n > 0 ? fl.n : 1), dim3(256), lds, s, m, L, frames, fl.act, fl.n, theta, lm,
                       use_landmarks, mode, obj, objb, H, g);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

template <typename TH>
static int launch_cholesky_t(int n, int K, const TH* H, const float* g, float* delta, int* status, int mode,
                             const int* mask, hipStream_t s)
{
    if (K > CMAXK) {
        set_error("cholesky: K=%d > %d", K, CMAXK);
        return TDI_ERR_ARG;
    }
    const size_t lds = (size_t)K * (K + 1) / 2 * sizeof(double);   // FP64 fallback of mode 1 included
    hipLaunchKernelGGL(k_cholesky<TH>, dim3(n), dim3(256), lds, s, K, H, g, delta, status, mask, mode);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_cholesky(int n, int K, const float* H, const float* g, float* delta, int* status, int mode, const int* mask,
                    hipStream_t s)
{
    // above one workgroup's LDS: the blocked FP64 factorisation (every listed matrix)
    if (K > CMAXK && !mask) return launch_cholesky_big(n, K, H, g, delta, status, s);
    return launch_cholesky_t(n, K, H, g, delta, status, mode, mask, s);
}

int launch_cholesky(int n, int K, const double* H, const float* g, float* delta, int* status, int mode, const int* mask,
                    hipStream_t s)
{
    return launch_cholesky_t(n, K, H, g, delta, status, mode, mask, s);
}

int launch_ls_init(int n, int K, FitState* st, const int* chol_status, const float* delta, const float* g,
                   const double* obj, int* searching, hipStream_t s)
{
    hipLaunchKernelGGL(k_ls_init, dim3((n + 255) / 256), dim3(256), 0, s, n, K, st, chol_status, delta, g, obj,
                       searching);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_ls_propose(int n, int K, int off_tz, float t_min, FitState* st, const float* theta, const float* delta,
                      float* theta_try, int* searching, int* evalmask, int hold0, int hold1, hipStream_t s)
{
    hipLaunchKernelGGL(k_ls_propose, dim3(n), dim3(64), 0, s, n, K, off_tz, t_min, st, theta, delta, theta_try,
                       searching, evalmask, hold0, hold1);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_ls_decide(int n, int K, FitState* st, const int* evalmask, int* searching, const double* obj,
                     const double* objb, const int* face_ok, float* theta, const float* theta_try,
                     const tdi_fit_options& o, hipStream_t s)
{
    hipLaunchKernelGGL(k_ls_decide, dim3(n), dim3(64), 0, s, n, K, st, evalmask, searching, obj, objb, face_ok, theta,
                       theta_try, o);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_face_check(int V, const int* lmk_ids, const float4* vproj, float ref_face_size, FrameList fl,
                      int* face_ok, hipStream_t s)
{
    hipLaunchKernelGGL(k_face_check, dim3((fl.n + 255) / 256 > 0 ? (fl.n + 255) / 256 : 1), dim3(256), 0, s, V,
                       lmk_ids, vproj, ref_face_size, fl.act, fl.n, face_ok);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_copy_mask(int n, const FitState* st, int* mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_mask, dim3((n + 255) / 256), dim3(256), 0, s, n, st, mask);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_split_fresh(int n, const FitState* st, int* mask_r, int* mask_u, hipStream_t s)
{
    hipLaunchKernelGGL(k_split_fresh, dim3((n + 255) / 256), dim3(256), 0, s, n, st, mask_r, mask_u);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

int launch_compact_frames(int n, const int* mask, int* act, hipStream_t s)
{
    hipLaunchKernelGGL(k_compact_frames, dim3(1), dim3(CMP_THREADS), 0, s, n, mask, act);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
