// init.hip — landmark-driven initialisation on gfx950: the orthographic fit of the 51
// landmarks and the pose it seeds, one wavefront per frame, in FP64.
//
// Reference: OrthographicInitializer::fit_model (orthographic_initializer.cu:243-377) fits
// x = (tau_x, tau_y, sigma_x, sigma_y, u) to the mean-centred landmarks with the residuals
//   e_x,i = sigma_x ((R p0L_i)_x + tau_x) - xl_i,   e_y,i = sigma_y ((R p0L_i)_y + tau_y) - yl_i
// and their Jacobian (compute_gradient_hessian_obj_orthographic, :56-131), by Gauss-Newton on
// J^T J (Solver, FP64 Cholesky) with an Armijo search (alpha 0.4, t halved from 1, the fit
// ends once t < 0.1; at most 100 outer iterations), from (0, 0, 15, 15, 0, 0, 0.05)
// (reset_orthographic, :133-150).  Logbarrier_Initializer::initialize_with_orthographic_t
// (logbarrier_initializer.cu:1862-1914) refits from the mirrored pose when sigma_x < 0 and
// seeds the perspective pose with initialize_kernel (logbarrier_initializer.cu:31-46).
//
// Here a frame is a wavefront: lane i holds landmark i, and the 28 + 7 + 1 sums of J^T J,
// J^T e and e.e are wavefront reductions whose lane-0 result is broadcast, so every lane takes
// the same decisions.  The reference runs each frame as a chain of cuBLAS calls and host
// copies; batching every frame into one launch is the point of this kernel.
#include "common.h"

namespace tdi {

namespace {

constexpr int OI_MAXITER = 100;
constexpr int OI_MAXINNER = 1000;
constexpr int OI_WAVES = 4;   // frames per workgroup

// Sum over the wavefront; every lane gets lane 0's result.
__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return __shfl(v, 0, 64);
}

// R = I + sin(th) [n]x + (1 - cos th) [n]x^2 and, if dR, dR/du_i = (u_i [u]x + [u x (I - R) e_i]x) R / th^2
// (rotation_header.cu:46-149), column-major, FP64.
__device__ void rot64(const double* u, double* R, double* dR)
{
    const double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double n0 = u[0] / th, n1 = u[1] / th, n2 = u[2] / th;
    const double N[9] = {0, n2, -n1, -n2, 0, n0, n1, -n0, 0};
    const double s = sin(th), c1 = 1.0 - cos(th);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += N[k * 3 + r] * N[c * 3 + k];
            R[c * 3 + r] = c1 * acc + s * N[c * 3 + r] + (r == c ? 1.0 : 0.0);
        }
    if (!dR) return;
    const double U[9] = {0, u[2], -u[1], -u[2], 0, u[0], u[1], -u[0], 0};
    const double it2 = 1.0 / (th * th);
    for (int i = 0; i < 3; ++i) {
        double e[3] = {-R[i * 3 + 0], -R[i * 3 + 1], -R[i * 3 + 2]};
        e[i] += 1.0;
        const double w[3] = {u[1] * e[2] - u[2] * e[1], u[2] * e[0] - u[0] * e[2], u[0] * e[1] - u[1] * e[0]};
        const double W[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
        double A[9];
        for (int k = 0; k < 9; ++k) A[k] = u[i] * U[k] + W[k];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += A[k * 3 + r] * R[c * 3 + k];
                dR[9 * i + c * 3 + r] = acc * it2;
            }
    }
}

// Delta = -H^-1 g for the 7x7 normal matrix held as its upper triangle (row by row).
__device__ bool chol7(const double* Hu, const double* g, double* d)
{
    double Lm[7][7];
    int k = 0;
    for (int a = 0; a < 7; ++a)
        for (int b = a; b < 7; ++b, ++k) Lm[b][a] = Hu[k];
    for (int j = 0; j < 7; ++j) {
        double dj = Lm[j][j];
        for (int q = 0; q < j; ++q) dj -= Lm[j][q] * Lm[j][q];
        if (!(dj > 0.0)) return false;
        dj = sqrt(dj);
        Lm[j][j] = dj;
        for (int i = j + 1; i < 7; ++i) {
            double s = Lm[i][j];
            for (int q = 0; q < j; ++q) s -= Lm[i][q] * Lm[j][q];
            Lm[i][j] = s / dj;
        }
    }
    double y[7];
    for (int i = 0; i < 7; ++i) {
        double s = -g[i];
        for (int q = 0; q < i; ++q) s -= Lm[i][q] * y[q];
        y[i] = s / Lm[i][i];
    }
    for (int i = 6; i >= 0; --i) {
        double s = y[i];
        for (int q = i + 1; q < 7; ++q) s -= Lm[q][i] * d[q];
        d[i] = s / Lm[i][i];
    }
    return true;
}

// This lane's residuals at x (zero for lanes without a landmark).
__device__ __forceinline__ void ortho_res(const double* x, const double* R, const double* p, double xlb, double ylb,
                                          bool on, double& ex, double& ey)
{
    const double rx = R[0] * p[0] + R[3] * p[1] + R[6] * p[2];
    const double ry = R[1] * p[0] + R[4] * p[1] + R[7] * p[2];
    ex = on ? x[2] * (rx + x[0]) - xlb : 0.0;
    ey = on ? x[3] * (ry + x[1]) - ylb : 0.0;
}

// One run of fit_model from x (wave-uniform, updated in place).  Returns the outer iterations;
// Rlast is the rotation of the last line-search trial evaluated, where the reference reads its
// angles (orthographic_initializer.cu:366-372).
__device__ int ortho_fit(double* x, const double* p, double xlb, double ylb, bool on, double* Rlast, int* status)
{
    bool terminate = false;
    int q = 0;
    rot64(x + 4, Rlast, nullptr);
    for (; q < OI_MAXITER && !terminate; ++q) {
        double R[9], dR[27];
        rot64(x + 4, R, dR);
        double ex, ey;
        ortho_res(x, R, p, xlb, ylb, on, ex, ey);
        double jx[7], jy[7];
        const double rx = R[0] * p[0] + R[3] * p[1] + R[6] * p[2];
        const double ry = R[1] * p[0] + R[4] * p[1] + R[7] * p[2];
        jx[0] = x[2]; jx[1] = 0.0;  jx[2] = x[0] + rx; jx[3] = 0.0;
        jy[0] = 0.0;  jy[1] = x[3]; jy[2] = 0.0;       jy[3] = x[1] + ry;
        for (int k = 0; k < 3; ++k) {
            const double* D = dR + 9 * k;
            jx[4 + k] = x[2] * (D[0] * p[0] + D[3] * p[1] + D[6] * p[2]);
            jy[4 + k] = x[3] * (D[1] * p[0] + D[4] * p[1] + D[7] * p[2]);
        }
        if (!on)
            for (int a = 0; a < 7; ++a) jx[a] = jy[a] = 0.0;
        double Hu[28], g[7];
        int k = 0;
        for (int a = 0; a < 7; ++a)
            for (int b = a; b < 7; ++b, ++k) Hu[k] = wave_sum(jx[a] * jx[b] + jy[a] * jy[b]);
        for (int a = 0; a < 7; ++a) g[a] = wave_sum(jx[a] * ex + jy[a] * ey);
        const double obj = wave_sum(ex * ex + ey * ey);
        double d[7];
        if (!chol7(Hu, g, d)) {
            // the reference goes on with whatever potrs returns after a failed potrf; the frame stops
            *status = TDI_NOT_PD;
            return q + 1;
        }
        double dd = 0.0;
        for (int a = 0; a < 7; ++a) dd += d[a] * g[a];
        double t = 1.0;
        for (int inner = 0; inner < OI_MAXINNER; ++inner) {
            if (t < 0.1) {
                terminate = true;
                break;
            }
            double xt[7];
            for (int a = 0; a < 7; ++a) xt[a] = x[a] + t * d[a];
            rot64(xt + 4, Rlast, nullptr);
            double etx, ety;
            ortho_res(xt, Rlast, p, xlb, ylb, on, etx, ety);
            const double ot = wave_sum(etx * etx + ety * ety);
            if (ot < obj + 0.4 * t * dd) {
                for (int a = 0; a < 7; ++a) x[a] = xt[a];
                break;
            }
            t *= 0.5;
        }
    }
    return q;
}

// RotationComputer::compute_euler_angles (rotation_header.cu:421-449), radians.
__device__ void euler64(const double* R, double& yaw, double& pitch, double& roll)
{
    const double cp = sqrt(R[5] * R[5] + R[8] * R[8]);
    if (!(cp < 1e-6)) {
        roll = atan2(R[1], R[0]);
        yaw = atan2(-R[2], cp);
        pitch = atan2(R[5], R[8]);
    } else {
        roll = atan2(-R[7], R[4]);
        yaw = atan2(-R[2], cp);
        pitch = 0.0;
    }
}

// Rotation vector of M = Rz(e2) Ry(e1) Rx(e0) (eulerAnglesToRotationMatrix, newfuncs.cu:720-747,
// then the matrix-to-vector direction of cv::Rodrigues, logbarrier_initializer.cu:1883): the
// log map theta / (2 sin theta) (M21 - M12, M02 - M20, M10 - M01), with its theta -> 0 and
// theta -> pi cases.  M row-major.
__device__ void rotvec_euler(const double* e, double* r)
{
    const double c0 = cos(e[0]), s0 = sin(e[0]), c1 = cos(e[1]), s1 = sin(e[1]), c2 = cos(e[2]), s2 = sin(e[2]);
    // Rz Ry Rx written out
    const double M[9] = {c2 * c1, c2 * s1 * s0 - s2 * c0, c2 * s1 * c0 + s2 * s0,
                         s2 * c1, s2 * s1 * s0 + c2 * c0, s2 * s1 * c0 - c2 * s0,
                         -s1,     c1 * s0,                c1 * c0};
    const double v[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
    const double sn = 0.5 * sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double cs = fmin(1.0, fmax(-1.0, 0.5 * (M[0] + M[4] + M[8] - 1.0)));
    const double th = acos(cs);
    if (sn < 1e-5) {
        if (cs > 0) {
            r[0] = r[1] = r[2] = 0.0;
            return;
        }
        const double rx = sqrt(fmax(0.5 * (M[0] + 1.0), 0.0));
        const double ry = sqrt(fmax(0.5 * (M[4] + 1.0), 0.0)) * (M[1] < 0 ? -1.0 : 1.0);
        double rz = sqrt(fmax(0.5 * (M[8] + 1.0), 0.0)) * (M[2] < 0 ? -1.0 : 1.0);
        if (fabs(rx) < fabs(ry) && fabs(rx) < fabs(rz) && ((M[5] > 0) != (ry * rz > 0))) rz = -rz;
        const double nr = sqrt(rx * rx + ry * ry + rz * rz);
        r[0] = rx * th / nr; r[1] = ry * th / nr; r[2] = rz * th / nr;
        return;
    }
    const double f = th / (2.0 * sn);
    r[0] = v[0] * f; r[1] = v[1] * f; r[2] = v[2] * f;
}

__global__ void __launch_bounds__(64 * OI_WAVES) k_ortho_init(const float* __restrict__ mean,
                                                               const int* __restrict__ lmk_ids, int n,
                                                               const float* __restrict__ lmk,
                                                               const float* __restrict__ cams,
                                                               float* __restrict__ x_orth, float* __restrict__ pose,
                                                               float* __restrict__ angles, int* __restrict__ info)
{
    const int lane = threadIdx.x & 63;
    const int f = blockIdx.x * OI_WAVES + (threadIdx.x >> 6);
    if (f >= n) return;   // whole wavefronts leave together
    const bool on = lane < NLMK;
    const int li = on ? lane : 0;
    const int vid = lmk_ids[li];
    const double p[3] = {on ? (double)mean[3 * vid] : 0.0, on ? (double)mean[3 * vid + 1] : 0.0,
                         on ? (double)mean[3 * vid + 2] : 0.0};
    const float* lf = lmk + (size_t)f * 2 * NLMK;
    const double lx = on ? (double)lf[li] : 0.0, ly = on ? (double)lf[NLMK + li] : 0.0;
    const double mx = wave_sum(lx) / NLMK, my = wave_sum(ly) / NLMK;
    const double xlb = lx - mx, ylb = ly - my;
    double x[7] = {0.0, 0.0, 15.0, 15.0, 0.0, 0.0, 0.05};
    double Rl[9];
    int status = TDI_OK;
    const int it0 = ortho_fit(x, p, xlb, ylb, on, Rl, &status);
    int it1 = 0, flipped = 0;
    double yaw, pitch, roll;
    euler64(Rl, yaw, pitch, roll);
    if (status == TDI_OK && x[2] < 0.0) {
        const double ev[3] = {roll, -pitch, -yaw};
        double u0[3];
        rotvec_euler(ev, u0);
        x[0] = -x[0];
        x[2] = -x[2];
        x[4] = u0[0]; x[5] = u0[1]; x[6] = u0[2];
        it1 = ortho_fit(x, p, xlb, ylb, on, Rl, &status);
        flipped = 1;
        euler64(Rl, yaw, pitch, roll);
    }
    if (lane != 0) return;
    for (int a = 0; a < 7; ++a) x_orth[7 * (size_t)f + a] = (float)x[a];
    const float* cf = cams + 4 * (size_t)f;
    const double tz = (double)cf[0] / ((fabs(x[2]) + fabs(x[3])) * 0.5);
    float* po = pose + 6 * (size_t)f;
    po[0] = (float)(tz * (mx - cf[2]) / cf[0]);
    po[1] = (float)(tz * (my - cf[3]) / cf[1]);
    po[2] = (float)tz;
    po[3] = (float)x[4]; po[4] = (float)x[5]; po[5] = (float)x[6];
    if (angles) {
        const double r2d = 180.0 / 3.14159265358979323846;
        angles[3 * (size_t)f + 0] = (float)(yaw * r2d);
        angles[3 * (size_t)f + 1] = (float)(pitch * r2d);
        angles[3 * (size_t)f + 2] = (float)(roll * r2d);
    }
    int* io = info + 4 * (size_t)f;
    io[0] = it0; io[1] = it1; io[2] = flipped; io[3] = status;
}

} // namespace

int launch_ortho_init(const ModelGPU& m, int n, const float* lmk, const float* cams, float* x_orth, float* pose,
                      float* angles, int* info, hipStream_t s)
{
    if (n <= 0) return TDI_OK;
    hipLaunchKernelGGL(k_ortho_init, dim3((n + OI_WAVES - 1) / OI_WAVES), dim3(64 * OI_WAVES), 0, s, m.mean, m.lmk_ids,
                       n, lmk, cams, x_orth, pose, angles, info);
    TDI_HIP(hipGetLastError());
    return TDI_OK;
}

} // namespace tdi
