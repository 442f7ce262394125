/*
 * tdi_oracle.cpp — FP64 OpenMP restatement of 3DI's per-frame 3DMM fitting step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py load it to compare against the HIP library
 * (3di_amd/csrc).  Nothing in the product links, loads or calls it.
 *
 * Everything is computed in double precision except the rasteriser's coverage test,
 * which follows the reference's FP32 operation sequence (funcs.cu:950-979) so that
 * triangle-id/visibility buffers can be compared bit for bit with the GPU given
 * identical FP32 projected vertices, and the frame preparation, which follows OpenCV's
 * fixed-point grey conversion and FP32 bilinear resize.  Build with -ffp-contract=off (oracle/Makefile).
 *
 * Intended divergences from the reference, shared with the HIP library:
 *   - visibility: packed (depth|triangle-id) minimum, so exact-depth ties resolve to the
 *     lowest triangle id instead of the reference's racy scatter (funcs.cu:816,1027);
 *   - no 26,000/52,000 pixel caps, no ushort ids, no slot-0 sentinel (renderer.cu:665-677),
 *     pixels outside the crop are dropped instead of wrapping a ushort index;
 *   - the NTMP=4 per-triangle cap (constants.h:19) is the `ntmp_cap` option (0 = off),
 *     counted over the bounding box clipped to the crop window;
 *   - the right-neighbour index at the crop's last column, a ushort underflow in the
 *     reference (funcs.cu:313), is clamped to sorted position 0.
 *
 * Parity status of each function is documented in the header of tests/test_oracle.py (the
 * landmark initialisation in tests/test_init_oracle.py, the frame preparation in
 * tests/test_frames_oracle.py).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <limits>
#include <omp.h>

#include "../include/tdi_api.h"

namespace {

constexpr int NL = TDI_NLANDMARKS;

struct Layout {
    int K = 0, off_tex = 0, off_id = 0, off_exp = 0, off_pose = 0, off_light = 0;
    int ntex = 0, nid = 0, nexp = 0, nlight = 0;
};

Layout make_layout(const tdi_model_desc* m, const tdi_fit_options* o)
{
    // OptimizationVariables (Optimizer.h:71-91): betas | alphas | epsilons | taus+us | (light)
    Layout L;
    L.ntex = o->fit_tex ? m->k_tex : 0;
    L.nid = o->fit_id ? m->k_id : 0;
    L.nexp = o->fit_exp ? m->k_exp : 0;
    L.nlight = o->fit_light == TDI_LIGHT_SH2 ? 9 : (o->fit_light ? 4 : 0);
    L.off_tex = 0;
    L.off_id = L.ntex;
    L.off_exp = L.off_id + L.nid;
    L.off_pose = L.off_exp + L.nexp;
    L.off_light = L.off_pose + 6;
    L.K = L.off_light + L.nlight;
    return L;
}

// ---------------------------------------------------------------- rotation
// RotationComputer::process -> fill_theta_etc (rotation_header.cu:46-149):
// R = I + sin(th)[n]x + (1-cos th)[n]x^2, dR/du_i = (u_i[u]x + [u x (I-R)e_i]x) R / th^2.
// Matrices column-major (R[col*3+row]) like the reference (funcs.cu:373-381).
void skew(const double* u, double* S)
{
    S[0] = 0;     S[3] = -u[2]; S[6] = u[1];
    S[1] = u[2];  S[4] = 0;     S[7] = -u[0];
    S[2] = -u[1]; S[5] = u[0];  S[8] = 0;
}

void matmul3(const double* A, const double* B, double* C)
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[k * 3 + r] * B[c * 3 + k];
            C[c * 3 + r] = s;
        }
}

void rodrigues(const double* u, double* R, double* dR /* [27] or null */)
{
    double th = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double n[3] = {u[0] / th, u[1] / th, u[2] / th};
    double N[9], N2[9];
    skew(n, N);
    matmul3(N, N, N2);
    double s = std::sin(th), c1 = 1.0 - std::cos(th);
    for (int i = 0; i < 9; ++i) R[i] = c1 * N2[i] + s * N[i];
    R[0] += 1; R[4] += 1; R[8] += 1;
    if (!dR) return;
    double U[9];
    skew(u, U);
    for (int i = 0; i < 3; ++i) {
        // w = u x ((I-R) e_i)
        double e[3] = {-R[i * 3 + 0], -R[i * 3 + 1], -R[i * 3 + 2]};
        e[i] += 1.0;
        double w[3] = {u[1] * e[2] - u[2] * e[1], u[2] * e[0] - u[0] * e[2], u[0] * e[1] - u[1] * e[0]};
        double W[9], A[9];
        skew(w, W);
        for (int k = 0; k < 9; ++k) A[k] = u[i] * U[k] + W[k];
        matmul3(A, R, dR + 9 * i);
        for (int k = 0; k < 9; ++k) dR[9 * i + k] /= th * th;
    }
}

// R*x for column-major R
inline void rot(const double* R, const double* x, double* y)
{
    y[0] = R[0] * x[0] + R[3] * x[1] + R[6] * x[2];
    y[1] = R[1] * x[0] + R[4] * x[1] + R[7] * x[2];
    y[2] = R[2] * x[0] + R[5] * x[1] + R[8] * x[2];
}
inline void rotT(const double* R, const double* x, double* y)
{
    y[0] = R[0] * x[0] + R[1] * x[1] + R[2] * x[2];
    y[1] = R[3] * x[0] + R[4] * x[1] + R[5] * x[2];
    y[2] = R[6] * x[0] + R[7] * x[1] + R[8] * x[2];
}

// ---------------------------------------------------------------- blur taps
// The reference hard-codes four 15-tap tables selected by KERNEL_SIGMA (video_fitter.cu:51-63),
// generated by the normalised Gaussian of video_fitter.cu:2836-2849.  Its "2.0" table is the
// sigma = 1.5 Gaussian; the other three use the stated sigma.
int blur_taps(double sigma_key, double* taps)
{
    double s;
    if (sigma_key == 0.5) s = 0.5;
    else if (sigma_key == 0.75) s = 0.75;
    else if (sigma_key == 1.0) s = 1.0;
    else if (sigma_key == 2.0) s = 1.5;
    else return TDI_ERR_ARG;
    double sum = 0;
    for (int i = 0; i < 15; ++i) {
        double x = i - 7;
        taps[i] = std::exp(-x * x / (2 * s * s));
        sum += taps[i];
    }
    for (int i = 0; i < 15; ++i) taps[i] /= sum;
    return TDI_OK;
}

// Separable 15-tap blur with zero padding (convolutionRowsKernel/ColumnsKernel,
// newfuncs.cu:539-715) on a column-major W x H crop.
void blur(const double* in, double* out, int W, int H, const double* k)
{
    std::vector<double> tmp((size_t)W * H);
    for (int x = 0; x < W; ++x)
        for (int y = 0; y < H; ++y) {
            double s = 0;
            for (int j = -7; j <= 7; ++j) {
                int yy = y + j;
                if (yy >= 0 && yy < H) s += k[7 - j] * in[(size_t)x * H + yy];
            }
            tmp[(size_t)x * H + y] = s;
        }
    for (int x = 0; x < W; ++x)
        for (int y = 0; y < H; ++y) {
            double s = 0;
            for (int j = -7; j <= 7; ++j) {
                int xx = x + j;
                if (xx >= 0 && xx < W) s += k[7 - j] * tmp[(size_t)xx * H + y];
            }
            out[(size_t)x * H + y] = s;
        }
}

// ---------------------------------------------------------------- raster
inline uint32_t orderable(float z)
{
    uint32_t u;
    std::memcpy(&u, &z, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// FP32 barycentric coverage test exactly as get_pixels_to_render (funcs.cu:940-979):
// sample at integer pixel coordinates, (b,c,a) = (tl0,tl1,tl2), alpha->b, beta->c.
inline bool cover(float bx, float by, float cx, float cy, float ax, float ay, int x, int y,
                  float* alpha_out, float* beta_out)
{
    float v0x = bx - ax, v1x = cx - ax, v2x = (float)x - ax;
    float v0y = by - ay, v1y = cy - ay, v2y = (float)y - ay;
    float d00 = v0x * v0x + v0y * v0y;
    float d01 = v0x * v1x + v0y * v1y;
    float d11 = v1x * v1x + v1y * v1y;
    float d20 = v0x * v2x + v0y * v2y;
    float d21 = v1x * v2x + v1y * v2y;
    float denom = d00 * d11 - d01 * d01;
    float alpha = (d11 * d20 - d01 * d21) / denom;
    float beta = (d00 * d21 - d01 * d20) / denom;
    *alpha_out = alpha;
    *beta_out = beta;
    return alpha > 0 && beta > 0 && (alpha + beta <= 1);
}

// Rasterise with a packed (orderable depth << 32 | triangle) minimum per pixel.
// Depth is the camera-space Z of the triangle's first vertex (funcs.cu:995,1000).
void rasterize(int NT, const int32_t* tris, const float* xp, const float* yp, const float* zf,
               int x0, int y0, int W, int H, int ntmp, uint64_t* keys, float* bary)
{
    const size_t A = (size_t)W * H;
    for (size_t i = 0; i < A; ++i) keys[i] = ~0ull;
    for (int pass = 0; pass < 2; ++pass) {
        for (int t = 0; t < NT; ++t) {
            int ib = tris[3 * t], ic = tris[3 * t + 1], ia = tris[3 * t + 2];
            float bx = xp[ib], cx = xp[ic], ax = xp[ia];
            float by = yp[ib], cy = yp[ic], ay = yp[ia];
            float xmin = std::floor(std::fmin(std::fmin(bx, cx), ax));
            float xmax = std::ceil(std::fmax(std::fmax(bx, cx), ax));
            float ymin = std::floor(std::fmin(std::fmin(by, cy), ay));
            float ymax = std::ceil(std::fmax(std::fmax(by, cy), ay));
            if (!(xmin == xmin) || !(ymin == ymin) || !(xmax == xmax) || !(ymax == ymax)) continue;
            // the scan is clipped to the crop window (the reference scans the whole box with
            // ushort counters and relies on the face lying inside its 256x256 crop)
            const float xlo = std::fmax(xmin, (float)x0), xhi = std::fmin(xmax, (float)(x0 + W));
            const float ylo = std::fmax(ymin, (float)y0), yhi = std::fmin(ymax, (float)(y0 + H));
            if (!(xlo < xhi) || !(ylo < yhi)) continue;
            uint64_t key = ((uint64_t)orderable(zf[ib]) << 32) | (uint32_t)t;
            int covered = 0;
            for (int x = (int)xlo; (float)x < xhi; ++x) {
                for (int y = (int)ylo; (float)y < yhi; ++y) {
                    float al, be;
                    if (!cover(bx, by, cx, cy, ax, ay, x, y, &al, &be)) continue;
                    if (ntmp > 0 && covered >= ntmp) continue;
                    covered++;
                    int xc = x - x0, yc = y - y0;
                    size_t pix = (size_t)xc * H + yc;
                    if (pass == 0) {
                        if (key < keys[pix]) keys[pix] = key;
                    } else if (keys[pix] == key) {
                        bary[2 * pix] = al;
                        bary[2 * pix + 1] = be;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- per-frame evaluation
struct Eval {
    // render products
    std::vector<uint64_t> keys;
    std::vector<float> bary;
    std::vector<int> vis;       // sorted crop indices of visible pixels
    std::vector<double> raw;    // shaded crop before the blur
    std::vector<double> img;    // blurred rendered crop
    double obj_data = 0, obj_barrier = 0, obj = 0;
    bool face_ok = true;
    // derivatives
    std::vector<double> H, g;
};

struct Frame {
    const tdi_model_desc* m;
    const tdi_frame_desc* f;
    const tdi_fit_options* o;
    Layout L;
    double taps[15];
    std::vector<double> obs_blur;   // blurred observation
    bool finetune = false;          // fit_3DMM_epsilon_finetune: widened expression bounds, pose held
};

struct Geometry {
    double R[9], dR[27];
    std::vector<double> P0, P;      // [V][3] model / camera space
    std::vector<float> xpf, ypf, zf;
    std::vector<double> xp, yp;
    std::vector<double> rho;        // per-vertex albedo
    double lam[3], Lint, tau[3];
    double a[3], sh[5];             // directional and second-order light terms (TDI_LIGHT_SH2)
};

// Second-order terms of the camera-space unit normal n (the light of TDI_LIGHT_SH2,
// include/tdi_api.h): Y = (nx ny, ny nz, nx nz, nx^2 - ny^2, 3 nz^2 - 1).  That light is
// I = rho (1 + n_c.a - L n_c.v_c + sum_k s_k Y_k(n_c)): the reference's point light
// (newfuncs.cu:122-137) written linearly in a = L lambda, plus the second-order terms.
void sh2_terms(const double* n, double* Y)
{
    Y[0] = n[0] * n[1];
    Y[1] = n[1] * n[2];
    Y[2] = n[0] * n[2];
    Y[3] = n[0] * n[0] - n[1] * n[1];
    Y[4] = 3 * n[2] * n[2] - 1;
}

void geometry(const Frame& F, const double* th, bool need_dR, Geometry& G)
{
    const tdi_model_desc* m = F.m;
    const Layout& L = F.L;
    const int V = m->n_vertices;
    const double* u = th + L.off_pose + 3;
    rodrigues(u, G.R, need_dR ? G.dR : nullptr);
    for (int i = 0; i < 3; ++i) G.tau[i] = th[L.off_pose + i];
    if (L.nlight == 4) {
        for (int i = 0; i < 3; ++i) G.lam[i] = th[L.off_light + i];
        G.Lint = th[L.off_light + 3];
    } else if (L.nlight == 9) {
        for (int i = 0; i < 3; ++i) G.lam[i] = 0.0;     // L Id = -L n_c.v_c
        G.Lint = th[L.off_light + 3];
    } else {
        for (int i = 0; i < 3; ++i) G.lam[i] = F.f->lambda[i];
        G.Lint = F.f->Lintensity;
    }
    for (int k = 0; k < 3; ++k) G.a[k] = L.nlight == 9 ? th[L.off_light + k] : 0.0;
    for (int k = 0; k < 5; ++k) G.sh[k] = L.nlight == 9 ? th[L.off_light + 4 + k] : 0.0;
    G.P0.assign((size_t)V * 3, 0);
    G.P.assign((size_t)V * 3, 0);
    G.xp.resize(V); G.yp.resize(V);
    G.xpf.resize(V); G.ypf.resize(V); G.zf.resize(V);
    // compute_nonrigid_shape2 (renderer.cu:806-861): P0 = mean + I alpha + E eps,
    // view_transform_3d_pts_and_render_2d (funcs.cu:440-452): P = R P0 + tau, pinhole.
    #pragma omp parallel for schedule(static) if (V > 4096 && !omp_in_parallel())
    for (int v = 0; v < V; ++v) {
        double p[3];
        for (int a = 0; a < 3; ++a) {
            double s = m->mean_shape[3 * v + a];
            const int r = 3 * v + a;
            for (int k = 0; k < L.nid; ++k) s += (double)m->id_basis[(size_t)r * m->k_id + k] * th[L.off_id + k];
            for (int k = 0; k < L.nexp; ++k) s += (double)m->exp_basis[(size_t)r * m->k_exp + k] * th[L.off_exp + k];
            p[a] = s;
            G.P0[3 * v + a] = s;
        }
        double q[3];
        rot(G.R, p, q);
        for (int a = 0; a < 3; ++a) { q[a] += G.tau[a]; G.P[3 * v + a] = q[a]; }
        G.xp[v] = F.f->phix * q[0] / q[2] + F.f->cx;
        G.yp[v] = F.f->phiy * q[1] / q[2] + F.f->cy;
        G.xpf[v] = (float)G.xp[v];
        G.ypf[v] = (float)G.yp[v];
        G.zf[v] = (float)q[2];
    }
    // compute_texture (renderer.cu:941-951): albedo = mean + T beta
    G.rho.assign(V, 0);
    for (int v = 0; v < V; ++v) {
        double s = m->tex_mean[v];
        for (int k = 0; k < L.ntex; ++k) s += (double)m->tex_basis[(size_t)v * m->k_tex + k] * th[L.off_tex + k];
        G.rho[v] = s;
    }
}

// Face-size sanity test of compute_nonrigid_shape2 (renderer.cu:836-857).
bool face_reasonable(const Frame& F, const Geometry& G)
{
    if (F.o->ref_face_size <= 0) return true;
    const int32_t* LI = F.m->landmark_ids;
    double iod = std::fabs(G.xp[LI[9]] - G.xp[LI[0]]);
    double lem = std::fabs(G.yp[LI[0]] - G.yp[LI[40]]);
    double rem = std::fabs(G.yp[LI[9]] - G.yp[LI[40]]);
    double lim = 2.0 * F.o->ref_face_size;
    return !(iod > lim || lem > lim || rem > lim);
}

// Landmark-box + coefficient-bound log barrier (logbarrier_initializer.cu:458-518,587-948,
// 1272-1842 with slack = 0; objective as evaluate_objective_function 1921-2081).
// The landmark Hessian is the Gauss-Newton sum of (dG/G)(dG/G)^T like the reference.
double barrier(const Frame& F, const double* th, const Geometry& G, double* grad, double* Hm)
{
    const tdi_model_desc* m = F.m;
    const Layout& L = F.L;
    const int K = L.K;
    double obj = 0;
    auto bound = [&](int off, int n, const float* lb, const float* ub) {
        for (int k = 0; k < n; ++k) {
            if (!(ub[k] < INFINITY) || !(lb[k] > -INFINITY)) continue;   // unbounded: no term
            double x = th[off + k];
            double fub = x - ub[k], flb = lb[k] - x;
            obj += -std::log(-fub) - std::log(-flb);
            if (grad) grad[off + k] += -1.0 / fub + 1.0 / flb;
            if (Hm) {
                double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
                Hm[(size_t)(off + k) * K + off + k] += a * a + b * b;
            }
        }
    };
    bound(L.off_id, L.nid, m->id_lb, m->id_ub);
    bound(L.off_tex, L.ntex, m->tex_lb, m->tex_ub);
    if (F.finetune) {
        // the finetune fits' bounds: FINETUNE_COEF times the regular ones (logbarrier_initializer.cu:285-286)
        std::vector<float> lb(L.nexp), ub(L.nexp);
        for (int k = 0; k < L.nexp; ++k) {
            lb[k] = m->exp_lb[k] * F.o->finetune_coef;
            ub[k] = m->exp_ub[k] * F.o->finetune_coef;
        }
        bound(L.off_exp, L.nexp, lb.data(), ub.data());
    } else {
        bound(L.off_exp, L.nexp, m->exp_lb, m->exp_ub);
    }
    if (!F.o->use_landmarks) return obj;

    const tdi_frame_desc* f = F.f;
    std::vector<double> dx(K), dy(K);
    for (int i = 0; i < NL; ++i) {
        const int v = m->landmark_ids[i];
        double p[3];
        for (int a = 0; a < 3; ++a) {
            double s = m->mean_shape[3 * v + a];
            const int r = 3 * v + a;
            for (int k = 0; k < L.nid; ++k) s += (double)m->id_basis[(size_t)r * m->k_id + k] * th[L.off_id + k];
            for (int k = 0; k < L.nexp; ++k) s += (double)m->exp_basis[(size_t)r * m->k_exp + k] * th[L.off_exp + k];
            p[a] = s;
        }
        double q[3];
        rot(G.R, p, q);
        double vx = q[0] + G.tau[0], vy = q[1] + G.tau[1], vz = q[2] + G.tau[2];
        double xproj = f->phix * vx / vz + f->cx;
        double yproj = f->phiy * vy / vz + f->cy;
        const double fs = f->face_size;
        double Gxm = xproj - (f->lmk_x[i] + f->box_lx[i] * fs);
        double Gym = yproj - (f->lmk_y[i] + f->box_ly[i] * fs);
        double Gxp = (f->lmk_x[i] + f->box_ux[i] * fs) - xproj;
        double Gyp = (f->lmk_y[i] + f->box_uy[i] * fs) - yproj;
        obj += -std::log(Gxm) - std::log(Gym) - std::log(Gxp) - std::log(Gyp);
        if (!grad && !Hm) continue;
        // gradient of the projected landmark (fill_optimization_dc_landmark & co.)
        std::fill(dx.begin(), dx.end(), 0.0);
        std::fill(dy.begin(), dy.end(), 0.0);
        auto put = [&](int col, const double* dv) {
            dx[col] = f->phix * (dv[0] * vz - dv[2] * vx) / (vz * vz);
            dy[col] = f->phiy * (dv[1] * vz - dv[2] * vy) / (vz * vz);
        };
        for (int a = 0; a < 3; ++a) {
            double e[3] = {0, 0, 0};
            e[a] = 1;
            put(L.off_pose + a, e);
        }
        for (int a = 0; a < 3; ++a) {
            double dv[3];
            rot(G.dR + 9 * a, p, dv);
            put(L.off_pose + 3 + a, dv);
        }
        for (int k = 0; k < L.nexp; ++k) {
            double e[3], dv[3];
            for (int a = 0; a < 3; ++a) e[a] = m->exp_basis[(size_t)(3 * v + a) * m->k_exp + k];
            rot(G.R, e, dv);
            put(L.off_exp + k, dv);
        }
        for (int k = 0; k < L.nid; ++k) {
            double e[3], dv[3];
            for (int a = 0; a < 3; ++a) e[a] = m->id_basis[(size_t)(3 * v + a) * m->k_id + k];
            rot(G.R, e, dv);
            put(L.off_id + k, dv);
        }
        const double* dd[4] = {dx.data(), dy.data(), dx.data(), dy.data()};
        const double sc[4] = {-1.0 / Gxm, -1.0 / Gym, 1.0 / Gxp, 1.0 / Gyp};
        for (int s = 0; s < 4; ++s) {
            for (int a = 0; a < K; ++a) {
                double pa = dd[s][a] * sc[s];
                if (pa == 0) continue;
                if (grad) grad[a] += pa;
                if (Hm)
                    for (int b = 0; b < K; ++b) Hm[(size_t)a * K + b] += pa * dd[s][b] * sc[s];
            }
        }
    }
    return obj;
}

// Full evaluation at theta: Renderer::render (renderer.cu:635-730) and, when derivs,
// DerivativeComputer::compute_hessian_and_gradient (derivative_computer.cu:22-213).
void evaluate(const Frame& F, const double* th, bool derivs, Eval& E, double* dI_out = nullptr,
              int* nbr_out = nullptr)
{
    const tdi_model_desc* m = F.m;
    const tdi_frame_desc* f = F.f;
    const Layout& L = F.L;
    const int K = L.K;
    const int W = f->width, Hh = f->height;
    const size_t A = (size_t)W * Hh;
    const double w = F.o->data_weight;

    Geometry G;
    geometry(F, th, derivs, G);
    E.face_ok = face_reasonable(F, G);

    E.keys.assign(A, ~0ull);
    E.bary.assign(2 * A, 0.f);
    rasterize(m->n_triangles, m->triangles, G.xpf.data(), G.ypf.data(), G.zf.data(), f->x0, f->y0, W, Hh,
              F.o->ntmp_cap, E.keys.data(), E.bary.data());
    E.vis.clear();
    for (size_t k = 0; k < A; ++k)
        if (E.keys[k] != ~0ull) E.vis.push_back((int)k);
    const int N = (int)E.vis.size();

    // shading (fill_diffuse_component_and_populate_texture_and_shape, newfuncs.cu:18-167)
    std::vector<double> I(A, 0.0);
    struct Px { double al, be, ga, v[3], p0[3], nrm[3], rho, Id, Y[5], S; int tb, tc, ta; };
    std::vector<Px> px(N);
    for (int n = 0; n < N; ++n) {
        const int k = E.vis[n];
        const int t = (int)(E.keys[k] & 0xffffffffu);
        Px& q = px[n];
        q.tb = m->triangles[3 * t]; q.tc = m->triangles[3 * t + 1]; q.ta = m->triangles[3 * t + 2];
        float alf = E.bary[2 * k], bef = E.bary[2 * k + 1];
        float gaf = 1.0f - alf - bef;
        q.al = alf; q.be = bef; q.ga = gaf;
        const double* B0 = &G.P0[3 * q.tb]; const double* C0 = &G.P0[3 * q.tc]; const double* A0 = &G.P0[3 * q.ta];
        double e1[3] = {B0[0] - A0[0], B0[1] - A0[1], B0[2] - A0[2]};
        double e2[3] = {C0[0] - A0[0], C0[1] - A0[1], C0[2] - A0[2]};
        double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
        for (int a = 0; a < 3; ++a) q.nrm[a] = nr[a] / nn;
        for (int a = 0; a < 3; ++a) {
            q.p0[a] = q.al * B0[a] + q.be * C0[a] + q.ga * A0[a];
            q.v[a] = q.al * G.P[3 * q.tb + a] + q.be * G.P[3 * q.tc + a] + q.ga * G.P[3 * q.ta + a];
        }
        double dlt[3] = {G.lam[0] - G.tau[0], G.lam[1] - G.tau[1], G.lam[2] - G.tau[2]};
        double rl[3];
        rotT(G.R, dlt, rl);
        q.Id = q.nrm[0] * rl[0] + q.nrm[1] * rl[1] + q.nrm[2] * rl[2]
             - (q.nrm[0] * q.p0[0] + q.nrm[1] * q.p0[1] + q.nrm[2] * q.p0[2]);
        q.rho = q.al * G.rho[q.tb] + q.be * G.rho[q.tc] + q.ga * G.rho[q.ta];
        // second-order light terms of the camera-space normal R n
        double nc[3];
        rot(G.R, q.nrm, nc);
        sh2_terms(nc, q.Y);
        q.S = nc[0] * G.a[0] + nc[1] * G.a[1] + nc[2] * G.a[2];
        for (int j = 0; j < 5; ++j) q.S += G.sh[j] * q.Y[j];
        I[k] = q.rho + G.Lint * q.rho * q.Id + q.rho * q.S;
    }
    E.img.assign(A, 0.0);
    blur(I.data(), E.img.data(), W, Hh, F.taps);
    E.raw.swap(I);

    // normalised gradients (fill_grads, funcs.cu:842-897); border neighbours read index 0
    std::vector<double> gx(N), gy(N), hr(N), gsx(N), gsy(N);
    double corr = 0;
    for (int n = 0; n < N; ++n) {
        const int k = E.vis[n];
        const int x = k / Hh, y = k % Hh;
        const int kr = (x < W - 1) ? k + Hh : 0, kl = (x > 0) ? k - Hh : 0;
        const int ka = (y > 0) ? k - 1 : 0, kb = (y < Hh - 1) ? k + 1 : 0;
        gx[n] = 0.5 * (E.img[kr] - E.img[kl]);
        gy[n] = 0.5 * (E.img[kb] - E.img[ka]);
        hr[n] = 1e-16 + std::sqrt(4 * gx[n] * gx[n] + 4 * gy[n] * gy[n]);
        const double sx = 0.5 * (F.obs_blur[kr] - F.obs_blur[kl]);
        const double sy = 0.5 * (F.obs_blur[kb] - F.obs_blur[ka]);
        const double hs = 1e-16 + std::sqrt(4 * sx * sx + 4 * sy * sy);
        gsx[n] = 2 * sx / hs;
        gsy[n] = 2 * sy / hs;
        corr += (2 * gx[n] / hr[n]) * gsx[n] + (2 * gy[n] / hr[n]) * gsy[n];
    }
    E.obj_data = -w * corr;

    if (derivs) {
        E.H.assign((size_t)K * K, 0.0);
        E.g.assign(K, 0.0);
    }
    E.obj_barrier = barrier(F, th, G, derivs ? E.g.data() : nullptr, derivs ? E.H.data() : nullptr);
    E.obj = E.obj_data + E.obj_barrier;
    if (!derivs || N == 0) return;

    // flow / direct Jacobian dI/dtheta per rendered pixel (newfuncs.cu:173-403)
    std::vector<double> dI((size_t)N * K, 0.0);
    for (int n = 0; n < N; ++n) {
        const Px& q = px[n];
        double* row = &dI[(size_t)n * K];
        const double vx = q.v[0], vy = q.v[1], vz = q.v[2];
        auto flow = [&](const double* dv) {
            double dWx = f->phix * (dv[0] * vz - dv[2] * vx) / (vz * vz);
            double dWy = f->phiy * (dv[1] * vz - dv[2] * vy) / (vz * vz);
            return -(gx[n] * dWx + gy[n] * dWy);
        };
        for (int a = 0; a < 3; ++a) {
            double e[3] = {0, 0, 0};
            e[a] = 1;
            row[L.off_pose + a] = flow(e);
            double dv[3];
            rot(G.dR + 9 * a, q.p0, dv);
            row[L.off_pose + 3 + a] = flow(dv);
        }
        auto basis_flow = [&](const float* B, int kb, int off, int cnt) {
            for (int k = 0; k < cnt; ++k) {
                double e[3], dv[3];
                for (int a = 0; a < 3; ++a)
                    e[a] = q.al * B[(size_t)(3 * q.tb + a) * kb + k] + q.be * B[(size_t)(3 * q.tc + a) * kb + k]
                         + q.ga * B[(size_t)(3 * q.ta + a) * kb + k];
                rot(G.R, e, dv);
                row[off + k] = flow(dv);
            }
        };
        basis_flow(m->exp_basis, m->k_exp, L.off_exp, L.nexp);
        if (L.nid) basis_flow(m->id_basis, m->k_id, L.off_id, L.nid);
        for (int k = 0; k < L.ntex; ++k) {
            double t = q.al * m->tex_basis[(size_t)q.tb * m->k_tex + k] + q.be * m->tex_basis[(size_t)q.tc * m->k_tex + k]
                     + q.ga * m->tex_basis[(size_t)q.ta * m->k_tex + k];
            row[L.off_tex + k] = t + G.Lint * q.Id * t + q.S * t;  // fill_optimization_dI_dbeta (newfuncs.cu:383-403)
        }
        if (L.nlight) {
            double rn[3];
            rot(G.R, q.nrm, rn);
            for (int a = 0; a < 3; ++a)                                                   // newfuncs.cu:151-153 x R
                row[L.off_light + a] = (L.nlight == 9 ? 1.0 : G.Lint) * q.rho * rn[a];
            row[L.off_light + 3] = q.rho * q.Id;                                          // newfuncs.cu:155
            for (int j = 0; j < L.nlight - 4; ++j) row[L.off_light + 4 + j] = q.rho * q.Y[j];
        }
    }

    // compacted neighbour positions (fill_krels/fill_krels2, funcs.cu:288-340;
    // derivative_computer.cu:43-72): sorted list, prefix count of occupied crop pixels.
    std::vector<int> cum(A);
    {
        int c = 0;
        for (size_t k = 0; k < A; ++k) { c += (E.keys[k] != ~0ull); cum[k] = c; }
    }
    std::vector<int> nb((size_t)4 * N, 0);
    for (int n = 0; n < N; ++n) {
        if (n == 0 || n == N - 1) continue;
        const int k = E.vis[n];
        const int x = k / Hh;
        const int kleft = (x > 0) ? k - Hh : 0;
        const int kright = (x < W - 1) ? k + Hh : 0;
        int kl = cum[kleft] > 1 ? cum[kleft] - 1 : 0;
        if (kl >= N - 1) kl = N - 1;
        int kr = cum[kright] - 1;
        if (kr < 0) kr = 0;
        nb[4 * n + 0] = kl;
        nb[4 * n + 1] = kr;
        nb[4 * n + 2] = n - 1;
        nb[4 * n + 3] = n + 1;
    }
    if (nbr_out) std::copy(nb.begin(), nb.end(), nbr_out);
    if (dI_out) std::copy(dI.begin(), dI.end(), dI_out);

    // normalised-gradient Jacobian (phase2_new, newfuncs.cu:417-463) and the normal
    // equations (derivative_computer.cu:159-164): H += w J^T J, g -= w J^T g_s.
    std::vector<double> Jx(K), Jy(K);
    for (int n = 0; n < N; ++n) {
        const double* rl = &dI[(size_t)nb[4 * n + 0] * K];
        const double* rr = &dI[(size_t)nb[4 * n + 1] * K];
        const double* ra = &dI[(size_t)nb[4 * n + 2] * K];
        const double* rb = &dI[(size_t)nb[4 * n + 3] * K];
        const double ddx = 2 * gx[n], ddy = 2 * gy[n];
        const double c = 1.0 / (0.001 + hr[n] * hr[n] * hr[n]);
        for (int j = 0; j < K; ++j) {
            double s = ddy * (rr[j] - rl[j]) - ddx * (rb[j] - ra[j]);
            Jx[j] = c * ddy * s;
            Jy[j] = -c * ddx * s;
        }
        for (int a = 0; a < K; ++a) {
            E.g[a] -= w * (gsx[n] * Jx[a] + gsy[n] * Jy[a]);
            const double xa = w * Jx[a], ya = w * Jy[a];
            double* Hrow = &E.H[(size_t)a * K];
            for (int b = 0; b < K; ++b) Hrow[b] += xa * Jx[b] + ya * Jy[b];
        }
    }
}

// Lower Cholesky + solve (LAPACK dpotrf/dpotrs semantics, solver.h:70-105).
bool cholesky_solve(int K, std::vector<double> Hm, const double* g, double* delta)
{
    for (int j = 0; j < K; ++j) {
        double d = Hm[(size_t)j * K + j];
        for (int k = 0; k < j; ++k) d -= Hm[(size_t)j * K + k] * Hm[(size_t)j * K + k];
        if (!(d > 0)) return false;
        d = std::sqrt(d);
        Hm[(size_t)j * K + j] = d;
        for (int i = j + 1; i < K; ++i) {
            double s = Hm[(size_t)i * K + j];
            for (int k = 0; k < j; ++k) s -= Hm[(size_t)i * K + k] * Hm[(size_t)j * K + k];
            Hm[(size_t)i * K + j] = s / d;
        }
    }
    std::vector<double> y(K);
    for (int i = 0; i < K; ++i) {
        double s = -g[i];
        for (int k = 0; k < i; ++k) s -= Hm[(size_t)i * K + k] * y[k];
        y[i] = s / Hm[(size_t)i * K + i];
    }
    for (int i = K - 1; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < K; ++k) s -= Hm[(size_t)k * K + i] * delta[k];
        delta[i] = s / Hm[(size_t)i * K + i];
    }
    return true;
}

void setup_frame(Frame& F, const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o)
{
    F.m = m; F.f = f; F.o = o;
    F.L = make_layout(m, o);
    blur_taps(o->blur_sigma, F.taps);
    const size_t A = (size_t)f->width * f->height;
    std::vector<double> raw(A);
    for (size_t i = 0; i < A; ++i) raw[i] = f->obs[i];
    F.obs_blur.assign(A, 0.0);
    blur(raw.data(), F.obs_blur.data(), f->width, f->height, F.taps);
}

// fit_3DMM_shape_rigid (model_fitter.cu:121-375) for T = 1, in FP64.
void fit_one(const Frame& F, double* th, tdi_fit_report* rep)
{
    const tdi_fit_options* o = F.o;
    const int K = F.L.K;
    std::memset(rep, 0, sizeof(*rep));
    int outer = 0, performed = 0, ls_evals = 0;
    bool terminate = false;
    Eval E;
    std::vector<double> delta(K), tt(K);
    for (int i = 0; i < o->max_iters; ++i) {
        outer++;
        if (terminate) break;
        evaluate(F, th, true, E);
        performed++;
        rep->n_pixels = (int)E.vis.size();
        if (!cholesky_solve(K, E.H, E.g.data(), delta.data())) {
            // the reference ignores a potrf failure here (model_fitter.cu:194); we stop.
            rep->stop |= TDI_STOP_NOT_PD;
            break;
        }
        if (std::isnan(delta[0])) {
            rep->stop |= TDI_STOP_NAN;
            break;
        }
        const double obj0 = E.obj;
        double dg = 0;
        for (int k = 0; k < K; ++k) dg += delta[k] * E.g[k];
        double t = 1.0;
        int inner = 0;
        Eval Es;
        while (inner < o->max_ls) {
            if (t < o->t_min) { terminate = true; rep->stop |= TDI_STOP_STEP; break; }
            // fit_3DMM_epsilon_finetune moves all but the six pose unknowns (model_fitter.cu:502)
            for (int k = 0; k < K; ++k) {
                const bool held = F.finetune && k >= F.L.off_pose && k < F.L.off_pose + 6;
                tt[k] = held ? th[k] : th[k] + t * delta[k];
            }
            if (tt[F.L.off_pose + 2] < 0) break;                       // model_fitter.cu:253-259
            evaluate(F, tt.data(), false, Es);
            if (!Es.face_ok || std::isnan(Es.obj_barrier)) { t *= o->armijo_beta; continue; }
            inner++;
            ls_evals++;
            if (-dg < o->decrement_tol) { terminate = true; rep->stop |= TDI_STOP_DECREMENT; break; }
            if (Es.obj < obj0 + o->armijo_alpha * t * dg) {
                std::copy(tt.begin(), tt.end(), th);
                break;
            }
            t *= o->armijo_beta;
        }
    }
    if (!terminate && !(rep->stop & (TDI_STOP_NAN | TDI_STOP_NOT_PD))) rep->stop |= TDI_STOP_MAXITER;
    evaluate(F, th, false, E);
    rep->objective = (float)E.obj;
    rep->iters = performed;
    rep->ls_evals = ls_evals;
    rep->success = outer > o->niters_success;
    rep->status = (rep->stop & TDI_STOP_NOT_PD) ? TDI_NOT_PD : (rep->stop & TDI_STOP_NAN) ? TDI_NAN : TDI_OK;
}

// Light sub-fits: fit_3DMM_lambdas (model_fitter.cu:981-1073) and fit_3DMM_Lintensity
// (model_fitter.cu:1084-1180) with compute_hessian_and_gradient_for_lambda / _for_Lintensity
// (derivative_computer.cu:228-312, 318-400).  Shape and pose stay fixed; the objective is
// -sum g^_r.g^_s without the data weight (render_for_illumination_only, renderer.cu:737-791)
// and the normal equations are J^T J and -J^T g^_s of the light columns.  They are taken from
// the joint evaluation with the light appended to the unknowns: its light block is exactly
// w (J^T J) and -w (J^T g^_s) (the barrier does not involve the light), so both are divided
// by w.
struct LightFit {
    bool intensity;
    double alpha, beta, t_min, dec_tol;
};

void fit_light_one(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o, const double* th,
                   double* lam, double* Lint, tdi_fit_report* rep, const LightFit& lf)
{
    tdi_fit_options ox = *o;
    ox.fit_light = 1;
    Frame F;
    setup_frame(F, m, f, &ox);
    const Layout Lx = F.L;
    const int K = make_layout(m, o).K, Kx = Lx.K, ol = Lx.off_light;
    const double w = o->data_weight;
    std::vector<double> x(Kx), xt(Kx);
    std::copy(th, th + K, x.begin());
    for (int a = 0; a < 3; ++a) x[ol + a] = lam[a];
    x[ol + 3] = *Lint;
    std::memset(rep, 0, sizeof(*rep));
    bool terminate = false;
    int outer = 0, ls = 0;
    Eval E, Es;
    for (int it = 0; it < 10000; ++it) {
        if (terminate) break;
        outer++;
        evaluate(F, x.data(), true, E);
        rep->n_pixels = (int)E.vis.size();
        const double obj0 = E.obj_data / w;
        double dir[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        int nd;
        if (lf.intensity) {
            nd = 1;
            const int j = ol + 3;
            g[0] = E.g[j] / w;
            dir[0] = -g[0] / (E.H[(size_t)j * Kx + j] / w);    // compute_Lintensity_search_dir
        } else {
            nd = 3;
            std::vector<double> H3(9);
            for (int a = 0; a < 3; ++a) {
                g[a] = E.g[ol + a] / w;
                for (int b = 0; b < 3; ++b) H3[a * 3 + b] = E.H[(size_t)(ol + a) * Kx + ol + b] / w;
            }
            if (!cholesky_solve(3, H3, g, dir)) { rep->stop |= TDI_STOP_NOT_PD; break; }
        }
        double dg = 0;
        for (int a = 0; a < nd; ++a) dg += dir[a] * g[a];
        double t = 1.0;
        int inner = 0;
        while (inner < 1000) {
            if (t < lf.t_min) { terminate = true; rep->stop |= TDI_STOP_STEP; break; }
            xt = x;
            for (int a = 0; a < nd; ++a) xt[ol + (lf.intensity ? 3 : a)] = x[ol + (lf.intensity ? 3 : a)] + t * dir[a];
            evaluate(F, xt.data(), false, Es);
            inner++;
            ls++;
            if (-dg < lf.dec_tol) { terminate = true; rep->stop |= TDI_STOP_DECREMENT; break; }
            if (Es.obj_data / w < obj0 + lf.alpha * t * dg) {
                x = xt;
                break;
            }
            t *= lf.beta;
        }
    }
    for (int a = 0; a < 3; ++a) lam[a] = x[ol + a];
    *Lint = x[ol + 3];
    evaluate(F, x.data(), false, E);
    rep->objective = (float)(E.obj_data / w);
    rep->iters = outer;
    rep->ls_evals = ls;
    rep->status = (rep->stop & TDI_STOP_NOT_PD) ? TDI_NOT_PD : TDI_OK;
}

} // namespace

// =====================================================================================
// C entry points (ctypes).  All theta/H/g arrays are double.
// =====================================================================================
namespace {

// ---------------------------------------------------------------- orthographic initialiser
// OrthographicInitializer (orthographic_initializer.cu): a 7-parameter Gauss-Newton fit of
// x = (tau_x, tau_y, sigma_x, sigma_y, u) to the 51 mean-centred landmarks,
//   e_x,i = sigma_x ((R p0L_i)_x + tau_x) - xl_i,   e_y,i = sigma_y ((R p0L_i)_y + tau_y) - yl_i
// (compute_gradient_hessian_obj_orthographic, orthographic_initializer.cu:56-131), started at
// (0, 0, 15, 15, 0, 0, 0.05) (reset_orthographic, :133-150), at most 100 outer iterations, an
// Armijo search with alpha 0.4 halving from t = 1 and stopping the fit once t < 0.1
// (fit_model, :243-364).  FP64 here; the reference is FP32 with cuBLAS sums.
constexpr int ORTHO_MAXITER = 100;
constexpr int ORTHO_MAXINNER = 1000;

// Residuals e[102] (x rows then y rows) and, if J, the Jacobian J[102][7] (row-major).
void ortho_residuals(const double* x, const double* xlb, const double* ylb, const double* p0L, double* e, double* J)
{
    double R[9], dR[27];
    rodrigues(x + 4, R, J ? dR : nullptr);
    for (int i = 0; i < NL; ++i) {
        double Rp[3];
        rot(R, p0L + 3 * i, Rp);
        e[i] = x[2] * (Rp[0] + x[0]) - xlb[i];
        e[NL + i] = x[3] * (Rp[1] + x[1]) - ylb[i];
        if (!J) continue;
        double* jx = J + 7 * i;
        double* jy = J + 7 * (NL + i);
        jx[0] = x[2]; jx[1] = 0;    jx[2] = x[0] + Rp[0]; jx[3] = 0;
        jy[0] = 0;    jy[1] = x[3]; jy[2] = 0;            jy[3] = x[1] + Rp[1];
        for (int k = 0; k < 3; ++k) {
            double d[3];
            rot(dR + 9 * k, p0L + 3 * i, d);
            jx[4 + k] = x[2] * d[0];
            jy[4 + k] = x[3] * d[1];
        }
    }
}

// RotationComputer::compute_euler_angles (rotation_header.cu:421-449), R column-major, radians.
void euler_angles(const double* R, double& yaw, double& pitch, double& roll)
{
    const double cp = std::sqrt(R[5] * R[5] + R[8] * R[8]);
    if (!(cp < 1e-6)) {
        roll = std::atan2(R[1], R[0]);
        yaw = std::atan2(-R[2], cp);
        pitch = std::atan2(R[5], R[8]);
    } else {
        roll = std::atan2(-R[7], R[4]);
        yaw = std::atan2(-R[2], cp);
        pitch = 0.0;
    }
}

// eulerAnglesToRotationMatrix (newfuncs.cu:720-747): M = Rz(e2) Ry(e1) Rx(e0), row-major,
// followed by the rotation vector of M (the matrix-to-vector direction of cv::Rodrigues, used at
// logbarrier_initializer.cu:1883): r = theta / (2 sin theta) (M21 - M12, M02 - M20, M10 - M01),
// with the theta -> 0 and theta -> pi cases of that published algorithm.
void rotvec_from_euler(const double* e, double* r)
{
    const double c0 = std::cos(e[0]), s0 = std::sin(e[0]);
    const double c1 = std::cos(e[1]), s1 = std::sin(e[1]);
    const double c2 = std::cos(e[2]), s2 = std::sin(e[2]);
    const double Rx[9] = {1, 0, 0, 0, c0, -s0, 0, s0, c0};
    const double Ry[9] = {c1, 0, s1, 0, 1, 0, -s1, 0, c1};
    const double Rz[9] = {c2, -s2, 0, s2, c2, 0, 0, 0, 1};
    double T[9], M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += Ry[3 * i + k] * Rx[3 * k + j];
            T[3 * i + j] = a;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += Rz[3 * i + k] * T[3 * k + j];
            M[3 * i + j] = a;
        }
    double v[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
    const double sn = 0.5 * std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double cs = std::min(1.0, std::max(-1.0, 0.5 * (M[0] + M[4] + M[8] - 1.0)));
    const double th = std::acos(cs);
    if (sn < 1e-5) {
        if (cs > 0) {
            r[0] = r[1] = r[2] = 0;
            return;
        }
        double rx = std::sqrt(std::max(0.5 * (M[0] + 1.0), 0.0));
        double ry = std::sqrt(std::max(0.5 * (M[4] + 1.0), 0.0)) * (M[1] < 0 ? -1.0 : 1.0);
        double rz = std::sqrt(std::max(0.5 * (M[8] + 1.0), 0.0)) * (M[2] < 0 ? -1.0 : 1.0);
        if (std::fabs(rx) < std::fabs(ry) && std::fabs(rx) < std::fabs(rz) && ((M[5] > 0) != (ry * rz > 0))) rz = -rz;
        const double nr = std::sqrt(rx * rx + ry * ry + rz * rz);
        r[0] = rx * th / nr; r[1] = ry * th / nr; r[2] = rz * th / nr;
        return;
    }
    const double f = th / (2.0 * sn);
    r[0] = v[0] * f; r[1] = v[1] * f; r[2] = v[2] * f;
}

// One run of fit_model from x (in place).  Returns the outer iterations (normal-equation
// solves); Rlast is the rotation of the last evaluated line-search trial, which is where the
// reference reads its yaw/pitch/roll (rc_linesearch, orthographic_initializer.cu:366-372).
int ortho_fit(double* x, const double* xlb, const double* ylb, const double* p0L, double* Rlast, int* status)
{
    double e[2 * NL], J[2 * NL * 7], et[2 * NL];
    bool terminate = false;
    int q = 0;
    rodrigues(x + 4, Rlast, nullptr);
    for (; q < ORTHO_MAXITER && !terminate; ++q) {
        ortho_residuals(x, xlb, ylb, p0L, e, J);
        std::vector<double> H(49, 0.0);
        double g[7] = {0, 0, 0, 0, 0, 0, 0}, obj = 0;
        for (int r = 0; r < 2 * NL; ++r) {
            for (int a = 0; a < 7; ++a) {
                g[a] += J[7 * r + a] * e[r];
                for (int b = 0; b < 7; ++b) H[7 * a + b] += J[7 * r + a] * J[7 * r + b];
            }
            obj += e[r] * e[r];
        }
        double d[7];
        if (!cholesky_solve(7, H, g, d)) {
            // the reference ignores a potrf failure here; the frame stops instead
            *status = TDI_NOT_PD;
            return q + 1;
        }
        double dd = 0;
        for (int a = 0; a < 7; ++a) dd += d[a] * g[a];
        double t = 1.0;
        for (int inner = 0; inner < ORTHO_MAXINNER; ++inner) {
            if (t < 0.1) {
                terminate = true;
                break;
            }
            double xt[7];
            for (int a = 0; a < 7; ++a) xt[a] = x[a] + t * d[a];
            ortho_residuals(xt, xlb, ylb, p0L, et, nullptr);
            rodrigues(xt + 4, Rlast, nullptr);
            double ot = 0;
            for (int r = 0; r < 2 * NL; ++r) ot += et[r] * et[r];
            if (ot < obj + 0.4 * t * dd) {
                for (int a = 0; a < 7; ++a) x[a] = xt[a];
                break;
            }
            t *= 0.5;
        }
    }
    return q;
}

// Logbarrier_Initializer::initialize_with_orthographic_t (logbarrier_initializer.cu:1862-1914)
// without its landmark-box part: the orthographic fit, the retry from the mirrored pose when
// sigma_x < 0, and initialize_kernel (logbarrier_initializer.cu:31-46):
//   tau_z = phix / ((|sigma_x| + |sigma_y|) / 2), tau_x = tau_z (mean_x - cx) / phix,
//   tau_y = tau_z (mean_y - cy) / phiy, u = u_orth.
void ortho_init_one(const float* lx, const float* ly, const float* cam, const double* p0L, double* xo, double* pose,
                    double* ang, int32_t* info)
{
    double mx = 0, my = 0;
    for (int i = 0; i < NL; ++i) { mx += lx[i]; my += ly[i]; }
    mx /= NL; my /= NL;
    double xlb[NL], ylb[NL];
    for (int i = 0; i < NL; ++i) { xlb[i] = lx[i] - mx; ylb[i] = ly[i] - my; }
    double x[7] = {0.0, 0.0, 15.0, 15.0, 0.0, 0.0, 0.05};
    double Rl[9];
    int status = TDI_OK;
    info[0] = ortho_fit(x, xlb, ylb, p0L, Rl, &status);
    info[1] = 0;
    info[2] = 0;
    double yaw, pitch, roll;
    euler_angles(Rl, yaw, pitch, roll);
    if (status == TDI_OK && x[2] < 0) {
        const double ev[3] = {roll, -pitch, -yaw};
        double u0[3];
        rotvec_from_euler(ev, u0);
        x[0] = -x[0];
        x[2] = -x[2];
        x[4] = u0[0]; x[5] = u0[1]; x[6] = u0[2];
        info[1] = ortho_fit(x, xlb, ylb, p0L, Rl, &status);
        info[2] = 1;
        euler_angles(Rl, yaw, pitch, roll);
    }
    info[3] = status;
    for (int a = 0; a < 7; ++a) xo[a] = x[a];
    const double tz = (double)cam[0] / ((std::fabs(x[2]) + std::fabs(x[3])) * 0.5);
    pose[0] = tz * (mx - cam[2]) / cam[0];
    pose[1] = tz * (my - cam[3]) / cam[1];
    pose[2] = tz;
    pose[3] = x[4]; pose[4] = x[5]; pose[5] = x[6];
    const double r2d = 180.0 / M_PI;
    ang[0] = yaw * r2d; ang[1] = pitch * r2d; ang[2] = roll * r2d;
}

// ---------------------------------------------------------------- render loss (differentiation check)
// The loss of the adjoint: over the visible pixels, the squared difference of the normalised
// gradients of the blurred render and of the blurred observation (the data term's residuals,
// fill_grads, funcs.cu:842-897; sum |g^_r - g^_s|^2 = const - (2 / w) * the fit's data objective
// when |g^| = 1), plus the squared reprojection error of the 51 landmark vertices in image pixels.
// The model-space vertices may carry per-vertex offsets dverts [V][3].  With tri == null or
// frozen == 0 the visibility is the rasteriser's (tri [W*H] receives the triangle per pixel, -1
// empty, when given); with frozen != 0 the triangle of each pixel is tri's, and the barycentrics
// are the exact FP64 ones of the pixel centre in the projected triangle (visibility held fixed,
// the render a smooth function of the vertices).
double render_loss(const Frame& F, const double* th, const double* dverts, int32_t* tri, int frozen)
{
    const tdi_model_desc* m = F.m;
    const tdi_frame_desc* f = F.f;
    const int W = f->width, Hh = f->height, V = m->n_vertices;
    const size_t A = (size_t)W * Hh;
    Geometry G;
    geometry(F, th, false, G);
    if (dverts) {
        for (int v = 0; v < V; ++v) {
            double q[3];
            for (int a = 0; a < 3; ++a) G.P0[3 * v + a] += dverts[3 * v + a];
            rot(G.R, &G.P0[3 * v], q);
            for (int a = 0; a < 3; ++a) { q[a] += G.tau[a]; G.P[3 * v + a] = q[a]; }
            G.xp[v] = f->phix * q[0] / q[2] + f->cx;
            G.yp[v] = f->phiy * q[1] / q[2] + f->cy;
            G.xpf[v] = (float)G.xp[v];
            G.ypf[v] = (float)G.yp[v];
            G.zf[v] = (float)q[2];
        }
    }
    std::vector<int> tmap(A, -1);
    std::vector<double> bal(A, 0.0), bbe(A, 0.0), bga(A, 0.0);
    if (tri && frozen) {
        for (size_t k = 0; k < A; ++k) {
            const int t = tri[k];
            tmap[k] = t;
            if (t < 0) continue;
            const int ib = m->triangles[3 * t], ic = m->triangles[3 * t + 1], ia = m->triangles[3 * t + 2];
            const double px = f->x0 + (double)(k / Hh), py = f->y0 + (double)(k % Hh);
            const double ux = G.xp[ib] - G.xp[ia], uy = G.yp[ib] - G.yp[ia];
            const double vx = G.xp[ic] - G.xp[ia], vy = G.yp[ic] - G.yp[ia];
            const double wx = px - G.xp[ia], wy = py - G.yp[ia];
            const double D = ux * vy - uy * vx;
            bal[k] = (wx * vy - wy * vx) / D;
            bbe[k] = (ux * wy - uy * wx) / D;
            bga[k] = 1.0 - bal[k] - bbe[k];
        }
    } else {
        std::vector<uint64_t> keys(A);
        std::vector<float> bary(2 * A);
        rasterize(m->n_triangles, m->triangles, G.xpf.data(), G.ypf.data(), G.zf.data(), f->x0, f->y0, W, Hh,
                  F.o->ntmp_cap, keys.data(), bary.data());
        for (size_t k = 0; k < A; ++k) {
            if (keys[k] == ~0ull) continue;
            tmap[k] = (int)(keys[k] & 0xffffffffu);
            bal[k] = bary[2 * k];
            bbe[k] = bary[2 * k + 1];
            bga[k] = (1.0f - bary[2 * k]) - bary[2 * k + 1];   // the renderer's FP32 gamma
        }
        if (tri) std::copy(tmap.begin(), tmap.end(), tri);
    }
    std::vector<double> I(A, 0.0);
    double dlt[3] = {G.lam[0] - G.tau[0], G.lam[1] - G.tau[1], G.lam[2] - G.tau[2]}, rl[3];
    rotT(G.R, dlt, rl);
    for (size_t k = 0; k < A; ++k) {
        const int t = tmap[k];
        if (t < 0) continue;
        const int ib = m->triangles[3 * t], ic = m->triangles[3 * t + 1], ia = m->triangles[3 * t + 2];
        const double al = bal[k], be = bbe[k], ga = bga[k];
        const double* B0 = &G.P0[3 * ib]; const double* C0 = &G.P0[3 * ic]; const double* A0 = &G.P0[3 * ia];
        double e1[3] = {B0[0] - A0[0], B0[1] - A0[1], B0[2] - A0[2]};
        double e2[3] = {C0[0] - A0[0], C0[1] - A0[1], C0[2] - A0[2]};
        double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
        double nrm[3], p0[3];
        for (int a = 0; a < 3; ++a) nrm[a] = nr[a] / nn;
        for (int a = 0; a < 3; ++a) p0[a] = al * B0[a] + be * C0[a] + ga * A0[a];
        const double Id = nrm[0] * rl[0] + nrm[1] * rl[1] + nrm[2] * rl[2] - (nrm[0] * p0[0] + nrm[1] * p0[1] + nrm[2] * p0[2]);
        const double rho = al * G.rho[ib] + be * G.rho[ic] + ga * G.rho[ia];
        double nc[3], Y[5];
        rot(G.R, nrm, nc);
        sh2_terms(nc, Y);
        double S = nc[0] * G.a[0] + nc[1] * G.a[1] + nc[2] * G.a[2];
        for (int j = 0; j < 5; ++j) S += G.sh[j] * Y[j];
        I[k] = rho + G.Lint * rho * Id + rho * S;
    }
    std::vector<double> B(A, 0.0);
    blur(I.data(), B.data(), W, Hh, F.taps);
    double loss = 0.0;
    for (size_t k = 0; k < A; ++k) {
        if (tmap[k] < 0) continue;
        const int x = (int)(k / Hh), y = (int)(k % Hh);
        const int kr = (x < W - 1) ? (int)k + Hh : 0, kl = (x > 0) ? (int)k - Hh : 0;
        const int ka = (y > 0) ? (int)k - 1 : 0, kb = (y < Hh - 1) ? (int)k + 1 : 0;
        const double gx = 0.5 * (B[kr] - B[kl]), gy = 0.5 * (B[kb] - B[ka]);
        const double h = 1e-16 + std::sqrt(4 * gx * gx + 4 * gy * gy);
        const double sx = 0.5 * (F.obs_blur[kr] - F.obs_blur[kl]), sy = 0.5 * (F.obs_blur[kb] - F.obs_blur[ka]);
        const double hs = 1e-16 + std::sqrt(4 * sx * sx + 4 * sy * sy);
        const double rx = 2 * gx / h - 2 * sx / hs, ry = 2 * gy / h - 2 * sy / hs;
        loss += rx * rx + ry * ry;
    }
    for (int i = 0; i < NL; ++i) {
        const int v = m->landmark_ids[i];
        const double dx = G.xp[v] - f->lmk_x[i], dy = G.yp[v] - f->lmk_y[i];
        loss += dx * dx + dy * dy;
    }
    return loss;
}

// ---------------------------------------------------------------- landmark-only fit with slack
// Logbarrier_Initializer::set_minimal_slack (logbarrier_initializer.cu:1041-1090) and fit_model
// with use_slack (1095-1267) for one frame (T = 1), no identity unknowns (li_init,
// video_fitter.cu:30), Kl expression coefficients.  Unknowns z = (eps[0..Kl), tau(3), u(3), s).
// Objective (compute_gradient_and_hessian / evaluate_objective_function, 1272-1842, 1921-2081):
//   tau_b s + sum_{i, 4 sides} -log G + sum_k -log(ub_k - eps_k) - log(eps_k - lb_k),
//   G_x- = s + xproj - (xl + blx fs), G_x+ = s + (xl + bux fs) - xproj (same for y) (compute_Gs,
//   458-518); gradient sum -dG/G over the sides (no epsilon) and tau_b - sum 1/(1e-6 + G) for s;
//   Gauss-Newton Hessian sum dG dG^T / G^2 (fill_optimization_*_landmark, 587-910), the s column
//   sum dG / G^2 (fill_for_nabla2F_dsdc), the s diagonal sum 1/(1e-6 + G)^2, the expression
//   bounds' diagonal 1/(ub-eps+1e-12)^2 + 1/(eps-lb+1e-12)^2 (update_diagonal_of_hessian_wbounds).
// The landmark positions are the mean shape plus the fixed identity terms of theta plus the
// first Kl expression columns times eps.
struct LmkFit {
    const tdi_model_desc* m;
    const tdi_frame_desc* f;
    int Kl = 0, n = 0;
    double base[NL][3];
};

void lmk_fit_setup(LmkFit& S, const tdi_model_desc* m, const tdi_frame_desc* f, const Layout& L, const double* th,
                   int Kl)
{
    S.m = m; S.f = f; S.Kl = Kl; S.n = Kl + 7;
    for (int i = 0; i < NL; ++i) {
        const int v = m->landmark_ids[i];
        for (int a = 0; a < 3; ++a) {
            const int r = 3 * v + a;
            double x = m->mean_shape[r];
            for (int k = 0; k < L.nid; ++k) x += (double)m->id_basis[(size_t)r * m->k_id + k] * th[L.off_id + k];
            S.base[i][a] = x;
        }
    }
}

// Objective at z; with g and H (n and n x n, overwritten) the gradient and Gauss-Newton Hessian.
// gmin: the least G (may be null).
double lmk_fit_terms(const LmkFit& S, const double* z, double tau_b, double* g, double* H, double* gmin)
{
    const tdi_model_desc* m = S.m;
    const tdi_frame_desc* f = S.f;
    const int Kl = S.Kl, n = S.n, np = n - 1;
    const double s = z[np];
    double R[9], dR[27];
    rodrigues(z + Kl + 3, R, g ? dR : nullptr);
    if (g) { std::fill(g, g + n, 0.0); std::fill(H, H + (size_t)n * n, 0.0); }
    double obj = tau_b * s, lo = INFINITY;
    for (int k = 0; k < Kl; ++k) {
        if (!(m->exp_ub[k] < INFINITY) || !(m->exp_lb[k] > -INFINITY)) continue;   // unbounded: no term
        const double x = z[k], fub = x - m->exp_ub[k], flb = m->exp_lb[k] - x;
        obj += -std::log(-fub) - std::log(-flb);
        if (g) {
            g[k] += -1.0 / fub + 1.0 / flb;
            const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
            H[(size_t)k * n + k] += a * a + b * b;
        }
    }
    std::vector<double> dx(np), dy(np);
    const double fs = f->face_size;
    for (int i = 0; i < NL; ++i) {
        const int v = m->landmark_ids[i];
        double p[3];
        for (int a = 0; a < 3; ++a) {
            double x = S.base[i][a];
            for (int k = 0; k < Kl; ++k) x += (double)m->exp_basis[(size_t)(3 * v + a) * m->k_exp + k] * z[k];
            p[a] = x;
        }
        double c[3];
        for (int a = 0; a < 3; ++a) c[a] = R[a] * p[0] + R[3 + a] * p[1] + R[6 + a] * p[2] + z[Kl + a];
        const double xp = f->phix * c[0] / c[2] + f->cx, yp = f->phiy * c[1] / c[2] + f->cy;
        const double Gxm = s + xp - (f->lmk_x[i] + f->box_lx[i] * fs);
        const double Gym = s + yp - (f->lmk_y[i] + f->box_ly[i] * fs);
        const double Gxp = s + (f->lmk_x[i] + f->box_ux[i] * fs) - xp;
        const double Gyp = s + (f->lmk_y[i] + f->box_uy[i] * fs) - yp;
        obj += -std::log(Gxm) - std::log(Gym) - std::log(Gxp) - std::log(Gyp);
        for (double G : {Gxm, Gym, Gxp, Gyp}) lo = (G != G || lo != lo) ? NAN : std::min(lo, G);   // NaN wins
        if (!g) continue;
        for (int j = 0; j < np; ++j) {
            double dv[3];
            if (j < Kl) {
                double e[3];
                for (int a = 0; a < 3; ++a) e[a] = m->exp_basis[(size_t)(3 * v + a) * m->k_exp + j];
                for (int a = 0; a < 3; ++a) dv[a] = R[a] * e[0] + R[3 + a] * e[1] + R[6 + a] * e[2];
            } else if (j < Kl + 3) {
                dv[0] = dv[1] = dv[2] = 0.0;
                dv[j - Kl] = 1.0;
            } else {
                const double* d = dR + 9 * (j - Kl - 3);
                for (int a = 0; a < 3; ++a) dv[a] = d[a] * p[0] + d[3 + a] * p[1] + d[6 + a] * p[2];
            }
            dx[j] = f->phix * (dv[0] * c[2] - dv[2] * c[0]) / (c[2] * c[2]);
            dy[j] = f->phiy * (dv[1] * c[2] - dv[2] * c[1]) / (c[2] * c[2]);
        }
        const double cx = -1.0 / Gxm + 1.0 / Gxp, cy = -1.0 / Gym + 1.0 / Gyp;
        const double wx = 1.0 / (Gxm * Gxm) + 1.0 / (Gxp * Gxp), wy = 1.0 / (Gym * Gym) + 1.0 / (Gyp * Gyp);
        const double hx = 1.0 / (Gxm * Gxm) - 1.0 / (Gxp * Gxp), hy = 1.0 / (Gym * Gym) - 1.0 / (Gyp * Gyp);
        for (int a = 0; a < np; ++a) {
            g[a] += cx * dx[a] + cy * dy[a];
            for (int b = 0; b < np; ++b) H[(size_t)a * n + b] += wx * dx[a] * dx[b] + wy * dy[a] * dy[b];
            const double hs = hx * dx[a] + hy * dy[a];
            H[(size_t)a * n + np] += hs;
            H[(size_t)np * n + a] += hs;
        }
        const double q[4] = {1.0 / (1e-6 + Gxm), 1.0 / (1e-6 + Gym), 1.0 / (1e-6 + Gxp), 1.0 / (1e-6 + Gyp)};
        for (int k = 0; k < 4; ++k) {
            g[np] -= q[k];
            H[(size_t)np * n + np] += q[k] * q[k];
        }
    }
    if (g) g[np] += tau_b;
    if (gmin) *gmin = lo;
    return obj;
}

// The whole fit; z in: the pose (z[Kl..Kl+6)); out: the final unknowns.  Returns the report.
tdi_lmk_report lmk_fit_run(const LmkFit& S, double* z)
{
    const int Kl = S.Kl, n = S.n;
    tdi_lmk_report rep{};
    for (int k = 0; k < Kl; ++k) z[k] = 0.0;
    // set_minimal_slack: s = 5 - min G at s = 0
    z[n - 1] = 0.0;
    double gmin = 0.0;
    lmk_fit_terms(S, z, 0.0, nullptr, nullptr, &gmin);
    z[n - 1] = 5.0 - gmin;
    rep.slack0 = (float)z[n - 1];
    std::vector<double> g(n), H((size_t)n * n), d(n), zt(n);
    double tau_b = 10.0;
    bool success = false;
    for (int round = 0; round < 5 && !success; ++round) {
        tau_b *= 10.0;
        rep.rounds++;
        for (int it = 0; it < 100; ++it) {
            if (z[n - 1] < -0.01) { success = true; break; }
            const double obj = lmk_fit_terms(S, z, tau_b, g.data(), H.data(), nullptr);
            rep.iterations++;
            if (!cholesky_solve(n, H, g.data(), d.data())) break;
            bool stop = false;
            double t = 1.0;
            for (int inner = 0; inner < 100; ++inner) {
                if (t < 0.01) { stop = true; break; }
                for (int j = 0; j < n; ++j) zt[j] = z[j] + t * d[j];
                const double ot = lmk_fit_terms(S, zt.data(), tau_b, nullptr, nullptr, nullptr);
                if (ot < obj) {
                    std::copy(zt.begin(), zt.end(), z);
                    rep.accepted++;
                    break;
                }
                t *= 0.5;
            }
            if (stop) break;
        }
    }
    rep.status = success ? TDI_OK : TDI_LMK_INFEASIBLE;
    rep.slack = (float)z[n - 1];
    return rep;
}

// ---------------------------------------------------------------- joint landmark-only fit over T frames
// Logbarrier_Initializer::fit_model over T frames with identity unknowns
// (fit_to_multi_images_landmarks_only, model_fitter.cu:1205-1247, and the two landmark fits at
// the start of fit_to_multi_images, 1291-1330: li_init with K_ALPHA_L = 60 identity
// coefficients and the slack, then li on all identity coefficients without it).  The frames share
// the first Ka identity coefficients (and the slack s); each has its own Ke expression
// coefficients and pose.  Joint order z = [alpha (Ka) | s (if slack) | frame 0 (eps Ke, tau 3,
// u 3) | frame 1 | ...].  compute_gradient_and_hessian (1272-1842) sums the frames' landmark
// terms (the identity columns of fill_optimization_dI_dalpha_landmark next to the expression and
// pose columns), adds the identity bounds once and the slack terms once (tau_b s, and per
// constraint the s derivatives as for one frame); without the slack s = 0 and is no unknown.
// Identity columns past Ka keep theta's values (frame 0's) in the landmark positions.
struct LmkJoint {
    const tdi_model_desc* m;
    const tdi_frame_desc* f;
    int T = 0, Ka = 0, Ke = 0, S1 = 0, n = 0;
    std::vector<double> base;   // [T][NL][3]
};

void lmk_joint_setup(LmkJoint& S, const tdi_model_desc* m, const tdi_frame_desc* frames, int T, const Layout& L,
                     const double* theta, int Ka, int Ke, int slack)
{
    S.m = m; S.f = frames; S.T = T; S.Ka = Ka; S.Ke = Ke; S.S1 = slack ? 1 : 0;
    S.n = Ka + S.S1 + T * (Ke + 6);
    S.base.assign((size_t)T * NL * 3, 0.0);
    for (int t = 0; t < T; ++t)
        for (int i = 0; i < NL; ++i) {
            const int v = m->landmark_ids[i];
            for (int a = 0; a < 3; ++a) {
                const int r = 3 * v + a;
                double x = m->mean_shape[r];
                for (int k = Ka; k < L.nid; ++k) x += (double)m->id_basis[(size_t)r * m->k_id + k] * theta[L.off_id + k];
                S.base[((size_t)t * NL + i) * 3 + a] = x;
            }
        }
}

// Objective at z; with g (n) and H (n x n) the gradient and Gauss-Newton Hessian.  gmin: the
// least G over the frames at z (may be null).
double lmk_joint_terms(const LmkJoint& S, const double* z, double tau_b, double* g, double* H, double* gmin)
{
    const tdi_model_desc* m = S.m;
    const int Ka = S.Ka, Ke = S.Ke, n = S.n, is = Ka;   // is: the slack's index (when S1)
    const double s = S.S1 ? z[is] : 0.0;
    if (g) { std::fill(g, g + n, 0.0); std::fill(H, H + (size_t)n * n, 0.0); }
    double obj = S.S1 ? tau_b * s : 0.0, lo = INFINITY;
    // the identity bounds, once
    for (int k = 0; k < Ka; ++k) {
        if (!(m->id_ub[k] < INFINITY) || !(m->id_lb[k] > -INFINITY)) continue;
        const double x = z[k], fub = x - m->id_ub[k], flb = m->id_lb[k] - x;
        obj += -std::log(-fub) - std::log(-flb);
        if (g) {
            g[k] += -1.0 / fub + 1.0 / flb;
            const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
            H[(size_t)k * n + k] += a * a + b * b;
        }
    }
    const int nl = Ka + Ke + 6;                 // a frame's columns: alpha, eps, tau, u
    std::vector<int> col(nl);
    std::vector<double> dx(nl), dy(nl);
    for (int t = 0; t < S.T; ++t) {
        const tdi_frame_desc* f = &S.f[t];
        const int zf = Ka + S.S1 + t * (Ke + 6);
        for (int j = 0; j < nl; ++j) col[j] = j < Ka ? j : zf + (j - Ka);
        const double* eps = z + zf;
        const double* tau = z + zf + Ke;
        const double* u = z + zf + Ke + 3;
        double R[9], dR[27];
        rodrigues(u, R, g ? dR : nullptr);
        for (int k = 0; k < Ke; ++k) {
            if (!(m->exp_ub[k] < INFINITY) || !(m->exp_lb[k] > -INFINITY)) continue;
            const double x = eps[k], fub = x - m->exp_ub[k], flb = m->exp_lb[k] - x;
            obj += -std::log(-fub) - std::log(-flb);
            if (g) {
                g[zf + k] += -1.0 / fub + 1.0 / flb;
                const double a = -1.0 / (fub + 1e-12), b = 1.0 / (flb + 1e-12);
                H[(size_t)(zf + k) * n + zf + k] += a * a + b * b;
            }
        }
        const double fs = f->face_size;
        for (int i = 0; i < NL; ++i) {
            const int v = m->landmark_ids[i];
            double p[3];
            for (int a = 0; a < 3; ++a) {
                const int r = 3 * v + a;
                double x = S.base[((size_t)t * NL + i) * 3 + a];
                for (int k = 0; k < Ka; ++k) x += (double)m->id_basis[(size_t)r * m->k_id + k] * z[k];
                for (int k = 0; k < Ke; ++k) x += (double)m->exp_basis[(size_t)r * m->k_exp + k] * eps[k];
                p[a] = x;
            }
            double c[3];
            for (int a = 0; a < 3; ++a) c[a] = R[a] * p[0] + R[3 + a] * p[1] + R[6 + a] * p[2] + tau[a];
            const double xp = f->phix * c[0] / c[2] + f->cx, yp = f->phiy * c[1] / c[2] + f->cy;
            const double Gxm = s + xp - (f->lmk_x[i] + f->box_lx[i] * fs);
            const double Gym = s + yp - (f->lmk_y[i] + f->box_ly[i] * fs);
            const double Gxp = s + (f->lmk_x[i] + f->box_ux[i] * fs) - xp;
            const double Gyp = s + (f->lmk_y[i] + f->box_uy[i] * fs) - yp;
            obj += -std::log(Gxm) - std::log(Gym) - std::log(Gxp) - std::log(Gyp);
            for (double G : {Gxm, Gym, Gxp, Gyp}) lo = (G != G || lo != lo) ? NAN : std::min(lo, G);
            if (!g) continue;
            for (int j = 0; j < nl; ++j) {
                double dv[3];
                if (j < Ka || j < Ka + Ke) {
                    const bool idc = j < Ka;
                    double e[3];
                    for (int a = 0; a < 3; ++a)
                        e[a] = idc ? m->id_basis[(size_t)(3 * v + a) * m->k_id + j]
                                   : m->exp_basis[(size_t)(3 * v + a) * m->k_exp + (j - Ka)];
                    for (int a = 0; a < 3; ++a) dv[a] = R[a] * e[0] + R[3 + a] * e[1] + R[6 + a] * e[2];
                } else if (j < Ka + Ke + 3) {
                    dv[0] = dv[1] = dv[2] = 0.0;
                    dv[j - Ka - Ke] = 1.0;
                } else {
                    const double* d = dR + 9 * (j - Ka - Ke - 3);
                    for (int a = 0; a < 3; ++a) dv[a] = d[a] * p[0] + d[3 + a] * p[1] + d[6 + a] * p[2];
                }
                dx[j] = f->phix * (dv[0] * c[2] - dv[2] * c[0]) / (c[2] * c[2]);
                dy[j] = f->phiy * (dv[1] * c[2] - dv[2] * c[1]) / (c[2] * c[2]);
            }
            const double cx = -1.0 / Gxm + 1.0 / Gxp, cy = -1.0 / Gym + 1.0 / Gyp;
            const double wx = 1.0 / (Gxm * Gxm) + 1.0 / (Gxp * Gxp), wy = 1.0 / (Gym * Gym) + 1.0 / (Gyp * Gyp);
            const double hx = 1.0 / (Gxm * Gxm) - 1.0 / (Gxp * Gxp), hy = 1.0 / (Gym * Gym) - 1.0 / (Gyp * Gyp);
            for (int a = 0; a < nl; ++a) {
                const int A = col[a];
                g[A] += cx * dx[a] + cy * dy[a];
                for (int b = 0; b < nl; ++b) H[(size_t)A * n + col[b]] += wx * dx[a] * dx[b] + wy * dy[a] * dy[b];
                if (S.S1) {
                    const double hs = hx * dx[a] + hy * dy[a];
                    H[(size_t)A * n + is] += hs;
                    H[(size_t)is * n + A] += hs;
                }
            }
            if (S.S1) {
                const double q[4] = {1.0 / (1e-6 + Gxm), 1.0 / (1e-6 + Gym), 1.0 / (1e-6 + Gxp), 1.0 / (1e-6 + Gyp)};
                for (int k = 0; k < 4; ++k) {
                    g[is] -= q[k];
                    H[(size_t)is * n + is] += q[k] * q[k];
                }
            }
        }
    }
    if (g && S.S1) g[is] += tau_b;
    if (gmin) *gmin = lo;
    return obj;
}

// The whole fit (fit_model): with the slack, s starts at max over the frames of 5 - min G at
// s = 0 (set_minimal_slack, 1041-1090) and the fit succeeds once s < -0.01; without it, all five
// rounds run until a line search fails.  z in: the start; out: the final unknowns.
tdi_lmk_report lmk_joint_run(const LmkJoint& S, double* z)
{
    const int n = S.n;
    tdi_lmk_report rep{};
    if (S.S1) {
        z[S.Ka] = 0.0;
        double hi = -INFINITY;
        for (int t = 0; t < S.T; ++t) {
            LmkJoint one = S;
            one.T = 1;
            one.f = S.f + t;
            one.n = S.Ka + 1 + (S.Ke + 6);
            one.base.assign(S.base.begin() + (size_t)t * NL * 3, S.base.begin() + (size_t)(t + 1) * NL * 3);
            std::vector<double> zt(one.n);
            for (int k = 0; k < S.Ka; ++k) zt[k] = z[k];
            zt[S.Ka] = 0.0;
            for (int k = 0; k < S.Ke + 6; ++k) zt[S.Ka + 1 + k] = z[S.Ka + 1 + (size_t)t * (S.Ke + 6) + k];
            double gmin = 0.0;
            lmk_joint_terms(one, zt.data(), 0.0, nullptr, nullptr, &gmin);
            const double st = 5.0 - gmin;
            hi = (st != st || hi != hi) ? NAN : std::max(hi, st);
        }
        z[S.Ka] = hi;
        rep.slack0 = (float)hi;
    }
    std::vector<double> g(n), H((size_t)n * n), d(n), zt(n);
    double tau_b = 10.0;
    bool success = false;
    for (int round = 0; round < 5 && !success; ++round) {
        tau_b *= 10.0;
        rep.rounds++;
        for (int it = 0; it < 100; ++it) {
            if (S.S1 && z[S.Ka] < -0.01) { success = true; break; }
            const double obj = lmk_joint_terms(S, z, tau_b, g.data(), H.data(), nullptr);
            rep.iterations++;
            if (!cholesky_solve(n, H, g.data(), d.data())) break;
            bool stop = false;
            double t = 1.0;
            for (int inner = 0; inner < 100; ++inner) {
                if (t < 0.01) { stop = true; break; }
                for (int j = 0; j < n; ++j) zt[j] = z[j] + t * d[j];
                const double ot = lmk_joint_terms(S, zt.data(), tau_b, nullptr, nullptr, nullptr);
                if (ot < obj) {
                    std::copy(zt.begin(), zt.end(), z);
                    rep.accepted++;
                    break;
                }
                t *= 0.5;
            }
            if (stop) break;
        }
    }
    rep.status = (!S.S1 || success) ? TDI_OK : TDI_LMK_INFEASIBLE;
    rep.slack = S.S1 ? (float)z[S.Ka] : 0.f;
    return rep;
}

// theta [T][K] -> z (alpha from frame 0) and back (alpha to every frame).
void lmk_joint_pack(const LmkJoint& S, const Layout& L, const double* theta, double* z)
{
    for (int k = 0; k < S.Ka; ++k) z[k] = theta[L.off_id + k];
    if (S.S1) z[S.Ka] = 0.0;
    for (int t = 0; t < S.T; ++t) {
        const double* th = theta + (size_t)t * L.K;
        double* zf = z + S.Ka + S.S1 + (size_t)t * (S.Ke + 6);
        for (int k = 0; k < S.Ke; ++k) zf[k] = th[L.off_exp + k];
        for (int a = 0; a < 6; ++a) zf[S.Ke + a] = th[L.off_pose + a];
    }
}

void lmk_joint_unpack(const LmkJoint& S, const Layout& L, const double* z, double* theta)
{
    for (int t = 0; t < S.T; ++t) {
        double* th = theta + (size_t)t * L.K;
        const double* zf = z + S.Ka + S.S1 + (size_t)t * (S.Ke + 6);
        for (int k = 0; k < S.Ka; ++k) th[L.off_id + k] = z[k];
        for (int k = 0; k < S.Ke; ++k) th[L.off_exp + k] = zf[k];
        for (int a = 0; a < 6; ++a) th[L.off_pose + a] = zf[S.Ke + a];
    }
}

// ---------------------------------------------------------------- frame preparation
// InputData::get_resized_frame (input_data.cu:44-53) and the crop + transpose of
// VideoFitter::fit_to_video_frame (video_fitter.cu:537-543), on 8-bit BGR frames.  The steps are
// OpenCV 4 calls (a dependency absent here), restated from OpenCV's published semantics in the
// FP32 / fixed-point arithmetic OpenCV uses, so this function is not FP64:
//   cvtColor(BGR2GRAY) on 8-bit data:  grey = (1868 B + 9617 G + 4899 R + 2^13) >> 14;
//   convertTo(CV_32F) and /255:        float(grey) * float(1/255.)  (convertTo's scale);
//   resize(fx = fy = coef, INTER_LINEAR): destination size round-half-even(src * coef); equal
//     sizes copy; otherwise each destination column d has f = float((d + 0.5) * (1/coef) - 0.5),
//     s = floor(f), w = f - s, with (s, w) = (0, 0) for s < 0 and (src-1, 0) for s >= src-1; each
//     row likewise, but rows s and s+1 are clamped to the image individually and keep w;
//     horizontal pass, then vertical, each a*(1-w) + b*w in FP32 with every operation rounded;
//   copyMakeBorder(0, DIMY, 0, DIMX, zero): pixels past the resized frame are 0;
//   the crop at (x0, y0), stored column-major.
int prep_resized_size(int n, float coef) { return (int)std::nearbyint((double)n * (double)coef); }

inline float prep_grey(const uint8_t* frame, int sw, int x, int y)
{
    const uint8_t* p = frame + 3 * ((size_t)y * sw + x);
    const int g = (1868 * (int)p[0] + 9617 * (int)p[1] + 4899 * (int)p[2] + 8192) >> 14;
    return (float)g * (float)(1.0 / 255.0);
}

inline float prep_lerp(float a, float b, float w)
{
    const float wa = 1.f - w;
    const float pa = a * wa;
    const float pb = b * w;
    return pa + pb;
}

void prep_crop(const uint8_t* frame, int sw, int sh, float coef, const int32_t* crop, float* out)
{
    const int x0 = crop[0], y0 = crop[1], W = crop[2], H = crop[3];
    const int dw = prep_resized_size(sw, coef), dh = prep_resized_size(sh, coef);
    const bool copy = dw == sw && dh == sh;
    const double scale = 1.0 / (double)coef;
    for (int c = 0; c < W; ++c) {
        const int X = x0 + c;
        int sx = 0;
        float wx = 0.f;
        if (!copy && X < dw) {
            const double t = ((double)X + 0.5) * scale;
            float f = (float)(t - 0.5);
            sx = (int)std::floor(f);
            wx = f - (float)sx;
            if (sx < 0) { sx = 0; wx = 0.f; }
            if (sx >= sw - 1) { sx = sw - 1; wx = 0.f; }
        }
        const int sx1 = std::min(sx + 1, sw - 1);
        for (int r = 0; r < H; ++r) {
            const int Y = y0 + r;
            float v = 0.f;
            if (X < dw && Y < dh) {
                if (copy) {
                    v = prep_grey(frame, sw, X, Y);
                } else {
                    const double t = ((double)Y + 0.5) * scale;
                    float f = (float)(t - 0.5);
                    const int sy = (int)std::floor(f);
                    const float wy = f - (float)sy;
                    const int r0 = std::min(std::max(sy, 0), sh - 1), r1 = std::min(std::max(sy + 1, 0), sh - 1);
                    const float h0 = prep_lerp(prep_grey(frame, sw, sx, r0), prep_grey(frame, sw, sx1, r0), wx);
                    const float h1 = prep_lerp(prep_grey(frame, sw, sx, r1), prep_grey(frame, sw, sx1, r1), wx);
                    v = prep_lerp(h0, h1, wy);
                }
            }
            out[(size_t)c * H + r] = v;
        }
    }
}

} // namespace

extern "C" {

// ---------------------------------------------------------------- inner-mouth masking
// paint_innermouth_black (preprocessing.cu:40-103), with the traced contour replaced by its
// geometry: the reference draws the inner-lip polygon (landmarks 43-50, 43 and 47 moved 3u
// inwards, u = face size / REF_FACE_SIZE) as 3-pixel-wide lines, takes the outer contour of that
// ring (findContours) and blackens every pixel of the mouth box (landmarks 31-50, clipped to the
// frame, (uint) bounds) whose signed distance to the contour exceeds u, when the mouth opening
// (mean of three lip distances, in REF_FACE_SIZE units) exceeds 3.  Here the contour is the
// polygon moved one pixel outwards: a pixel is blackened when it is inside the polygon by more
// than u - 1, or outside by less than 1 - u.  Distances are compared squared, in FP64 without
// contraction (the library's kernel uses the same operations).  A frame with a non-finite
// landmark among 31-50 (or a non-finite face size) is left as it is; the box is clamped to the
// frame before its (uint) conversion.
struct Mouth {
    double px[8], py[8];
    double u;
    int x0, x1, y0, y1;   // pixel box [x0, x1) x [y0, y1)
    int active;
};

Mouth mouth_setup(const float* x, const float* y, int w, int h, float ref_face_size)
{
    Mouth M{};
    float minx = x[0], maxx = x[0], miny = y[0], maxy = y[0];
    for (int i = 1; i < NL; ++i) {
        minx = std::min(minx, x[i]); maxx = std::max(maxx, x[i]);
        miny = std::min(miny, y[i]); maxy = std::max(maxy, y[i]);
    }
    const float face = std::max(maxx - minx, maxy - miny);       // compute_face_size (funcs.cu:44-53)
    const float unit = face / ref_face_size;
    for (int k = 0; k < 8; ++k) {
        const int i = 43 + k;
        float off = 0.f;
        if (i == 43) off += 3 * unit;
        if (i == 47) off -= 3 * unit;
        M.px[k] = (double)(x[i] + off);
        M.py[k] = (double)y[i];
    }
    float bx0 = x[31], bx1 = x[31], by0 = y[31], by1 = y[31];
    for (int i = 32; i < NL; ++i) {
        bx0 = std::min(bx0, x[i]); bx1 = std::max(bx1, x[i]);
        by0 = std::min(by0, y[i]); by1 = std::max(by1, y[i]);
    }
    bx0 = std::min((float)w, std::max(0.f, bx0)); bx1 = std::max(0.f, std::min((float)w, bx1));
    by0 = std::min((float)h, std::max(0.f, by0)); by1 = std::max(0.f, std::min((float)h, by1));
    auto dist = [&](int a, int b) { return std::sqrt((x[a] - x[b]) * (x[a] - x[b]) + (y[a] - y[b]) * (y[a] - y[b])); };
    const float open = (ref_face_size / face) * (dist(44, 50) + dist(45, 49) + dist(46, 48)) / 3.0f;
    bool finite = std::isfinite(face);
    for (int i = 31; i < NL; ++i) finite = finite && std::isfinite(x[i]) && std::isfinite(y[i]);
    M.u = unit;
    M.active = finite && face > 0.f && open > 3.0f;
    if (!M.active) return M;
    M.x0 = (int)(unsigned)bx0; M.x1 = (int)(unsigned)bx1;
    M.y0 = (int)(unsigned)by0; M.y1 = (int)(unsigned)by1;
    return M;
}

bool mouth_paint(const Mouth& M, int j, int i)
{
    const double X = j, Y = i;
    double best = 1e300;
    bool inside = false;
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
        if (d2 < best) best = d2;
        // even-odd crossing of the ray towards +x
        if ((ay > Y) != (by > Y)) {
            const double xc = ax + (Y - ay) / (by - ay) * ex;
            if (X < xc) inside = !inside;
        }
    }
    const double m = M.u - 1.0;
    if (inside) return m < 0.0 || best > m * m;
    return m < 0.0 && best < m * m;
}

// frames: n frames of [sh][sw][3] bytes, frame_stride bytes apart; coef [n]; crop [n][4] =
// (x0, y0, width, height); out: the crops column-major, frame i at the prefix sum of width*height.
void orc_prepare_crops(int n, const uint8_t* frames, int sw, int sh, long long frame_stride, const float* coef,
                       const int32_t* crop, float* out)
{
    std::vector<size_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (size_t)crop[4 * i + 2] * crop[4 * i + 3];
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i)
        prep_crop(frames + (size_t)i * frame_stride, sw, sh, coef[i], crop + 4 * i, out + off[i]);
}

int orc_resized_size(int n, float coef) { return prep_resized_size(n, coef); }

// n frames of [h][w][3] bytes (B, G, R) frame_stride bytes apart, painted in place; lmk_x, lmk_y
// [n][51] in frame pixels.
void orc_paint_innermouth_black(int n, uint8_t* frames, int w, int h, long long frame_stride, const float* lmk_x,
                                const float* lmk_y, float ref_face_size)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < n; ++f) {
        const Mouth M = mouth_setup(lmk_x + (size_t)f * NL, lmk_y + (size_t)f * NL, w, h, ref_face_size);
        if (!M.active) continue;
        uint8_t* fr = frames + (size_t)f * frame_stride;
        for (int i = M.y0; i < M.y1; ++i)
            for (int j = M.x0; j < M.x1; ++j)
                if (mouth_paint(M, j, i)) {
                    uint8_t* p = fr + 3 * ((size_t)i * w + j);
                    p[0] = p[1] = p[2] = 0;
                }
    }
}

void orc_ortho_residuals(const double* x, const double* xlb, const double* ylb, const double* p0L, double* e,
                         double* J)
{
    ortho_residuals(x, xlb, ylb, p0L, e, J);
}

void orc_rotvec_from_euler(const double* e, double* r) { rotvec_from_euler(e, r); }

void orc_euler_angles(const double* R, double* ypr)
{
    euler_angles(R, ypr[0], ypr[1], ypr[2]);
}

// lmk [n][2][51] (x row, y row), cams [n][4] (phix, phiy, cx, cy), p0L [51][3];
// outputs x_orth [n][7], pose [n][6], angles [n][3] (degrees), info [n][4].
void orc_ortho_init(int n, const float* lmk, const float* cams, const double* p0L, double* x_orth, double* pose,
                    double* angles, int32_t* info)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < n; ++f)
        ortho_init_one(lmk + (size_t)f * 2 * NL, lmk + (size_t)f * 2 * NL + NL, cams + 4 * (size_t)f, p0L,
                       x_orth + 7 * (size_t)f, pose + 6 * (size_t)f, angles + 3 * (size_t)f, info + 4 * (size_t)f);
}

int orc_num_unknowns(const tdi_model_desc* m, const tdi_fit_options* o) { return make_layout(m, o).K; }

// The render loss of one frame (render_loss above): theta [K], dverts [V][3] (may be null),
// tri [W*H] (may be null; see render_loss), frozen.
double orc_render_loss(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o, const double* th,
                       const double* dverts, int32_t* tri, int frozen)
{
    Frame F;
    setup_frame(F, m, f, o);
    return render_loss(F, th, dverts, tri, frozen);
}

// Landmark-only fit with slack for n frames: theta [n][K] double in/out (the pose is the start; on
// success the pose and the first Kl expression coefficients are overwritten), reports [n].
int orc_fit_landmarks_slack(const tdi_model_desc* m, int n, const tdi_frame_desc* frames, const tdi_fit_options* o,
                            int Kl, double* theta, tdi_lmk_report* reports)
{
    const Layout L = make_layout(m, o);
    if (Kl < 1 || Kl > L.nexp) return TDI_ERR_ARG;
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < n; ++f) {
        double* th = theta + (size_t)f * L.K;
        LmkFit S;
        lmk_fit_setup(S, m, &frames[f], L, th, Kl);
        std::vector<double> z(S.n);
        for (int a = 0; a < 6; ++a) z[Kl + a] = th[L.off_pose + a];
        reports[f] = lmk_fit_run(S, z.data());
        if (reports[f].status == TDI_OK) {
            for (int k = 0; k < Kl; ++k) th[L.off_exp + k] = z[k];
            for (int a = 0; a < 6; ++a) th[L.off_pose + a] = z[Kl + a];
        }
    }
    return TDI_OK;
}

// The terms of that fit at z [Kl + 7] for one frame (theta supplies fixed identity terms):
// g [Kl + 7], H [(Kl + 7)^2] (may be null together); returns the objective.
double orc_landmark_slack_terms(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o, int Kl,
                                const double* theta, const double* z, double tau_b, double* g, double* H)
{
    const Layout L = make_layout(m, o);
    LmkFit S;
    lmk_fit_setup(S, m, f, L, theta, Kl);
    return lmk_fit_terms(S, z, tau_b, g, H, nullptr);
}

// The joint landmark-only fit over the n frames of one group: theta [n][K] double in/out (frame
// 0's identity is the start of the shared block; every frame's pose and expressions its own
// start; with the slack, on success, and always without it, the first Ka identity coefficients
// of every frame, the first Ke expression coefficients and the pose are overwritten).
int orc_fit_landmarks_joint(const tdi_model_desc* m, int n, const tdi_frame_desc* frames, const tdi_fit_options* o,
                            int Ka, int Ke, int slack, double* theta, tdi_lmk_report* report)
{
    const Layout L = make_layout(m, o);
    if (n < 1 || Ka < 0 || Ka > L.nid || Ke < 0 || Ke > L.nexp) return TDI_ERR_ARG;
    LmkJoint S;
    lmk_joint_setup(S, m, frames, n, L, theta, Ka, Ke, slack);
    std::vector<double> z(S.n);
    lmk_joint_pack(S, L, theta, z.data());
    *report = lmk_joint_run(S, z.data());
    if (report->status == TDI_OK) lmk_joint_unpack(S, L, z.data(), theta);
    return TDI_OK;
}

// The terms of that fit at z [joint order] (theta: frame 0's identity past Ka is fixed);
// g [N], H [N^2] may be null together; returns the objective.  N = Ka + slack + n (Ke + 6).
double orc_landmark_joint_terms(const tdi_model_desc* m, int n, const tdi_frame_desc* frames,
                                const tdi_fit_options* o, int Ka, int Ke, int slack, const double* theta,
                                const double* z, double tau_b, double* g, double* H)
{
    const Layout L = make_layout(m, o);
    LmkJoint S;
    lmk_joint_setup(S, m, frames, n, L, theta, Ka, Ke, slack);
    return lmk_joint_terms(S, z, tau_b, g, H, nullptr);
}

int orc_blur_taps(double sigma, double* taps) { return blur_taps(sigma, taps); }

void orc_rodrigues(const double* u, double* R, double* dR) { rodrigues(u, R, dR); }

void orc_blur(const double* in, double* out, int W, int H, const double* taps) { blur(in, out, W, H, taps); }

void orc_rasterize(int NT, const int32_t* tris, const float* xp, const float* yp, const float* z,
                   int x0, int y0, int W, int H, int ntmp, uint64_t* keys, float* bary)
{
    rasterize(NT, tris, xp, yp, z, x0, y0, W, H, ntmp, keys, bary);
}

// Geometry at theta: camera-space vertices [V][3] and pixel coordinates [V][2] (double).
void orc_geometry(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o,
                  const double* th, double* xyz, double* xyp, double* R, double* dR)
{
    Frame F;
    F.m = m; F.f = f; F.o = o; F.L = make_layout(m, o);
    Geometry G;
    geometry(F, th, true, G);
    for (int v = 0; v < m->n_vertices; ++v) {
        for (int a = 0; a < 3; ++a) xyz[3 * v + a] = G.P[3 * v + a];
        xyp[2 * v] = G.xp[v];
        xyp[2 * v + 1] = G.yp[v];
    }
    if (R) std::memcpy(R, G.R, sizeof(G.R));
    if (dR) std::memcpy(dR, G.dR, sizeof(G.dR));
}

// Render at theta: keys [W*H], raw and blurred images [W*H], objective pieces [3] = (data, barrier, total).
// Returns the rendered pixel count.
int orc_render(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o,
               const double* th, uint64_t* keys, float* bary, double* raw, double* img, double* obj3)
{
    Frame F;
    setup_frame(F, m, f, o);
    Eval E;
    evaluate(F, th, false, E);
    const size_t A = (size_t)f->width * f->height;
    if (keys) std::copy(E.keys.begin(), E.keys.end(), keys);
    if (bary) std::copy(E.bary.begin(), E.bary.end(), bary);
    if (raw) std::copy(E.raw.begin(), E.raw.end(), raw);
    if (img) std::copy(E.img.begin(), E.img.end(), img);
    (void)A;
    if (obj3) { obj3[0] = E.obj_data; obj3[1] = E.obj_barrier; obj3[2] = E.obj; }
    return (int)E.vis.size();
}

// Normal equations at theta: H [K*K], g [K], obj3 [3].  Optional dI [N*K] and neighbour
// positions [N*4] (kl, kr, ka, kb) when the caller passes buffers of sufficient size
// (capacity given by cap_pixels).  Returns the rendered pixel count.
int orc_normal_equations(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o,
                         const double* th, double* Hm, double* g, double* obj3,
                         double* dI, int32_t* nbr, int cap_pixels)
{
    Frame F;
    setup_frame(F, m, f, o);
    Eval E;
    const int K = F.L.K;
    // count first so that optional outputs are only written when they fit
    std::vector<double> dIv;
    std::vector<int> nbv;
    evaluate(F, th, true, E, nullptr, nullptr);
    const int N = (int)E.vis.size();
    if ((dI || nbr) && N <= cap_pixels) {
        dIv.resize((size_t)N * K);
        nbv.resize((size_t)4 * N);
        evaluate(F, th, true, E, dIv.data(), nbv.data());
        if (dI) std::copy(dIv.begin(), dIv.end(), dI);
        if (nbr) for (size_t i = 0; i < nbv.size(); ++i) nbr[i] = nbv[i];
    }
    std::copy(E.H.begin(), E.H.end(), Hm);
    std::copy(E.g.begin(), E.g.end(), g);
    if (obj3) { obj3[0] = E.obj_data; obj3[1] = E.obj_barrier; obj3[2] = E.obj; }
    return N;
}

// Barrier alone: objective, gradient [K], Hessian [K*K] (any may be null except obj).
double orc_barrier(const tdi_model_desc* m, const tdi_frame_desc* f, const tdi_fit_options* o,
                   const double* th, double* g, double* Hm)
{
    Frame F;
    F.m = m; F.f = f; F.o = o; F.L = make_layout(m, o);
    Geometry G;
    geometry(F, th, true, G);
    const int K = F.L.K;
    if (g) std::fill(g, g + K, 0.0);
    if (Hm) std::fill(Hm, Hm + (size_t)K * K, 0.0);
    return barrier(F, th, G, g, Hm);
}

int orc_cholesky_solve(int K, const double* Hm, const double* g, double* delta)
{
    std::vector<double> H(Hm, Hm + (size_t)K * K);
    return cholesky_solve(K, H, g, delta) ? TDI_OK : TDI_NOT_PD;
}

// Fit n frames independently (OpenMP over frames, dynamic schedule for uneven crops).
// theta: [n][K] double in/out.  Returns TDI_OK.
int orc_fit(const tdi_model_desc* m, int n, const tdi_frame_desc* frames, const tdi_fit_options* o,
            double* theta, tdi_fit_report* reports, int n_threads, int finetune)
{
    const int K = make_layout(m, o).K;
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        Frame F;
        setup_frame(F, m, &frames[i], o);
        F.finetune = finetune != 0;
        fit_one(F, theta + (size_t)i * K, &reports[i]);
    }
    return TDI_OK;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// Light sub-fits for n frames (light not among the unknowns of o): theta [n][K] fixed,
// lam [n][3] and Lint [n] in/out.  which = 0: fit_3DMM_lambdas, 1: fit_3DMM_Lintensity.
int orc_fit_light(const tdi_model_desc* m, int n, const tdi_frame_desc* frames, const tdi_fit_options* o,
                  const double* theta, double* lam, double* Lint, tdi_fit_report* reports, int which, int n_threads)
{
    if (o->fit_light) return TDI_ERR_ARG;
    const int K = make_layout(m, o).K;
    const LightFit lf = which ? LightFit{true, 0.3, 0.5, 0.1, 0.0025} : LightFit{false, 0.5, 0.5, 0.01, 0.0003};
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i)
        fit_light_one(m, &frames[i], o, theta + (size_t)i * K, lam + 3 * i, Lint + i, &reports[i], lf);
    return TDI_OK;
}

} // extern "C"
