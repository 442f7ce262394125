// tdi_host.cpp — C++ host layer above the C ABI of libtdi.so.
//
// Restates the reference's host-side code around the per-frame fitting step:
//   * VideoOutput (video_fitter.cu:2603-2828): per-frame results of several fits, averaged
//     (expressions over the NRES fits, poses over all fits; the illumination keeps the first
//     fit) and written as text by write_2d_vector (funcs.h:112-124): every value with
//     setprecision(7) followed by one space, one row per frame.  save_expressions,
//     save_poses and save_illums write the frames t < max(frame id), so the last frame is not
//     written (video_fitter.cu:2754,2787,2821); save_expressions_all writes t <= max.
//   * compute_euler_angles (rotation_header.cu:421-449): yaw, pitch, roll of R in FP32.
//   * fit_to_video_frame (video_fitter.cu:554-615): L = 0.005 and a lambda fit per frame,
//     the shape fit, and with TWOSTAGE_ILLUM two more lambda + shape rounds.
//   * the config-file keys of this path (config.cu:300-560).
// Only the C ABI of include/tdi_api.h is called; the batched shape fit takes a device copy of
// the unknowns, made with the HIP runtime.
#include <random>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/tdi_api.h"
#include "../../include/tdi_host.h"
#include "../../include/tdi_model_fitter.h"

namespace tdi_host {

// write_2d_vector (funcs.h:112-124): the row width is that of the first row.
static void write_2d_vector(const std::string& path, const std::vector<std::vector<float>>& rows)
{
    std::ofstream out(path.c_str());
    for (size_t i = 0; i < rows.size(); ++i) {
        for (size_t j = 0; j < rows[0].size(); ++j)
            out << std::setprecision(std::numeric_limits<float>::digits10 + 1) << rows[i][j] << " ";
        out << std::endl;
    }
}

// R (column-major, R[col*3+row]) of the axis-angle vector u, evaluated in FP64 and rounded
// to FP32 like the rotation the GPU path uses (R = I + sin(t)[n]x + (1-cos(t))[n]x^2).
static void rotation_f32(const float* uf, float* R)
{
    const double u[3] = {uf[0], uf[1], uf[2]};
    const double th = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double n[3] = {u[0] / th, u[1] / th, u[2] / th};
    const double N[9] = {0, n[2], -n[1], -n[2], 0, n[0], n[1], -n[0], 0};
    const double s = std::sin(th), c1 = 1.0 - std::cos(th);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double nn = 0;
            for (int k = 0; k < 3; ++k) nn += N[k * 3 + r] * N[c * 3 + k];
            R[c * 3 + r] = (float)(c1 * nn + s * N[c * 3 + r] + (r == c ? 1.0 : 0.0));
        }
}

void compute_euler_angles(const float* R, float& yaw, float& pitch, float& roll)
{
    const float cosine_for_pitch = std::sqrt(R[5] * R[5] + R[8] * R[8]);
    if (!(cosine_for_pitch < 1e-6f)) {
        roll = std::atan2(R[1], R[0]);
        yaw = std::atan2(-R[2], cosine_for_pitch);
        pitch = std::atan2(R[5], R[8]);
    } else {
        roll = std::atan2(-R[7], R[4]);
        yaw = std::atan2(-R[2], cosine_for_pitch);
        pitch = 0.0f;
    }
}

class VideoOutput {
public:
    VideoOutput(int k_exp, int nres) : Kepsilon(k_exp), nres_coefs(nres) {}

    void add_exp_coeffs(size_t frame_id, const std::vector<float>& e)
    {
        exp_coefs[frame_id].push_back(e);
        abs_frame_ids.insert(frame_id);
    }
    void add_pose(size_t frame_id, const std::vector<float>& p) { poses[frame_id].push_back(p); }
    void add_illum(size_t frame_id, const std::vector<float>& il)
    {
        if (illum_coeffs.find(frame_id) == illum_coeffs.end()) illum_coeffs.insert(std::make_pair(frame_id, il));
    }

    // Every fit added, in order, so that a long run can be saved and replayed exactly.
    struct Added {
        size_t frame_id;
        std::vector<float> exp, pose, illum;   // pose, illum empty when not given
    };
    std::vector<Added> log;
    size_t dumped = 0;   // entries of `log` already written by dump()

    void add(size_t frame_id, const std::vector<float>& e, const std::vector<float>& p, const std::vector<float>& il)
    {
        add_exp_coeffs(frame_id, e);
        if (!p.empty()) add_pose(frame_id, p);
        if (!il.empty()) add_illum(frame_id, il);
        log.push_back(Added{frame_id, e, p, il});
    }

    // The entries added since the last dump, appended to `path` as text with every float in
    // hexadecimal (exact): "frame_id n_exp n_pose n_illum values...".
    bool dump(const std::string& path)
    {
        std::ofstream out(path.c_str(), std::ios::app);
        if (!out) return false;
        out << std::hexfloat;
        for (; dumped < log.size(); ++dumped) {
            const Added& a = log[dumped];
            out << a.frame_id << ' ' << a.exp.size() << ' ' << a.pose.size() << ' ' << a.illum.size();
            for (float v : a.exp) out << ' ' << v;
            for (float v : a.pose) out << ' ' << v;
            for (float v : a.illum) out << ' ' << v;
            out << '\n';
        }
        return (bool)out;
    }

    // Replays a dump, entries of frames below `frame_limit` only (the entries count as written).
    bool load(const std::string& path, size_t frame_limit = (size_t)-1)
    {
        std::ifstream in(path.c_str());
        if (!in) return false;
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty()) continue;
            std::istringstream ls(line);
            size_t fid = 0, ne = 0, np = 0, ni = 0;
            if (!(ls >> fid >> ne >> np >> ni) || ne != (size_t)Kepsilon || (np != 0 && np != 6) || (ni != 0 && ni != 4))
                return false;
            std::vector<float> e(ne), p(np), il(ni);
            for (auto* v : {&e, &p, &il})
                for (float& x : *v) {
                    std::string tok;
                    if (!(ls >> tok)) return false;
                    x = std::strtof(tok.c_str(), nullptr);
                }
            if (fid < frame_limit) add(fid, e, p, il);
        }
        dumped = log.size();
        return true;
    }

    // Rewrites `path` with every entry of the log.
    bool rewrite(const std::string& path)
    {
        { std::ofstream trunc(path.c_str(), std::ios::trunc); if (!trunc) return false; }
        dumped = 0;
        return dump(path);
    }

    std::vector<float> compute_avg_exp_coeffs(size_t frame_id, bool ignore_first)
    {
        auto it = exp_coefs.find(frame_id);
        if (it == exp_coefs.end()) return std::vector<float>(Kepsilon, NAN);
        std::vector<float> avgs(Kepsilon, 0);
        const int K = (int)it->second.size();
        const int k0 = ignore_first ? 1 : 0;
        for (int k = k0; k < K; ++k)
            for (int ei = 0; ei < Kepsilon; ++ei) avgs[ei] += it->second[k][ei] / ((float)(K - k0));
        return avgs;
    }
    std::vector<float> combine_exp_coeffs(size_t frame_id)
    {
        auto it = exp_coefs.find(frame_id);
        if (it == exp_coefs.end()) return std::vector<float>((size_t)Kepsilon * nres_coefs, NAN);
        std::vector<float> exps;
        for (int k = 0; k < nres_coefs; ++k)
            for (int ei = 0; ei < Kepsilon; ++ei)
                exps.push_back(k >= (int)it->second.size() ? NAN : it->second[k][ei]);
        return exps;
    }
    std::vector<float> compute_avg_pose(size_t frame_id)
    {
        auto it = poses.find(frame_id);
        if (it == poses.end()) return std::vector<float>(6, NAN);
        std::vector<float> avgs(6, 0);
        const int K = (int)it->second.size();
        for (int k = 0; k < K; ++k)
            for (int ei = 0; ei < 6; ++ei) avgs[ei] += it->second[k][ei] / ((float)K);
        return avgs;
    }
    std::vector<float> get_illum(size_t frame_id)
    {
        auto it = illum_coeffs.find(frame_id);
        return it == illum_coeffs.end() ? std::vector<float>(4, NAN) : it->second;
    }

    int max_frame_id() const { return abs_frame_ids.empty() ? -1 : (int)*abs_frame_ids.rbegin(); }

    void save_expressions(const std::string& path, bool ignore_first)
    {
        std::vector<std::vector<float>> rows;
        for (int t = 0; t < max_frame_id(); ++t) rows.push_back(compute_avg_exp_coeffs(t, ignore_first));
        write_2d_vector(path, rows);
    }
    void save_expressions_all(const std::string& path)
    {
        std::vector<std::vector<float>> rows;
        for (int t = 0; t <= max_frame_id(); ++t) rows.push_back(combine_exp_coeffs(t));
        write_2d_vector(path, rows);
    }
    void save_poses(const std::string& path, bool with_euler)
    {
        std::vector<std::vector<float>> rows;
        for (int t = 0; t < max_frame_id(); ++t) {
            std::vector<float> p = compute_avg_pose(t);
            if (with_euler) {
                float R[9], yaw, pitch, roll;
                rotation_f32(&p[3], R);
                compute_euler_angles(R, yaw, pitch, roll);
                p.push_back(yaw);
                p.push_back(pitch);
                p.push_back(roll);
            }
            rows.push_back(p);
        }
        write_2d_vector(path, rows);
    }
    void save_illums(const std::string& path)
    {
        std::vector<std::vector<float>> rows;
        for (int t = 0; t < max_frame_id(); ++t) rows.push_back(get_illum(t));
        write_2d_vector(path, rows);
    }

private:
    int Kepsilon, nres_coefs;
    std::map<size_t, std::vector<std::vector<float>>> exp_coefs, poses;
    std::map<size_t, std::vector<float>> illum_coeffs;
    std::set<size_t> abs_frame_ids;
};

// ---------------------------------------------------------------------------- fits
// fit_3DMM_shape_rigid (model_fitter.cu:121-375) or fit_3DMM_epsilon_finetune (397-661) for every
// frame of the batch; the batched path takes a device copy of the unknowns.
static int shape_fit(tdi_batch* b, int n, int K, bool batched, bool finetune, float* theta, tdi_fit_report* reports)
{
    if (!batched) return finetune ? tdi_fit_plain_finetune(b, theta, reports) : tdi_fit_plain(b, theta, reports);
    float* d = nullptr;
    const size_t bytes = sizeof(float) * (size_t)n * K;
    if (hipMalloc((void**)&d, bytes) != hipSuccess) return TDI_ERR_NOMEM;
    int rc = TDI_OK;
    if (hipMemcpy(d, theta, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc) rc = finetune ? tdi_fit_batched_finetune(b, d, reports, nullptr) : tdi_fit_batched(b, d, reports, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc && hipMemcpy(theta, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = TDI_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

// fit_3DMM_lambdas / fit_3DMM_Lintensity (model_fitter.cu:981-1180) followed by the light update
// the next shape fit reads.
static int light_fit(tdi_batch* b, int n, const float* theta, float* lam, float* Lint, bool intensity,
                     tdi_fit_report* reports)
{
    std::vector<tdi_fit_report> r(n);
    tdi_fit_report* rp = reports ? reports : r.data();
    int rc = intensity ? tdi_fit_lintensity(b, theta, lam, Lint, rp) : tdi_fit_lambdas(b, theta, lam, Lint, rp);
    if (rc) return rc;
    return tdi_batch_set_light(b, lam, Lint);
}

static int fit_3DMM_lambdas(tdi_batch* b, int n, const float* theta, float* lam, float* Lint)
{
    return light_fit(b, n, theta, lam, Lint, false, nullptr);
}

static bool all_success(const tdi_fit_report* r, int n)
{
    for (int i = 0; i < n; ++i)
        if (!r[i].success) return false;
    return true;
}

// The landmark-only fit with slack of every frame of b (tdi_fit_landmarks_slack) on host theta
// [n][K], in/out.
static int landmark_fit(tdi_batch* b, int n, int K, int k_exp, float* theta, tdi_lmk_report* reports)
{
    float* d = nullptr;
    const size_t bytes = sizeof(float) * (size_t)n * K;
    if (hipMalloc((void**)&d, bytes) != hipSuccess) return TDI_ERR_NOMEM;
    int rc = hipMemcpy(d, theta, bytes, hipMemcpyHostToDevice) == hipSuccess ? TDI_OK : TDI_ERR_HIP;
    if (!rc) rc = tdi_fit_landmarks_slack(b, d, k_exp, reports, nullptr);
    if (!rc && hipMemcpy(theta, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = TDI_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

// Logbarrier_Initializer::initialize_with_orthographic_t (logbarrier_initializer.cu:1862-1914),
// orthographic fit and pose seed, for n frames at once: host arrays in and out.
static int initialize_with_orthographic_t(tdi_ctx* ctx, int n, const float* lmk_x, const float* lmk_y,
                                          const float* cams, float* pose, float* angles, int32_t* info)
{
    const int NL = TDI_NLANDMARKS;
    std::vector<float> lmk((size_t)n * 2 * NL);
    for (int f = 0; f < n; ++f)
        for (int i = 0; i < NL; ++i) {
            lmk[(size_t)f * 2 * NL + i] = lmk_x[(size_t)f * NL + i];
            lmk[(size_t)f * 2 * NL + NL + i] = lmk_y[(size_t)f * NL + i];
        }
    // one device block: landmarks, cameras, x_orth, pose, angles, info
    const size_t nl = lmk.size(), nc = 4 * (size_t)n, nx = 7 * (size_t)n, np = 6 * (size_t)n, na = 3 * (size_t)n;
    const size_t bytes = sizeof(float) * (nl + nc + nx + np + na) + sizeof(int32_t) * 4 * (size_t)n;
    char* d = nullptr;
    if (hipMalloc((void**)&d, bytes) != hipSuccess) return TDI_ERR_NOMEM;
    float* dl = (float*)d;
    float* dc = dl + nl;
    float* dx = dc + nc;
    float* dp = dx + nx;
    float* da = dp + np;
    int32_t* di = (int32_t*)(da + na);
    int rc = TDI_OK;
    if (hipMemcpy(dl, lmk.data(), sizeof(float) * nl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dc, cams, sizeof(float) * nc, hipMemcpyHostToDevice) != hipSuccess)
        rc = TDI_ERR_HIP;
    if (!rc) rc = tdi_ortho_init(ctx, n, dl, dc, dx, dp, da, di, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc && hipMemcpy(pose, dp, sizeof(float) * np, hipMemcpyDeviceToHost) != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc && angles && hipMemcpy(angles, da, sizeof(float) * na, hipMemcpyDeviceToHost) != hipSuccess) rc = TDI_ERR_HIP;
    if (!rc && info && hipMemcpy(info, di, sizeof(int32_t) * 4 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = TDI_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

// ---------------------------------------------------------------------------- landmark files
// read2DVectorFromFile_unknown_size (funcs.h:352-413): the first line's count of numbers (with
// "nan" counted as a number) fixes the columns; the file's line count, at most 32767, the rows;
// the values are then read in order, one that does not parse reading as NaN and skipping three
// characters (the text "nan").  LandmarkData::fill_xpypvec (input_data.cu:105-121) takes
// x_i = row[2i], y_i = row[2i+1].
static int count_numbers(std::string line)
{
    for (size_t at = line.find("nan"); at != std::string::npos; at = line.find("nan", at + 1)) line.replace(at, 3, "0");
    std::istringstream in(line);
    int cnt = 0;
    float v;
    while (in >> v) ++cnt;
    return cnt;
}

static int read_landmark_rows(const char* path, std::vector<std::vector<float>>& rows)
{
    std::ifstream shape(path);
    if (!shape.is_open()) return TDI_ERR_ARG;
    std::string line;
    std::getline(shape, line);
    const int cols = count_numbers(line);
    int nrows = 1;
    while (nrows < 32767 && std::getline(shape, line)) ++nrows;
    shape.close();
    if (cols < 2 * TDI_NLANDMARKS) return TDI_ERR_ARG;
    std::ifstream in(path);
    rows.assign(nrows, std::vector<float>());
    for (int r = 0; r < nrows; ++r)
        for (int c = 0; c < cols; ++c) {
            float v;
            if (!(in >> v)) {
                rows[r].push_back(std::numeric_limits<float>::quiet_NaN());
                in.clear();
                in.ignore(3);
                continue;
            }
            rows[r].push_back(v);
        }
    return TDI_OK;
}

// ---------------------------------------------------------------------------- frame geometry
// LandmarkData::get_face_size (input_data.cu:124-138): extents of the landmarks truncated to int.
static float landmark_face_size(const float* x, const float* y)
{
    const int NL = TDI_NLANDMARKS;
    const int x0 = (int)*std::min_element(x, x + NL), x1 = (int)*std::max_element(x, x + NL);
    const int y0 = (int)*std::min_element(y, y + NL), y1 = (int)*std::max_element(y, y + NL);
    return (float)std::max(x1 - x0, y1 - y0);
}

// compute_face_size (funcs.cu:44-53): the larger extent of the landmarks, in float.
static float compute_face_size(const float* x, const float* y)
{
    const int NL = TDI_NLANDMARKS;
    const float w = *std::max_element(x, x + NL) - *std::min_element(x, x + NL);
    const float h = *std::max_element(y, y + NL) - *std::min_element(y, y + NL);
    return std::max(w, h);
}

// Renderer::set_x0_short_y0_short (renderer.cu:585-606): the crop origin is the landmarks' top
// left corner less a margin of 1/1.5 of the face's larger extent (pad) or 20 pixels, clamped at 0
// and truncated to an unsigned short.
static void set_x0_short_y0_short(const float* x, const float* y, bool pad, int32_t* x0, int32_t* y0)
{
    const int NL = TDI_NLANDMARKS;
    const float xmin = *std::min_element(x, x + NL), xmax = *std::max_element(x, x + NL);
    const float ymin = *std::min_element(y, y + NL), ymax = *std::max_element(y, y + NL);
    if (pad) {
        const float sz = std::max(xmax - xmin, ymax - ymin);
        *x0 = (uint16_t)std::max<float>(xmin - sz / 1.5, 0);
        *y0 = (uint16_t)std::max<float>(ymin - sz / 1.5, 0);
    } else {
        *x0 = (uint16_t)std::max<float>(xmin - 20, 0);
        *y0 = (uint16_t)std::max<float>(ymin - 20, 0);
    }
}

// One frame of the video loop before its fit: the in-frame test (video_fitter.cu:180-186), the
// resize coefficient ref_face_size / face size (video_fitter.cu:357-360) applied to the camera
// (Camera::update_camera, camera.h:162-174) and to the landmarks (get_resized_landmarks,
// input_data.cu:33-41), the crop origin of the resized landmarks and their face size
// (video_fitter.cu:466).
static int frame_geometry(const float* x, const float* y, const float* cam, int fw, int fh, float ref_face_size,
                          bool pad, int cw, int ch, float* coef, float* xo, float* yo, float* cam_out, int32_t* crop,
                          float* face_size)
{
    const int NL = TDI_NLANDMARKS;
    const float minx = *std::min_element(x, x + NL), maxx = *std::max_element(x, x + NL);
    const float miny = *std::min_element(y, y + NL), maxy = *std::max_element(y, y + NL);
    if (!(minx > 0.0f && miny > 0.0f && maxx < fw && maxy < fh)) return TDI_OUT_OF_FRAME;
    const float k = ref_face_size / landmark_face_size(x, y);
    if (!(k > 0.f) || !std::isfinite(k)) return TDI_OUT_OF_FRAME;
    *coef = k;
    for (int i = 0; i < 4; ++i) cam_out[i] = cam[i] * k;
    for (int i = 0; i < NL; ++i) { xo[i] = k * x[i]; yo[i] = k * y[i]; }
    set_x0_short_y0_short(xo, yo, pad, crop, crop + 1);
    crop[2] = cw;
    crop[3] = ch;
    *face_size = compute_face_size(xo, yo);
    return TDI_OK;
}

} // namespace tdi_host

// ---------------------------------------------------------------------------- C entry points
struct tdi_video_output {
    tdi_host::VideoOutput vo;
    int k_exp;
};

extern "C" {

tdi_video_output* tdi_vo_create(int32_t k_exp, int32_t nres_coefs)
{
    if (k_exp <= 0 || nres_coefs <= 0) return nullptr;
    return new tdi_video_output{tdi_host::VideoOutput(k_exp, nres_coefs), k_exp};
}

void tdi_vo_destroy(tdi_video_output* vo) { delete vo; }

int tdi_vo_add(tdi_video_output* vo, int64_t frame_id, const float* exp, const float* pose, const float* illum)
{
    if (!vo || !exp || frame_id < 0) return TDI_ERR_ARG;
    vo->vo.add((size_t)frame_id, std::vector<float>(exp, exp + vo->k_exp),
               pose ? std::vector<float>(pose, pose + 6) : std::vector<float>(),
               illum ? std::vector<float>(illum, illum + 4) : std::vector<float>());
    return TDI_OK;
}

int tdi_vo_dump(tdi_video_output* vo, const char* path)
{
    if (!vo || !path) return TDI_ERR_ARG;
    return vo->vo.dump(path) ? TDI_OK : TDI_ERR_ARG;
}

int tdi_vo_load(tdi_video_output* vo, const char* path)
{
    if (!vo || !path) return TDI_ERR_ARG;
    return vo->vo.load(path) ? TDI_OK : TDI_ERR_ARG;
}

int tdi_vo_save(tdi_video_output* vo, int32_t kind, const char* path, int32_t flag)
{
    if (!vo || !path) return TDI_ERR_ARG;
    if (vo->vo.max_frame_id() < 0) return TDI_ERR_ARG;
    switch (kind) {
    case 0: vo->vo.save_expressions(path, flag != 0); break;
    case 1: vo->vo.save_expressions_all(path); break;
    case 2: vo->vo.save_poses(path, flag != 0); break;
    case 3: vo->vo.save_illums(path); break;
    default: return TDI_ERR_ARG;
    }
    return TDI_OK;
}

void tdi_euler_angles(const float* u, float* ypr)
{
    float R[9];
    tdi_host::rotation_f32(u, R);
    tdi_host::compute_euler_angles(R, ypr[0], ypr[1], ypr[2]);
}

// OpenCV FileStorage YAML subset: "KEY: value" lines, '#' comments, a "%YAML" header and "---".
int tdi_read_config(const char* path, tdi_fit_options* opts, int32_t* twostage_illum, int32_t* nres_coefs)
{
    std::ifstream in(path);
    if (!in || !opts) return TDI_ERR_ARG;
    int two = 0, nres = 1;
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        const size_t colon = line.find(':');
        if (colon == std::string::npos || line.empty() || line[0] == '%') continue;
        std::string key = line.substr(0, colon), val = line.substr(colon + 1);
        key.erase(0, key.find_first_not_of(" \t"));
        key.erase(key.find_last_not_of(" \t\r") + 1);
        val.erase(0, val.find_first_not_of(" \t\""));
        val.erase(val.find_last_not_of(" \t\r\"") + 1);
        if (key.empty() || val.empty()) continue;
        const double v = std::atof(val.c_str());
        if (key == "KERNEL_SIGMA" && v > 0) opts->blur_sigma = (float)v;       // config.cu:553-556
        else if (key == "REF_FACE_SIZE") opts->ref_face_size = (float)v;
        else if (key == "NITERS_SUCCESS") opts->niters_success = (int)v;
        else if (key == "TWOSTAGE_ILLUM") two = (int)v;
        else if (key == "NRES_COEFS") nres = (int)v;
    }
    if (twostage_illum) *twostage_illum = two;
    if (nres_coefs) *nres_coefs = nres;
    return TDI_OK;
}

int tdi_initialize_with_orthographic(tdi_ctx* ctx, int32_t n, const float* lmk_x, const float* lmk_y,
                                     const float* cams, float* pose, float* angles, int32_t* info)
{
    if (!ctx || n <= 0 || !lmk_x || !lmk_y || !cams || !pose) return TDI_ERR_ARG;
    return tdi_host::initialize_with_orthographic_t(ctx, n, lmk_x, lmk_y, cams, pose, angles, info);
}

int tdi_fit_video_frames(tdi_batch* b, int32_t twostage_illum, int32_t batched, float* theta, float* lam,
                         float* Lint, tdi_fit_report* reports)
{
    if (!b || !theta || !lam || !Lint || !reports) return TDI_ERR_ARG;
    const int n = tdi_batch_size(b), K = tdi_batch_unknowns(b);
    // video_fitter.cu:558-565: L = 0.005, then the lambda fit of every frame
    for (int i = 0; i < n; ++i) Lint[i] = 0.005f;
    int rc = tdi_batch_set_light(b, lam, Lint);
    if (!rc) rc = tdi_host::fit_3DMM_lambdas(b, n, theta, lam, Lint);
    if (!rc) rc = tdi_host::shape_fit(b, n, K, batched != 0, false, theta, reports);
    // TWOSTAGE_ILLUM (video_fitter.cu:577-615): lambda + shape, twice more
    for (int round = 0; !rc && twostage_illum && round < 2; ++round) {
        rc = tdi_host::fit_3DMM_lambdas(b, n, theta, lam, Lint);
        if (!rc) rc = tdi_host::shape_fit(b, n, K, batched != 0, false, theta, reports);
    }
    return rc;
}

int tdi_read_landmarks(const char* path, int32_t max_frames, float* lmk_x, float* lmk_y, int32_t* n_frames)
{
    if (!path || !n_frames || (lmk_x && (!lmk_y || max_frames < 0))) return TDI_ERR_ARG;
    std::vector<std::vector<float>> rows;
    const int rc = tdi_host::read_landmark_rows(path, rows);
    if (rc) return rc;
    const int NL = TDI_NLANDMARKS;
    const int n = lmk_x ? std::min<int>(max_frames, (int)rows.size()) : (int)rows.size();
    for (int t = 0; lmk_x && t < n; ++t)
        for (int i = 0; i < NL; ++i) {
            lmk_x[(size_t)t * NL + i] = rows[t][2 * i];
            lmk_y[(size_t)t * NL + i] = rows[t][2 * i + 1];
        }
    *n_frames = n;
    return TDI_OK;
}

int tdi_frame_geometry(int32_t n, const float* lmk_x, const float* lmk_y, const float* cams, int32_t frame_w,
                       int32_t frame_h, float ref_face_size, int32_t pad, int32_t crop_w, int32_t crop_h, float* coef,
                       float* lmk_x_out, float* lmk_y_out, float* cams_out, int32_t* crop, float* face_size,
                       int32_t* status)
{
    if (n < 0 || (n > 0 && (!lmk_x || !lmk_y || !cams || !coef || !lmk_x_out || !lmk_y_out || !cams_out || !crop ||
                            !face_size || !status)) ||
        frame_w < 1 || frame_h < 1 || crop_w < 1 || crop_h < 1 || !(ref_face_size > 0.f))
        return TDI_ERR_ARG;
    const int NL = TDI_NLANDMARKS;
    for (int f = 0; f < n; ++f) {
        const size_t o = (size_t)f * NL;
        status[f] = tdi_host::frame_geometry(lmk_x + o, lmk_y + o, cams + 4 * (size_t)f, frame_w, frame_h,
                                             ref_face_size, pad != 0, crop_w, crop_h, coef + f, lmk_x_out + o,
                                             lmk_y_out + o, cams_out + 4 * (size_t)f, crop + 4 * (size_t)f,
                                             face_size + f);
        if (status[f]) {
            coef[f] = 0.f;
            face_size[f] = 0.f;
            for (int i = 0; i < 4; ++i) { cams_out[4 * (size_t)f + i] = 0.f; crop[4 * (size_t)f + i] = 0; }
            for (int i = 0; i < NL; ++i) lmk_x_out[o + i] = lmk_y_out[o + i] = 0.f;
        }
    }
    return TDI_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------- the reference's names
namespace tdi {

bool fit_3DMM_shape_rigid(FitContext& ctx, float* theta, tdi_fit_report* reports)
{
    if (!ctx.batch || !theta || !reports) return false;
    const int n = tdi_batch_size(ctx.batch), K = tdi_batch_unknowns(ctx.batch);
    if (tdi_host::shape_fit(ctx.batch, n, K, ctx.batched, false, theta, reports)) return false;
    return tdi_host::all_success(reports, n);
}

bool fit_3DMM_epsilon_finetune(FitContext& ctx, float* theta, tdi_fit_report* reports)
{
    if (!ctx.batch || !theta || !reports) return false;
    const int n = tdi_batch_size(ctx.batch), K = tdi_batch_unknowns(ctx.batch);
    if (tdi_host::shape_fit(ctx.batch, n, K, ctx.batched, true, theta, reports)) return false;
    return tdi_host::all_success(reports, n);
}

bool fit_3DMM_lambdas(FitContext& ctx, const float* theta, float* lambdas, float* Lintensity, tdi_fit_report* reports)
{
    if (!ctx.batch || !theta || !lambdas || !Lintensity || !reports) return false;
    const int n = tdi_batch_size(ctx.batch);
    if (tdi_host::light_fit(ctx.batch, n, theta, lambdas, Lintensity, false, reports)) return false;
    return tdi_host::all_success(reports, n);
}

bool fit_3DMM_Lintensity(FitContext& ctx, const float* theta, float* lambdas, float* Lintensity,
                         tdi_fit_report* reports)
{
    if (!ctx.batch || !theta || !lambdas || !Lintensity || !reports) return false;
    const int n = tdi_batch_size(ctx.batch);
    if (tdi_host::light_fit(ctx.batch, n, theta, lambdas, Lintensity, true, reports)) return false;
    return tdi_host::all_success(reports, n);
}

bool fit_model_with_slack(FitContext& ctx, float* theta, int k_exp, tdi_lmk_report* reports)
{
    if (!ctx.batch || !theta || !reports) return false;
    const int n = tdi_batch_size(ctx.batch), K = tdi_batch_unknowns(ctx.batch);
    if (tdi_host::landmark_fit(ctx.batch, n, K, k_exp, theta, reports)) return false;
    for (int i = 0; i < n; ++i)
        if (reports[i].status != TDI_OK) return false;
    return true;
}

} // namespace tdi

// ---------------------------------------------------------------------------- per-frame stage of a video
namespace tdi_host {

// update_shape_single_ref_size's sizes (video_fitter.cu:273-280): REF_EXTRA = 5 (145).
static std::vector<float> reference_sizes(float R, int nres)
{
    const float RE = 5.0f, RE2 = RE / 2.0f;
    std::vector<float> v({R, R + RE + 5.0f, R + RE + 3.0f, R + RE + 9.0f, R + RE + 7.0f, R + RE - 9.0f,
                          R + RE - 6.0f, R + RE - 3.0f, R + RE2 + 5.0f, R + RE2 + 3.0f, R + RE2 + 9.0f,
                          R + RE2 + 7.0f, R + RE2 - 9.0f, R + RE2 - 6.0f, R + RE2 - 3.0f});
    v.resize(nres);
    return v;
}

} // namespace tdi_host

extern "C" {

void tdi_default_video_params(tdi_video_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->ref_face_size = 86.f;      // config.cu:18
    p->nres_coefs = 4;            // config.cu:41
    p->twostage_illum = 0;
    p->finetune_expressions = 0;  // config.cu:95
    p->pad = 1;                   // config.cu:43
    p->crop_w = 256;              // constants.h:55-56
    p->crop_h = 256;
    p->batched = 1;
    p->init_pose = 1;
    p->box_half_width = 0.15f;
    p->paint_innermouth_black = 1; // config.cu:92
    p->lmk_k_exp = 79;             // K_EPSILON_L, config.cu:112
}

int tdi_read_video_config(const char* path, tdi_fit_options* opts, tdi_video_params* p)
{
    if (!path || !opts || !p) return TDI_ERR_ARG;
    int two = p->twostage_illum, nres = p->nres_coefs;
    int rc = tdi_read_config(path, opts, &two, &nres);
    if (rc) return rc;
    std::ifstream in(path);
    std::string line;
    bool nres_set = false;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        const size_t colon = line.find(':');
        if (colon == std::string::npos || line.empty() || line[0] == '%') continue;
        std::string key = line.substr(0, colon), val = line.substr(colon + 1);
        key.erase(0, key.find_first_not_of(" \t"));
        key.erase(key.find_last_not_of(" \t\r") + 1);
        val.erase(0, val.find_first_not_of(" \t\""));
        val.erase(val.find_last_not_of(" \t\r\"") + 1);
        if (key.empty() || val.empty()) continue;
        const double v = std::atof(val.c_str());
        if (key == "FINETUNE_EXPRESSIONS") p->finetune_expressions = (int)v;   // config.cu:512-513
        else if (key == "FINETUNE_COEF" && v > 0) opts->finetune_coef = (float)v; // 518-519
        else if (key == "PAD_RENDERING") p->pad = (int)v;                      // 434-435
        else if (key == "PAINT_INNERMOUTH_BLACK") p->paint_innermouth_black = (int)v; // 509-510
        else if (key == "REF_FACE_SIZE" && v > 0) p->ref_face_size = (float)v; // 383-384
        else if (key == "NRES_COEFS") nres_set = true;
    }
    p->twostage_illum = two;
    if (nres_set) p->nres_coefs = nres;
    return TDI_OK;
}

int tdi_fit_video_frames_nres(tdi_ctx* ctx, const tdi_video_params* p, int32_t n, uint8_t* bgr, int32_t src_w,
                              int32_t src_h, int64_t frame_stride, const float* lmk_x, const float* lmk_y,
                              const float* cams, const float* boxes, const int64_t* frame_ids, float* theta, float* lam,
                              float* Lint, tdi_video_output* vo, int32_t* status, tdi_fit_report* reports)
{
    if (!ctx || !p || n <= 0 || !bgr || !lmk_x || !lmk_y || !cams || !frame_ids || !theta || !lam || !Lint ||
        !status || p->nres_coefs < 1 || p->nres_coefs > 15 || !(p->ref_face_size > 0.f))
        return TDI_ERR_ARG;
    const int NL = TDI_NLANDMARKS;
    // pose offsets of the unknowns: the pose follows the coefficient blocks (tdi_api.h); the
    // light is not an unknown here, so the pose is the last six
    int K = 0;
    std::vector<int> ok_idx;
    int rc = TDI_OK;
    std::vector<float> coef(n), lx((size_t)n * NL), ly((size_t)n * NL), cam2((size_t)n * 4), fsz(n);
    std::vector<int32_t> crop((size_t)n * 4), st(n);
    std::vector<float> box_default(NL), box_neg(NL);
    for (int i = 0; i < NL; ++i) { box_default[i] = p->box_half_width; box_neg[i] = -p->box_half_width; }
    for (int i = 0; i < n; ++i) status[i] = TDI_OK;
    const std::vector<float> sizes = tdi_host::reference_sizes(p->ref_face_size, p->nres_coefs);
    if (p->paint_innermouth_black) {
        // the frames whose landmarks lie inside the frame (video_fitter.cu:183-191), painted in one call;
        // the landmarks of the others are replaced by NaN, which the masking skips
        rc = tdi_frame_geometry(n, lmk_x, lmk_y, cams, src_w, src_h, sizes[0], p->pad, p->crop_w, p->crop_h,
                                coef.data(), lx.data(), ly.data(), cam2.data(), crop.data(), fsz.data(), st.data());
        if (rc) return rc;
        std::vector<float> px(lmk_x, lmk_x + (size_t)n * NL), py(lmk_y, lmk_y + (size_t)n * NL);
        for (int i = 0; i < n; ++i)
            if (st[i] != TDI_OK)
                for (int k = 0; k < NL; ++k) px[(size_t)i * NL + k] = py[(size_t)i * NL + k] = NAN;
        rc = tdi_paint_innermouth_black(n, bgr, src_w, src_h, frame_stride, px.data(), py.data(), p->ref_face_size,
                                        nullptr);
        if (rc) return rc;
    }
    for (size_t ri = 0; ri < sizes.size(); ++ri) {
        rc = tdi_frame_geometry(n, lmk_x, lmk_y, cams, src_w, src_h, sizes[ri], p->pad, p->crop_w, p->crop_h,
                                coef.data(), lx.data(), ly.data(), cam2.data(), crop.data(), fsz.data(), st.data());
        if (rc) return rc;
        ok_idx.clear();
        for (int i = 0; i < n; ++i)
            if (st[i] == TDI_OK) ok_idx.push_back(i); else status[i] = st[i];
        const int m = (int)ok_idx.size();
        if (m == 0) return TDI_OK;
        // the frames at this size: resized, grey, cropped (InputData::get_resized_frame and the crop of
        // fit_to_video_frame, video_fitter.cu:537-549), on the device
        std::vector<float> mcoef(m);
        std::vector<int32_t> mcrop((size_t)m * 4);
        long long pixels = 0;
        for (int j = 0; j < m; ++j) {
            const int i = ok_idx[j];
            mcoef[j] = coef[i];
            for (int q = 0; q < 4; ++q) mcrop[(size_t)j * 4 + q] = crop[(size_t)i * 4 + q];
            pixels += (long long)crop[(size_t)i * 4 + 2] * crop[(size_t)i * 4 + 3];
        }
        // the ok list's frames, one call per run of consecutive frames in bgr
        float* dcrops = nullptr;
        if (hipMalloc((void**)&dcrops, sizeof(float) * (size_t)pixels) != hipSuccess) return TDI_ERR_NOMEM;
        long long off = 0;
        for (int j = 0; j < m && !rc;) {
            int e = j + 1;
            while (e < m && ok_idx[e] == ok_idx[e - 1] + 1) ++e;
            rc = tdi_prepare_crops(e - j, bgr + (size_t)ok_idx[j] * frame_stride, src_w, src_h, frame_stride, &mcoef[j],
                                   &mcrop[(size_t)j * 4], dcrops + off, nullptr);
            for (; j < e; ++j) off += (long long)mcrop[(size_t)j * 4 + 2] * mcrop[(size_t)j * 4 + 3];
        }
        std::vector<tdi_frame_desc> fd(m);
        off = 0;
        for (int j = 0; j < m && !rc; ++j) {
            const int i = ok_idx[j];
            tdi_frame_desc& f = fd[j];
            std::memset(&f, 0, sizeof(f));
            f.phix = cam2[(size_t)i * 4]; f.phiy = cam2[(size_t)i * 4 + 1];
            f.cx = cam2[(size_t)i * 4 + 2]; f.cy = cam2[(size_t)i * 4 + 3];
            f.x0 = mcrop[(size_t)j * 4]; f.y0 = mcrop[(size_t)j * 4 + 1];
            f.width = mcrop[(size_t)j * 4 + 2]; f.height = mcrop[(size_t)j * 4 + 3];
            f.face_size = fsz[i];
            f.obs = dcrops + off;
            off += (long long)f.width * f.height;
            f.lmk_x = &lx[(size_t)i * NL];
            f.lmk_y = &ly[(size_t)i * NL];
            const float* bx = boxes ? boxes + (size_t)i * 4 * NL : nullptr;
            f.box_lx = bx ? bx : box_neg.data();
            f.box_ux = bx ? bx + NL : box_default.data();
            f.box_ly = bx ? bx + 2 * NL : box_neg.data();
            f.box_uy = bx ? bx + 3 * NL : box_default.data();
            for (int a = 0; a < 3; ++a) f.lambda[a] = lam[(size_t)i * 3 + a];
            f.Lintensity = Lint[i];
        }
        tdi_batch* b = nullptr;
        if (!rc) rc = tdi_batch_create(ctx, m, fd.data(), &b);
        if (!rc) {
            K = tdi_batch_unknowns(b);
            // the per-frame stage's unknowns are the expressions and the pose (the per-frame app fits
            // neither identity nor albedo, video_from_saved_identity.cu:154-157)
            if (vo && K - 6 != vo->k_exp) rc = TDI_ERR_ARG;
        }
        if (!rc) {
            std::vector<float> th((size_t)m * K), ml((size_t)m * 3), mL(m);
            std::vector<tdi_fit_report> rep(m);
            for (int j = 0; j < m; ++j) {
                const int i = ok_idx[j];
                std::memcpy(&th[(size_t)j * K], &theta[(size_t)i * K], sizeof(float) * K);
                for (int a = 0; a < 3; ++a) ml[(size_t)j * 3 + a] = lam[(size_t)i * 3 + a];
                mL[j] = Lint[i];
            }
            if (p->init_pose) {
                // fit_to_video_frame (video_fitter.cu:459-486): the orthographic start from this size's
                // landmarks and camera, then the landmark-only fit with slack from it
                std::vector<float> mx((size_t)m * NL), my((size_t)m * NL), mc((size_t)m * 4), pose((size_t)m * 6);
                for (int j = 0; j < m; ++j) {
                    const int i = ok_idx[j];
                    std::memcpy(&mx[(size_t)j * NL], &lx[(size_t)i * NL], sizeof(float) * NL);
                    std::memcpy(&my[(size_t)j * NL], &ly[(size_t)i * NL], sizeof(float) * NL);
                    std::memcpy(&mc[(size_t)j * 4], &cam2[(size_t)i * 4], sizeof(float) * 4);
                }
                rc = tdi_host::initialize_with_orthographic_t(ctx, m, mx.data(), my.data(), mc.data(), pose.data(),
                                                              nullptr, nullptr);
                const int kl = std::min(p->lmk_k_exp, K - 6);
                if (!rc && kl > 0) {
                    std::vector<float> tt(th);
                    std::vector<tdi_lmk_report> lrep(m);
                    for (int j = 0; j < m; ++j)
                        std::memcpy(&tt[(size_t)j * K + K - 6], &pose[(size_t)j * 6], sizeof(float) * 6);
                    rc = tdi_host::landmark_fit(b, m, K, kl, tt.data(), lrep.data());
                    for (int j = 0; !rc && j < m; ++j)
                        if (lrep[j].status == TDI_OK)
                            std::memcpy(&th[(size_t)j * K], &tt[(size_t)j * K], sizeof(float) * K);
                } else if (!rc) {
                    for (int j = 0; j < m; ++j)
                        std::memcpy(&th[(size_t)j * K + K - 6], &pose[(size_t)j * 6], sizeof(float) * 6);
                }
            }
            rc = tdi_fit_video_frames(b, p->twostage_illum, p->batched, th.data(), ml.data(), mL.data(), rep.data());
            // FINETUNE_EXPRESSIONS (video_fitter.cu:627-628)
            if (!rc && p->finetune_expressions)
                rc = tdi_host::shape_fit(b, m, K, p->batched != 0, true, th.data(), rep.data());
            for (int j = 0; !rc && j < m; ++j) {
                const int i = ok_idx[j];
                std::memcpy(&theta[(size_t)i * K], &th[(size_t)j * K], sizeof(float) * K);
                for (int a = 0; a < 3; ++a) lam[(size_t)i * 3 + a] = ml[(size_t)j * 3 + a];
                Lint[i] = mL[j];
                status[i] = rep[j].status;
                if (reports) reports[i] = rep[j];
                if (vo) {
                    // update_shape_single_ref_size (video_fitter.cu:365-389): expressions, pose, light
                    const float il[4] = {ml[(size_t)j * 3], ml[(size_t)j * 3 + 1], ml[(size_t)j * 3 + 2], mL[j]};
                    tdi_vo_add(vo, frame_ids[i], &th[(size_t)j * K], &th[(size_t)j * K + K - 6], il);
                }
            }
        }
        if (b) tdi_batch_destroy(b);
        (void)hipFree(dcrops);
        if (rc) return rc;
    }
    return TDI_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------- long videos
namespace tdi_host {

struct LongState {
    int n_frames = 0, K = 0, frames_done = 0, chunks_done = 0;
    std::vector<float> theta, lam, Lint;
};

static bool read_state(const std::string& path, LongState& st)
{
    std::ifstream in(path.c_str());
    if (!in) return false;
    std::string tag;
    int version = 0;
    if (!(in >> tag >> version) || tag != "tdi_long_video" || version != 1) return false;
    auto floats = [&](const char* name, size_t n, std::vector<float>& v) {
        std::string key;
        if (!(in >> key) || key != name) return false;
        v.resize(n);
        for (float& x : v) {
            std::string tok;
            if (!(in >> tok)) return false;
            x = std::strtof(tok.c_str(), nullptr);
        }
        return true;
    };
    std::string k1, k2, k3, k4;
    if (!(in >> k1 >> st.n_frames >> k2 >> st.K >> k3 >> st.frames_done >> k4 >> st.chunks_done)) return false;
    if (k1 != "n_frames" || k2 != "K" || k3 != "frames_done" || k4 != "chunks_done") return false;
    return floats("theta", (size_t)st.K, st.theta) && floats("lam", 3, st.lam) && floats("Lint", 1, st.Lint);
}

// Written to a temporary file and renamed over the old one, so that a run stopped while writing
// leaves the previous state whole.
static bool write_state(const std::string& path, const LongState& st)
{
    const std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp.c_str(), std::ios::trunc);
        if (!out) return false;
        out << "tdi_long_video 1\n";
        out << "n_frames " << st.n_frames << " K " << st.K << " frames_done " << st.frames_done << " chunks_done "
            << st.chunks_done << "\n" << std::hexfloat;
        out << "theta";
        for (float v : st.theta) out << ' ' << v;
        out << "\nlam";
        for (float v : st.lam) out << ' ' << v;
        out << "\nLint";
        for (float v : st.Lint) out << ' ' << v;
        out << "\n";
        if (!out) return false;
    }
    return std::rename(tmp.c_str(), path.c_str()) == 0;
}

} // namespace tdi_host

extern "C" {

int tdi_fit_video_long(tdi_ctx* ctx, const tdi_video_params* p, const tdi_long_video* lv, tdi_frame_source src,
                       void* user, const float* lmk_x, const float* lmk_y, const float* cams, float* theta, float* lam,
                       float* Lint, tdi_video_output* vo, int32_t* status, int32_t* frames_done)
{
    if (!ctx || !p || !lv || !src || !lmk_x || !lmk_y || !cams || !theta || !lam || !Lint || !vo || !status ||
        lv->n_frames <= 0 || lv->src_w < 1 || lv->src_h < 1 || lv->chunk_frames < 0 || lv->max_chunks < 0)
        return TDI_ERR_ARG;
    const int NL = TDI_NLANDMARKS, n = lv->n_frames;
    const int K = vo->k_exp + 6;
    const int64_t stride = (int64_t)lv->src_w * lv->src_h * 3;
    const int chunk = lv->chunk_frames > 0 ? std::min(lv->chunk_frames, n) : std::min(n, 1024);
    const std::string state_path = lv->state_path ? lv->state_path : "";
    const std::string fits_path = state_path.empty() ? "" : state_path + ".fits";
    tdi_host::LongState st;
    st.n_frames = n;
    st.K = K;
    st.theta.assign(theta, theta + K);
    st.lam.assign(lam, lam + 3);
    st.Lint.assign(Lint, Lint + 1);
    for (int i = 0; i < n; ++i) status[i] = -1;
    if (lv->resume && !state_path.empty()) {
        tdi_host::LongState prev;
        if (tdi_host::read_state(state_path, prev)) {
            if (prev.n_frames != n || prev.K != K || !vo->vo.log.empty()) return TDI_ERR_ARG;
            // the fits of the chunks the state records (a chunk whose state was not written is redone)
            if (!vo->vo.load(fits_path, (size_t)prev.frames_done) || !vo->vo.rewrite(fits_path)) return TDI_ERR_ARG;
            st = prev;
        } else {
            std::remove(fits_path.c_str());   // no state: nothing finished, start afresh
        }
    } else if (!state_path.empty()) {
        std::remove(fits_path.c_str());
        std::remove(state_path.c_str());
    }
    int ran = 0, rc = TDI_OK;
    std::vector<float> th, ml, mL;
    std::vector<int64_t> ids;
    while (st.frames_done < n && (lv->max_chunks == 0 || ran < lv->max_chunks)) {
        const int first = st.frames_done, m = std::min(chunk, n - first);
        uint8_t* bgr = const_cast<uint8_t*>(src(user, first, m));
        if (!bgr) { rc = TDI_ERR_ARG; break; }
        // every frame of the chunk starts from the carried state (the last fit of the previous chunk)
        th.resize((size_t)m * K);
        ml.resize((size_t)m * 3);
        mL.resize(m);
        ids.resize(m);
        for (int j = 0; j < m; ++j) {
            std::memcpy(&th[(size_t)j * K], st.theta.data(), sizeof(float) * K);
            std::memcpy(&ml[(size_t)j * 3], st.lam.data(), sizeof(float) * 3);
            mL[j] = st.Lint[0];
            ids[j] = first + j;
        }
        rc = tdi_fit_video_frames_nres(ctx, p, m, bgr, lv->src_w, lv->src_h, stride, lmk_x + (size_t)first * NL,
                                       lmk_y + (size_t)first * NL, cams + (size_t)first * 4, nullptr, ids.data(),
                                       th.data(), ml.data(), mL.data(), vo, status + first, nullptr);
        if (rc) break;
        for (int j = m - 1; j >= 0; --j)
            if (status[first + j] == TDI_OK) {
                std::memcpy(st.theta.data(), &th[(size_t)j * K], sizeof(float) * K);
                std::memcpy(st.lam.data(), &ml[(size_t)j * 3], sizeof(float) * 3);
                st.Lint[0] = mL[j];
                break;
            }
        st.frames_done += m;
        st.chunks_done += 1;
        ++ran;
        if (!state_path.empty() && (!vo->vo.dump(fits_path) || !tdi_host::write_state(state_path, st))) {
            rc = TDI_ERR_ARG;
            break;
        }
    }
    std::memcpy(theta, st.theta.data(), sizeof(float) * K);
    std::memcpy(lam, st.lam.data(), sizeof(float) * 3);
    Lint[0] = st.Lint[0];
    if (frames_done) *frames_done = st.frames_done;
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------------------- identity stage
namespace tdi_host {

// check_if_bb_OK (funcs.cu:582-612) with PAD_RENDERING: the box extended by half its size to the
// left and above must stay inside the frame.
static bool bb_ok(const float* x, const float* y, bool pad)
{
    const int NL = TDI_NLANDMARKS;
    const float x0 = *std::min_element(x, x + NL), y0 = *std::min_element(y, y + NL);
    const float xs = *std::max_element(x, x + NL) - x0, ys = *std::max_element(y, y + NL) - y0;
    if (!pad) return true;
    return !(x0 - xs / 2 <= 0) && !(y0 - ys / 2 <= 0);
}

// check_if_face_in_frame (funcs.cu:614-626)
static bool face_in_frame(const float* x, const float* y, float cols, float rows)
{
    const int NL = TDI_NLANDMARKS;
    return !(*std::min_element(x, x + NL) < 0 || *std::max_element(x, x + NL) >= cols ||
             *std::min_element(y, y + NL) < 0 || *std::max_element(y, y + NL) >= rows);
}

// Nearest centroid index (the first on a tie, as std::min_element finds it).
static size_t nearest(const std::vector<int>& centroids, double deg)
{
    size_t best = 0;
    double bd = std::fabs(centroids[0] - deg);
    for (size_t i = 1; i < centroids.size(); ++i) {
        const double d = std::fabs(centroids[i] - deg);
        if (d < bd) { bd = d; best = i; }
    }
    return best;
}

// A batch of frames whose crops are 1 x 1 (the landmark fits read no pixels).
static int landmark_batch(tdi_ctx* ctx, int n, const float* lx, const float* ly, const float* cams,
                          const float* face_size, const float* box_lo, const float* box_hi, const float* one_pixel,
                          tdi_batch** b)
{
    const int NL = TDI_NLANDMARKS;
    std::vector<tdi_frame_desc> fd(n);
    for (int i = 0; i < n; ++i) {
        tdi_frame_desc& f = fd[i];
        std::memset(&f, 0, sizeof(f));
        f.phix = cams[4 * (size_t)i]; f.phiy = cams[4 * (size_t)i + 1];
        f.cx = cams[4 * (size_t)i + 2]; f.cy = cams[4 * (size_t)i + 3];
        f.width = 1; f.height = 1;
        f.face_size = face_size[i];
        f.obs = one_pixel;
        f.lmk_x = lx + (size_t)i * NL;
        f.lmk_y = ly + (size_t)i * NL;
        f.box_lx = box_lo; f.box_ux = box_hi; f.box_ly = box_lo; f.box_uy = box_hi;
        f.lambda[0] = -7.3627f; f.lambda[1] = 51.1364f; f.lambda[2] = 100.1784f;
        f.Lintensity = 0.005f;
    }
    return tdi_batch_create(ctx, n, fd.data(), b);
}

// The binning of fit_video_frames_landmarks_sparse (video_fitter.cu:2380-2550): bins, shuffles,
// permutations and groups; accept(group, keep) decides whether a full group is kept (the
// reference's joint landmark-only fit of the group, 2505-2545).  out: the kept groups' frames.
template <typename Accept>
static int bin_groups(const tdi_identity_params* p, int n, const float* yaw, const float* pitch, Accept accept,
                      std::vector<int>& out)
{
    // 13 x 13 (yaw, pitch) centroids at -55 .. 65 degrees (video_fitter.cu:2380-2400)
    std::vector<int> yc, pc;
    for (int c = -60; c <= 60; c += 10) yc.push_back(c + 5);
    for (int c = -60; c <= 60; c += 10) pc.push_back(c + 5);
    std::vector<std::vector<std::vector<int>>> bins(yc.size(), std::vector<std::vector<int>>(pc.size()));
    const double r2d = 180.0 / M_PI;
    for (int fi = 0; fi < n; ++fi) {
        if (std::isnan(yaw[fi])) continue;
        bins[tdi_host::nearest(yc, r2d * yaw[fi])][tdi_host::nearest(pc, r2d * pitch[fi])].push_back(fi);
    }
    std::default_random_engine e_seed(1907);
    std::vector<int> chosen;
    for (size_t yi = 0; yi < yc.size(); ++yi)
        for (size_t pi = 0; pi < pc.size(); ++pi) {
            std::vector<int>& v = bins[yi][pi];
            std::shuffle(v.begin(), v.end(), e_seed);
            for (size_t ci = 0; ci < std::min<size_t>(v.size(), (size_t)p->nframes_per_angle_bin); ++ci)
                chosen.push_back(v[ci]);
        }
    // the chosen frames in frame order (the video is read again from its start, 2455-2478)
    std::sort(chosen.begin(), chosen.end());
    chosen.erase(std::unique(chosen.begin(), chosen.end()), chosen.end());
    std::vector<int> order;
    for (int k = 0; k < p->nperms; ++k) {
        std::vector<int> iv(chosen.size());
        for (size_t i = 0; i < iv.size(); ++i) iv[i] = (int)i;
        std::shuffle(iv.begin(), iv.end(), e_seed);
        order.insert(order.end(), iv.begin(), iv.end());
    }
    std::vector<int> group;
    for (int ri : order) {
        group.push_back(chosen[ri]);
        if ((int)group.size() == p->nframes) {
            bool keep = false;
            const int rc = accept(group, keep);
            if (rc) return rc;
            if (keep) out.insert(out.end(), group.begin(), group.end());
            group.clear();
        }
        if ((long long)out.size() >= p->nmax_frames) break;
        if ((long long)out.size() >= (long long)p->ntot_reconstrs * p->nframes) break;
    }
    return TDI_OK;
}

} // namespace tdi_host

extern "C" {

void tdi_default_identity_params(tdi_identity_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nframes = 7;                 // config.cu:45
    p->nframes_per_angle_bin = 14;  // config.cu:79
    p->nperms = 10;                 // config.cu:27
    p->ntot_reconstrs = 6;          // config.cu:29
    p->nmax_frames = 432000;        // config.cu:28
    p->max_vid_frames = 50000;      // config.cu:59
    p->skip_first_n_secs = 0.f;     // config.cu:55
    p->fps = 30.f;
    p->ref_face_size = 86.f;        // config.cu:18
    p->pad = 1;                     // config.cu:43
    p->crop_w = 256;                // constants.h:55-56
    p->crop_h = 256;
    p->paint_innermouth_black = 1;  // config.cu:92
    p->lmk_k_exp = 79;              // config.cu:112
    p->box_half_width = 0.15f;
    p->lmk_k_id = 60;               // K_ALPHA_L, config.cu:160
}

int tdi_bin_identity_frames(const tdi_identity_params* p, int32_t n, const float* yaw, const float* pitch,
                            int32_t* selected, int32_t cap, int32_t* n_selected)
{
    if (!p || n < 0 || (n > 0 && (!yaw || !pitch)) || !n_selected || cap < 0 || (cap > 0 && !selected) ||
        p->nframes < 1 || p->nperms < 0 || p->nframes_per_angle_bin < 0 || p->ntot_reconstrs < 0)
        return TDI_ERR_ARG;
    std::vector<int> out;
    const int rc = tdi_host::bin_groups(p, n, yaw, pitch, [](const std::vector<int>&, bool& keep) {
        keep = true;
        return TDI_OK;
    }, out);
    if (rc) return rc;
    *n_selected = (int32_t)out.size();
    if ((int)out.size() > cap) return TDI_ERR_ARG;
    for (size_t i = 0; i < out.size(); ++i) selected[i] = out[i];
    return TDI_OK;
}

int tdi_select_identity_frames(tdi_ctx* ctx, const tdi_identity_params* p, int32_t n, const float* lmk_x,
                               const float* lmk_y, int32_t frame_w, int32_t frame_h, const float* cam, float* angles,
                               int32_t* selected, int32_t cap, int32_t* n_selected, float* resize_coef)
{
    if (!ctx || !p || n <= 0 || !lmk_x || !lmk_y || !cam || !n_selected || frame_w < 1 || frame_h < 1)
        return TDI_ERR_ARG;
    const int NL = TDI_NLANDMARKS;
    // first pass (video_fitter.cu:2118-2240): the frames considered and their median face size
    int every = 1;
    if (n > 800) every = n / 800;
    std::vector<float> sizes;
    for (int idx = 1; idx - 1 < n; ++idx) {
        if (idx < p->fps * p->skip_first_n_secs) continue;
        if (idx % every != 0) continue;
        if (idx >= p->max_vid_frames) break;
        const float* x = lmk_x + (size_t)(idx - 1) * NL;
        const float* y = lmk_y + (size_t)(idx - 1) * NL;
        bool finite = true;
        for (int i = 0; i < NL; ++i) finite = finite && std::isfinite(x[i]) && std::isfinite(y[i]);
        if (!finite) continue;
        const int xmn = (int)*std::min_element(x, x + NL), xmx = (int)*std::max_element(x, x + NL);
        const int ymn = (int)*std::min_element(y, y + NL), ymx = (int)*std::max_element(y, y + NL);
        if (xmn <= 0 || ymn <= 0 || xmx >= frame_w || ymx >= frame_h) continue;
        sizes.push_back(tdi_host::compute_face_size(x, y));
    }
    if (sizes.empty()) { *n_selected = 0; return TDI_OK; }
    const size_t ni = sizes.size() / 2;
    std::nth_element(sizes.begin(), sizes.begin() + ni, sizes.end());
    const float coef = 76.0f / sizes[ni];
    if (resize_coef) *resize_coef = coef;
    // second pass (2300-2372): every frame's resized landmarks, checks, landmark fits, angles
    std::vector<float> lx((size_t)n * NL), ly((size_t)n * NL), cams((size_t)n * 4), fs(n);
    std::vector<char> ok(n, 0);
    for (int f = 0; f < n; ++f) {
        bool finite = true;
        for (int i = 0; i < NL; ++i) {
            lx[(size_t)f * NL + i] = coef * lmk_x[(size_t)f * NL + i];
            ly[(size_t)f * NL + i] = coef * lmk_y[(size_t)f * NL + i];
            finite = finite && std::isfinite(lx[(size_t)f * NL + i]) && std::isfinite(ly[(size_t)f * NL + i]);
        }
        for (int a = 0; a < 4; ++a) cams[4 * (size_t)f + a] = coef * cam[a];
        if (!finite) { fs[f] = 1.f; continue; }
        const float* x = &lx[(size_t)f * NL];
        const float* y = &ly[(size_t)f * NL];
        fs[f] = tdi_host::compute_face_size(x, y);
        ok[f] = x[0] != 0.0f && tdi_host::bb_ok(x, y, p->pad != 0) &&
                tdi_host::face_in_frame(x, y, coef * (float)frame_w, coef * (float)frame_h);
    }
    std::vector<float> pose((size_t)n * 6), yaw(n, NAN), pitch(n, NAN), roll(n, NAN);
    int rc = tdi_host::initialize_with_orthographic_t(ctx, n, lx.data(), ly.data(), cams.data(), pose.data(), nullptr,
                                                      nullptr);
    if (rc) return rc;
    std::vector<float> lo(NL, -p->box_half_width), hi(NL, p->box_half_width);
    float* one = nullptr;
    if (hipMalloc((void**)&one, sizeof(float)) != hipSuccess) return TDI_ERR_NOMEM;
    (void)hipMemset(one, 0, sizeof(float));
    int32_t lay[6];
    tdi_ctx_layout(ctx, lay);
    const int op = lay[3], K = lay[5];
    const int kl = std::max(0, std::min(p->lmk_k_exp, op - lay[2]));
    const int ka = std::max(0, std::min(p->lmk_k_id, lay[2] - lay[1]));
    // a group of frames (or one) with identity, albedo and expressions at zero and the orthographic
    // poses, fitted jointly with the slack (li_init of video_fitter.cu:2253-2275, reset per frame,
    // and fit_to_multi_images_landmarks_only, model_fitter.cu:1205-1247)
    auto fit_group = [&](const std::vector<int>& idx, std::vector<float>& th, tdi_lmk_report& rep) -> int {
        const int m = (int)idx.size();
        std::vector<float> gx((size_t)m * NL), gy((size_t)m * NL), gc((size_t)m * 4), gf(m);
        for (int t = 0; t < m; ++t) {
            const int f = idx[t];
            std::memcpy(&gx[(size_t)t * NL], &lx[(size_t)f * NL], sizeof(float) * NL);
            std::memcpy(&gy[(size_t)t * NL], &ly[(size_t)f * NL], sizeof(float) * NL);
            std::memcpy(&gc[(size_t)t * 4], &cams[(size_t)f * 4], sizeof(float) * 4);
            gf[t] = fs[f];
        }
        th.assign((size_t)m * K, 0.f);
        for (int t = 0; t < m; ++t)
            for (int a = 0; a < 6; ++a) th[(size_t)t * K + op + a] = pose[(size_t)idx[t] * 6 + a];
        tdi_batch* b = nullptr;
        int rc = tdi_host::landmark_batch(ctx, m, gx.data(), gy.data(), gc.data(), gf.data(), lo.data(), hi.data(), one,
                                          &b);
        if (!rc) rc = tdi_fit_landmarks_joint(b, th.data(), ka, kl, 1, &rep);
        if (b) tdi_batch_destroy(b);
        return rc;
    };
    std::vector<float> th;
    {
        // every checked frame's own fit, all frames together
        std::vector<int> idx;
        for (int f = 0; f < n; ++f)
            if (ok[f]) idx.push_back(f);
        const int m = (int)idx.size();
        std::vector<float> gx((size_t)m * NL), gy((size_t)m * NL), gc((size_t)m * 4), gf(m);
        for (int t = 0; t < m; ++t) {
            const int f = idx[t];
            std::memcpy(&gx[(size_t)t * NL], &lx[(size_t)f * NL], sizeof(float) * NL);
            std::memcpy(&gy[(size_t)t * NL], &ly[(size_t)f * NL], sizeof(float) * NL);
            std::memcpy(&gc[(size_t)t * 4], &cams[(size_t)f * 4], sizeof(float) * 4);
            gf[t] = fs[f];
        }
        th.assign((size_t)m * K, 0.f);
        for (int t = 0; t < m; ++t)
            for (int a = 0; a < 6; ++a) th[(size_t)t * K + op + a] = pose[(size_t)idx[t] * 6 + a];
        std::vector<tdi_lmk_report> rep(m);
        tdi_batch* b = nullptr;
        if (m > 0) {
            rc = tdi_host::landmark_batch(ctx, m, gx.data(), gy.data(), gc.data(), gf.data(), lo.data(), hi.data(), one,
                                          &b);
            if (!rc) rc = tdi_fit_landmarks_each(b, th.data(), ka, kl, rep.data());
            if (b) tdi_batch_destroy(b);
        }
        for (int t = 0; !rc && t < m; ++t) {
            if (rep[t].status != TDI_OK) continue;
            float R[9], ypr[3];
            tdi_host::rotation_f32(&th[(size_t)t * K + op + 3], R);
            tdi_host::compute_euler_angles(R, ypr[0], ypr[1], ypr[2]);
            const int f = idx[t];
            yaw[f] = ypr[0]; pitch[f] = ypr[1]; roll[f] = ypr[2];
        }
    }
    if (angles)
        for (int f = 0; f < n; ++f) { angles[3 * (size_t)f] = yaw[f]; angles[3 * (size_t)f + 1] = pitch[f]; angles[3 * (size_t)f + 2] = roll[f]; }
    std::vector<int> out;
    if (!rc)
        rc = tdi_host::bin_groups(p, n, yaw.data(), pitch.data(), [&](const std::vector<int>& group, bool& keep) {
            tdi_lmk_report rep{};
            const int e = fit_group(group, th, rep);
            keep = !e && rep.status == TDI_OK;
            return e;
        }, out);
    (void)hipFree(one);
    if (rc) return rc;
    *n_selected = (int32_t)out.size();
    if ((int)out.size() > cap) return TDI_ERR_ARG;
    for (size_t i = 0; i < out.size(); ++i) selected[i] = out[i];
    return TDI_OK;
}

} // extern "C"

extern "C" {

int tdi_learn_identity(tdi_ctx* ctx, const tdi_identity_params* p, int32_t n, uint8_t* bgr, int32_t frame_w,
                       int32_t frame_h, int64_t frame_stride, const float* lmk_x, const float* lmk_y, const float* cam,
                       float* alpha, float* beta, int32_t* n_used)
{
    if (!ctx || !p || n < 0 || (n > 0 && (!bgr || !lmk_x || !lmk_y)) || !cam || !n_used || p->nframes < 1 ||
        frame_w < 1 || frame_h < 1)
        return TDI_ERR_ARG;
    const int NL = TDI_NLANDMARKS, nf = p->nframes;
    int32_t lay[6];
    int rc = tdi_ctx_layout(ctx, lay);
    if (rc) return rc;
    const int off_tex = lay[0], off_id = lay[1], off_exp = lay[2], off_pose = lay[3], K = lay[5];
    const int ntex = off_id - off_tex, nid = off_exp - off_id;
    if ((ntex && !beta) || (nid && !alpha)) return TDI_ERR_ARG;
    // learn_identity's sums in FP32: the first used fit assigns, the others add (1962-1982)
    std::vector<float> sa(nid, 0.f), sb(ntex, 0.f);
    int used = 0;
    std::vector<float> cams_in((size_t)nf * 4);
    for (int f = 0; f < nf; ++f)
        for (int a = 0; a < 4; ++a) cams_in[4 * (size_t)f + a] = cam[a];
    std::vector<float> box_lo(NL, -p->box_half_width), box_hi(NL, p->box_half_width);
    for (int g = 0; g < p->ntot_reconstrs && (g + 1) * nf <= n; ++g) {
        const int g0 = g * nf;
        const float* gx = lmk_x + (size_t)g0 * NL;
        const float* gy = lmk_y + (size_t)g0 * NL;
        uint8_t* gb = bgr + (size_t)g0 * frame_stride;
        // fit_multiframe (video_fitter.cu:2023-2083): each frame's camera scaled to REF_FACE_SIZE;
        // fit_to_multi_images (model_fitter.cu:1252-1479): grey, inner mouth, resize, crop
        std::vector<float> coef(nf), lx((size_t)nf * NL), ly((size_t)nf * NL), cam2((size_t)nf * 4), fsz(nf);
        std::vector<int32_t> crop((size_t)nf * 4), st(nf);
        rc = tdi_frame_geometry(nf, gx, gy, cams_in.data(), frame_w, frame_h, p->ref_face_size, p->pad, p->crop_w,
                                p->crop_h, coef.data(), lx.data(), ly.data(), cam2.data(), crop.data(), fsz.data(),
                                st.data());
        if (rc) return rc;
        bool all_in = true;
        for (int f = 0; f < nf; ++f) all_in = all_in && st[f] == TDI_OK;
        if (!all_in) continue;
        if (p->paint_innermouth_black &&
            (rc = tdi_paint_innermouth_black(nf, gb, frame_w, frame_h, frame_stride, gx, gy, p->ref_face_size, nullptr)))
            return rc;
        long long pixels = 0;
        for (int f = 0; f < nf; ++f) pixels += (long long)crop[4 * (size_t)f + 2] * crop[4 * (size_t)f + 3];
        float* dcrops = nullptr;
        if (hipMalloc((void**)&dcrops, sizeof(float) * (size_t)pixels) != hipSuccess) return TDI_ERR_NOMEM;
        rc = tdi_prepare_crops(nf, gb, frame_w, frame_h, frame_stride, coef.data(), crop.data(), dcrops, nullptr);
        std::vector<tdi_frame_desc> fd(nf);
        long long off = 0;
        for (int f = 0; f < nf; ++f) {
            tdi_frame_desc& d = fd[f];
            std::memset(&d, 0, sizeof(d));
            d.phix = cam2[4 * (size_t)f]; d.phiy = cam2[4 * (size_t)f + 1];
            d.cx = cam2[4 * (size_t)f + 2]; d.cy = cam2[4 * (size_t)f + 3];
            d.x0 = crop[4 * (size_t)f]; d.y0 = crop[4 * (size_t)f + 1];
            d.width = crop[4 * (size_t)f + 2]; d.height = crop[4 * (size_t)f + 3];
            d.face_size = fsz[f];
            d.obs = dcrops + off;
            off += (long long)d.width * d.height;
            d.lmk_x = &lx[(size_t)f * NL];
            d.lmk_y = &ly[(size_t)f * NL];
            d.box_lx = box_lo.data(); d.box_ux = box_hi.data(); d.box_ly = box_lo.data(); d.box_uy = box_hi.data();
            d.lambda[0] = -7.3627f; d.lambda[1] = 51.1364f; d.lambda[2] = 100.1784f;
            d.Lintensity = 0.005f;
        }
        tdi_batch* b = nullptr;
        if (!rc) rc = tdi_batch_create(ctx, nf, fd.data(), &b);
        bool success = false, start_ok = false;
        std::vector<float> th((size_t)nf * K, 0.f);
        if (!rc) {
            // the start of fit_to_multi_images (model_fitter.cu:1291-1330): the orthographic poses
            // with identity, albedo and expressions at zero; the joint landmark-only fit with the
            // slack over K_ALPHA_L identity coefficients (the group is not used if it fails); from
            // its result the joint landmark-only fit without the slack over all identity coefficients
            std::vector<float> pose((size_t)nf * 6);
            rc = tdi_host::initialize_with_orthographic_t(ctx, nf, lx.data(), ly.data(), cam2.data(), pose.data(), nullptr,
                                                          nullptr);
            for (int f = 0; !rc && f < nf; ++f)
                for (int a = 0; a < 6; ++a) th[(size_t)f * K + off_pose + a] = pose[(size_t)f * 6 + a];
            const int kl = std::max(0, std::min(p->lmk_k_exp, off_pose - off_exp));
            const int ka = std::max(0, std::min(p->lmk_k_id, nid));
            tdi_lmk_report lrep{};
            if (!rc) rc = tdi_fit_landmarks_joint(b, th.data(), ka, kl, 1, &lrep);
            start_ok = !rc && lrep.status == TDI_OK;
            if (start_ok) rc = tdi_fit_landmarks_joint(b, th.data(), nid, kl, 0, &lrep);
        }
        if (!rc && start_ok) {
            // the light sub-fits of fit_to_multi_images (model_fitter.cu:1374-1386)
            std::vector<float> lam((size_t)nf * 3), Lint(nf, 0.005f);
            for (int f = 0; f < nf; ++f) { lam[3 * (size_t)f] = -7.3627f; lam[3 * (size_t)f + 1] = 51.1364f; lam[3 * (size_t)f + 2] = 100.1784f; }
            if (!rc) rc = tdi_batch_set_light(b, lam.data(), Lint.data());
            for (int k = 0; !rc && k < 2; ++k) {
                rc = tdi_host::light_fit(b, nf, th.data(), lam.data(), Lint.data(), true, nullptr);
                if (!rc) rc = tdi_host::light_fit(b, nf, th.data(), lam.data(), Lint.data(), false, nullptr);
            }
            // the joint shape fit over the group's frames
            tdi_fit_report rep;
            if (!rc) rc = tdi_fit_multiframe(b, th.data(), &rep);
            success = !rc && rep.success && rep.status == TDI_OK;
        }
        if (b) tdi_batch_destroy(b);
        (void)hipFree(dcrops);
        if (rc) return rc;
        if (success) {
            ++used;
            for (int k = 0; k < nid; ++k) sa[k] = used == 1 ? th[off_id + k] : sa[k] + th[off_id + k];
            for (int k = 0; k < ntex; ++k) sb[k] = used == 1 ? th[off_tex + k] : sb[k] + th[off_tex + k];
        }
    }
    *n_used = used;
    if (used) {
        // the average over the groups (learn_identity's running sum times 1 / num_used_recs)
        const float w = (float)1. / (float)used;
        for (int k = 0; k < nid; ++k) alpha[k] = sa[k] * w;
        for (int k = 0; k < ntex; ++k) beta[k] = sb[k] * w;
    }
    return TDI_OK;
}

} // extern "C"
