// capi.hip — the C ABI of include/tdi_api.h: context, batch, building blocks and the two fits.
//
// Ownership follows the reference's split (Renderer/Optimizer own device memory,
// renderer.cu:28-72, Optimizer.h:249-405) but everything lives in two objects:
//   tdi_ctx   — the model in HBM (mean, bases, triangles, landmark ids, bounds, blur taps);
//   tdi_batch — the frames in HBM (normalised observation gradients, landmarks) plus one
//               workspace sized for a chunk of frames.  No allocation happens inside a fit.
#include "common.h"

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace tdi {

static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

// Reference taps (video_fitter.cu:51-63): KERNEL_SIGMA 0.5/0.75/1 use that sigma, 2.0 is the
// sigma = 1.5 table; normalised Gaussian of width 15 (video_fitter.cu:2836-2849).
static int taps_for(float key, float* taps)
{
    double s;
    if (key == 0.5f) s = 0.5;
    else if (key == 0.75f) s = 0.75;
    else if (key == 1.0f) s = 1.0;
    else if (key == 2.0f) s = 1.5;
    else return TDI_ERR_ARG;
    double t[15], sum = 0;
    for (int i = 0; i < 15; ++i) {
        const double x = i - 7;
        t[i] = std::exp(-x * x / (2 * s * s));
        sum += t[i];
    }
    for (int i = 0; i < 15; ++i) taps[i] = (float)(t[i] / sum);
    return TDI_OK;
}

int launch_misc_pack(int n, int V, const float* xyz, const float* xyp, float4* vproj, hipStream_t s);
int launch_misc_unpack(int n, int V, const float* P, const float4* vproj, float* xyz, float* xyp, hipStream_t s);
int launch_misc_fill_f32(float* p, long long count, float v, hipStream_t s);
int launch_misc_f64_to_f32(const double* in, float* out, int n, hipStream_t s);
int launch_misc_rotation(int n, const float* u, float* R, float* dR, hipStream_t s);
// the light sub-fits' per-frame results packed into 16 doubles per slot (light block of H, light
// gradient, objective, barrier objective, pixel count) for one device-to-host copy
int launch_misc_light_pack(int n, int K, int j0, int nd, const double* H, const float* g, const double* obj,
                           const double* objb, const int* npix, double* out, hipStream_t s);

} // namespace tdi

using namespace tdi;

// Device blocks that a context keeps for its later batches: creating and destroying batches (the
// video drivers make one per chunk and reference face size) then reuses workspaces of the same
// sizes instead of calling hipMalloc/hipFree for gigabytes each time.  Blocks are rounded to 2 MiB
// (256 B below that); a request takes the smallest cached block that is at most 25% + 2 MiB larger.
// The pool holds at most a quarter of the device's memory; what would exceed it is freed, and the
// whole pool is freed when an allocation fails (then retried) and with the context.
struct DevPool {
    std::mutex mu;
    std::multimap<size_t, void*> blocks;
    size_t cached = 0, limit = 0;
    DevPool()
    {
        size_t fb = 0, tb = 0;
        limit = hipMemGetInfo(&fb, &tb) == hipSuccess ? tb / 4 : 0;
    }
    ~DevPool() { trim(); }
    void* take(size_t bytes)
    {
        std::lock_guard<std::mutex> lock(mu);
        const auto it = blocks.lower_bound(bytes);
        if (it == blocks.end() || it->first > bytes + bytes / 4 + (2u << 20)) return nullptr;
        void* p = it->second;
        cached -= it->first;
        blocks.erase(it);
        return p;
    }
    void put(size_t bytes, void* p)
    {
        std::lock_guard<std::mutex> lock(mu);
        if (cached + bytes > limit) { (void)hipFree(p); return; }
        blocks.emplace(bytes, p);
        cached += bytes;
    }
    void trim()
    {
        std::lock_guard<std::mutex> lock(mu);
        for (auto& kv : blocks) (void)hipFree(kv.second);
        blocks.clear();
        cached = 0;
    }
    size_t cached_bytes()
    {
        std::lock_guard<std::mutex> lock(mu);
        return cached;
    }
};

// Device allocations of one owner (a context's model arrays, a batch's buffers); a failed
// allocation returns nullptr and is recorded as a null entry.  With a pool, released blocks go
// back to it once the device has finished every queued use of them.
struct Alloc {
    std::vector<void*> ptrs;
    std::vector<size_t> sizes;
    std::shared_ptr<DevPool> pool;
    template <typename T>
    T* get(size_t count)
    {
        if (count == 0) count = 1;
        size_t bytes = count * sizeof(T);
        bytes = bytes >= (2u << 20) ? (bytes + (2u << 20) - 1) / (2u << 20) * (2u << 20) : (bytes + 255) / 256 * 256;
        void* p = pool ? pool->take(bytes) : nullptr;
        if (!p && hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            if (pool) {
                pool->trim();
                if (hipMalloc(&p, bytes) != hipSuccess) { (void)hipGetLastError(); p = nullptr; }
            }
        }
        ptrs.push_back(p);
        sizes.push_back(bytes);
        return (T*)p;
    }
    void release()
    {
        if (pool && !ptrs.empty()) (void)hipDeviceSynchronize();
        for (size_t i = 0; i < ptrs.size(); ++i) {
            if (!ptrs[i]) continue;
            if (pool) pool->put(sizes[i], ptrs[i]);
            else (void)hipFree(ptrs[i]);
        }
        ptrs.clear();
        sizes.clear();
    }
};

struct tdi_ctx {
    int device = 0;
    tdi_fit_options opts{};
    Layout L{};
    Layout Lx{};       // L with the light appended when it is not an unknown (light sub-fits)
    ModelGPU m{};
    ModelGPU m_fine{};  // m with the finetune fits' expression bounds (finetune_coef times wider)
    ModelGPU m_joint{}; // m without the albedo and identity bounds (the joint fit adds them once)
    float taps[15];
    Alloc mem;
    std::shared_ptr<DevPool> pool;   // blocks for this context's batches
};

struct Workspace {
    int C = 0;                 // frame slots
    long long cap = 0;         // crop pixels
    FrameGPU* frames = nullptr;
    double* P0 = nullptr;      // [C][V][3] model-space vertices (FP64)
    float* P = nullptr;
    float4* vproj = nullptr;
    uint64_t* keys = nullptr;
    int *cum = nullptr, *vis = nullptr;
    float* bary = nullptr;
    float *img = nullptr, *tmp = nullptr, *blurred = nullptr;
    float4* pixg = nullptr;
    int4* pnb = nullptr;
    float* pcoef = nullptr;
    int4* pvert = nullptr;       // per visible position: vertices of its triangle, crop index
    float2* pbary = nullptr;     // per visible position: (alpha, beta)
    float* dI = nullptr;
    float *theta = nullptr, *theta_try = nullptr, *delta = nullptr, *g = nullptr;
    double* H = nullptr;       // [C][K][K] normal matrices (FP64)
    double *obj = nullptr, *objb = nullptr;
    double* objp = nullptr;    // [C][256] per-workgroup partials of the data objective
    int *npix = nullptr, *status = nullptr, *face_ok = nullptr, *mask = nullptr, *searching = nullptr,
        *evalmask = nullptr;
    int *act_m = nullptr, *act_e = nullptr;   // frame lists of mask / evalmask ([0] = count)
    int *mask_r = nullptr, *mask_u = nullptr;  // active frames with a stale / reusable render
    int *act_r = nullptr, *act_u = nullptr;    // their frame lists
    int* tcount = nullptr;                     // [C][tmax] resolve tile counts
    int tmax = 0;
    double* rt = nullptr;                      // [C][12] FP64 rotation (column-major) + translation
    int4* box = nullptr;                       // [C] visible bounding box of the last render
    FitState* st = nullptr;
    float* partials = nullptr;
    float* lmk = nullptr;      // [C][6][51]
    float* alb = nullptr;      // [C][V] vertex albedos of the current render (fitted albedo only)
    double* lpack = nullptr;   // [C][16] the light sub-fits' per-frame results, packed
    double* idshape = nullptr; // [V][3] the shared identity shape of an identity-stage evaluation
    int S = 1;                 // J^T J splits per frame (the heuristic's, or the autotuner's)
    int S_cap = 1;             // splits the partials buffer holds
    int di_area = 0;           // the area that sizes k_dI's workgroups per frame (0: the mean crop)
    // chunk bookkeeping (host)
    std::vector<FrameGPU> hframes;
    std::vector<int> hlist;    // staging of a frame list (count, then slots)
    int n_cur = 0;
    int max_area = 0;
};

struct tdi_batch {
    tdi_ctx* ctx = nullptr;
    ModelGPU m{};                     // the context's model with this batch's vertex albedos (alb)
    int n = 0;
    long long total = 0;
    int max_area = 0;
    std::vector<FrameGPU> hframes;    // obs_off = batch offsets
    float2* obs_g = nullptr;
    float* lmk = nullptr;             // [n][6][51]
    Workspace ws;
    Alloc mem;
    hipGraphExec_t graph = nullptr;
    hipGraphExec_t graph_fine = nullptr;   // the finetune fit's schedule
    int finetune = 0;                      // set while a finetune fit runs: bounds and pose hold
    int joint = 0;                         // set while the identity stage's joint fit evaluates
    // scratch of the normal equations above 111 unknowns (allocated on first use): the Jacobian
    // rows U, the J^T J partials and the barrier's landmark-derivative rows
    float* big_U = nullptr;
    double* big_partials = nullptr;
    double* big_D = nullptr;
    int big_S = 0;
    hipStream_t stream = nullptr;     // capture-capable stream owned by the batch (default priority)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

extern "C" {

const char* tdi_last_error(void) { return g_err.c_str(); }

void tdi_default_options(tdi_fit_options* o)
{
    std::memset(o, 0, sizeof(*o));
    o->fit_exp = 1;
    o->max_iters = 100;
    o->max_ls = 100;
    o->ntmp_cap = 0;
    o->use_landmarks = 1;
    o->niters_success = 6;
    o->data_weight = 25.f;
    o->armijo_alpha = 0.4f;
    o->armijo_beta = 0.8f;
    o->t_min = 0.075f;
    o->decrement_tol = 3e-4f;
    o->blur_sigma = 0.75f;
    o->ref_face_size = 0.f;
    o->mixed_flags = TDI_MIXED_SYNTH_MFMA | TDI_MIXED_JTJ_BF16 | TDI_MIXED_CHOL_F32;
    o->finetune_coef = 1.5f;
}

int32_t tdi_num_unknowns(const tdi_model_desc* m, const tdi_fit_options* o) { return make_layout(m, o).K; }

int tdi_blur_taps(float sigma, float taps[15]) { return taps_for(sigma, taps); }

int tdi_ctx_create(const tdi_model_desc* md, const tdi_fit_options* opts, int device, tdi_ctx** out)
{
    if (!md || !opts || !out) { set_error("null argument"); return TDI_ERR_ARG; }
    if (md->n_vertices <= 0 || md->n_triangles <= 0 || !md->mean_shape || !md->triangles || !md->landmark_ids ||
        !md->tex_mean) {
        set_error("model: empty mesh or missing arrays");
        return TDI_ERR_ARG;
    }
    if (!(opts->finetune_coef > 0.f)) {
        set_error("finetune_coef %g must be positive", opts->finetune_coef);
        return TDI_ERR_ARG;
    }
    if (opts->fit_light < 0 || opts->fit_light > TDI_LIGHT_SH2) {
        set_error("fit_light %d: 0, TDI_LIGHT_POINT or TDI_LIGHT_SH2", opts->fit_light);
        return TDI_ERR_ARG;
    }
    if ((opts->fit_exp && (md->k_exp <= 0 || !md->exp_basis || !md->exp_lb || !md->exp_ub)) ||
        (opts->fit_id && (md->k_id <= 0 || !md->id_basis || !md->id_lb || !md->id_ub)) ||
        (opts->fit_tex && (md->k_tex <= 0 || !md->tex_basis || !md->tex_lb || !md->tex_ub))) {
        set_error("model: fitted block without basis or bounds");
        return TDI_ERR_ARG;
    }
    for (int i = 0; i < NLMK; ++i)
        if (md->landmark_ids[i] < 0 || md->landmark_ids[i] >= md->n_vertices) {
            set_error("landmark id %d out of range", md->landmark_ids[i]);
            return TDI_ERR_ARG;
        }
    for (long long i = 0; i < 3LL * md->n_triangles; ++i)
        if (md->triangles[i] < 0 || md->triangles[i] >= md->n_vertices) {
            set_error("triangle index out of range at %lld", i);
            return TDI_ERR_ARG;
        }
    tdi_ctx* c = new tdi_ctx;
    c->device = device;
    c->opts = *opts;
    c->L = make_layout(md, opts);
    if (taps_for(opts->blur_sigma, c->taps) != TDI_OK) {
        delete c;
        set_error("blur_sigma %g has no reference table (0.5, 0.75, 1, 2)", opts->blur_sigma);
        return TDI_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        set_error("hipSetDevice(%d) failed", device);
        return TDI_ERR_HIP;
    }
    const int V = md->n_vertices, NT = md->n_triangles;
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* d = c->mem.get<char>(bytes);
        if (d && src) (void)hipMemcpy(d, src, bytes, hipMemcpyHostToDevice);
        return d;
    };
    // bases with rows padded to a multiple of 16 columns (zeros), see ModelGPU
    auto up_rows = [&](const float* src, size_t rows, int cols, int ld) -> const float* {
        std::vector<float> h(rows * (size_t)ld, 0.f);
        for (size_t r = 0; r < rows; ++r) std::memcpy(h.data() + r * ld, src + r * cols, sizeof(float) * cols);
        return (const float*)up(h.data(), sizeof(float) * h.size());
    };
    auto ld16 = [](int k) { return (k + 15) & ~15; };
    ModelGPU& m = c->m;
    m.V = V; m.NT = NT; m.k_id = md->k_id; m.k_tex = md->k_tex; m.k_exp = md->k_exp;
    m.id_ld = ld16(md->k_id); m.exp_ld = ld16(md->k_exp); m.tex_ld = ld16(md->k_tex);
    m.mean = (const float*)up(md->mean_shape, sizeof(float) * 3 * V);
    m.exp_basis = md->k_exp ? up_rows(md->exp_basis, 3 * (size_t)V, md->k_exp, m.exp_ld) : nullptr;
    m.id_basis = md->k_id ? up_rows(md->id_basis, 3 * (size_t)V, md->k_id, m.id_ld) : nullptr;
    m.tex_mean = (const float*)up(md->tex_mean, sizeof(float) * V);
    m.tex_basis = md->k_tex ? up_rows(md->tex_basis, (size_t)V, md->k_tex, m.tex_ld) : nullptr;
    m.tris = (const int*)up(md->triangles, sizeof(int) * 3 * (size_t)NT);
    m.lmk_ids = (const int*)up(md->landmark_ids, sizeof(int) * NLMK);
    const Layout& L = c->L;
    c->Lx = L;
    if (!L.nlight) {
        c->Lx.nlight = 4;
        c->Lx.K = L.K + 4;
    }
    std::vector<float> lb(c->Lx.K, -INFINITY), ub(c->Lx.K, INFINITY);
    for (int k = 0; k < L.ntex; ++k) { lb[L.off_tex + k] = md->tex_lb[k]; ub[L.off_tex + k] = md->tex_ub[k]; }
    for (int k = 0; k < L.nid; ++k) { lb[L.off_id + k] = md->id_lb[k]; ub[L.off_id + k] = md->id_ub[k]; }
    for (int k = 0; k < L.nexp; ++k) { lb[L.off_exp + k] = md->exp_lb[k]; ub[L.off_exp + k] = md->exp_ub[k]; }
    m.lb = (const float*)up(lb.data(), sizeof(float) * c->Lx.K);
    m.ub = (const float*)up(ub.data(), sizeof(float) * c->Lx.K);
    // the finetune fits' bounds (logbarrier_initializer.cu:285-286): expression bounds scaled
    for (int k = 0; k < L.nexp; ++k) {
        lb[L.off_exp + k] *= opts->finetune_coef;
        ub[L.off_exp + k] *= opts->finetune_coef;
    }
    c->m_fine = m;
    c->m_fine.lb = (const float*)up(lb.data(), sizeof(float) * c->Lx.K);
    c->m_fine.ub = (const float*)up(ub.data(), sizeof(float) * c->Lx.K);
    // the identity stage's frames: regular expression bounds, no albedo or identity bounds
    for (int k = 0; k < L.nexp; ++k) { lb[L.off_exp + k] = md->exp_lb[k]; ub[L.off_exp + k] = md->exp_ub[k]; }
    for (int k = 0; k < L.ntex + L.nid; ++k) { lb[k] = -INFINITY; ub[k] = INFINITY; }
    c->m_joint = m;
    c->m_joint.lb = (const float*)up(lb.data(), sizeof(float) * c->Lx.K);
    c->m_joint.ub = (const float*)up(ub.data(), sizeof(float) * c->Lx.K);
    for (void* p : c->mem.ptrs)
        if (!p) { c->mem.release(); delete c; set_error("hipMalloc failed for the model"); return TDI_ERR_NOMEM; }
    if (hipDeviceSynchronize() != hipSuccess) { c->mem.release(); delete c; set_error("upload failed"); return TDI_ERR_HIP; }
    *out = c;
    return TDI_OK;
}

int tdi_ctx_release_cached(tdi_ctx* c)
{
    if (!c) { set_error("tdi_ctx_release_cached: null context"); return TDI_ERR_ARG; }
    (void)hipSetDevice(c->device);
    if (c->pool) c->pool->trim();
    return TDI_OK;
}

int tdi_ctx_destroy(tdi_ctx* c)
{
    if (!c) return TDI_OK;
    (void)hipSetDevice(c->device);
    c->mem.release();
    c->pool.reset();   // freed here unless a batch still holds it
    delete c;
    return TDI_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------- autotuned launch shapes
// Choices of tdi_batch_autotune, per (unknowns, frame slots, crop pixels, largest crop, mixed
// flags): a batch created later in the process with the same key starts from them.
using TuneKey = std::tuple<int, int, long long, int, int>;
struct TuneChoice {
    int S, di_area;
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tuned;

// ---------------------------------------------------------------------------- workspace
static int ws_alloc(tdi_batch* b, int C, long long cap)
{
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    Alloc& A = b->mem;
    const int V = c->m.V, K = c->Lx.K;   // room for the light unknowns of the light sub-fits
    w.C = C;
    w.cap = cap;
    w.max_area = b->max_area;
    w.frames = A.get<FrameGPU>(C);
    w.P0 = A.get<double>((size_t)C * V * 3);
    w.P = A.get<float>((size_t)C * V * 3);
    w.vproj = A.get<float4>((size_t)C * V);
    w.keys = A.get<uint64_t>(cap);
    w.cum = A.get<int>(cap);
    w.vis = A.get<int>(cap);
    w.bary = A.get<float>(2 * cap);
    w.img = A.get<float>(cap);
    w.tmp = A.get<float>(cap);
    w.blurred = A.get<float>(cap);
    w.pixg = A.get<float4>(cap);
    w.pnb = A.get<int4>(cap);
    w.pcoef = A.get<float>(cap);
    w.pvert = A.get<int4>(cap);
    w.pbary = A.get<float2>(cap);
    w.dI = A.get<float>((size_t)cap * K);
    w.theta = A.get<float>((size_t)C * K);
    w.theta_try = A.get<float>((size_t)C * K);
    w.delta = A.get<float>((size_t)C * K);
    w.H = A.get<double>((size_t)C * K * K);
    w.g = A.get<float>((size_t)C * K);
    w.obj = A.get<double>(C);
    w.objb = A.get<double>(C);
    w.objp = A.get<double>((size_t)C * 256);
    w.npix = A.get<int>(C);
    w.status = A.get<int>(C);
    w.face_ok = A.get<int>(C);
    w.mask = A.get<int>(C);
    w.mask_r = A.get<int>(C);
    w.mask_u = A.get<int>(C);
    w.act_r = A.get<int>(C + 1);
    w.act_u = A.get<int>(C + 1);
    w.searching = A.get<int>(C);
    w.evalmask = A.get<int>(C);
    w.act_m = A.get<int>(C + 1);
    w.act_e = A.get<int>(C + 1);
    w.tmax = resolve_tiles(b->max_area);
    w.tcount = A.get<int>((size_t)C * (w.tmax > 0 ? w.tmax : 1));
    w.rt = A.get<double>((size_t)C * 12);
    w.box = A.get<int4>(C);
    w.st = A.get<FitState>(C);
    w.S = jtj_splits(C, b->max_area);
    // room for the autotuner's candidates up to twice the heuristic's split count
    w.S_cap = std::min(64, 2 * w.S);
    const size_t pf = std::max(jtj_partials_floats(C, w.S_cap, K, false), jtj_partials_floats(C, w.S_cap, K, true));
    w.partials = A.get<float>(pf);
    w.lmk = A.get<float>((size_t)C * 6 * NLMK);
    // per-frame vertex albedos when the albedo is fitted (read by the shading and the Jacobian)
    w.alb = c->L.ntex ? A.get<float>((size_t)C * V) : nullptr;
    w.lpack = A.get<double>((size_t)C * 16);
    w.idshape = c->L.nid ? A.get<double>((size_t)V * 3) : nullptr;
    for (void* p : A.ptrs)
        if (!p) { set_error("workspace allocation failed (C=%d, cap=%lld)", C, cap); return TDI_ERR_NOMEM; }
    return TDI_OK;
}

// Load frames [i0, i0+n) into the workspace: rebased frame descriptors and landmarks.
static int ws_load(tdi_batch* b, int i0, int n, hipStream_t s)
{
    Workspace& w = b->ws;
    if (n > w.C) { set_error("chunk of %d frames exceeds workspace (%d)", n, w.C); return TDI_ERR_ARG; }
    w.hframes.assign(w.C, FrameGPU{});
    long long off = 0;
    for (int i = 0; i < n; ++i) {
        FrameGPU f = b->hframes[i0 + i];
        f.pix_off = off;
        off += (long long)f.W * f.H;
        w.hframes[i] = f;
    }
    if (off > w.cap) { set_error("chunk needs %lld crop pixels, workspace has %lld", off, w.cap); return TDI_ERR_ARG; }
    // padding slots: a 1x1 crop at offset 0, never active
    for (int i = n; i < w.C; ++i) {
        FrameGPU f = b->hframes[i0];
        f.W = 1; f.H = 1; f.pix_off = 0;
        w.hframes[i] = f;
    }
    w.n_cur = n;
    TDI_HIP(hipMemcpyAsync(w.frames, w.hframes.data(), sizeof(FrameGPU) * w.C, hipMemcpyHostToDevice, s));
    TDI_HIP(hipMemcpyAsync(w.lmk, b->lmk + (size_t)i0 * 6 * NLMK, sizeof(float) * 6 * NLMK * n, hipMemcpyDeviceToDevice, s));
    // the keys left by the previous chunk lie anywhere in the new crops
    return launch_box_full(w.frames, w.C, w.box, s);
}

// Mean crop area of the workspace's chunk (its crop pixels over its frame slots).
static int mean_crop(const Workspace& w)
{
    return (int)(w.cap / (w.C > 0 ? w.C : 1));
}

// The area k_dI's grid is sized for: the autotuner's choice, else the mean crop.
static int dI_area(const Workspace& w)
{
    return w.di_area > 0 ? w.di_area : mean_crop(w);
}

static LmkArrays lmk_arrays(const Workspace& w)
{
    LmkArrays a;
    a.base = w.lmk;
    return a;
}

extern "C" int tdi_batch_create(tdi_ctx* c, int32_t n, const tdi_frame_desc* fd, tdi_batch** out)
{
    if (!c || !fd || !out || n <= 0) { set_error("null argument or empty batch"); return TDI_ERR_ARG; }
    (void)hipSetDevice(c->device);
    tdi_batch* b = new tdi_batch;
    b->ctx = c;
    b->n = n;
    long long off = 0;
    int maxA = 0;
    for (int i = 0; i < n; ++i) {
        const tdi_frame_desc& d = fd[i];
        if (d.width <= 0 || d.height <= 0 || !d.obs || !d.lmk_x || !d.lmk_y || !d.box_lx || !d.box_ux || !d.box_ly ||
            !d.box_uy) {
            delete b;
            set_error("frame %d: empty crop or missing arrays", i);
            return TDI_ERR_ARG;
        }
        if ((long long)d.width * d.height > (1LL << 30)) { delete b; set_error("frame %d: crop too large", i); return TDI_ERR_ARG; }
        FrameGPU f{};
        f.phix = d.phix; f.phiy = d.phiy; f.cx = d.cx; f.cy = d.cy;
        f.x0 = d.x0; f.y0 = d.y0; f.W = d.width; f.H = d.height;
        f.face_size = d.face_size;
        for (int a = 0; a < 3; ++a) f.lam[a] = d.lambda[a];
        f.Lint = d.Lintensity;
        f.obs_off = off;
        f.pix_off = off;
        off += (long long)d.width * d.height;
        maxA = std::max(maxA, d.width * d.height);
        b->hframes.push_back(f);
    }
    b->total = off;
    b->max_area = maxA;
    hipStream_t s = 0;
    // observations: upload, blur, normalised gradients (once per batch)
    if (!c->pool) c->pool = std::make_shared<DevPool>();
    b->mem.pool = c->pool;
    b->obs_g = b->mem.get<float2>(off);
    b->lmk = b->mem.get<float>((size_t)n * 6 * NLMK);
    FrameGPU* dfr = b->mem.get<FrameGPU>(n);
    if (!b->obs_g || !b->lmk || !dfr) { b->mem.release(); delete b; set_error("batch allocation failed"); return TDI_ERR_NOMEM; }
    {
        Alloc stage;
        stage.pool = c->pool;
        float* raw = stage.get<float>(off);
        float* tmp = stage.get<float>(off);
        float* blr = stage.get<float>(off);
        if (!raw || !tmp || !blr) {
            stage.release();
            b->mem.release();
            delete b;
            set_error("observation staging allocation failed");
            return TDI_ERR_NOMEM;
        }
        // an observation is host memory or, as tdi_prepare_crops leaves it, device memory
        int rc = TDI_OK;
        for (int i = 0; i < n && !rc; ++i) {
            hipPointerAttribute_t at{};
            hipMemcpyKind kind = hipMemcpyHostToDevice;
            if (hipPointerGetAttributes(&at, fd[i].obs) == hipSuccess && at.type == hipMemoryTypeDevice)
                kind = hipMemcpyDeviceToDevice;
            else
                (void)hipGetLastError();
            if (hipMemcpy(raw + b->hframes[i].obs_off, fd[i].obs, sizeof(float) * fd[i].width * fd[i].height, kind) !=
                hipSuccess) {
                set_error("frame %d: observation copy failed", i);
                rc = TDI_ERR_HIP;
            }
        }
        (void)hipMemcpy(dfr, b->hframes.data(), sizeof(FrameGPU) * n, hipMemcpyHostToDevice);
        if (!rc) rc = launch_obs_prepare(dfr, n, c->taps, raw, tmp, blr, b->obs_g, maxA, s);
        (void)hipDeviceSynchronize();
        stage.release();
        if (rc) { b->mem.release(); delete b; return rc; }
    }
    {
        std::vector<float> h((size_t)n * 6 * NLMK);
        const size_t S = (size_t)n * NLMK;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < NLMK; ++j) {
                h[0 * S + i * NLMK + j] = fd[i].lmk_x[j];
                h[1 * S + i * NLMK + j] = fd[i].lmk_y[j];
                h[2 * S + i * NLMK + j] = fd[i].box_lx[j];
                h[3 * S + i * NLMK + j] = fd[i].box_ux[j];
                h[4 * S + i * NLMK + j] = fd[i].box_ly[j];
                h[5 * S + i * NLMK + j] = fd[i].box_uy[j];
            }
        // per-frame layout [n][6][51] so that a chunk is one contiguous copy
        std::vector<float> hp((size_t)n * 6 * NLMK);
        for (int i = 0; i < n; ++i)
            for (int q = 0; q < 6; ++q)
                for (int j = 0; j < NLMK; ++j) hp[((size_t)i * 6 + q) * NLMK + j] = h[q * S + i * NLMK + j];
        (void)hipMemcpy(b->lmk, hp.data(), sizeof(float) * hp.size(), hipMemcpyHostToDevice);
    }
    // workspace: the whole batch if it fits the budget, else chunks of C consecutive frames;
    // the per-pixel buffers are sized by the largest actual crop-pixel sum over the chunks.
    // Budget: TDI_WORKSPACE_GB, default half of the device memory free at this point.
    const char* env = getenv("TDI_WORKSPACE_GB");
    double budget_gb = 48.0;
    if (env) {
        budget_gb = atof(env);
    } else {
        size_t fb = 0, tb = 0;
        // the context's cached blocks count as free: the workspace reuses them
        if (hipMemGetInfo(&fb, &tb) == hipSuccess && fb > 0) budget_gb = 0.5 * (double)(fb + c->pool->cached_bytes()) / 1e9;
    }
    const int V = c->m.V, K = c->Lx.K;
    const double per_frame = (double)V * 40 + 64.0 * K * K + 8.0 * resolve_tiles(maxA);
    const double per_pixel = 108.0 + 4.0 * K;
    auto chunk_cap = [&](int Cc) {
        long long best = 0;
        for (int i0 = 0; i0 < n; i0 += Cc) {
            const int i1 = std::min(n, i0 + Cc);
            const long long e = (i1 < n ? b->hframes[i1].obs_off : off) - b->hframes[i0].obs_off;
            best = std::max(best, e);
        }
        return best;
    };
    int C = n;
    long long cap = off;
    while (C > 1 && (per_frame * C + per_pixel * (double)cap) > budget_gb * 1e9) {
        C = (C + 1) / 2;
        cap = chunk_cap(C);
    }
    int rc = ws_alloc(b, C, cap);
    if (rc) { b->mem.release(); delete b; return rc; }
    b->m = c->m;
    b->m.alb = b->ws.alb;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        const auto it = g_tuned.find(TuneKey(c->L.K, C, cap, b->max_area, c->opts.mixed_flags));
        if (it != g_tuned.end() && it->second.S <= b->ws.S_cap) {
            b->ws.S = it->second.S;
            b->ws.di_area = it->second.di_area;
        }
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
        b->mem.release();
        delete b;
        set_error("stream/event creation failed");
        return TDI_ERR_HIP;
    }
    *out = b;
    return TDI_OK;
}

extern "C" int tdi_batch_destroy(tdi_batch* b)
{
    if (!b) return TDI_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_fine) (void)hipGraphExecDestroy(b->graph_fine);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    (void)hipFree(b->big_U);
    (void)hipFree(b->big_partials);
    (void)hipFree(b->big_D);
    b->mem.release();
    delete b;
    return TDI_OK;
}

// ---------------------------------------------------------------------------- evaluation
// The model whose coefficient bounds the barrier uses: the finetune fits' widened expression
// bounds while a finetune fit runs (model_fitter.cu:409-410), the bounds without the shared
// albedo and identity blocks while the joint fit evaluates its frames, else the regular ones.
static const ModelGPU& bounds_model(const tdi_batch* b)
{
    return b->joint ? b->ctx->m_joint : b->finetune ? b->ctx->m_fine : b->ctx->m;
}

constexpr int SMALL_JTJ_MAXK = 111;   // the LDS-staged J^T J kernels and k_barrier's LDS rows

// Scratch of the normal equations above SMALL_JTJ_MAXK unknowns, allocated once per batch.
static int big_scratch(tdi_batch* b)
{
    if (b->big_U) return TDI_OK;
    const Workspace& w = b->ws;
    const int K = b->ctx->Lx.K;      // the light sub-fits evaluate the layout with the light appended
    int S = w.max_area / 512 + 1;    // pixel splits per frame: about 512 pixels of a full crop or more each
    if (S > 16) S = 16;
    const size_t nu = (size_t)w.cap * jtj_big_row_stride(K), np = jtj_big_partials_doubles(w.C, S, K);
    const size_t nd = (size_t)w.C * 2 * NLMK * K;
    if (hipMalloc((void**)&b->big_U, sizeof(float) * nu) != hipSuccess ||
        hipMalloc((void**)&b->big_partials, sizeof(double) * np) != hipSuccess ||
        hipMalloc((void**)&b->big_D, sizeof(double) * nd) != hipSuccess) {
        (void)hipFree(b->big_U);
        (void)hipFree(b->big_partials);
        (void)hipFree(b->big_D);
        b->big_U = nullptr;
        b->big_partials = nullptr;
        b->big_D = nullptr;
        set_error("normal equations at K=%d: %zu bytes of scratch not available", K,
                  sizeof(float) * nu + sizeof(double) * (np + nd));
        return TDI_ERR_NOMEM;
    }
    b->big_S = S;
    return TDI_OK;
}

// Objective at theta for the frames in `mask`: synthesis, face check, raster, resolve,
// barrier objective, shading, blur, gradients (Renderer::render + evaluate_objective_function).
// full_crop: blur the whole crops (the rendered image is an output), else only the region
// the normalised gradients read.
// dverts (optional, device [fl.n][V][3]): per-vertex model-space offsets added after the synthesis
// (all frames of the chunk listed, fl.act == nullptr).
static int eval_render(tdi_batch* b, FrameList fl, const float* theta, int mixed, hipStream_t s,
                       const Layout* Lp = nullptr, bool full_crop = false, const float* dverts = nullptr)
{
    tdi_ctx* c = b->ctx;
    const Layout& L = Lp ? *Lp : c->L;
    Workspace& w = b->ws;
    int rc;
    if (b->joint && L.nid && w.idshape)   // the identity stage's frames share alpha
        rc = launch_synthesize_shared_id(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, w.idshape, s);
    else if (mixed & TDI_MIXED_SYNTH_MFMA)
        rc = launch_synthesize_gemm(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, s);
    else
        rc = launch_synthesize(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, s);
    if (rc) return rc;
    if (dverts && (rc = launch_add_offsets(c->m.V, w.frames, fl.n, w.rt, dverts, w.P0, w.vproj, s))) return rc;
    if ((rc = launch_face_check(c->m.V, c->m.lmk_ids, w.vproj, c->opts.ref_face_size, fl, w.face_ok, s))) return rc;
    if ((rc = launch_clear_keys_frames(w.frames, fl, w.keys, w.box, w.max_area, s))) return rc;
    if ((rc = launch_raster(c->m, w.frames, fl, w.vproj, c->opts.ntmp_cap, w.keys, w.box, s))) return rc;
    if ((rc = launch_resolve(c->m, w.frames, fl, w.vproj, w.keys, w.cum, w.vis, w.npix, nullptr, w.pvert, w.pbary, w.max_area, w.tcount, w.tmax, w.box, s)))
        return rc;
    if ((rc = launch_barrier(bounds_model(b), L, w.frames, fl, theta, lmk_arrays(w), c->opts.use_landmarks, 0, w.obj,
                             w.objb, w.H, w.g, s)))
        return rc;
    if ((rc = launch_shade(b->m, L, w.frames, fl, theta, w.P0, w.pvert, w.pbary, w.npix,
                           full_crop ? nullptr : w.box, w.img, w.max_area, s)))
        return rc;
    if ((rc = launch_blur(w.frames, fl, c->taps, full_crop ? nullptr : w.box, w.img, w.tmp, w.blurred, w.max_area, s))) return rc;
    return launch_grads(w.frames, fl, w.vis, w.npix, w.cum, w.blurred, b->obs_g, c->opts.data_weight, w.pixg, w.pnb,
                        w.pcoef, w.objp, w.obj, w.max_area, s);
}

// Objective only, for the line search (the F(theta + t delta) of fit_3DMM_shape_rigid's Armijo
// loop, model_fitter.cu:219-260): the same synthesis, raster and per-pixel terms as
// eval_render, but shading and the objective work from the keys over the visible box, so
// there is no visible list, no scatter and no per-pixel Jacobian records.
static int eval_objective(tdi_batch* b, FrameList fl, const float* theta, int mixed, hipStream_t s,
                          const Layout* Lp = nullptr)
{
    tdi_ctx* c = b->ctx;
    const Layout& L = Lp ? *Lp : c->L;
    Workspace& w = b->ws;
    int rc;
    if (b->joint && L.nid && w.idshape)   // the identity stage's frames share alpha
        rc = launch_synthesize_shared_id(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, w.idshape, s);
    else if (mixed & TDI_MIXED_SYNTH_MFMA)
        rc = launch_synthesize_gemm(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, s);
    else
        rc = launch_synthesize(b->m, L, w.frames, fl, theta, w.rt, w.P0, nullptr, w.vproj, s);
    if (rc) return rc;
    if ((rc = launch_face_check(c->m.V, c->m.lmk_ids, w.vproj, c->opts.ref_face_size, fl, w.face_ok, s))) return rc;
    if ((rc = launch_clear_keys_frames(w.frames, fl, w.keys, w.box, w.max_area, s))) return rc;
    if ((rc = launch_raster(c->m, w.frames, fl, w.vproj, c->opts.ntmp_cap, w.keys, w.box, s))) return rc;
    if ((rc = launch_visible_box(w.frames, fl, w.keys, w.max_area, w.tcount, w.tmax, w.box, s))) return rc;
    if ((rc = launch_barrier(bounds_model(b), L, w.frames, fl, theta, lmk_arrays(w), c->opts.use_landmarks, 0, w.obj,
                             w.objb, w.H, w.g, s)))
        return rc;
    if ((rc = launch_shade_keys(b->m, L, w.frames, fl, theta, w.P0, w.vproj, w.keys, w.box, w.img, s))) return rc;
    if ((rc = launch_blur(w.frames, fl, c->taps, w.box, w.img, w.tmp, w.blurred, w.max_area, s))) return rc;
    return launch_objective(w.frames, fl, w.keys, w.box, w.blurred, b->obs_g, c->opts.data_weight, w.objp, w.obj, s);
}

static int launch_normal_eq(tdi_batch* b, FrameList fl, int mixed, hipStream_t s, const Layout* Lp = nullptr)
{
    tdi_ctx* c = b->ctx;
    const Layout& L = Lp ? *Lp : c->L;
    Workspace& w = b->ws;
    if (mixed & TDI_MIXED_JTJ_BF16)
        return launch_jtj_mfma(L, w.frames, fl, w.npix, w.pnb, w.pcoef, w.pixg, w.dI, c->opts.data_weight, w.H,
                               w.g, w.partials, w.S, (mixed & TDI_MIXED_JTR_BF16) ? 1 : 0,
                               (mixed & TDI_MIXED_TWO_TERM) ? 1 : 0, s);
    return launch_jtj_f32(L, w.frames, fl, w.npix, w.pnb, w.pcoef, w.pixg, w.dI, c->opts.data_weight, w.H, w.g,
                          w.partials, w.S, s);
}

// The visible list, barrier value, per-pixel records and objective of frames whose render of
// theta is already in the workspace (synthesis, keys, visible box, shaded and blurred images).
static int eval_reused(tdi_batch* b, FrameList fl, const float* theta, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    const Layout& L = c->L;
    Workspace& w = b->ws;
    int rc;
    if ((rc = launch_resolve(c->m, w.frames, fl, w.vproj, w.keys, w.cum, w.vis, w.npix, nullptr, w.pvert, w.pbary,
                             w.max_area, w.tcount, w.tmax, w.box, s)))
        return rc;
    if ((rc = launch_barrier(bounds_model(b), L, w.frames, fl, theta, lmk_arrays(w), c->opts.use_landmarks, 0, w.obj,
                             w.objb, w.H, w.g, s)))
        return rc;
    return launch_grads(w.frames, fl, w.vis, w.npix, w.cum, w.blurred, b->obs_g, c->opts.data_weight, w.pixg, w.pnb,
                        w.pcoef, w.objp, w.obj, w.max_area, s);
}

// dI, the normal equations and the barrier's gradient and Hessian terms, after a render.
static int eval_jacobian(tdi_batch* b, FrameList fl, const float* theta, int mixed, hipStream_t s,
                         const Layout* Lp = nullptr)
{
    tdi_ctx* c = b->ctx;
    const Layout& L = Lp ? *Lp : c->L;
    Workspace& w = b->ws;
    int rc;
    if ((rc = launch_dI(b->m, L, w.frames, fl, theta, w.P0, w.pvert, w.pbary, w.npix, w.pixg, w.dI,
                        dI_area(w), s)))
        return rc;
    if (L.K > SMALL_JTJ_MAXK) {
        // hundreds of unknowns (the identity stage): J^T J on FP64 MFMA over 64 x 64 tiles and
        // the barrier's rows in global scratch
        if ((rc = big_scratch(b))) return rc;
        if ((rc = launch_jtj_big(L, w.frames, fl, w.npix, w.pnb, w.pcoef, w.pixg, w.dI, c->opts.data_weight, w.H, w.g,
                                 b->big_U, b->big_partials, b->big_S, w.max_area, s)))
            return rc;
        return launch_barrier_big(bounds_model(b), L, w.frames, fl, theta, lmk_arrays(w), c->opts.use_landmarks,
                                  b->big_D, w.H, w.g, s);
    }
    if ((rc = launch_normal_eq(b, fl, mixed, s, &L))) return rc;
    return launch_barrier(bounds_model(b), L, w.frames, fl, theta, lmk_arrays(w), c->opts.use_landmarks, 1, w.obj,
                          w.objb, w.H, w.g, s);
}

// Objective, gradient and Gauss-Newton matrix at theta (get_obj_hessian_and_gradient_multiframe).
static int eval_full(tdi_batch* b, FrameList fl, const float* theta, int mixed, hipStream_t s,
                     const Layout* Lp = nullptr)
{
    int rc = eval_render(b, fl, theta, mixed, s, Lp);
    if (rc) return rc;
    return eval_jacobian(b, fl, theta, mixed, s, Lp);
}

// active mask for real frames of the current chunk
static int set_mask_all(tdi_batch* b, hipStream_t s)
{
    std::vector<int> h(b->ws.C, 0);
    for (int i = 0; i < b->ws.n_cur; ++i) h[i] = 1;
    TDI_HIP(hipMemcpyAsync(b->ws.mask, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
    TDI_HIP(hipStreamSynchronize(s));
    return TDI_OK;
}

extern "C" {

int tdi_workspace_frames(tdi_batch* b) { return b->ws.C; }

int tdi_rotation(int32_t n, const float* u, float* R, float* dR, void* stream)
{
    return launch_misc_rotation(n, u, R, dR, (hipStream_t)stream);
}

// The building blocks run chunk by chunk over the workspace; outputs are in batch order.
int tdi_synthesize(tdi_batch* b, const float* theta, float* xyz, float* xyp, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    tdi_ctx* c = b->ctx;
    const int K = c->L.K, V = c->m.V;
    for (int i0 = 0; i0 < b->n; i0 += b->ws.C) {
        const int n = std::min(b->ws.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        if ((rc = launch_synthesize(b->m, c->L, b->ws.frames, all_frames(n), theta + (size_t)i0 * K, b->ws.rt, b->ws.P0, b->ws.P,
                                    b->ws.vproj, s)))
            return rc;
        if ((rc = launch_misc_unpack(n, V, b->ws.P, b->ws.vproj, xyz + (size_t)i0 * V * 3, xyp + (size_t)i0 * V * 2, s)))
            return rc;
    }
    return TDI_OK;
}

int tdi_rasterize(tdi_batch* b, const float* xyz, const float* xyp, uint64_t* keys, float* bary, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    const int V = c->m.V;
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        const long long off = b->hframes[i0].obs_off;
        const long long len = (i0 + n < b->n ? b->hframes[i0 + n].obs_off : b->total) - off;
        if ((rc = launch_misc_pack(n, V, xyz + (size_t)i0 * V * 3, xyp + (size_t)i0 * V * 2, w.vproj, s))) return rc;
        if ((rc = launch_clear_keys(w.keys, len, s))) return rc;
        if ((rc = launch_misc_fill_f32(w.bary, 2 * len, 0.f, s))) return rc;
        if ((rc = launch_raster(c->m, w.frames, all_frames(n), w.vproj, c->opts.ntmp_cap, w.keys, w.box, s))) return rc;
        if ((rc = launch_resolve(c->m, w.frames, all_frames(n), w.vproj, w.keys, w.cum, w.vis, w.npix, w.bary, w.pvert, w.pbary, w.max_area, w.tcount, w.tmax, w.box, s)))
            return rc;
        if (keys) TDI_HIP(hipMemcpyAsync(keys + off, w.keys, sizeof(uint64_t) * len, hipMemcpyDeviceToDevice, s));
        if (bary) TDI_HIP(hipMemcpyAsync(bary + 2 * off, w.bary, sizeof(float) * 2 * len, hipMemcpyDeviceToDevice, s));
    }
    return TDI_OK;
}

int tdi_render(tdi_batch* b, const float* theta, float* img, float* blurred, float* obj, int32_t* npix,
               void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    Workspace& w = b->ws;
    const int K = b->ctx->L.K;
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        const long long off = b->hframes[i0].obs_off;
        const long long len = (i0 + n < b->n ? b->hframes[i0 + n].obs_off : b->total) - off;
        if ((rc = eval_render(b, all_frames(n), theta + (size_t)i0 * K, 0, s, nullptr, true))) return rc;
        if (img) TDI_HIP(hipMemcpyAsync(img + off, w.img, sizeof(float) * len, hipMemcpyDeviceToDevice, s));
        if (blurred) TDI_HIP(hipMemcpyAsync(blurred + off, w.blurred, sizeof(float) * len, hipMemcpyDeviceToDevice, s));
        if (obj && (rc = launch_misc_f64_to_f32(w.obj, obj + i0, n, s))) return rc;
        if (npix) TDI_HIP(hipMemcpyAsync(npix + i0, w.npix, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    }
    return TDI_OK;
}

// The render loss and its adjoint, chunk by chunk; the render (rasterisation included) is
// recomputed from theta and dverts.
int tdi_render_loss_grad(tdi_batch* b, const float* theta, const float* dverts, const double* gloss, double* loss,
                         float* gtheta, float* gverts, void* stream)
{
    if (!b || !theta) { set_error("tdi_render_loss_grad: null argument"); return TDI_ERR_ARG; }
    hipStream_t s = (hipStream_t)stream;
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    const int K = c->L.K, V = c->m.V;
    if (!c->m.lmk_ids) { set_error("tdi_render_loss_grad: the model has no landmark ids"); return TDI_ERR_ARG; }
    const bool grad = gtheta || gverts;
    AdjScratch a{};
    a.cap = w.cap;
    const size_t C = (size_t)w.C;
    const size_t nd = (grad ? (size_t)w.cap + C * V * 3 + C * 15 + C * c->L.nexp + C * c->L.nid : 0) + C;
    double* blk = nullptr;
    if (hipMalloc((void**)&blk, sizeof(double) * nd) != hipSuccess) {
        set_error("tdi_render_loss_grad: scratch allocation of %zu bytes failed", sizeof(double) * nd);
        return TDI_ERR_NOMEM;
    }
    double* q = blk;
    a.loss = q; q += C;
    if (grad) {
        a.dB = q; q += w.cap;
        a.dP0 = q; q += C * V * 3;
        a.gpl = q; q += C * 15;
        a.gexp = q; q += C * c->L.nexp;
        a.gid = q;
    }
    int rc = TDI_OK;
    for (int i0 = 0; i0 < b->n && !rc; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        if ((rc = ws_load(b, i0, n, s))) break;
        const float* th = theta + (size_t)i0 * K;
        if ((rc = eval_render(b, all_frames(n), th, 0, s, nullptr, false, dverts ? dverts + (size_t)i0 * V * 3 : nullptr)))
            break;
        rc = launch_adjoint(c->m, c->L, w.frames, n, th, w.P0, w.pvert, w.pbary, w.vis, w.npix, w.pixg, w.lmk, c->taps,
                            w.max_area, w.img, w.tmp, w.blurred, a, gloss ? gloss + i0 : nullptr,
                            loss ? loss + i0 : nullptr, gtheta ? gtheta + (size_t)i0 * K : nullptr,
                            gverts ? gverts + (size_t)i0 * V * 3 : nullptr, s);
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) { set_error("tdi_render_loss_grad: %s", hipGetErrorString(e)); rc = TDI_ERR_HIP; }
    (void)hipFree(blk);
    return rc;
}

static int normal_eq(tdi_batch* b, const float* theta, float* H, float* g, float* obj, void* stream, bool mixed)
{
    hipStream_t s = (hipStream_t)stream;
    Workspace& w = b->ws;
    const int K = b->ctx->L.K;
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        if ((rc = eval_full(b, all_frames(n), theta + (size_t)i0 * K, mixed ? TDI_MIXED_JTJ_BF16 : 0, s))) return rc;
        if (H && (rc = launch_misc_f64_to_f32(w.H, H + (size_t)i0 * K * K, n * K * K, s))) return rc;
        if (g) TDI_HIP(hipMemcpyAsync(g + (size_t)i0 * K, w.g, sizeof(float) * n * K, hipMemcpyDeviceToDevice, s));
        if (obj && (rc = launch_misc_f64_to_f32(w.obj, obj + i0, n, s))) return rc;
    }
    return TDI_OK;
}

int tdi_normal_equations(tdi_batch* b, const float* theta, float* H, float* g, float* obj, void* stream)
{
    return normal_eq(b, theta, H, g, obj, stream, false);
}

int tdi_normal_equations_mixed(tdi_batch* b, const float* theta, float* H, float* g, float* obj, void* stream)
{
    return normal_eq(b, theta, H, g, obj, stream, true);
}

// ---------------------------------------------------------------------------- identity stage
// The joint fit of frames that share the albedo and identity coefficients (fit_to_multi_images'
// shape fit, fit_3DMM_shape_rigid over T frames, model_fitter.cu:121-375, with the frames'
// products summed as in get_obj_hessian_and_gradient_multiframe, model_fitter.cu:21-116).
// Per-frame layout [albedo | identity | expression | pose | light]; the shared block is the
// first Ks = ntex + nid unknowns; joint vector [shared | frame 0's rest | frame 1's rest | ...].
static void joint_sizes(const tdi_batch* b, int& Ks, long long& KJ)
{
    const Layout& L = b->ctx->L;
    Ks = L.ntex + L.nid;
    KJ = Ks + (long long)b->n * (L.K - Ks);
}

// Every frame evaluated at its theta (device [n][K]) without the shared bounds, then the joint
// matrix HJ and gradient gJ (device FP64) and the shared bounds' objective (*objs, device).
static int joint_eval(tdi_batch* b, const float* dtheta, double* HJ, double* gJ, double* objs, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    int Ks;
    long long KJ;
    joint_sizes(b, Ks, KJ);
    b->joint = 1;
    int rc = eval_full(b, all_frames(b->n), dtheta, 0, s);
    b->joint = 0;
    if (rc) return rc;
    if ((rc = launch_assemble_joint(b->n, c->L.K, Ks, w.H, w.g, HJ, gJ, s))) return rc;
    if ((rc = launch_shared_bounds(Ks, KJ, dtheta, c->m.lb, c->m.ub, 1, nullptr, HJ, gJ, s))) return rc;
    return launch_shared_bounds(Ks, KJ, dtheta, c->m.lb, c->m.ub, 0, objs, nullptr, nullptr, s);
}

static int joint_objective(tdi_batch* b, const float* dtheta, double* objs, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    int Ks;
    long long KJ;
    joint_sizes(b, Ks, KJ);
    b->joint = 1;
    int rc = eval_objective(b, all_frames(b->n), dtheta, 0, s);
    b->joint = 0;
    if (rc) return rc;
    return launch_shared_bounds(Ks, KJ, dtheta, c->m.lb, c->m.ub, 0, objs, nullptr, nullptr, s);
}

static int joint_check(tdi_batch* b, const char* who)
{
    if (b->n > b->ws.C) {
        set_error("%s: %d frames do not fit one workspace chunk (%d)", who, b->n, b->ws.C);
        return TDI_ERR_ARG;
    }
    return TDI_OK;
}

// Joint unknowns of the batch's frames (Ks + n (K - Ks)).
int64_t tdi_joint_unknowns(tdi_batch* b)
{
    if (!b) return 0;
    int Ks;
    long long KJ;
    joint_sizes(b, Ks, KJ);
    return KJ;
}

int tdi_normal_equations_joint(tdi_batch* b, const float* theta, double* HJ, double* gJ, double* obj, void* stream)
{
    if (!b || !theta || !HJ || !gJ) { set_error("tdi_normal_equations_joint: null argument"); return TDI_ERR_ARG; }
    int rc = joint_check(b, "tdi_normal_equations_joint");
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    Workspace& w = b->ws;
    if ((rc = ws_load(b, 0, b->n, s))) return rc;
    double* objs = nullptr;
    if (hipMalloc((void**)&objs, sizeof(double)) != hipSuccess) { set_error("joint: allocation failed"); return TDI_ERR_NOMEM; }
    rc = joint_eval(b, theta, HJ, gJ, objs, s);
    if (!rc && obj) {
        std::vector<double> h(b->n + 1);
        if (hipMemcpyAsync(h.data(), w.obj, sizeof(double) * b->n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(h.data() + b->n, objs, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = TDI_ERR_HIP;
        double sum = 0;
        for (double v : h) sum += v;
        if (!rc && (hipMemcpyAsync(obj, &sum, sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess))
            rc = TDI_ERR_HIP;
        if (rc) set_error("tdi_normal_equations_joint: copy failed");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(objs);
    return rc;
}

int tdi_fit_multiframe(tdi_batch* b, float* theta_host, tdi_fit_report* report)
{
    if (!b || !theta_host || !report) { set_error("tdi_fit_multiframe: null argument"); return TDI_ERR_ARG; }
    int rc = joint_check(b, "tdi_fit_multiframe");
    if (rc) return rc;
    tdi_ctx* c = b->ctx;
    const tdi_fit_options& o = c->opts;
    const int n = b->n, K = c->L.K, off_tz = c->L.off_pose + 2;
    int Ks;
    long long KJl;
    joint_sizes(b, Ks, KJl);
    const int KJ = (int)KJl, Kf = K - Ks;
    if (KJl > CHOL_BIG_MAXK) { set_error("tdi_fit_multiframe: %lld joint unknowns > %d", KJl, CHOL_BIG_MAXK); return TDI_ERR_ARG; }
    hipStream_t s = 0;
    Workspace& w = b->ws;
    // device scratch of the call, released on every return
    Alloc A;
    struct Release {
        Alloc& a;
        hipStream_t s;
        ~Release() { (void)hipStreamSynchronize(s); a.release(); }
    } release{A, s};
    float* dth = A.get<float>((size_t)n * K);
    double* HJ = A.get<double>((size_t)KJ * KJ);
    double* gJ = A.get<double>(KJ);
    float* dJ = A.get<float>(KJ);
    int* st = A.get<int>(1);
    double* objs = A.get<double>(1);
    if (!dth || !HJ || !gJ || !dJ || !st || !objs) { set_error("tdi_fit_multiframe: allocation failed"); return TDI_ERR_NOMEM; }
    auto fail = [](int e) { return e; };
    if ((rc = ws_load(b, 0, n, s))) return fail(rc);
    // joint vector from the frames' thetas (shared block from frame 0)
    std::vector<float> x(KJ), xt(KJ), th((size_t)n * K), delta(KJ);
    std::vector<double> g(KJ), ob(n), obb(n);
    std::vector<int> fok(n), np(n);
    for (int k = 0; k < Ks; ++k) x[k] = theta_host[k];
    for (int t = 0; t < n; ++t)
        for (int k = 0; k < Kf; ++k) x[Ks + (size_t)t * Kf + k] = theta_host[(size_t)t * K + Ks + k];
    auto expand = [&](const std::vector<float>& v) {
        for (int t = 0; t < n; ++t) {
            for (int k = 0; k < Ks; ++k) th[(size_t)t * K + k] = v[k];
            for (int k = 0; k < Kf; ++k) th[(size_t)t * K + Ks + k] = v[Ks + (size_t)t * Kf + k];
        }
    };
    auto upload = [&](const std::vector<float>& v) -> int {
        expand(v);
        TDI_HIP(hipMemcpyAsync(dth, th.data(), sizeof(float) * th.size(), hipMemcpyHostToDevice, s));
        TDI_HIP(hipStreamSynchronize(s));
        return TDI_OK;
    };
    // the reference's step threshold: 0.075 for one frame, 0.001 for several (model_fitter.cu:149-151)
    const float t_min = n > 1 ? 0.001f : o.t_min;
    tdi_fit_report& rep = *report;
    std::memset(&rep, 0, sizeof(rep));
    int outer = 0, performed = 0, ls_evals = 0;
    bool terminate = false;
    for (int it = 0; it < o.max_iters; ++it) {
        outer++;
        if (terminate) break;
        if ((rc = upload(x))) return fail(rc);
        if ((rc = joint_eval(b, dth, HJ, gJ, objs, s))) return fail(rc);
        double os = 0;
        TDI_HIP(hipMemcpyAsync(g.data(), gJ, sizeof(double) * KJ, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(ob.data(), w.obj, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(np.data(), w.npix, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(&os, objs, sizeof(double), hipMemcpyDeviceToHost, s));
        if ((rc = launch_cholesky_big_f64(1, KJ, HJ, gJ, dJ, st, s))) return fail(rc);
        int sst = 0;
        TDI_HIP(hipMemcpyAsync(&sst, st, sizeof(int), hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(delta.data(), dJ, sizeof(float) * KJ, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        performed++;
        rep.n_pixels = 0;
        for (int t = 0; t < n; ++t) rep.n_pixels += np[t];
        if (sst != TDI_OK) { rep.stop |= TDI_STOP_NOT_PD; break; }
        if (std::isnan(delta[0])) { rep.stop |= TDI_STOP_NAN; break; }
        double obj0 = os, dg = 0;
        for (int t = 0; t < n; ++t) obj0 += ob[t];
        for (int k = 0; k < KJ; ++k) dg += (double)delta[k] * g[k];
        float tc = 1.f;
        int inner = 0;
        while (inner < o.max_ls) {
            if (tc < t_min) { terminate = true; rep.stop |= TDI_STOP_STEP; break; }
            for (int k = 0; k < KJ; ++k) xt[k] = x[k] + tc * delta[k];
            // the reference checks the depth of the frame its variables point at (frame 0)
            if (xt[off_tz] < 0.f) break;
            if ((rc = upload(xt))) return fail(rc);
            if ((rc = joint_objective(b, dth, objs, s))) return fail(rc);
            double ost = 0;
            TDI_HIP(hipMemcpyAsync(ob.data(), w.obj, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(obb.data(), w.objb, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(fok.data(), w.face_ok, sizeof(int) * n, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(&ost, objs, sizeof(double), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipStreamSynchronize(s));
            bool skip = std::isnan(ost);
            double objt = ost;
            for (int t = 0; t < n; ++t) {
                if (!fok[t] || std::isnan(obb[t])) skip = true;
                objt += ob[t];
            }
            if (skip) { tc *= o.armijo_beta; continue; }
            inner++;
            ls_evals++;
            if (-dg < o.decrement_tol) { terminate = true; rep.stop |= TDI_STOP_DECREMENT; break; }
            if (objt < obj0 + (double)(o.armijo_alpha * tc) * dg) {
                x = xt;
                break;
            }
            tc *= o.armijo_beta;
        }
    }
    if (!terminate && !(rep.stop & (TDI_STOP_NAN | TDI_STOP_NOT_PD))) rep.stop |= TDI_STOP_MAXITER;
    if ((rc = upload(x))) return fail(rc);
    if ((rc = joint_objective(b, dth, objs, s))) return fail(rc);
    double fin = 0;
    TDI_HIP(hipMemcpyAsync(ob.data(), w.obj, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TDI_HIP(hipMemcpyAsync(&fin, objs, sizeof(double), hipMemcpyDeviceToHost, s));
    TDI_HIP(hipStreamSynchronize(s));
    for (int t = 0; t < n; ++t) fin += ob[t];
    rep.objective = (float)fin;
    rep.iters = performed;
    rep.ls_evals = ls_evals;
    rep.success = outer > o.niters_success;
    rep.status = (rep.stop & TDI_STOP_NOT_PD) ? TDI_NOT_PD : (rep.stop & TDI_STOP_NAN) ? TDI_NAN : TDI_OK;
    expand(x);
    std::memcpy(theta_host, th.data(), sizeof(float) * th.size());
    return TDI_OK;
}

int tdi_cholesky_solve(int32_t n, int32_t K, const float* H, const float* g, float* delta, int32_t* status,
                       int32_t mode, void* stream)
{
    if (n <= 0 || K <= 0 || (mode != 0 && mode != 1)) { set_error("bad cholesky arguments"); return TDI_ERR_ARG; }
    return launch_cholesky(n, K, H, g, delta, status, mode, nullptr, (hipStream_t)stream);
}

int tdi_ortho_init(tdi_ctx* c, int32_t n, const float* lmk, const float* cams, float* x_orth, float* pose,
                   float* angles, int32_t* info, void* stream)
{
    if (!c || n < 0 || (n > 0 && (!lmk || !cams || !x_orth || !pose || !info))) {
        set_error("tdi_ortho_init: null argument or negative frame count");
        return TDI_ERR_ARG;
    }
    if (!c->m.lmk_ids) { set_error("tdi_ortho_init: the model has no landmark ids"); return TDI_ERR_ARG; }
    (void)hipSetDevice(c->device);
    return launch_ortho_init(c->m, n, lmk, cams, x_orth, pose, angles, info, (hipStream_t)stream);
}

// The landmark-only fit with slack and its terms: every frame of the batch in one launch.
static int lmk_slack_run(tdi_batch* b, float* theta, int32_t k_exp, int mode, const double* z, double tau, double* H,
                         double* g, double* obj, tdi_lmk_report* reports, hipStream_t s, const char* who)
{
    if (!b || !theta || (mode == 1 && !z)) { set_error("%s: null argument", who); return TDI_ERR_ARG; }
    tdi_ctx* c = b->ctx;
    if (!c->m.lmk_ids) { set_error("%s: the model has no landmark ids", who); return TDI_ERR_ARG; }
    if (b->n == 0) return TDI_OK;
    (void)hipSetDevice(c->device);
    static_assert(sizeof(tdi_lmk_report) == 24, "tdi_lmk_report layout");
    FrameGPU* dfr = nullptr;
    tdi_lmk_report* drep = nullptr;
    int rc = TDI_OK;
    if (hipMalloc(&dfr, sizeof(FrameGPU) * b->n) != hipSuccess ||
        hipMalloc(&drep, sizeof(tdi_lmk_report) * b->n) != hipSuccess) {
        (void)hipFree(dfr);
        set_error("%s: allocation failed", who);
        return TDI_ERR_NOMEM;
    }
    if (hipMemcpyAsync(dfr, b->hframes.data(), sizeof(FrameGPU) * b->n, hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("%s: upload failed", who);
        rc = TDI_ERR_HIP;
    }
    if (!rc) rc = launch_lmk_slack(c->m, c->L, dfr, b->lmk, b->n, theta, k_exp, mode, z, tau, H, g, obj, drep, s);
    if (!rc && reports &&
        hipMemcpyAsync(reports, drep, sizeof(tdi_lmk_report) * b->n, hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("%s: report download failed", who);
        rc = TDI_ERR_HIP;
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) { set_error("%s: %s", who, hipGetErrorString(e)); rc = TDI_ERR_HIP; }
    (void)hipFree(dfr);
    (void)hipFree(drep);
    return rc;
}

// The joint landmark-only fit over the batch's frames (one group) and its terms: the per-frame
// terms in one launch, the frames' matrices placed on the joint one with the once-counted terms,
// the FP64 large-K Cholesky for the Newton step, the line searches from the host (few trials per
// step, each one launch over the frames).  theta host [n][K]; zin/Hout/gout/objout for the terms.
static int lmk_joint(tdi_batch* b, float* theta, int Ka, int Ke, int slack, const double* zin, double tau,
                     double* Hout, double* gout, double* objout, tdi_lmk_report* report, const char* who)
{
    if (!b || !theta || (zin && (!Hout || !gout || !objout)) || (!zin && !report)) {
        set_error("%s: null argument", who);
        return TDI_ERR_ARG;
    }
    tdi_ctx* c = b->ctx;
    if (!c->m.lmk_ids) { set_error("%s: the model has no landmark ids", who); return TDI_ERR_ARG; }
    const int n = b->n, K = c->L.K, S1 = slack ? 1 : 0;
    const int Ks = Ka + S1, nf = Ks + Ke + 6;
    const long long KJl = Ks + (long long)n * (Ke + 6);
    if (n < 1 || KJl > CHOL_BIG_MAXK) {
        set_error("%s: %d frames, %lld joint unknowns (at most %d)", who, n, KJl, CHOL_BIG_MAXK);
        return TDI_ERR_ARG;
    }
    const int KJ = (int)KJl;
    (void)hipSetDevice(c->device);
    hipStream_t s = 0;
    Alloc A;
    A.pool = c->pool;
    struct Release {
        Alloc& a;
        ~Release() { (void)hipDeviceSynchronize(); a.release(); }
    } release{A};
    FrameGPU* dfr = A.get<FrameGPU>(n);
    float* dth = A.get<float>(K);
    double* dz = A.get<double>(KJ);
    double* objf = A.get<double>(n + 1);   // the frames' objectives, then the once-counted terms'
    double* gminf = A.get<double>(n);
    double* Hf = A.get<double>((size_t)n * nf * nf);
    double* gf = A.get<double>((size_t)n * nf);
    double* J = A.get<double>(lmk_joint_scratch_doubles(n, Ka, Ke, S1));
    double* HJ = A.get<double>((size_t)KJ * KJ);
    double* gJ = A.get<double>(KJ);
    double* dd = A.get<double>(KJ);
    int* dst = A.get<int>(1);
    for (void* p : A.ptrs)
        if (!p) { set_error("%s: allocation failed", who); return TDI_ERR_NOMEM; }
    TDI_HIP(hipMemcpyAsync(dfr, b->hframes.data(), sizeof(FrameGPU) * n, hipMemcpyHostToDevice, s));
    TDI_HIP(hipMemcpyAsync(dth, theta, sizeof(float) * K, hipMemcpyHostToDevice, s));
    std::vector<double> hobj(n + 1), hgmin(n);
    // objective (and with mode 1 the joint normal equations in HJ, gJ) at zv
    auto eval = [&](const std::vector<double>& zv, double tau_b, int mode, double& total, bool want_gmin) -> int {
        TDI_HIP(hipMemcpyAsync(dz, zv.data(), sizeof(double) * KJ, hipMemcpyHostToDevice, s));
        int rc = launch_lmk_joint(c->m, c->L, dfr, b->lmk, dth, n, Ka, Ke, S1, dz, mode, 0, nullptr, objf, gminf, Hf,
                                  gf, J, s);
        if (!rc && mode) rc = launch_assemble_joint(n, nf, Ks, Hf, gf, HJ, gJ, s);
        if (!rc)
            rc = launch_lmk_joint_shared(c->m, c->L, Ka, S1, KJ, dz, tau_b, nullptr, mode, 1, nullptr, objf + n, HJ, gJ, s);
        if (rc) return rc;
        TDI_HIP(hipMemcpyAsync(hobj.data(), objf, sizeof(double) * (n + 1), hipMemcpyDeviceToHost, s));
        if (want_gmin) TDI_HIP(hipMemcpyAsync(hgmin.data(), gminf, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        total = hobj[n];
        for (int f = 0; f < n; ++f) total += hobj[f];
        return TDI_OK;
    };
    std::vector<double> z(KJ), zt(KJ), d(KJ);
    if (zin) {
        std::copy(zin, zin + KJ, z.begin());
        double obj = 0.0;
        int rc = eval(z, tau, 1, obj, false);
        if (rc) return rc;
        TDI_HIP(hipMemcpyAsync(Hout, HJ, sizeof(double) * (size_t)KJ * KJ, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(gout, gJ, sizeof(double) * KJ, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        *objout = obj;
        return TDI_OK;
    }
    const Layout& L = c->L;
    for (int k = 0; k < Ka; ++k) z[k] = theta[L.off_id + k];
    for (int t = 0; t < n; ++t) {
        const float* th = theta + (size_t)t * K;
        double* zf = &z[(size_t)Ks + (size_t)t * (Ke + 6)];
        for (int k = 0; k < Ke; ++k) zf[k] = th[L.off_exp + k];
        for (int a = 0; a < 6; ++a) zf[Ke + a] = th[L.off_pose + a];
    }
    tdi_lmk_report rep{};
    int rc = TDI_OK;
    if (S1) {
        // set_minimal_slack: s = max over the frames of 5 - min G at s = 0
        z[Ka] = 0.0;
        double obj = 0.0;
        if ((rc = eval(z, 0.0, 0, obj, true))) return rc;
        double hi = -INFINITY;
        for (int t = 0; t < n; ++t) {
            const double st = 5.0 - hgmin[t];
            hi = (st != st || hi != hi) ? NAN : std::max(hi, st);
        }
        z[Ka] = hi;
        rep.slack0 = (float)hi;
    }
    double tau_b = 10.0;
    bool success = false;
    for (int round = 0; round < 5 && !success; ++round) {
        tau_b *= 10.0;
        rep.rounds++;
        for (int it = 0; it < 100; ++it) {
            if (S1 && z[Ka] < -0.01) { success = true; break; }
            double obj = 0.0;
            if ((rc = eval(z, tau_b, 1, obj, false))) return rc;
            rep.iterations++;
            if ((rc = launch_cholesky_big_f64(1, KJ, HJ, gJ, dd, dst, s))) return rc;
            int hst = 0;
            TDI_HIP(hipMemcpyAsync(d.data(), dd, sizeof(double) * KJ, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(&hst, dst, sizeof(int), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipStreamSynchronize(s));
            if (hst != TDI_OK) break;
            bool stop = false;
            double t = 1.0;
            for (int inner = 0; inner < 100; ++inner) {
                if (t < 0.01) { stop = true; break; }
                for (int j = 0; j < KJ; ++j) zt[j] = z[j] + t * d[j];
                double ot = 0.0;
                if ((rc = eval(zt, tau_b, 0, ot, false))) return rc;
                if (ot < obj) {
                    z = zt;
                    rep.accepted++;
                    break;
                }
                t *= 0.5;
            }
            if (stop) break;
        }
    }
    rep.status = (!S1 || success) ? TDI_OK : TDI_LMK_INFEASIBLE;
    rep.slack = S1 ? (float)z[Ka] : 0.f;
    if (rep.status == TDI_OK)
        for (int t = 0; t < n; ++t) {
            float* th = theta + (size_t)t * K;
            const double* zf = &z[(size_t)Ks + (size_t)t * (Ke + 6)];
            for (int k = 0; k < Ka; ++k) th[L.off_id + k] = (float)z[k];
            for (int k = 0; k < Ke; ++k) th[L.off_exp + k] = (float)zf[k];
            for (int a = 0; a < 6; ++a) th[L.off_pose + a] = (float)zf[Ke + a];
        }
    *report = rep;
    return TDI_OK;
}

// Every frame of the batch a group of its own, with the slack (the per-frame landmark fits of the
// identity stage's selection): each frame keeps its own state of the loop above (barrier round,
// Newton step, line-search trial), and each evaluation round is one launch over the frames that
// need it; the Newton steps are one batched large-K Cholesky.
int tdi_fit_landmarks_each(tdi_batch* b, float* theta, int32_t k_id, int32_t k_exp, tdi_lmk_report* reports)
{
    const char* who = "tdi_fit_landmarks_each";
    if (!b || !theta || !reports) { set_error("%s: null argument", who); return TDI_ERR_ARG; }
    tdi_ctx* c = b->ctx;
    if (!c->m.lmk_ids) { set_error("%s: the model has no landmark ids", who); return TDI_ERR_ARG; }
    const int n = b->n, K = c->L.K, Ka = k_id, Ke = k_exp, Ks = Ka + 1, nf = Ks + Ke + 6;
    if (n < 1) return TDI_OK;
    (void)hipSetDevice(c->device);
    hipStream_t s = 0;
    Alloc A;
    A.pool = c->pool;
    struct Release {
        Alloc& a;
        ~Release() { (void)hipDeviceSynchronize(); a.release(); }
    } release{A};
    FrameGPU* dfr = A.get<FrameGPU>(n);
    float* dth = A.get<float>((size_t)n * K);
    double* dz = A.get<double>((size_t)n * nf);
    double* objf = A.get<double>(2 * (size_t)n);   // the frames' landmark terms, then their once-counted terms
    double* gminf = A.get<double>(n);
    double* Hf = A.get<double>((size_t)n * nf * nf);
    double* gf = A.get<double>((size_t)n * nf);
    double* J = A.get<double>(lmk_joint_scratch_doubles(n, Ka, Ke, 1));
    double* dd = A.get<double>((size_t)n * nf);
    int* dst = A.get<int>(n);
    int* dmask = A.get<int>(n);
    double* dtau = A.get<double>(n);
    for (void* p : A.ptrs)
        if (!p) { set_error("%s: allocation failed", who); return TDI_ERR_NOMEM; }
    TDI_HIP(hipMemcpyAsync(dfr, b->hframes.data(), sizeof(FrameGPU) * n, hipMemcpyHostToDevice, s));
    TDI_HIP(hipMemcpyAsync(dth, theta, sizeof(float) * (size_t)n * K, hipMemcpyHostToDevice, s));
    const Layout& L = c->L;
    std::vector<double> z((size_t)n * nf), zt, d((size_t)n * nf), hobj(2 * (size_t)n), hgmin(n), htau(n, 0.0);
    std::vector<int> mask(n, 1), hst(n);
    for (int f = 0; f < n; ++f) {
        const float* th = theta + (size_t)f * K;
        double* zf = &z[(size_t)f * nf];
        for (int k = 0; k < Ka; ++k) zf[k] = th[L.off_id + k];
        zf[Ka] = 0.0;
        for (int k = 0; k < Ke; ++k) zf[Ks + k] = th[L.off_exp + k];
        for (int a = 0; a < 6; ++a) zf[Ks + Ke + a] = th[L.off_pose + a];
    }
    // objective (mode 1: also each frame's system in Hf, gf) of the masked frames at zv
    auto eval = [&](const std::vector<double>& zv, int mode, bool want_gmin) -> int {
        TDI_HIP(hipMemcpyAsync(dz, zv.data(), sizeof(double) * zv.size(), hipMemcpyHostToDevice, s));
        TDI_HIP(hipMemcpyAsync(dmask, mask.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        TDI_HIP(hipMemcpyAsync(dtau, htau.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        int rc = launch_lmk_joint(c->m, L, dfr, b->lmk, dth, n, Ka, Ke, 1, dz, mode, 1, dmask, objf, gminf, Hf, gf, J, s);
        if (!rc) rc = launch_lmk_joint_shared(c->m, L, Ka, 1, nf, dz, 0.0, dtau, mode, n, dmask, objf + n, Hf, gf, s);
        if (rc) return rc;
        TDI_HIP(hipMemcpyAsync(hobj.data(), objf, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
        if (want_gmin) TDI_HIP(hipMemcpyAsync(hgmin.data(), gminf, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        return TDI_OK;
    };
    auto total = [&](int f) { return hobj[(size_t)n + f] + hobj[f]; };
    enum Phase { TOP, SEARCH, DONE };
    struct St {
        Phase phase = TOP;
        int round = 0, it = 0;
        double t = 1.0, obj = 0.0;
        bool success = false;
    };
    std::vector<St> st(n);
    std::vector<tdi_lmk_report> rep(n);
    // set_minimal_slack: s = 5 - min G at s = 0
    int rc = eval(z, 0, true);
    if (rc) return rc;
    for (int f = 0; f < n; ++f) {
        z[(size_t)f * nf + Ka] = 5.0 - hgmin[f];
        rep[f].slack0 = (float)z[(size_t)f * nf + Ka];
        htau[f] = 100.0;
        rep[f].rounds = 1;
    }
    // the round of frame f ends: the next one starts, or the fit ends after five
    auto next_round = [&](int f) {
        St& q = st[f];
        if (++q.round >= 5) { q.phase = DONE; return; }
        htau[f] *= 10.0;
        rep[f].rounds++;
        q.it = 0;
        q.phase = TOP;
    };
    for (;;) {
        // the top of an iteration: success, the round's step limit, else a Newton step
        bool any = false;
        for (int f = 0; f < n; ++f) {
            St& q = st[f];
            while (q.phase == TOP) {
                if (z[(size_t)f * nf + Ka] < -0.01) { q.success = true; q.phase = DONE; break; }
                if (q.it >= 100) { next_round(f); continue; }
                break;
            }
            mask[f] = q.phase == TOP;
            any = any || mask[f];
        }
        if (!any) break;
        if ((rc = eval(z, 1, false))) return rc;
        if ((rc = launch_cholesky_big_f64(n, nf, Hf, gf, dd, dst, s))) return rc;
        TDI_HIP(hipMemcpyAsync(d.data(), dd, sizeof(double) * d.size(), hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(hst.data(), dst, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        for (int f = 0; f < n; ++f) {
            if (!mask[f]) continue;
            St& q = st[f];
            rep[f].iterations++;
            if (hst[f] != TDI_OK) { next_round(f); continue; }
            q.obj = total(f);
            q.t = 1.0;
            q.phase = SEARCH;
        }
        // the line searches, one objective evaluation per round for every frame still searching
        for (;;) {
            zt = z;
            bool trial = false;
            for (int f = 0; f < n; ++f) {
                St& q = st[f];
                mask[f] = 0;
                if (q.phase != SEARCH) continue;
                if (q.t < 0.01) { next_round(f); continue; }
                for (int j = 0; j < nf; ++j) zt[(size_t)f * nf + j] = z[(size_t)f * nf + j] + q.t * d[(size_t)f * nf + j];
                mask[f] = 1;
                trial = true;
            }
            if (!trial) break;
            if ((rc = eval(zt, 0, false))) return rc;
            for (int f = 0; f < n; ++f) {
                if (!mask[f]) continue;
                St& q = st[f];
                if (total(f) < q.obj) {
                    std::copy(zt.begin() + (size_t)f * nf, zt.begin() + (size_t)(f + 1) * nf, z.begin() + (size_t)f * nf);
                    rep[f].accepted++;
                    q.it++;
                    q.phase = TOP;
                } else {
                    q.t *= 0.5;
                }
            }
        }
    }
    for (int f = 0; f < n; ++f) {
        const double* zf = &z[(size_t)f * nf];
        rep[f].status = st[f].success ? TDI_OK : TDI_LMK_INFEASIBLE;
        rep[f].slack = (float)zf[Ka];
        reports[f] = rep[f];
        if (!st[f].success) continue;
        float* th = theta + (size_t)f * K;
        for (int k = 0; k < Ka; ++k) th[L.off_id + k] = (float)zf[k];
        for (int k = 0; k < Ke; ++k) th[L.off_exp + k] = (float)zf[Ks + k];
        for (int a = 0; a < 6; ++a) th[L.off_pose + a] = (float)zf[Ks + Ke + a];
    }
    return TDI_OK;
}

int tdi_fit_landmarks_joint(tdi_batch* b, float* theta, int32_t k_id, int32_t k_exp, int32_t slack,
                            tdi_lmk_report* report)
{
    return lmk_joint(b, theta, k_id, k_exp, slack, nullptr, 0.0, nullptr, nullptr, nullptr, report,
                     "tdi_fit_landmarks_joint");
}

int tdi_landmark_joint_terms(tdi_batch* b, const float* theta, int32_t k_id, int32_t k_exp, int32_t slack,
                             const double* z, double tau, double* H, double* g, double* obj)
{
    if (!z) { set_error("tdi_landmark_joint_terms: null argument"); return TDI_ERR_ARG; }
    // the terms only read theta
    return lmk_joint(b, const_cast<float*>(theta), k_id, k_exp, slack, z, tau, H, g, obj, nullptr,
                     "tdi_landmark_joint_terms");
}

int tdi_fit_landmarks_slack(tdi_batch* b, float* theta, int32_t k_exp, tdi_lmk_report* reports, void* stream)
{
    return lmk_slack_run(b, theta, k_exp, 0, nullptr, 0.0, nullptr, nullptr, nullptr, reports, (hipStream_t)stream,
                         "tdi_fit_landmarks_slack");
}

int tdi_landmark_slack_terms(tdi_batch* b, const float* theta, int32_t k_exp, const double* z, double tau, double* H,
                             double* g, double* obj, void* stream)
{
    // mode 1 only reads theta
    return lmk_slack_run(b, const_cast<float*>(theta), k_exp, 1, z, tau, H, g, obj, nullptr, (hipStream_t)stream,
                         "tdi_landmark_slack_terms");
}

// Stage timers: one full evaluation (the work of one GN iteration before the solve) with
// HIP events around each stage, on the batch's stream.  Mirrors the reference's compile-time
// MEASURE_TIME timers (model_fitter.cu:43-58, derivative_computer.cu:32-37,141-172).
int tdi_fit_buffers(tdi_batch* b, const float* theta, int32_t mixed_flags, float* vproj, int32_t* nbr,
                    int32_t* npix, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    Workspace& w = b->ws;
    const int K = b->ctx->L.K, V = b->ctx->m.V;
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        const long long off = b->hframes[i0].obs_off;
        const long long len = (i0 + n < b->n ? b->hframes[i0 + n].obs_off : b->total) - off;
        if ((rc = eval_render(b, all_frames(n), theta + (size_t)i0 * K, mixed_flags, s))) return rc;
        if (vproj)
            TDI_HIP(hipMemcpyAsync(vproj + (size_t)i0 * V * 4, w.vproj, sizeof(float4) * (size_t)n * V,
                                   hipMemcpyDeviceToDevice, s));
        // pnb holds each frame's positions at the frame's crop offset: the same layout as the output
        if (nbr) TDI_HIP(hipMemcpyAsync(nbr + 4 * off, w.pnb, sizeof(int4) * len, hipMemcpyDeviceToDevice, s));
        if (npix) TDI_HIP(hipMemcpyAsync(npix + i0, w.npix, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    }
    return TDI_OK;
}

int tdi_barrier_terms(tdi_batch* b, const float* theta, double* H, double* g, double* obj, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    Workspace& w = b->ws;
    tdi_ctx* c = b->ctx;
    const int K = c->L.K;
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        const float* th = theta + (size_t)i0 * K;
        TDI_HIP(hipMemsetAsync(w.H, 0, sizeof(double) * (size_t)n * K * K, s));
        TDI_HIP(hipMemsetAsync(w.g, 0, sizeof(float) * (size_t)n * K, s));
        if ((rc = launch_barrier(c->m, c->L, w.frames, all_frames(n), th, lmk_arrays(w), c->opts.use_landmarks, 0,
                                 w.obj, w.objb, w.H, w.g, s)))
            return rc;
        if ((rc = launch_barrier(c->m, c->L, w.frames, all_frames(n), th, lmk_arrays(w), c->opts.use_landmarks, 1,
                                 w.obj, w.objb, w.H, w.g, s)))
            return rc;
        if (H) TDI_HIP(hipMemcpyAsync(H + (size_t)i0 * K * K, w.H, sizeof(double) * (size_t)n * K * K,
                                      hipMemcpyDeviceToDevice, s));
        if (g) {
            std::vector<float> hg((size_t)n * K);
            std::vector<double> dg((size_t)n * K);
            TDI_HIP(hipMemcpyAsync(hg.data(), w.g, sizeof(float) * hg.size(), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipStreamSynchronize(s));
            for (size_t k = 0; k < hg.size(); ++k) dg[k] = hg[k];
            TDI_HIP(hipMemcpyAsync(g + (size_t)i0 * K, dg.data(), sizeof(double) * dg.size(), hipMemcpyHostToDevice, s));
            TDI_HIP(hipStreamSynchronize(s));
        }
        if (obj) TDI_HIP(hipMemcpyAsync(obj + i0, w.objb, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    return TDI_OK;
}

// Launch-shape autotuner: on the first workspace chunk, rendered at theta, time the per-pixel
// Jacobian kernel for four grid sizes (workgroups per frame sized for half, one, two and four
// times the mean crop) and the J^T J kernel for the split counts 1, 2, 4, ... up to the partials
// buffer's capacity, `reps` launches each after one untimed launch (HIP events on the batch's
// stream); keep the fastest of each for this batch, drop its captured fit schedules (they hold
// the old shapes) and remember the choice for later batches of the same shape.
int tdi_batch_autotune(tdi_batch* b, const float* theta, int32_t reps, int32_t* choice, float* report)
{
    if (!b || !theta || reps < 1) { set_error("tdi_batch_autotune: null argument or reps < 1"); return TDI_ERR_ARG; }
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    hipStream_t s = b->stream;
    (void)hipSetDevice(c->device);
    if (b->n == 0) return TDI_OK;
    const int n = std::min(w.C, b->n);
    const int mixed = c->opts.mixed_flags;
    int rc = ws_load(b, 0, n, s);
    if (!rc) rc = eval_render(b, all_frames(n), theta, mixed, s);
    if (rc) return rc;
    hipEvent_t e0, e1;
    TDI_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("tdi_batch_autotune: event creation failed");
        return TDI_ERR_HIP;
    }
    float rep[11];
    for (float& r : rep) r = -1.f;
    auto timed = [&](auto&& launch, float& out) -> int {
        int r = launch();
        if (r) return r;
        if (hipEventRecord(e0, s) != hipSuccess) return TDI_ERR_HIP;
        for (int i = 0; i < reps && !r; ++i) r = launch();
        if (r) return r;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return TDI_ERR_HIP;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return TDI_ERR_HIP;
        out = ms / reps;
        return TDI_OK;
    };
    const int a0 = mean_crop(w);
    const int areas[4] = {std::max(1, a0 / 2), a0, 2 * a0, 4 * a0};
    int best_area = a0;
    float best = 1e30f;
    for (int i = 0; i < 4 && !rc; ++i) {
        rc = timed([&] { return launch_dI(b->m, c->L, w.frames, all_frames(n), theta, w.P0, w.pvert, w.pbary, w.npix,
                                          w.pixg, w.dI, areas[i], s); },
                   rep[i]);
        if (!rc && rep[i] < best) { best = rep[i]; best_area = areas[i]; }
    }
    const int S0 = w.S;
    int best_S = S0;
    best = 1e30f;
    for (int lg = 0; lg <= 6 && !rc; ++lg) {
        const int S = 1 << lg;
        if (S > w.S_cap) break;
        w.S = S;
        rc = timed([&] { return launch_normal_eq(b, all_frames(n), mixed, s); }, rep[4 + lg]);
        if (!rc && rep[4 + lg] < best) { best = rep[4 + lg]; best_S = S; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) { w.S = S0; return rc; }
    w.S = best_S;
    w.di_area = best_area;
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    if (b->graph_fine) { (void)hipGraphExecDestroy(b->graph_fine); b->graph_fine = nullptr; }
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[TuneKey(c->L.K, w.C, w.cap, b->max_area, mixed)] = TuneChoice{best_S, best_area};
    }
    if (choice) { choice[0] = best_S; choice[1] = best_area; choice[2] = a0; choice[3] = S0; }
    if (report) std::memcpy(report, rep, sizeof(rep));
    return TDI_OK;
}

int tdi_batch_launch_shape(tdi_batch* b, int32_t* shape)
{
    if (!b || !shape) { set_error("tdi_batch_launch_shape: null argument"); return TDI_ERR_ARG; }
    shape[0] = b->ws.S;
    shape[1] = dI_area(b->ws);
    return TDI_OK;
}

// Stages: 0 synthesis, 1 raster (clear+raster+resolve), 2 shade+blur+grads, 3 barrier,
// 4 dI, 5 J^T J (+reduce), 6 Cholesky.  ms[7] receives the mean per repetition.
int tdi_stage_times(tdi_batch* b, const float* theta, int32_t mixed, int32_t reps, float* ms)
{
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    hipStream_t s = b->stream;
    const int n = std::min(b->n, w.C);   // the first workspace chunk
    int rc;
    if ((rc = ws_load(b, 0, n, s))) return rc;
    hipEvent_t ev[8];
    for (int i = 0; i < 8; ++i) TDI_HIP(hipEventCreate(&ev[i]));
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = 0; r < reps; ++r) {
        TDI_HIP(hipEventRecord(ev[0], s));
        if (mixed & TDI_MIXED_SYNTH_MFMA)
            rc = launch_synthesize_gemm(b->m, c->L, w.frames, all_frames(n), theta, w.rt, w.P0, nullptr, w.vproj, s);
        else
            rc = launch_synthesize(b->m, c->L, w.frames, all_frames(n), theta, w.rt, w.P0, nullptr, w.vproj, s);
        if (rc) return rc;
        TDI_HIP(hipEventRecord(ev[1], s));
        if ((rc = launch_clear_keys_frames(w.frames, all_frames(n), w.keys, w.box, w.max_area, s))) return rc;
        if ((rc = launch_raster(c->m, w.frames, all_frames(n), w.vproj, c->opts.ntmp_cap, w.keys, w.box, s))) return rc;
        if ((rc = launch_resolve(c->m, w.frames, all_frames(n), w.vproj, w.keys, w.cum, w.vis, w.npix, nullptr, w.pvert, w.pbary, w.max_area, w.tcount, w.tmax, w.box, s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[2], s));
        if ((rc = launch_shade(b->m, c->L, w.frames, all_frames(n), theta, w.P0, w.pvert, w.pbary, w.npix, w.box, w.img,
                               w.max_area, s)))
            return rc;
        if ((rc = launch_blur(w.frames, all_frames(n), c->taps, w.box, w.img, w.tmp, w.blurred, w.max_area, s))) return rc;
        if ((rc = launch_grads(w.frames, all_frames(n), w.vis, w.npix, w.cum, w.blurred, b->obs_g, c->opts.data_weight, w.pixg,
                               w.pnb, w.pcoef, w.objp, w.obj, w.max_area, s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[3], s));
        if ((rc = launch_barrier(c->m, c->L, w.frames, all_frames(n), theta, lmk_arrays(w), c->opts.use_landmarks, 0, w.obj,
                                 w.objb, w.H, w.g, s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[4], s));
        if ((rc = launch_dI(b->m, c->L, w.frames, all_frames(n), theta, w.P0, w.pvert, w.pbary, w.npix, w.pixg, w.dI,
                            dI_area(w), s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[5], s));
        if ((rc = launch_normal_eq(b, all_frames(n), mixed, s))) return rc;
        if ((rc = launch_barrier(c->m, c->L, w.frames, all_frames(n), theta, lmk_arrays(w), c->opts.use_landmarks, 1, w.obj,
                                 w.objb, w.H, w.g, s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[6], s));
        if ((rc = launch_cholesky(n, c->L.K, w.H, w.g, w.delta, w.status, (mixed & TDI_MIXED_CHOL_F32) ? 1 : 0, nullptr, s)))
            return rc;
        TDI_HIP(hipEventRecord(ev[7], s));
        TDI_HIP(hipEventSynchronize(ev[7]));
        for (int k = 0; k < 7; ++k) {
            float t = 0;
            TDI_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            acc[k] += t;
        }
    }
    for (int i = 0; i < 8; ++i) (void)hipEventDestroy(ev[i]);
    for (int k = 0; k < 7; ++k) ms[k] = (float)(acc[k] / reps);
    return TDI_OK;
}

// Visible-pixel counts of the last evaluation (host copy), for algorithmic-work accounting.
int tdi_last_pixel_counts(tdi_batch* b, int32_t* npix)
{
    TDI_HIP(hipStreamSynchronize(b->stream));
    TDI_HIP(hipMemcpy(npix, b->ws.npix, sizeof(int) * b->ws.n_cur, hipMemcpyDeviceToHost));
    return TDI_OK;
}

// ---------------------------------------------------------------------------- plain fit
// fit_3DMM_shape_rigid (model_fitter.cu:121-375) one frame at a time: the host owns the
// loop and reads the objective, the direction and the status back after every step, like
// the reference's blocking cudaMemcpy pattern; FP32 J^T J, FP64 Cholesky, no graph.
static int fit_plain(tdi_batch* b, float* theta_host, tdi_fit_report* reports)
{
    tdi_ctx* c = b->ctx;
    const tdi_fit_options& o = c->opts;
    const int K = c->L.K, off_tz = c->L.off_pose + 2;
    const int hold0 = b->finetune ? c->L.off_pose : 0, hold1 = b->finetune ? c->L.off_pose + 6 : 0;
    hipStream_t s = 0;
    Workspace& w = b->ws;
    std::vector<float> th(K), tt(K), delta(K), g(K);
    for (int i = 0; i < b->n; ++i) {
        int rc = ws_load(b, i, 1, s);
        if (rc) return rc;
        if ((rc = set_mask_all(b, s))) return rc;
        tdi_fit_report& rep = reports[i];
        std::memset(&rep, 0, sizeof(rep));
        std::memcpy(th.data(), theta_host + (size_t)i * K, sizeof(float) * K);
        TDI_HIP(hipMemcpyAsync(w.theta, th.data(), sizeof(float) * K, hipMemcpyHostToDevice, s));
        int outer = 0, performed = 0, ls_evals = 0;
        bool terminate = false;
        for (int it = 0; it < o.max_iters; ++it) {
            outer++;
            if (terminate) break;
            if ((rc = eval_full(b, all_frames(1), w.theta, 0, s))) return rc;
            if ((rc = launch_cholesky(1, K, w.H, w.g, w.delta, w.status, 0, w.mask, s))) return rc;
            int st = 0, np = 0;
            double obj0 = 0;
            TDI_HIP(hipMemcpyAsync(&st, w.status, sizeof(int), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(delta.data(), w.delta, sizeof(float) * K, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(g.data(), w.g, sizeof(float) * K, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(&obj0, w.obj, sizeof(double), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipMemcpyAsync(&np, w.npix, sizeof(int), hipMemcpyDeviceToHost, s));
            TDI_HIP(hipStreamSynchronize(s));
            performed++;
            rep.n_pixels = np;
            if (st != TDI_OK) { rep.stop |= TDI_STOP_NOT_PD; break; }
            if (std::isnan(delta[0])) { rep.stop |= TDI_STOP_NAN; break; }
            double dg = 0;
            for (int k = 0; k < K; ++k) dg += (double)delta[k] * g[k];
            float t = 1.f;
            int inner = 0;
            while (inner < o.max_ls) {
                if (t < o.t_min) { terminate = true; rep.stop |= TDI_STOP_STEP; break; }
                for (int k = 0; k < K; ++k) tt[k] = (k >= hold0 && k < hold1) ? th[k] : th[k] + t * delta[k];
                if (tt[off_tz] < 0.f) break;
                TDI_HIP(hipMemcpyAsync(w.theta_try, tt.data(), sizeof(float) * K, hipMemcpyHostToDevice, s));
                if ((rc = eval_objective(b, all_frames(1), w.theta_try, 0, s))) return rc;
                double objt = 0, objb = 0;
                int fok = 1;
                TDI_HIP(hipMemcpyAsync(&objt, w.obj, sizeof(double), hipMemcpyDeviceToHost, s));
                TDI_HIP(hipMemcpyAsync(&objb, w.objb, sizeof(double), hipMemcpyDeviceToHost, s));
                TDI_HIP(hipMemcpyAsync(&fok, w.face_ok, sizeof(int), hipMemcpyDeviceToHost, s));
                TDI_HIP(hipStreamSynchronize(s));
                if (!fok || std::isnan(objb)) { t *= o.armijo_beta; continue; }
                inner++;
                ls_evals++;
                if (-dg < o.decrement_tol) { terminate = true; rep.stop |= TDI_STOP_DECREMENT; break; }
                if (objt < obj0 + (double)(o.armijo_alpha * t) * dg) {
                    th = tt;
                    TDI_HIP(hipMemcpyAsync(w.theta, th.data(), sizeof(float) * K, hipMemcpyHostToDevice, s));
                    break;
                }
                t *= o.armijo_beta;
            }
        }
        if (!terminate && !(rep.stop & (TDI_STOP_NAN | TDI_STOP_NOT_PD))) rep.stop |= TDI_STOP_MAXITER;
        if ((rc = eval_objective(b, all_frames(1), w.theta, 0, s))) return rc;
        double fin = 0;
        TDI_HIP(hipMemcpyAsync(&fin, w.obj, sizeof(double), hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        rep.objective = (float)fin;
        rep.iters = performed;
        rep.ls_evals = ls_evals;
        rep.success = outer > o.niters_success;
        rep.status = (rep.stop & TDI_STOP_NOT_PD) ? TDI_NOT_PD : (rep.stop & TDI_STOP_NAN) ? TDI_NAN : TDI_OK;
        std::memcpy(theta_host + (size_t)i * K, th.data(), sizeof(float) * K);
    }
    return TDI_OK;
}

int tdi_fit_plain(tdi_batch* b, float* theta_host, tdi_fit_report* reports)
{
    if (!b || !theta_host || !reports) { set_error("tdi_fit_plain: null argument"); return TDI_ERR_ARG; }
    b->finetune = 0;
    return fit_plain(b, theta_host, reports);
}

int tdi_fit_plain_finetune(tdi_batch* b, float* theta_host, tdi_fit_report* reports)
{
    if (!b || !theta_host || !reports) { set_error("tdi_fit_plain_finetune: null argument"); return TDI_ERR_ARG; }
    b->finetune = 1;
    const int rc = fit_plain(b, theta_host, reports);
    b->finetune = 0;
    return rc;
}

// ---------------------------------------------------------------------------- light sub-fits
// fit_3DMM_lambdas (model_fitter.cu:981-1073) and fit_3DMM_Lintensity (model_fitter.cu:1084-1180):
// every frame takes the reference's own sequence of decisions (its outer iterations, line-search
// steps and stops), but the evaluations of a workspace chunk run together: one evaluation of the
// objective, gradient and normal matrix for the frames starting an outer iteration, then rounds
// of objective evaluations for the frames still in their line search.  Shape and pose stay at
// theta; the light is appended to the unknowns (layout Lx), so the light block of the joint
// normal equations is w J^T J and -w J^T g^_s of the light columns (the barrier does not involve
// the light): divided by w they are the reference's JTJ_lambda / dG_dtheta_lambda
// (derivative_computer.cu:228-312) and their L counterparts (318-400).  The objective
// -sum g^_r.g^_s (render_for_illumination_only, renderer.cu:737-791) is the data part of
// the evaluation divided by w.  lambda: 3x3 FP64 Cholesky, Armijo 0.5/0.5, stop below
// t = 0.01 or -d.g < 3e-4; L: d = -g/H, Armijo 0.3/0.5, stop below t = 0.1 or -d.g < 2.5e-3.
namespace {

struct LightFrame {
    enum Phase { OUTER, SEARCH, DONE };
    Phase phase = OUTER;
    int outer = 0, inner = 0;
    double obj0 = 0, dg = 0, t = 1;
    double d[3] = {0, 0, 0};
};

// 3x3 lower Cholesky in FP64 (Solver(3), solver.h:70-105): d = -A^-1 g; false when A is not PD.
bool solve3(double A[9], const double g[3], double d[3])
{
    for (int jj = 0; jj < 3; ++jj) {
        double dd = A[jj * 3 + jj];
        for (int k = 0; k < jj; ++k) dd -= A[jj * 3 + k] * A[jj * 3 + k];
        if (!(dd > 0)) return false;
        dd = std::sqrt(dd);
        A[jj * 3 + jj] = dd;
        for (int ii = jj + 1; ii < 3; ++ii) {
            double v = A[ii * 3 + jj];
            for (int k = 0; k < jj; ++k) v -= A[ii * 3 + k] * A[jj * 3 + k];
            A[ii * 3 + jj] = v / dd;
        }
    }
    double y[3];
    for (int ii = 0; ii < 3; ++ii) {
        double v = -g[ii];
        for (int k = 0; k < ii; ++k) v -= A[ii * 3 + k] * y[k];
        y[ii] = v / A[ii * 3 + ii];
    }
    for (int ii = 2; ii >= 0; --ii) {
        double v = y[ii];
        for (int k = ii + 1; k < 3; ++k) v -= A[k * 3 + ii] * d[k];
        d[ii] = v / A[ii * 3 + ii];
    }
    return true;
}

} // namespace

// Evaluations of the light sub-fits after the first: the shape and pose are those of the last full
// evaluation of each listed frame's slot, so its synthesis, keys, visible box, visible list and
// barrier value in the workspace hold; only the shading and what follows it are redone.
// Objective, gradient and normal matrix (eval_full's shading, blur, normalised gradients and
// Jacobian stages):
// The light block of the normal equations: the light columns of the Jacobian and their J^T J,
// H [n][4][4] and g [n][4] over the four appended light unknowns (lambda, Lintensity).  The
// barrier adds nothing to this block: the light has no bounds and moves no landmark.
static int eval_light_jacobian(tdi_batch* b, FrameList fl, const float* theta, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    const Layout& Lx = c->Lx;
    Workspace& w = b->ws;
    int rc;
    if ((rc = launch_dI_light(b->m, Lx, w.frames, fl, theta, w.P0, w.pvert, w.pbary, w.npix, w.dI, w.max_area, s)))
        return rc;
    return launch_jtj_light(w.frames, fl, w.npix, w.pnb, w.pcoef, w.pixg, w.dI, c->opts.data_weight, w.H, w.g, s);
}

static int eval_light_full(tdi_batch* b, FrameList fl, const float* theta, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    const Layout& Lx = c->Lx;
    Workspace& w = b->ws;
    int rc;
    if ((rc = launch_obj_from_barrier(fl, w.objb, w.obj, s))) return rc;
    if ((rc = launch_shade(b->m, Lx, w.frames, fl, theta, w.P0, w.pvert, w.pbary, w.npix, w.box, w.img, w.max_area, s)))
        return rc;
    if ((rc = launch_blur(w.frames, fl, c->taps, w.box, w.img, w.tmp, w.blurred, w.max_area, s))) return rc;
    if ((rc = launch_grads(w.frames, fl, w.vis, w.npix, w.cum, w.blurred, b->obs_g, c->opts.data_weight, w.pixg, w.pnb,
                           w.pcoef, w.objp, w.obj, w.max_area, s)))
        return rc;
    return eval_light_jacobian(b, fl, theta, s);
}

// Objective only (eval_objective's shading, blur and objective stages):
static int eval_light_objective(tdi_batch* b, FrameList fl, const float* theta, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    Workspace& w = b->ws;
    int rc;
    if ((rc = launch_obj_from_barrier(fl, w.objb, w.obj, s))) return rc;
    if ((rc = launch_shade_keys(b->m, c->Lx, w.frames, fl, theta, w.P0, w.vproj, w.keys, w.box, w.img, s))) return rc;
    if ((rc = launch_blur(w.frames, fl, c->taps, w.box, w.img, w.tmp, w.blurred, w.max_area, s))) return rc;
    return launch_objective(w.frames, fl, w.keys, w.box, w.blurred, b->obs_g, c->opts.data_weight, w.objp, w.obj, s);
}

// The frames of `list` as the device frame list act (count, then slots).
static int upload_list(Workspace& W, int* act, const std::vector<int>& list, hipStream_t s)
{
    W.hlist.assign(1, (int)list.size());
    W.hlist.insert(W.hlist.end(), list.begin(), list.end());
    TDI_HIP(hipMemcpyAsync(act, W.hlist.data(), sizeof(int) * W.hlist.size(), hipMemcpyHostToDevice, s));
    return TDI_OK;
}

static int light_fit(tdi_batch* b, const float* theta_host, float* lam_host, float* Lint_host,
                     tdi_fit_report* reports, bool intensity)
{
    tdi_ctx* c = b->ctx;
    if (c->L.nlight) {
        set_error("light sub-fits need a context whose light is not an unknown (fit_light = 0)");
        return TDI_ERR_ARG;
    }
    const Layout& Lx = c->Lx;
    const int K = c->L.K, Kx = Lx.K, ol = Lx.off_light, Kl = Kx - ol;   // Kl: the light block's rows in H, g
    const double w = c->opts.data_weight;
    const double alpha = intensity ? 0.3 : 0.5, beta = 0.5, t_min = intensity ? 0.1 : 0.01,
                 dec_tol = intensity ? 0.0025 : 0.0003;
    const int nd = intensity ? 1 : 3, j0 = intensity ? ol + 3 : ol;
    hipStream_t s = 0;
    Workspace& W = b->ws;
    int rc = TDI_OK;
    for (int i0 = 0; i0 < b->n; i0 += W.C) {
        const int n = std::min(W.C, b->n - i0);
        if ((rc = ws_load(b, i0, n, s))) return rc;
        std::vector<float> x((size_t)n * Kx), xt;
        std::vector<LightFrame> st(n);
        for (int i = 0; i < n; ++i) {
            tdi_fit_report& rep = reports[i0 + i];
            std::memset(&rep, 0, sizeof(rep));
            float* xi = &x[(size_t)i * Kx];
            std::memcpy(xi, theta_host + (size_t)(i0 + i) * K, sizeof(float) * K);
            for (int a = 0; a < 3; ++a) xi[ol + a] = lam_host[3 * (i0 + i) + a];
            xi[ol + 3] = Lint_host[i0 + i];
        }
        std::vector<double> pk((size_t)16 * n), ob0(n), ob1(n);
        std::vector<int> list;
        bool first = true;
        for (;;) {
            // frames starting an outer iteration (10000 at most)
            list.clear();
            for (int i = 0; i < n; ++i) {
                if (st[i].phase == LightFrame::OUTER && st[i].outer >= 10000) st[i].phase = LightFrame::DONE;
                if (st[i].phase == LightFrame::OUTER) list.push_back(i);
            }
            if (list.empty()) break;
            TDI_HIP(hipMemcpyAsync(W.theta, x.data(), sizeof(float) * (size_t)n * Kx, hipMemcpyHostToDevice, s));
            if ((rc = upload_list(W, W.act_e, list, s))) return rc;
            const FrameList fl{W.act_e, n};
            // the first evaluation renders every frame of the chunk; the light changes nothing else
            if (first) {
                if ((rc = eval_render(b, fl, W.theta, 0, s, &Lx)) || (rc = eval_light_jacobian(b, fl, W.theta, s))) return rc;
            } else if ((rc = eval_light_full(b, fl, W.theta, s))) {
                return rc;
            }
            first = false;
            // the light block, gradient, objectives and pixel counts, one copy
            if ((rc = launch_misc_light_pack(n, Kl, j0 - ol, nd, W.H, W.g, W.obj, W.objb, W.npix, W.lpack, s))) return rc;
            TDI_HIP(hipMemcpyAsync(pk.data(), W.lpack, sizeof(double) * 16 * n, hipMemcpyDeviceToHost, s));
            TDI_HIP(hipStreamSynchronize(s));
            for (int i : list) {
                LightFrame& f = st[i];
                tdi_fit_report& rep = reports[i0 + i];
                const double* P = &pk[(size_t)i * 16];
                f.outer++;
                rep.n_pixels = (int)P[14];
                f.obj0 = (P[12] - P[13]) / w;
                double g[3] = {0, 0, 0};
                f.d[0] = f.d[1] = f.d[2] = 0;
                for (int a = 0; a < nd; ++a) g[a] = P[9 + a] / w;
                if (intensity) {
                    f.d[0] = -g[0] / (P[0] / w);     // compute_Lintensity_search_dir
                } else {
                    double A[9];
                    for (int a = 0; a < 3; ++a)
                        for (int bb = 0; bb < 3; ++bb) A[a * 3 + bb] = P[3 * a + bb] / w;
                    if (!solve3(A, g, f.d)) {
                        rep.stop |= TDI_STOP_NOT_PD;
                        f.phase = LightFrame::DONE;
                        continue;
                    }
                }
                f.dg = 0;
                for (int a = 0; a < nd; ++a) f.dg += f.d[a] * g[a];
                f.t = 1.0;
                f.inner = 0;
                f.phase = LightFrame::SEARCH;
            }
            // the line searches, one objective evaluation per round for every frame still searching
            for (;;) {
                list.clear();
                xt = x;
                for (int i = 0; i < n; ++i) {
                    LightFrame& f = st[i];
                    if (f.phase != LightFrame::SEARCH) continue;
                    if (f.inner >= 1000) { f.phase = LightFrame::OUTER; continue; }
                    if (f.t < t_min) {
                        reports[i0 + i].stop |= TDI_STOP_STEP;
                        f.phase = LightFrame::DONE;
                        continue;
                    }
                    float* xi = &xt[(size_t)i * Kx];
                    for (int a = 0; a < nd; ++a) xi[j0 + a] = (float)((double)x[(size_t)i * Kx + j0 + a] + f.t * f.d[a]);
                    list.push_back(i);
                }
                if (list.empty()) break;
                TDI_HIP(hipMemcpyAsync(W.theta_try, xt.data(), sizeof(float) * xt.size(), hipMemcpyHostToDevice, s));
                if ((rc = upload_list(W, W.act_e, list, s))) return rc;
                if ((rc = eval_light_objective(b, FrameList{W.act_e, n}, W.theta_try, s))) return rc;
                if ((rc = launch_misc_light_pack(n, Kl, j0 - ol, nd, W.H, W.g, W.obj, W.objb, W.npix, W.lpack, s))) return rc;
                TDI_HIP(hipMemcpyAsync(pk.data(), W.lpack, sizeof(double) * 16 * n, hipMemcpyDeviceToHost, s));
                TDI_HIP(hipStreamSynchronize(s));
                for (int i : list) {
                    ob0[i] = pk[(size_t)i * 16 + 12];
                    ob1[i] = pk[(size_t)i * 16 + 13];
                }
                for (int i : list) {
                    LightFrame& f = st[i];
                    tdi_fit_report& rep = reports[i0 + i];
                    f.inner++;
                    rep.ls_evals++;
                    if (-f.dg < dec_tol) {
                        rep.stop |= TDI_STOP_DECREMENT;
                        f.phase = LightFrame::DONE;
                    } else if ((ob0[i] - ob1[i]) / w < f.obj0 + alpha * f.t * f.dg) {
                        std::memcpy(&x[(size_t)i * Kx], &xt[(size_t)i * Kx], sizeof(float) * Kx);
                        f.phase = LightFrame::OUTER;
                    } else {
                        f.t *= beta;
                    }
                }
            }
        }
        // the objective at the result, every frame of the chunk
        TDI_HIP(hipMemcpyAsync(W.theta, x.data(), sizeof(float) * (size_t)n * Kx, hipMemcpyHostToDevice, s));
        if ((rc = eval_light_objective(b, all_frames(n), W.theta, s))) return rc;
        TDI_HIP(hipMemcpyAsync(ob0.data(), W.obj, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(ob1.data(), W.objb, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) {
            tdi_fit_report& rep = reports[i0 + i];
            rep.objective = (float)((ob0[i] - ob1[i]) / w);
            rep.iters = st[i].outer;
            rep.status = (rep.stop & TDI_STOP_NOT_PD) ? TDI_NOT_PD : TDI_OK;
            for (int a = 0; a < 3; ++a) lam_host[3 * (i0 + i) + a] = x[(size_t)i * Kx + ol + a];
            Lint_host[i0 + i] = x[(size_t)i * Kx + ol + 3];
        }
    }
    return TDI_OK;
}

int tdi_ctx_layout(tdi_ctx* c, int32_t* layout)
{
    if (!c || !layout) { set_error("tdi_ctx_layout: null argument"); return TDI_ERR_ARG; }
    const Layout& L = c->L;
    layout[0] = L.off_tex; layout[1] = L.off_id; layout[2] = L.off_exp;
    layout[3] = L.off_pose; layout[4] = L.off_light; layout[5] = L.K;
    return TDI_OK;
}

int32_t tdi_batch_size(tdi_batch* b) { return b ? b->n : 0; }

int32_t tdi_batch_unknowns(tdi_batch* b) { return b ? b->ctx->L.K : 0; }

int tdi_batch_set_light(tdi_batch* b, const float* lam, const float* Lint)
{
    if (!b || !lam || !Lint) { set_error("set_light: null argument"); return TDI_ERR_ARG; }
    for (int i = 0; i < b->n; ++i) {
        for (int a = 0; a < 3; ++a) b->hframes[i].lam[a] = lam[3 * i + a];
        b->hframes[i].Lint = Lint[i];
    }
    return TDI_OK;
}

int tdi_fit_lambdas(tdi_batch* b, const float* theta, float* lam, float* Lint, tdi_fit_report* reports)
{
    return light_fit(b, theta, lam, Lint, reports, false);
}

int tdi_fit_lintensity(tdi_batch* b, const float* theta, float* lam, float* Lint, tdi_fit_report* reports)
{
    return light_fit(b, theta, lam, Lint, reports, true);
}

} // extern "C"

// ---------------------------------------------------------------------------- batched fit
// One GN iteration for every frame of the workspace, all decisions on the device.
static int record_fit(tdi_batch* b, hipStream_t s)
{
    tdi_ctx* c = b->ctx;
    const tdi_fit_options& o = c->opts;
    Workspace& w = b->ws;
    const int n = w.C, K = c->L.K, off_tz = c->L.off_pose + 2;
    // the finetune fit's trials keep the pose (model_fitter.cu:502)
    const int hold0 = b->finetune ? c->L.off_pose : 0, hold1 = b->finetune ? c->L.off_pose + 6 : 0;
    // evaluation rounds of the line search: steps t = beta^k >= t_min, capped by max_ls
    int rounds = 0;
    for (double t = 1.0; t >= o.t_min && rounds < o.max_ls; t *= o.armijo_beta) ++rounds;
    int rc;
    for (int it = 0; it < o.max_iters; ++it) {
        if ((rc = launch_copy_mask(n, w.st, w.mask, s))) return rc;
        if ((rc = launch_compact_frames(n, w.mask, w.act_m, s))) return rc;
        // A frame whose last line-search evaluation was accepted has the render of its new theta
        // in its own workspace slots (later rounds render only the frames still searching): its
        // synthesis, keys, visible box, shaded and blurred images are reused.  The others (the
        // first iteration, a search that ended on tz < 0 or on the round cap) are rendered.
        if ((rc = launch_split_fresh(n, w.st, w.mask_r, w.mask_u, s))) return rc;
        if ((rc = launch_compact_frames(n, w.mask_r, w.act_r, s))) return rc;
        if ((rc = launch_compact_frames(n, w.mask_u, w.act_u, s))) return rc;
        if ((rc = eval_render(b, FrameList{w.act_r, w.C}, w.theta, o.mixed_flags, s))) return rc;
        if ((rc = eval_reused(b, FrameList{w.act_u, w.C}, w.theta, s))) return rc;
        if ((rc = eval_jacobian(b, FrameList{w.act_m, w.C}, w.theta, o.mixed_flags, s))) return rc;
        if ((rc = launch_cholesky(n, K, w.H, w.g, w.delta, w.status, (o.mixed_flags & TDI_MIXED_CHOL_F32) ? 1 : 0,
                                  w.mask, s)))
            return rc;
        if ((rc = launch_ls_init(n, K, w.st, w.status, w.delta, w.g, w.obj, w.searching, s))) return rc;
        for (int r = 0; r < rounds; ++r) {
            if ((rc = launch_ls_propose(n, K, off_tz, o.t_min, w.st, w.theta, w.delta, w.theta_try, w.searching,
                                        w.evalmask, hold0, hold1, s)))
                return rc;
            if ((rc = launch_compact_frames(n, w.evalmask, w.act_e, s))) return rc;
            if ((rc = eval_objective(b, FrameList{w.act_e, w.C}, w.theta_try, o.mixed_flags, s))) return rc;
            if ((rc = launch_ls_decide(n, K, w.st, w.evalmask, w.searching, w.obj, w.objb, w.face_ok, w.theta,
                                       w.theta_try, o, s)))
                return rc;
        }
        if ((rc = launch_ls_propose(n, K, off_tz, o.t_min, w.st, w.theta, w.delta, w.theta_try, w.searching, w.evalmask,
                                    hold0, hold1, s)))
            return rc;
    }
    return TDI_OK;
}

static int fit_batched(tdi_batch* b, float* theta, tdi_fit_report* reports, void* stream)
{
    tdi_ctx* c = b->ctx;
    const tdi_fit_options& o = c->opts;
    Workspace& w = b->ws;
    const int K = c->L.K;
    // All work runs on the batch's own stream (graph capture is not allowed on the legacy
    // default stream); it is ordered after the caller's stream and the caller's stream is
    // ordered after it.
    hipStream_t caller = (hipStream_t)stream;
    hipStream_t s = b->stream;
    TDI_HIP(hipEventRecord(b->ev_in, caller));
    TDI_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    const bool use_graph = getenv("TDI_NO_GRAPH") == nullptr;
    std::vector<FitState> hst(w.C);
    for (int i0 = 0; i0 < b->n; i0 += w.C) {
        const int n = std::min(w.C, b->n - i0);
        int rc = ws_load(b, i0, n, s);
        if (rc) return rc;
        for (int i = 0; i < w.C; ++i) {
            FitState f{};
            f.active = i < n;
            hst[i] = f;
        }
        TDI_HIP(hipMemcpyAsync(w.st, hst.data(), sizeof(FitState) * w.C, hipMemcpyHostToDevice, s));
        TDI_HIP(hipMemcpyAsync(w.theta, theta + (size_t)i0 * K, sizeof(float) * n * K, hipMemcpyDeviceToDevice, s));
        if (use_graph) {
            // one captured schedule per kind of fit (the finetune fit's bounds and pose hold differ)
            hipGraphExec_t& ge = b->finetune ? b->graph_fine : b->graph;
            if (!ge) {
                hipGraph_t g;
                TDI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                rc = record_fit(b, s);
                hipError_t ec = hipStreamEndCapture(s, &g);
                if (rc) return rc;
                if (ec != hipSuccess) { set_error("graph capture failed: %s", hipGetErrorString(ec)); return TDI_ERR_HIP; }
                TDI_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
            }
            TDI_HIP(hipGraphLaunch(ge, s));
        } else {
            if ((rc = record_fit(b, s))) return rc;
        }
        TDI_HIP(hipMemcpyAsync(theta + (size_t)i0 * K, w.theta, sizeof(float) * n * K, hipMemcpyDeviceToDevice, s));
        TDI_HIP(hipMemcpyAsync(hst.data(), w.st, sizeof(FitState) * w.C, hipMemcpyDeviceToHost, s));
        if (reports) {
            // objective at the returned parameters
            if ((rc = set_mask_all(b, s))) return rc;
            if ((rc = eval_render(b, all_frames(w.n_cur), w.theta, c->opts.mixed_flags, s))) return rc;
        }
        std::vector<double> hobj(w.C);
        std::vector<int> hnp(w.C);
        TDI_HIP(hipMemcpyAsync(hobj.data(), w.obj, sizeof(double) * w.C, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipMemcpyAsync(hnp.data(), w.npix, sizeof(int) * w.C, hipMemcpyDeviceToHost, s));
        TDI_HIP(hipStreamSynchronize(s));
        if (reports) {
            for (int i = 0; i < n; ++i) {
                const FitState& f = hst[i];
                tdi_fit_report& r = reports[i0 + i];
                std::memset(&r, 0, sizeof(r));
                r.stop = f.stop | (f.active ? TDI_STOP_MAXITER : 0);
                r.iters = f.iters;
                r.ls_evals = f.ls_evals;
                r.n_pixels = hnp[i];
                const bool term = (f.stop & (TDI_STOP_STEP | TDI_STOP_DECREMENT)) != 0;
                r.success = (f.iters + (term && f.iters < o.max_iters ? 1 : 0)) > o.niters_success;
                r.status = f.status;
                r.objective = (float)hobj[i];
            }
        }
    }
    TDI_HIP(hipEventRecord(b->ev_out, s));
    TDI_HIP(hipStreamWaitEvent(caller, b->ev_out, 0));
    return TDI_OK;
}

extern "C" int tdi_fit_batched(tdi_batch* b, float* theta, tdi_fit_report* reports, void* stream)
{
    if (!b || !theta) { set_error("tdi_fit_batched: null argument"); return TDI_ERR_ARG; }
    b->finetune = 0;
    return fit_batched(b, theta, reports, stream);
}

extern "C" int tdi_fit_batched_finetune(tdi_batch* b, float* theta, tdi_fit_report* reports, void* stream)
{
    if (!b || !theta) { set_error("tdi_fit_batched_finetune: null argument"); return TDI_ERR_ARG; }
    b->finetune = 1;
    const int rc = fit_batched(b, theta, reports, stream);
    b->finetune = 0;
    return rc;
}
