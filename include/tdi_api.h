/*
 * tdi_api.h — C ABI of the MI355X (gfx950) 3DMM per-frame fitting step.
 *
 * The drop-in boundary.  The reference has no FFI for this path: its fitting
 * step is the C++ library `open3DI` (src/CMakeLists.txt:41) whose entry points
 * take cuBLAS/cuSOLVER handles and raw device pointers (src/lib/model_fitter.h:29-79).
 * This header is the thin, type-clean layer beneath the C++ host layer
 * (include/tdi_host.h, 3di_amd/host/tdi_host.cpp): plain pointers, explicit sizes, int status
 * codes, a hipStream_t passed as void*, no exceptions across the ABI, no torch
 * types.  Each entry point names the reference interface it replaces.
 *
 * Conventions (shared by the HIP library and by the CPU oracle in oracle/):
 *   - vertices are AoS float [V][3]; bases are row-major [3V][K] with row 3*v+axis;
 *   - triangles are int32 [NT][3] in the reference's (tl0,tl1,tl2) = (b,c,a) order,
 *     0-based (renderer.cu:194-205);
 *   - images are column-major crops, index (x-x0)*H + (y-y0)
 *     (the reference transposes its crop, video_fitter.cu:542-549);
 *   - the per-frame unknown vector follows OptimizationVariables
 *     (Optimizer.h:71-91, T=1): [beta | alpha | epsilon | tau_x tau_y tau_z u1 u2 u3 | light],
 *     a block present only when its fit_* flag is set; light = (lambda1..3, L) with
 *     fit_light = TDI_LIGHT_POINT, or (a1..3, L, s1..s5) with fit_light = TDI_LIGHT_SH2.
 *
 * Shading (newfuncs.cu:122-137), with n_c = R n the camera-space unit normal and v_c the
 * camera-space surface point:
 *   TDI_LIGHT_POINT (the reference): I = rho (1 + L n_c.(lambda - v_c));
 *   TDI_LIGHT_SH2: I = rho (1 + n_c.a - L n_c.v_c + sum_k s_k Y_k(n_c)), with
 *     Y_1 = nx ny, Y_2 = ny nz, Y_3 = nx nz, Y_4 = nx^2 - ny^2, Y_5 = 3 nz^2 - 1
 *   the five second-order spherical-harmonic terms.  With s = 0 this is the reference's light
 *   with a = L lambda: the same model written linearly in its unknowns.  Fitted jointly, the
 *   reference's (lambda, L) is degenerate, since a distant light (|lambda| -> inf, L -> 0 with
 *   L lambda fixed) shades like a directional one and the normal matrix becomes singular;
 *   (a, L) has no such direction.  The constant term stays 1: the normalised-gradient data
 *   term is blind to the scale of the image.
 */
#ifndef TDI_API_H_
#define TDI_API_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDI_NLANDMARKS 51
#define TDI_ABI_VERSION 1

/* Status codes: per call, and per frame in batched fits. */
enum tdi_status {
    TDI_OK = 0,
    TDI_ERR_ARG = 1,          /* bad size / null pointer / unsupported option             */
    TDI_ERR_HIP = 2,          /* a HIP runtime call failed                                */
    TDI_ERR_NOMEM = 3,        /* device allocation failed                                 */
    TDI_NOT_PD = 4,           /* Cholesky met a non-positive pivot (solver.h:86-92)        */
    TDI_NAN = 5,              /* NaN search direction (model_fitter.cu:206-213)            */
    TDI_OUT_OF_FRAME = 6,     /* landmarks outside the image (video_fitter.cu:187-208)     */
    TDI_TOO_FEW_PIXELS = 7,   /* nothing rendered inside the crop                          */
    TDI_LMK_INFEASIBLE = 8    /* the landmark-only fit did not bring its slack below -0.01
                                 (Logbarrier_Initializer::fit_model returns false)         */
};

/* Per-frame termination reasons reported by the fits (bit flags). */
enum tdi_stop {
    TDI_STOP_MAXITER = 1,     /* iteration budget used up                                  */
    TDI_STOP_DECREMENT = 2,   /* -delta.g < decrement_tol (model_fitter.cu:341)            */
    TDI_STOP_STEP = 4,        /* step t fell below t_min (model_fitter.cu:241)             */
    TDI_STOP_NAN = 8,         /* NaN direction                                             */
    TDI_STOP_NOT_PD = 16      /* Cholesky failure                                          */
};

/* Shared model: mesh, bases and coefficient bounds (Renderer ctor, renderer.cu:22-442;
 * Logbarrier_Initializer bounds, logbarrier_initializer.cu:270-286). */
typedef struct tdi_model_desc {
    int32_t n_vertices;
    int32_t n_triangles;
    int32_t k_id, k_tex, k_exp;
    const float* mean_shape;     /* [V][3]                                   */
    const float* id_basis;       /* [3V][k_id]  (NULL if k_id == 0)          */
    const float* exp_basis;      /* [3V][k_exp]                              */
    const float* tex_mean;       /* [V]                                      */
    const float* tex_basis;      /* [V][k_tex] (NULL if k_tex == 0)          */
    const int32_t* triangles;    /* [NT][3]                                  */
    const int32_t* landmark_ids; /* [51] vertex ids of the 51 landmarks      */
    const float* id_lb;  const float* id_ub;    /* [k_id]  */
    const float* tex_lb; const float* tex_ub;   /* [k_tex] */
    const float* exp_lb; const float* exp_ub;   /* [k_exp] */
} tdi_model_desc;

/* One frame: camera, crop, observation, landmarks, landmark box, fixed light. */
typedef struct tdi_frame_desc {
    float phix, phiy, cx, cy;          /* pinhole intrinsics in image pixels (camera.h:15-60)   */
    int32_t x0, y0, width, height;     /* crop origin and size in image pixels                   */
    float face_size;                   /* compute_face_size of the landmarks (funcs.cu:44-53)    */
    const float* obs;                  /* [width][height] raw grey crop, column-major; host or
                                          device memory (tdi_prepare_crops output)               */
    const float* lmk_x;                /* [51] detected landmarks, image pixels                  */
    const float* lmk_y;
    const float* box_lx;               /* [51] landmark box, face-size units                     */
    const float* box_ux;               /*      (logbarrier_initializer.cu:501-507)               */
    const float* box_ly;
    const float* box_uy;
    float lambda[3];                   /* light position when light is held fixed                 */
    float Lintensity;                  /* (video_fitter.cu:115-116 initial values)                */
} tdi_frame_desc;

/* Fit options.  Defaults (tdi_default_options) are the reference's constants. */
typedef struct tdi_fit_options {
    int32_t fit_id, fit_tex, fit_exp;  /* unknown blocks                                           */
    int32_t fit_light;          /* 0: light fixed per frame (tdi_frame_desc.lambda, Lintensity);
                                   TDI_LIGHT_POINT: (lambda, L) are 4 unknowns;
                                   TDI_LIGHT_SH2: (lambda, L, s1..s5) are 9 unknowns             */
    int32_t max_iters;          /* GN iterations (model_fitter.cu:163: 100)                        */
    int32_t max_ls;             /* line-search evaluations per iteration (model_fitter.cu:224: 100) */
    int32_t ntmp_cap;           /* pixels kept per triangle, 0 = all (constants.h:19 uses 4)       */
    int32_t use_landmarks;      /* landmark-box log barrier on/off                                  */
    int32_t niters_success;     /* success if iterations > this (config.cu:99: 6)                  */
    float data_weight;          /* w = tau_logbarrier = 5*5 (model_fitter.cu:137-159)              */
    float armijo_alpha;         /* 0.4 (model_fitter.cu:219)                                       */
    float armijo_beta;          /* 0.8 (model_fitter.cu:220)                                       */
    float t_min;                /* 0.075 for T=1 (model_fitter.cu:149-151)                          */
    float decrement_tol;        /* 3e-4 (model_fitter.cu:341)                                      */
    float blur_sigma;           /* KERNEL_SIGMA config key (config.cu:53; cfg files use 0.75)      */
    float ref_face_size;        /* REF_FACE_SIZE; "face too large" if > 2x (renderer.cu:845-857), 0 = off */
    int32_t mixed_flags;        /* batched path arithmetic (no reference counterpart), bit set:
                                   1 = FP64-MFMA basis GEMM over frames, 2 = BF16 hi/lo MFMA J^T J,
                                   4 = FP32 Cholesky + FP64 refinement, 8 = J^T r from the BF16
                                   MFMA tiles (else FP32; with 8 a NaN observed pixel contributes
                                   zero instead of stopping the frame on a NaN direction),
                                   16 = two-term hi/lo split of J
                                   (else three-term hi/mid/lo).  Default 7.               */
    float finetune_coef;        /* FINETUNE_COEF (config.cu:100: 1.5): the expression bounds of the
                                   finetune fits are these times the regular ones
                                   (logbarrier_initializer.cu:285-286)                              */
} tdi_fit_options;

#define TDI_LIGHT_POINT 1
#define TDI_LIGHT_SH2 2

#define TDI_MIXED_SYNTH_MFMA 1
#define TDI_MIXED_JTJ_BF16 2
#define TDI_MIXED_CHOL_F32 4
#define TDI_MIXED_JTR_BF16 8
#define TDI_MIXED_TWO_TERM 16

/* Per-frame fit report. */
typedef struct tdi_fit_report {
    int32_t status;        /* tdi_status                        */
    int32_t stop;          /* tdi_stop flags                    */
    int32_t iters;         /* outer GN iterations started       */
    int32_t ls_evals;      /* line-search objective evaluations */
    int32_t n_pixels;      /* rendered pixels at the last render */
    int32_t success;       /* iters > niters_success (model_fitter.cu:374) */
    float objective;       /* F at the returned theta           */
    float pad;
} tdi_fit_report;

typedef struct tdi_ctx tdi_ctx;       /* model resident in HBM                        */
typedef struct tdi_batch tdi_batch;   /* frames resident in HBM + per-frame scratch    */

/* Number of unknowns per frame for a model/options pair. */
int32_t tdi_num_unknowns(const tdi_model_desc* model, const tdi_fit_options* opts);
/* The unknown-vector layout of a context: layout[0..5] = offsets of the albedo, identity,
 * expression, pose and light blocks, then K. */
int tdi_ctx_layout(tdi_ctx* ctx, int32_t* layout);

/* Fill *opts with the reference defaults (config.cu:15-140, model_fitter.cu:137-345). */
void tdi_default_options(tdi_fit_options* opts);

/* Blur taps, 15 entries, for a KERNEL_SIGMA value (video_fitter.cu:51-63).  Returns
 * TDI_ERR_ARG for a sigma the reference has no table for. */
int tdi_blur_taps(float sigma, float taps[15]);

/* ---- Context: replaces Renderer/Optimizer/Solver construction (renderer.cu:22-442,
 *      Optimizer.h:249-405, solver.h:20-68, video_fitter.cu:16-81). ---- */
int tdi_ctx_create(const tdi_model_desc* model, const tdi_fit_options* opts, int device, tdi_ctx** out);
int tdi_ctx_destroy(tdi_ctx* ctx);
/* A context keeps the device blocks of its destroyed batches for its later batches (workspaces of
 * the same sizes are taken back instead of allocated again; at most a quarter of the device's
 * memory is kept, and all of it is freed with the context or when an allocation fails).  This
 * frees the kept blocks now, for a long-lived context between workloads; batches in use keep
 * theirs.  No counterpart in the reference (its objects allocate once per run). */
int tdi_ctx_release_cached(tdi_ctx* ctx);
const char* tdi_last_error(void);

/* ---- Batch: frames uploaded once; the observed crop is blurred once and its normalised
 *      gradients precomputed (video_fitter.cu:537-550, funcs.cu:876-896). ---- */
int tdi_batch_create(tdi_ctx* ctx, int32_t n_frames, const tdi_frame_desc* frames, tdi_batch** out);
int tdi_batch_destroy(tdi_batch* batch);

/* ---- Building blocks (all on device pointers, stream-ordered) ---- */

/* R(u) and dR/du_i, column-major 3x3 — RotationComputer::process (rotation_header.cu:46-149).
 * u: [n][3], R: [n][9], dR: [n][27] (may be NULL). */
int tdi_rotation(int32_t n, const float* u, float* R, float* dR, void* stream);

/* Vertex synthesis + view transform + projection for every frame —
 * Renderer::compute_nonrigid_shape2 (renderer.cu:806-861), view_transform_3d_pts_and_render_2d
 * (funcs.cu:395-453).  theta: [n][K] device.  xyz: [n][V][3] camera space, xyp: [n][V][2] pixels. */
int tdi_synthesize(tdi_batch* b, const float* theta, float* xyz, float* xyp, void* stream);

/* Rasterise projected vertices into packed depth|triangle keys —
 * get_pixels_to_render + atomicMinFloat + keep_only_minZ + remove_if (funcs.cu:900-1029,
 * renderer.cu:657-667).  Stand-alone form used by the visibility parity tests: xyz/xyp as
 * produced by tdi_synthesize (or any float input), keys: [sum W*H] u64 in crop order,
 * bary: [sum W*H][2] (alpha,beta) of the winning triangle.  Pixel offsets per frame are
 * the exclusive prefix sum of W*H over the batch. */
int tdi_rasterize(tdi_batch* b, const float* xyz, const float* xyp, uint64_t* keys, float* bary, void* stream);

/* Render + objective at theta: Renderer::render (renderer.cu:635-730) plus the barrier
 * objective (logbarrier_initializer.cu:1921-2081).  img: [sum W*H] rendered crop before the
 * blur (fill_tex_im1, funcs.cu:800-817), blurred: [sum W*H] the same after the 15-tap blur
 * (convolutionRowsGPU/ColumnsGPU, newfuncs.cu:539-715), both zero-padded crops in the batch's
 * crop order (either may be NULL); obj: [n]; npix: [n] rendered pixel counts. */
int tdi_render(tdi_batch* b, const float* theta, float* img, float* blurred, float* obj, int32_t* npix,
               void* stream);

/* Normal equations at theta, FP32 accumulation: H [n][K][K], g [n][K], obj [n] —
 * get_obj_hessian_and_gradient_multiframe + DerivativeComputer::compute_hessian_and_gradient
 * (model_fitter.cu:21-116, derivative_computer.cu:22-213) with the barrier terms merged. */
int tdi_normal_equations(tdi_batch* b, const float* theta, float* H, float* g, float* obj, void* stream);

/* Same, with J split into BF16 hi/lo parts and J^T J, J^T r on MFMA (FP32 accumulate). */
int tdi_normal_equations_mixed(tdi_batch* b, const float* theta, float* H, float* g, float* obj, void* stream);

/* Above 111 unknowns per frame (the identity stage: albedo + identity + expression + pose +
 * light) the normal equations take J^T J and J^T r on MFMA over 64 x 64 tiles (BF16 hi/mid/lo
 * parts, FP32 accumulation, FP64 sums; both entry points above) and the barrier's derivative
 * rows in scratch the batch allocates on first use.
 *
 * Identity stage: the joint fit of the batch's frames, which share the albedo and identity
 * coefficients — the shape fit of fit_to_multi_images (model_fitter.cu:1252-1479), i.e.
 * fit_3DMM_shape_rigid over T frames (model_fitter.cu:121-375) with the frames' products summed
 * into one matrix as get_obj_hessian_and_gradient_multiframe does (model_fitter.cu:21-116).
 * The shared block is the first Ks = k_tex + k_id unknowns of every frame's layout; the joint
 * vector is [shared | frame 0's expression, pose, light | frame 1's | ...], KJ = Ks + n (K - Ks).
 * The coefficient bounds of the shared block count once, every frame's landmark barrier and
 * expression bounds once per frame.  All frames must fit one workspace chunk. */
int64_t tdi_joint_unknowns(tdi_batch* b);

/* Joint normal matrix HJ [KJ][KJ] FP64, gradient gJ [KJ] FP64 and objective obj [1] FP64 (may be
 * NULL) at the frames' theta [n][K] (DEVICE; the shared block is read from frame 0). */
int tdi_normal_equations_joint(tdi_batch* b, const float* theta, double* HJ, double* gJ, double* obj, void* stream);

/* The joint Gauss-Newton fit: theta [n][K] HOST in/out (the shared block is read from frame 0 and
 * written to every frame).  Per iteration the joint system is assembled on the device and solved
 * by the blocked FP64 Cholesky; Armijo line search on the summed objective with the reference's
 * constants (ALPHA 0.4, BETA 0.8, -d.g < 3e-4, step threshold 0.001 for several frames, a
 * negative depth of frame 0 ends the search).  report: one tdi_fit_report for the whole fit
 * (n_pixels summed over the frames). */
int tdi_fit_multiframe(tdi_batch* b, float* theta, tdi_fit_report* report);

/* Batched Cholesky solve of H delta = -g.  mode 0: FP64 factorisation (Solver::solve,
 * solver.h:70-105, cusolverDnDpotrf/Dpotrs); mode 1: FP32 factorisation of the equilibrated
 * matrix D^-1/2 H D^-1/2 + one FP64 residual-refinement step, with the FP64 factorisation
 * as the fallback of a frame whose FP32 factorisation breaks down.  status: [n] tdi_status.
 * K <= 112: one workgroup per matrix in LDS.  112 < K <= 4096 (the joint identity-stage system,
 * 199 + 199 + 94 per frame): blocked FP64 factorisation of an FP64 copy in HBM (32-column panels,
 * the trailing update over a grid of tiles), mode ignored. */
int tdi_cholesky_solve(int32_t n, int32_t K, const float* H, const float* g, float* delta,
                       int32_t* status, int32_t mode, void* stream);

/* Landmark-driven initialisation of n frames, one wavefront per frame in FP64 —
 * OrthographicInitializer::fit_model (orthographic_initializer.cu:243-377): a 7-parameter
 * Gauss-Newton fit of x_orth = (tau_x, tau_y, sigma_x, sigma_y, u) to the mean-centred
 * landmarks against the model's mean shape at its 51 landmark vertices; then, as in
 * Logbarrier_Initializer::initialize_with_orthographic_t (logbarrier_initializer.cu:1862-1914),
 * a refit from the mirrored pose when sigma_x < 0 and the perspective pose of initialize_kernel
 * (logbarrier_initializer.cu:31-46).  Device pointers: lmk [n][2][51] (x row, then y row, image
 * pixels); cams [n][4] (phix, phiy, cx, cy); outputs x_orth [n][7], pose [n][6] = (tau_x,
 * tau_y, tau_z, u), angles [n][3] = (yaw, pitch, roll) in degrees of the last line-search
 * rotation (may be NULL), info [n][4] = (outer iterations, refit iterations, refit from the
 * mirrored pose, tdi_status).  A frame whose 7x7 normal matrix is not positive definite stops
 * with TDI_NOT_PD (the reference carries on with the failed factorisation). */
int tdi_ortho_init(tdi_ctx* ctx, int32_t n, const float* lmk, const float* cams, float* x_orth, float* pose,
                   float* angles, int32_t* info, void* stream);

/* Landmark-only log-barrier fit with slack, the step between the orthographic start and the
 * first shape fit of every frame (video_fitter.cu:470-486: Logbarrier_Initializer::
 * set_minimal_slack, logbarrier_initializer.cu:1041-1090, then fit_model with use_slack,
 * 1095-1267; the object li_init of video_fitter.cu:30: no identity, k_exp expression
 * coefficients, T = 1).  Unknowns z = (eps[0..k_exp), tau_x, tau_y, tau_z, u, s); objective
 *   tau s + sum over the 51 landmarks and 4 box sides of -log G + sum_k -log(ub_k - eps_k) - log(eps_k - lb_k),
 * G = s + the signed distance of the projected landmark inside its box side (compute_Gs,
 * logbarrier_initializer.cu:458-518), with the Gauss-Newton Hessian of
 * compute_gradient_and_hessian (1272-1842).  Start: eps = 0, the pose of theta, s = 5 - min G
 * at s = 0.  Rounds tau = 100, 1e3, ..., 1e6, each up to 100 Newton steps (Cholesky; a failed
 * factorisation ends the round) with a backtracking search (t = 1, halved, a step accepted when
 * the objective decreases, the round ends below t = 0.01); the fit succeeds when s < -0.01 is
 * seen at the top of an iteration: every landmark then lies 0.01 px inside its box.
 * One workgroup per frame runs the whole fit in FP64 in one launch.
 * theta: DEVICE [n][K], in/out: the pose is the start; on success the pose and the first k_exp
 * expression coefficients are overwritten (the others are not fitted and stay), otherwise theta
 * is left as it is.  Identity coefficients of theta (if the layout has them) are held fixed.
 * The expression bounds are the model's.  0 < k_exp <= the layout's expression count, k_exp <= 89.
 * reports: HOST [n] (may be NULL).  Returns when the fit is done. */
typedef struct tdi_lmk_report {
    int32_t status;        /* TDI_OK (s < -0.01 reached) or TDI_LMK_INFEASIBLE         */
    int32_t iterations;    /* Newton steps computed, over all rounds                   */
    int32_t rounds;        /* barrier rounds started                                   */
    int32_t accepted;      /* steps the line search accepted                           */
    float slack0;          /* the start of s (set_minimal_slack)                       */
    float slack;           /* s at the end                                             */
} tdi_lmk_report;

int tdi_fit_landmarks_slack(tdi_batch* b, float* theta, int32_t k_exp, tdi_lmk_report* reports, void* stream);

/* The terms of that fit at given unknowns, for checks: z DEVICE [n][k_exp + 7] (FP64), theta
 * DEVICE [n][K] (its identity coefficients, if any, enter the landmark positions), tau the
 * barrier weight; outputs (DEVICE, any may be NULL) H [n][k_exp + 7]^2 (Gauss-Newton Hessian,
 * FP64), g [n][k_exp + 7], obj [n] (NaN or +inf outside the feasible set). */
int tdi_landmark_slack_terms(tdi_batch* b, const float* theta, int32_t k_exp, const double* z, double tau, double* H,
                             double* g, double* obj, void* stream);

/* The joint landmark-only fit over a group of frames (the batch's n frames), replacing
 * Logbarrier_Initializer::fit_model over T frames with identity unknowns as the identity stage
 * runs it: fit_to_multi_images_landmarks_only (model_fitter.cu:1205-1247; with the slack, 60
 * identity coefficients, the group kept when it succeeds) and the start of fit_to_multi_images
 * (1291-1330: the same, then a fit without the slack over all identity coefficients from its
 * result).  The frames share the first k_id identity coefficients (and the slack); each has its
 * own first k_exp expression coefficients and pose.  Joint unknowns
 *   z = [alpha (k_id) | s (if slack) | frame 0 (eps, tau, u) | frame 1 | ...],
 * objective: tau_b s (with the slack) + the identity bounds + for every frame its landmark terms
 * (as tdi_fit_landmarks_slack's, with the identity columns in the landmark positions) and its
 * expression bounds; Gauss-Newton Hessian and gradient as compute_gradient_and_hessian
 * (logbarrier_initializer.cu:1272-1842) sums them over the frames.  The loop is the single-frame
 * fit's (rounds tau_b = 100 .. 1e6 of up to 100 Newton steps, halving line search from t = 1);
 * with the slack, s starts at the largest over the frames of 5 - min G at s = 0 and the fit
 * succeeds when s < -0.01; without it s = 0 throughout, and every round runs until its line
 * search fails (the reference keeps that fit's result whatever happens).  Identity
 * coefficients past k_id enter the landmark positions at frame 0's values of theta.
 * theta: HOST [n][K] in/out: the start (frame 0's identity, every frame's expressions and pose);
 * with the slack on success, and always without it, every frame's first k_id identity
 * coefficients (the shared values), first k_exp expression coefficients and pose are
 * overwritten.  k_id <= the layout's identity count, k_exp <= min(the layout's expression
 * count, 192), k_id + slack + k_exp + 6 <= 320 and the joint count <= 4096.  report: HOST,
 * status TDI_OK (without the slack always) or TDI_LMK_INFEASIBLE.  Synchronous. */
int tdi_fit_landmarks_joint(tdi_batch* b, float* theta, int32_t k_id, int32_t k_exp, int32_t slack,
                            tdi_lmk_report* report);

/* That fit with the slack for every frame of the batch on its own (each frame a group of one: the
 * per-frame landmark fits of the identity stage's selection, video_fitter.cu:2253-2275,
 * 2330-2345, where li_init carries K_ALPHA_L identity coefficients and is reset for every frame).
 * The same steps as tdi_fit_landmarks_joint on a one-frame batch, for all frames together: each
 * frame keeps its own barrier round, Newton step and line-search trial, each evaluation round is
 * one launch over the frames that need it, the Newton steps one batched large-K Cholesky.
 * theta: HOST [n][K] in/out (each frame's own identity, expressions and pose are its start;
 * overwritten for the frames that succeed); reports: HOST [n].  Synchronous. */
int tdi_fit_landmarks_each(tdi_batch* b, float* theta, int32_t k_id, int32_t k_exp, tdi_lmk_report* reports);

/* The terms of that fit at joint unknowns z (HOST, FP64, the order above), for checks: theta
 * HOST [n][K] (frame 0's identity past k_id), tau the barrier weight; outputs HOST: H [N][N]
 * (Gauss-Newton, FP64), g [N], obj (NaN or +inf outside the feasible set);
 * N = k_id + slack + n (k_exp + 6). */
int tdi_landmark_joint_terms(tdi_batch* b, const float* theta, int32_t k_id, int32_t k_exp, int32_t slack,
                             const double* z, double tau, double* H, double* g, double* obj);

/* The integer buffers the fit itself computes, for bit-for-bit checks: a render at theta through
 * the same launches as a fit with these mixed_flags (TDI_MIXED_SYNTH_MFMA selects the FP64-MFMA
 * synthesis the batched fit uses), then
 *   vproj [n][V][4] the projected vertices (x, y in image pixels, camera z, 0) the rasteriser
 *   reads -- view_transform_3d_pts_and_render_2d (funcs.cu:440-452);
 *   nbr   [sum W*H][4] per rendered pixel, in the frame's crop order from its crop offset, the
 *   sorted positions of its left, right, above and below neighbours -- fill_krels/fill_krels2
 *   (funcs.cu:288-340, derivative_computer.cu:43-72); entries past the frame's pixel count are
 *   unspecified;
 *   npix  [n] rendered pixel counts.
 * Device pointers; any output may be NULL. */
int tdi_fit_buffers(tdi_batch* b, const float* theta, int32_t mixed_flags, float* vproj, int32_t* nbr,
                    int32_t* npix, void* stream);

/* The log-barrier terms alone at theta, as the fits add them -- Logbarrier_Initializer::
 * compute_gradient_and_hessian and evaluate_objective_function (logbarrier_initializer.cu:
 * 1272-1842, 1921-2081): H [n][K][K] (FP64), g [n][K] (the fits keep g in FP32; returned
 * widened), obj [n].  Device pointers; any may be NULL. */
int tdi_barrier_terms(tdi_batch* b, const float* theta, double* H, double* g, double* obj, void* stream);

/* The render loss and its adjoint (vector-Jacobian product), for differentiation (the repo's
 * Python driver binds it behind a torch.autograd.Function).  No counterpart entry point in the
 * reference, whose fit forms the Jacobian itself (newfuncs.cu:173-463, derivative_computer.cu:
 * 141-172); this is its transpose applied to a vector, without J.  Per frame:
 *   loss = sum over the visible pixels of |g^_r - g^_s|^2 (the normalised gradients of the
 *          blurred render and of the blurred observation, fill_grads, funcs.cu:842-897; equal to
 *          (2/w) times the fit's data objective plus a constant while |g^| = 1)
 *        + sum over the 51 landmarks of the squared image-pixel distance between the projected
 *          landmark vertex and the detected landmark;
 * the model-space vertices are the synthesis plus optional per-vertex offsets dverts.  The
 * render, rasterisation included, is recomputed from theta (and dverts) on every call; the
 * gradients hold the visibility (the triangle of each pixel) fixed and are exact derivatives of
 * the rendered values, with the barycentrics of each pixel centre following the projected
 * triangle.  Device pointers: theta [n][K]; dverts [n][V][3] (may be NULL); gloss [n] upstream
 * gradient of each frame's loss (NULL = 1); outputs (any may be NULL) loss [n], gtheta [n][K]
 * (pose, expression, identity and light unknowns; albedo unknowns get 0), gverts [n][V][3]
 * (d loss / d model-space vertices).  With gtheta and gverts NULL only the loss is computed.
 * Returns when done. */
int tdi_render_loss_grad(tdi_batch* b, const float* theta, const float* dverts, const double* gloss, double* loss,
                         float* gtheta, float* gverts, void* stream);

/* Launch-shape autotuner (no counterpart in the reference, whose launch shapes are fixed in the
 * source): on the batch's first workspace chunk rendered at theta (DEVICE [n][K]), times the
 * per-pixel Jacobian kernel with its grid sized for 1/2, 1, 2 and 4 times the mean crop, and the
 * J^T J kernel with 1, 2, 4, ... splits per frame (up to twice the default), `reps` launches each,
 * and keeps the fastest of each for the batch (its captured fit schedules are rebuilt on the next
 * fit); later batches of the same shape in the process start from the same choice.  choice [4]
 * (may be NULL): J^T J splits, Jacobian area, mean crop, the default splits; report [11] (may be
 * NULL): ms per launch for the four areas, then for 1, 2, 4, ..., 64 splits (-1 where not timed). */
int tdi_batch_autotune(tdi_batch* b, const float* theta, int32_t reps, int32_t* choice, float* report);
/* The batch's current launch shape: shape[0] J^T J splits per frame, shape[1] the area k_dI's
 * grid is sized for. */
int tdi_batch_launch_shape(tdi_batch* b, int32_t* shape);

/* Frames per workspace chunk (the batch is processed in chunks of this many frames). */
int tdi_workspace_frames(tdi_batch* b);

/* Stage timers (HIP events on the batch's stream) for one full evaluation + solve of the
 * first workspace chunk, mean over `reps`: ms[0] synthesis, [1] raster, [2] shade+blur+gradients, [3] barrier objective,
 * [4] per-pixel Jacobian, [5] J^T J / J^T r (+ barrier derivatives), [6] Cholesky.
 * Mirrors the reference's MEASURE_TIME stage timers (model_fitter.cu:43-58). */
int tdi_stage_times(tdi_batch* b, const float* theta, int32_t mixed_flags, int32_t reps, float ms[7]);

/* ---- Frame preparation: replaces InputData::get_resized_frame (input_data.cu:44-53; grey
 *      conversion, /255, cv::resize by Camera::resize_coef, zero border at the bottom and right)
 *      and the crop + transpose of VideoFitter::fit_to_video_frame (video_fitter.cu:537-543).
 * bgr: DEVICE, n frames of src_h rows x src_w pixels x 3 bytes (B, G, R: OpenCV's CV_8UC3
 * layout), frame i at bgr + i*frame_stride bytes (frame_stride >= 3*src_w*src_h).
 * coef: HOST [n] resize coefficients; crop: HOST [n][4] (x0, y0, width, height) in pixels of the
 * resized frame (Renderer::set_x0_short_y0_short, renderer.cu:585-606, gives x0, y0).
 * out: DEVICE, the grey crops, column-major, frame i at the exclusive prefix sum of
 * width*height: the raw observation that tdi_frame_desc.obs takes.  Pixels beyond the resized
 * frame are 0.  Returns when the crops are written. */
int tdi_prepare_crops(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                      const float* coef, const int32_t* crop, float* out, void* stream);

/* tdi_prepare_crops without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as tdi_prepare_crops, at
 * most TDI_PREP_NOALLOC_MAX_FRAMES frames per call (TDI_ERR_ARG beyond, before any device call).
 * The per-frame descriptors go to a fixed device array of the library (under 1 MB per GPU, held
 * until the process ends) instead of a per-call allocation.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (from several host threads or streams) race. */
#define TDI_PREP_NOALLOC_MAX_FRAMES 16384
int tdi_prepare_crops_noalloc(int32_t n, const uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                              const float* coef, const int32_t* crop, float* out, void* stream);

/* Inner-mouth masking of the frames before the crops are taken (config PAINT_INNERMOUTH_BLACK,
 * default 1; paint_innermouth_black, preprocessing.cu:40-103, called on every frame at
 * video_fitter.cu:531-535).  bgr: DEVICE, n frames as for tdi_prepare_crops, painted in place;
 * lmk_x, lmk_y: HOST [n][51] landmarks in frame pixels.  A frame whose mouth is open by more
 * than 3 ref_face_size units (mean of the distances of landmarks 44-50, 45-49, 46-48 times
 * ref_face_size / face size) has the pixels of its mouth box (landmarks 31-50, clipped to the
 * frame) set to 0 in all three channels when they lie inside the inner-lip polygon (landmarks
 * 43-50, 43 and 47 moved 3u inwards, u = face size / ref_face_size) deeper than u - 1 pixels,
 * or outside it by less than 1 - u: the reference's contour of the 3-pixel stroke of that
 * polygon taken as the polygon moved one pixel outwards (an approximation of OpenCV's
 * rasterised contour; the per-frame quantities keep the reference's FP32 operations).  Frames
 * with a non-finite landmark among 31-50 are left as they are.  Returns when the frames are
 * written. */
int tdi_paint_innermouth_black(int32_t n, uint8_t* bgr, int32_t src_w, int32_t src_h, int64_t frame_stride,
                               const float* lmk_x, const float* lmk_y, float ref_face_size, void* stream);

/* Rendered pixel counts of the batch's last evaluation (host array of n). */
int tdi_last_pixel_counts(tdi_batch* b, int32_t* npix);

/* ---- Fits ---- */

/* Plain path: fit_3DMM_shape_rigid (model_fitter.cu:121-375) for one frame at a time,
 * host-driven loop, FP32 J^T J, FP64 Cholesky, no graph.  theta: [n][K] HOST, in/out. */
int tdi_fit_plain(tdi_batch* b, float* theta, tdi_fit_report* reports);

/* Batched mixed-precision path: every frame of the batch in one device-resident loop,
 * BF16 hi/lo J^T J on MFMA, FP32 Cholesky + FP64 refinement, per-frame line-search masks,
 * iterations captured in a hipGraph.  theta: [n][K] DEVICE, in/out; reports: HOST [n]. */
int tdi_fit_batched(tdi_batch* b, float* theta, tdi_fit_report* reports, void* stream);

/* Expression finetune: fit_3DMM_epsilon_finetune (model_fitter.cu:397-661).  The same
 * Gauss-Newton fit with the expression bounds widened by finetune_coef (restored afterwards,
 * 409-410, 657-658) and line-search trials that move every unknown but the six pose unknowns
 * (set_xtmp over Ktotal - 6, 502): the normal equations and the decrement test still cover all
 * of them.  The pose stays at theta's (the reference's trials carry the pose its previous line
 * search left in ov_linesearch, which is theta's after an accepted step).  _plain: host theta,
 * one frame at a time as tdi_fit_plain; _batched: device theta, every frame on the device as
 * tdi_fit_batched. */
int tdi_fit_plain_finetune(tdi_batch* b, float* theta, tdi_fit_report* reports);
int tdi_fit_batched_finetune(tdi_batch* b, float* theta, tdi_fit_report* reports, void* stream);

/* Frames in the batch and unknowns per frame. */
int32_t tdi_batch_size(tdi_batch* b);
int32_t tdi_batch_unknowns(tdi_batch* b);

/* Replace the fixed light (lambda [n][3], L [n]) of every frame of the batch: the light a
 * shape fit uses when the light is not an unknown (the reference's ov.lambdas/ov.Lintensity
 * between its sub-fits, video_fitter.cu:554-615). */
int tdi_batch_set_light(tdi_batch* b, const float* lam, const float* Lint);

/* Light sub-fits, one frame at a time (host arrays), for a context whose light is not an
 * unknown (fit_light = 0): shape and pose stay at theta [n][K]; lam [n][3] and Lint [n] are
 * updated in place.
 *   tdi_fit_lambdas    replaces fit_3DMM_lambdas (model_fitter.cu:981-1073; normal equations
 *                      DerivativeComputer::compute_hessian_and_gradient_for_lambda,
 *                      derivative_computer.cu:228-312): 3x3 Gauss-Newton on lambda.
 *   tdi_fit_lintensity replaces fit_3DMM_Lintensity (model_fitter.cu:1084-1180;
 *                      compute_hessian_and_gradient_for_Lintensity, derivative_computer.cu:318-400):
 *                      scalar Newton on L.
 * reports[i].objective is -sum g^_r.g^_s at the result (no data weight, no barrier). */
int tdi_fit_lambdas(tdi_batch* b, const float* theta, float* lam, float* Lint, tdi_fit_report* reports);
int tdi_fit_lintensity(tdi_batch* b, const float* theta, float* lam, float* Lint, tdi_fit_report* reports);

#ifdef __cplusplus
}
#endif

#endif /* TDI_API_H_ */
