/*
 * tdi_host.h — C entry points of the C++ host layer (3di_amd/host/tdi_host.cpp).
 *
 * The host layer sits above the C ABI of include/tdi_api.h and restates the reference's
 * host-side code for the per-frame fitting step: the orchestration of one frame's fits
 * (VideoFitter::fit_to_video_frame, video_fitter.cu:432-648), the output files
 * (VideoOutput, video_fitter.cu:2603-2828, written by write_2d_vector, funcs.h:112-124),
 * Euler angles of a pose (RotationComputer::compute_euler_angles, rotation_header.cu:421-449)
 * and the config-file keys that reach this path (config.cu:300-560).  It is C++; these
 * functions exist for bindings (ctypes) and tests.  Library: 3di_amd/libtdi_host.so.
 */
#ifndef TDI_HOST_H
#define TDI_HOST_H

#include <stdint.h>
#include "tdi_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- VideoOutput (video_fitter.h:38-75, video_fitter.cu:2603-2828) ---- */
typedef struct tdi_video_output tdi_video_output;

/* k_exp: expression coefficients per fit; nres_coefs: columns groups of save_expressions_all
 * (config NRES_COEFS). */
tdi_video_output* tdi_vo_create(int32_t k_exp, int32_t nres_coefs);
void tdi_vo_destroy(tdi_video_output* vo);
/* One fit of frame `frame_id` (add_exp_coeffs, add_pose, add_illum; pose = tau(3), u(3);
 * illum = lambda(3), L).  pose / illum may be NULL (FINETUNE_ONLY adds expressions only). */
int tdi_vo_add(tdi_video_output* vo, int64_t frame_id, const float* exp, const float* pose, const float* illum);
/* kind 0: save_expressions (flag = ignore_first), 1: save_expressions_all,
 * 2: save_poses (flag = append yaw, pitch, roll), 3: save_illums. */
int tdi_vo_save(tdi_video_output* vo, int32_t kind, const char* path, int32_t flag);

/* yaw, pitch, roll of the rotation with axis-angle vector u (compute_euler_angles). */
void tdi_euler_angles(const float* u, float* yaw_pitch_roll);

/* Reads the keys of a 3DI config file (OpenCV YAML "KEY: value") that act on this path into
 * opts (KERNEL_SIGMA -> blur_sigma, REF_FACE_SIZE -> ref_face_size, NITERS_SUCCESS ->
 * niters_success) and returns TWOSTAGE_ILLUM and NRES_COEFS (defaults 0 and 1, config.cu). */
int tdi_read_config(const char* path, tdi_fit_options* opts, int32_t* twostage_illum, int32_t* nres_coefs);

/* Starting poses from the landmarks, as Logbarrier_Initializer::initialize_with_orthographic_t
 * (logbarrier_initializer.cu:1862-1914) computes them before each frame's fit
 * (video_fitter.cu:216-224,479-485), for n frames at once on the GPU (tdi_ortho_init).  Host
 * arrays: lmk_x, lmk_y [n][51] image pixels, cams [n][4] (phix, phiy, cx, cy); outputs pose
 * [n][6] (tau, u), angles [n][3] (yaw, pitch, roll, degrees; may be NULL), info [n][4] (may be
 * NULL; see tdi_ortho_init). */
int tdi_initialize_with_orthographic(tdi_ctx* ctx, int32_t n, const float* lmk_x, const float* lmk_y,
                                     const float* cams, float* pose, float* angles, int32_t* info);

/* The per-frame sequence of fit_to_video_frame (video_fitter.cu:554-615) for every frame of the
 * batch: L = 0.005, lambda fit, shape fit; with twostage_illum two more lambda + shape rounds.
 * theta [n][K], lam [n][3], Lint [n] are host arrays updated in place; reports are those of
 * the last shape fit.  batched = 0 runs tdi_fit_plain, 1 tdi_fit_batched. */
int tdi_fit_video_frames(tdi_batch* b, int32_t twostage_illum, int32_t batched, float* theta, float* lam,
                         float* Lint, tdi_fit_report* reports);

/* ---- The per-frame stage of a video (VideoFitter::fit_video_frames_auto after landmark
 *      detection, video_fitter.cu:160-393): for every frame, the landmark-driven starting pose,
 *      then for each of NRES_COEFS reference face sizes (update_shape_single_ref_size, 355-393)
 *      the frame resized and cropped at that size and fit_to_video_frame's fits (554-628),
 *      the results added to a VideoOutput (averaged there over the sizes). ---- */
typedef struct tdi_video_params {
    float ref_face_size;        /* REF_FACE_SIZE (config.cu:18: 86)                                 */
    int32_t nres_coefs;         /* NRES_COEFS (config.cu:41: 4), 1..15                               */
    int32_t twostage_illum;     /* TWOSTAGE_ILLUM (two more light + shape rounds)                    */
    int32_t finetune_expressions; /* FINETUNE_EXPRESSIONS (config.cu:95): the expression finetune     */
    int32_t pad;                /* PAD_RENDERING (config.cu:43: 1)                                   */
    int32_t crop_w, crop_h;     /* DIMX, DIMY (constants.h:55-56: 256)                               */
    int32_t batched;            /* 1: the batched fits, 0: frame by frame                            */
    int32_t init_pose;          /* 1: start each frame's pose at tdi_initialize_with_orthographic's  */
    float box_half_width;       /* landmark boxes (face-size units) when none are given              */
    int32_t paint_innermouth_black; /* PAINT_INNERMOUTH_BLACK (config.cu:92: 1): tdi_paint_innermouth_black
                                     * on every frame inside the frame, once, before the sizes loop   */
    int32_t lmk_k_exp;          /* K_EPSILON_L (config.cu:112: 79): expression coefficients of the
                                 * landmark-only fit with slack that follows the orthographic start
                                 * at every size (at most the layout's); 0: no landmark fit        */
} tdi_video_params;

void tdi_default_video_params(tdi_video_params* p);

/* The config keys of the per-frame stage: tdi_read_config's plus FINETUNE_EXPRESSIONS,
 * FINETUNE_COEF (-> opts.finetune_coef), PAD_RENDERING, PAINT_INNERMOUTH_BLACK and REF_FACE_SIZE
 * (both the face-size test of opts and the base of the reference sizes). */
int tdi_read_video_config(const char* path, tdi_fit_options* opts, tdi_video_params* p);
/* (K_EPSILON_L is set by the reference's model choice, config.cu:300-340, not by a key.) */

/* Per size, as fit_to_video_frame (video_fitter.cu:459-486): with init_pose, the orthographic
 * start from that size's resized landmarks and camera, then (lmk_k_exp > 0) the landmark-only fit
 * with slack from it (tdi_fit_landmarks_slack); a frame whose landmark fit succeeds starts its
 * fits from that pose and expression, the others from theta as it stands (the reference keeps
 * its previous unknowns; without the landmark fit, from the orthographic pose).
 * n frames of one video: bgr DEVICE [n] frames of src_h x src_w x 3 bytes at frame_stride bytes
 * (painted in place when p->paint_innermouth_black, as the reference paints its decoded frames,
 * video_fitter.cu:189-191);
 * lmk_x, lmk_y [n][51] landmarks and cams [n][4] (phix, phiy, cx, cy) of the unresized frames,
 * host; boxes [n][4][51] landmark boxes (lx, ux, ly, uy; face-size units; NULL = +-box_half_width);
 * frame_ids [n] the frames' numbers in the video (VideoOutput rows).  The context's light must
 * not be an unknown (the reference fits it with its own sub-fits).  theta [n][K], lam [n][3],
 * Lint [n] host, in/out: the start (pose overwritten when init_pose) and the last fit.
 * status [n]: TDI_OUT_OF_FRAME for a frame whose landmarks leave the frame (skipped, as the
 * reference's frame loop skips it, video_fitter.cu:187-208), else the last shape fit's status;
 * reports [n] (may be NULL): the last shape fit's reports.  vo (may be NULL) receives every fit. */
int tdi_fit_video_frames_nres(tdi_ctx* ctx, const tdi_video_params* p, int32_t n, uint8_t* bgr, int32_t src_w,
                              int32_t src_h, int64_t frame_stride, const float* lmk_x, const float* lmk_y,
                              const float* cams, const float* boxes, const int64_t* frame_ids, float* theta, float* lam,
                              float* Lint, tdi_video_output* vo, int32_t* status, tdi_fit_report* reports);

/* ---- Long videos: the per-frame stage over more frames than one batch should hold (the
 *      reference's frame loop of fit_video_frames_auto runs to the end of the video and writes
 *      its VideoOutput at the end, video_fitter.cu:160-393).  The frames are taken chunk by
 *      chunk from a source, each chunk runs tdi_fit_video_frames_nres, and every fit goes to vo;
 *      with a state file the run can stop after any chunk and resume there. ---- */

/* Frames [first, first + count) of the video: a DEVICE pointer to count consecutive BGR8 frames
 * of src_h x src_w x 3 bytes, valid until the next call; the driver paints them in place
 * (inner-mouth masking).  NULL aborts the run with TDI_ERR_ARG. */
typedef const uint8_t* (*tdi_frame_source)(void* user, int32_t first, int32_t count);

typedef struct tdi_long_video {
    int32_t n_frames;           /* frames of the video                                          */
    int32_t src_w, src_h;       /* frame size                                                   */
    int32_t chunk_frames;       /* frames per chunk (0: 1024)                                   */
    int32_t max_chunks;         /* chunks to run in this call (0: to the end)                   */
    const char* state_path;     /* state file (NULL: none); the fits so far go to <state>.fits  */
    int32_t resume;             /* 1: continue after the chunks the state file records          */
} tdi_long_video;

/* Landmarks lmk_x, lmk_y [n_frames][51] and cams [n_frames][4] of the unresized frames, host.
 * theta [K], lam [3], Lint [1] (host, in/out): the start of every frame of the first chunk; each
 * later chunk starts from the last successful fit of the one before (the state carried, and
 * returned).  status [n_frames]: as tdi_fit_video_frames_nres's for the frames run in this call,
 * -1 for the others.  vo receives every fit (on resume, first the saved fits of the chunks the
 * state records; vo must then be empty).  frames_done (may be NULL): frames finished so far.
 * The state file records the frames done and the carried state, written after each chunk by
 * renaming a temporary file; <state>.fits holds every fit in exact (hexadecimal) form, so that a
 * resumed run writes the same output files as an uninterrupted one. */
int tdi_fit_video_long(tdi_ctx* ctx, const tdi_video_params* p, const tdi_long_video* lv, tdi_frame_source src,
                       void* user, const float* lmk_x, const float* lmk_y, const float* cams, float* theta, float* lam,
                       float* Lint, tdi_video_output* vo, int32_t* status, int32_t* frames_done);

/* Every fit added to vo since its last dump, appended to path in exact (hexadecimal) form; and
 * the fits of such a file replayed into vo. */
int tdi_vo_dump(tdi_video_output* vo, const char* path);
int tdi_vo_load(tdi_video_output* vo, const char* path);

/* ---- Identity stage: frame selection and identity learning (VideoFitter::learn_identity,
 *      video_fitter.cu:1925-2016; fit_video_frames_landmarks_sparse, 2086-2583) ---- */
typedef struct tdi_identity_params {
    int32_t nframes;                /* NFRAMES (config.cu:45: 7): frames of one multi-frame fit        */
    int32_t nframes_per_angle_bin;  /* NFRAMES_PER_ANGLE_BIN (config.cu:79: 14)                        */
    int32_t nperms;                 /* NPERMS (config.cu:27: 10): shuffled passes over the binned frames */
    int32_t ntot_reconstrs;         /* NTOT_RECONSTRS (config.cu:29: 6): multi-frame fits averaged       */
    int32_t nmax_frames;            /* NMAX_FRAMES (config.cu:28: 432000)                              */
    int32_t max_vid_frames;         /* MAX_VID_FRAMES_TO_PROCESS (config.cu:59: 50000)                 */
    float skip_first_n_secs;        /* SKIP_FIRST_N_SECS (config.cu:55: 0)                             */
    float fps;                      /* the video's frame rate (for skip_first_n_secs)                  */
    float ref_face_size;            /* REF_FACE_SIZE of the multi-frame fits' frames (config.cu:18: 86) */
    int32_t pad;                    /* PAD_RENDERING (config.cu:43: 1)                                 */
    int32_t crop_w, crop_h;         /* DIMX, DIMY (constants.h:55-56: 256)                             */
    int32_t paint_innermouth_black; /* PAINT_INNERMOUTH_BLACK (config.cu:92: 1)                        */
    int32_t lmk_k_exp;              /* K_EPSILON_L (config.cu:112: 79): expression coefficients of the
                                     * landmark fits                                                   */
    float box_half_width;           /* landmark boxes (face-size units): 0.15                          */
    int32_t lmk_k_id;               /* K_ALPHA_L (config.cu:160: 60): identity coefficients of the
                                     * landmark fits with the slack                                    */
} tdi_identity_params;

void tdi_default_identity_params(tdi_identity_params* p);

/* The binning of fit_video_frames_landmarks_sparse (video_fitter.cu:2380-2550), host arithmetic:
 * yaw, pitch [n] (radians; NaN: the frame has no angle) are binned to the nearest of the 13 x 13
 * centroids -55, -45, ..., 65 degrees; a std::default_random_engine seeded with 1907 shuffles every
 * bin (yaw-major order) and at most nframes_per_angle_bin frames of each are kept; the kept frames
 * in frame order are then shuffled nperms times by the same engine and concatenated; groups of
 * nframes consecutive entries of that order are taken until ntot_reconstrs groups (or nmax_frames
 * frames).  selected [cap] receives the frame indices in group order, *n_selected their count (a
 * multiple of nframes).  Every group is taken: tdi_select_identity_frames adds the reference's
 * acceptance test, a joint landmark-only fit of the group's frames. */
int tdi_bin_identity_frames(const tdi_identity_params* p, int32_t n, const float* yaw, const float* pitch,
                            int32_t* selected, int32_t cap, int32_t* n_selected);

/* The whole selection of fit_video_frames_landmarks_sparse from a video's landmarks: the frames
 * considered (every max(1, n / 800)-th, after skip_first_n_secs, before max_vid_frames, with all
 * 51 landmarks inside the frame), the median face size m of those (resize coefficient 76 / m,
 * returned in *resize_coef), then for every frame the resized landmarks' bounding-box checks
 * (check_if_bb_OK, check_if_face_in_frame, funcs.cu:582-626), the orthographic start (all frames
 * in one launch) and each checked frame's landmark-only fit with the slack over lmk_k_id identity
 * and lmk_k_exp expression coefficients (tdi_fit_landmarks_joint on the frame alone, video_fitter.cu:
 * 2253-2275, 2330-2345) giving yaw, pitch, roll (angles [n][3], radians, NaN where there is none;
 * may be NULL), then the binning of tdi_bin_identity_frames, where a group is kept only when the
 * joint landmark-only fit of its frames with the slack succeeds (fit_to_multi_images_landmarks_only,
 * model_fitter.cu:1205-1247, from zero identity and expressions and the orthographic poses;
 * video_fitter.cu:2505-2545).
 * ctx: the identity stage's context (fit_id = fit_tex = 1).  lmk_x, lmk_y [n][51] image pixels of
 * frame_w x frame_h frames, cam [4] (phix, phiy, cx, cy). */
int tdi_select_identity_frames(tdi_ctx* ctx, const tdi_identity_params* p, int32_t n, const float* lmk_x,
                               const float* lmk_y, int32_t frame_w, int32_t frame_h, const float* cam, float* angles,
                               int32_t* selected, int32_t cap, int32_t* n_selected, float* resize_coef);

/* learn_identity (video_fitter.cu:1925-2016) over the selected frames: for each group of nframes
 * consecutive frames (at most ntot_reconstrs), fit_multiframe (2023-2083) — each frame resized to
 * ref_face_size, masked and cropped as in the per-frame stage; the start of fit_to_multi_images
 * (model_fitter.cu:1291-1330): the orthographic poses, the joint landmark-only fit with the slack
 * over lmk_k_id identity coefficients (a group whose fit fails is not used), then from its result
 * the joint landmark-only fit without the slack over all identity coefficients
 * (tdi_fit_landmarks_joint); the light sub-fits of fit_to_multi_images (L, lambda, L, lambda per
 * frame, 1374-1386), then the joint fit (tdi_fit_multiframe); alpha [k_id] and beta [k_tex] are
 * the averages over the groups whose joint fit succeeded (more than NITERS_SUCCESS iterations),
 * *n_used their count (0: alpha
 * and beta untouched, TDI_OK returned).  ctx: fit_id = fit_tex = 1, fit_light = 0 (the light is held
 * by the sub-fits, as in the reference).  bgr: DEVICE [n][frame_h][frame_w][3] the selected frames
 * in selection order (masked in place when paint_innermouth_black); lmk_x, lmk_y [n][51]. */
int tdi_learn_identity(tdi_ctx* ctx, const tdi_identity_params* p, int32_t n, uint8_t* bgr, int32_t frame_w,
                       int32_t frame_h, int64_t frame_stride, const float* lmk_x, const float* lmk_y, const float* cam,
                       float* alpha, float* beta, int32_t* n_used);

/* ---- Frame preparation (host side; the pixels are tdi_prepare_crops in tdi_api.h) ---- */

/* Landmark text file, one frame per line "x0 y0 x1 y1 ... x50 y50", as
 * LandmarkData::init_from_txtfile reads it (input_data.cu:98-121, read2DVectorFromFile_unknown_size,
 * funcs.h:352-413): the first line's count of numbers fixes the columns (at least 102), at most
 * 32767 lines, a value that does not parse reads as NaN.  With lmk_x == NULL only *n_frames is
 * set; otherwise lmk_x, lmk_y [max_frames][51] receive the first min(max_frames, lines) frames. */
int tdi_read_landmarks(const char* path, int32_t max_frames, float* lmk_x, float* lmk_y, int32_t* n_frames);

/* Per-frame geometry of the video loop before each fit, for n frames: the landmarks-inside-the-frame
 * test (video_fitter.cu:180-186), the resize coefficient ref_face_size / face size with the face
 * size of LandmarkData::get_face_size (input_data.cu:124-138; video_fitter.cu:357-360), the camera
 * scaled by it (Camera::update_camera, camera.h:162-174), the resized landmarks
 * (InputData::get_resized_landmarks, input_data.cu:33-41), the crop origin
 * (Renderer::set_x0_short_y0_short, renderer.cu:585-606; pad = PAD_RENDERING) and the face size
 * of the resized landmarks (compute_face_size, funcs.cu:44-53).  Host arrays: lmk_x, lmk_y [n][51]
 * and cams [n][4] (phix, phiy, cx, cy) of the unresized frame of frame_w x frame_h pixels.
 * Outputs: coef [n], resized landmarks [n][51], cams_out [n][4], crop [n][4] = (x0, y0, crop_w,
 * crop_h) (the reference's crop is DIMX x DIMY = 256 x 256, constants.h:55-56), face_size [n],
 * status [n] = TDI_OK or TDI_OUT_OF_FRAME (outputs of such a frame are zero). */
int tdi_frame_geometry(int32_t n, const float* lmk_x, const float* lmk_y, const float* cams, int32_t frame_w,
                       int32_t frame_h, float ref_face_size, int32_t pad, int32_t crop_w, int32_t crop_h, float* coef,
                       float* lmk_x_out, float* lmk_y_out, float* cams_out, int32_t* crop, float* face_size,
                       int32_t* status);

#ifdef __cplusplus
}
#endif

#endif
