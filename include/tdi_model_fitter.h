/*
 * tdi_model_fitter.h — the reference's fitting entry points, by name, in C++ (libtdi_host.so).
 *
 * The reference declares them in src/lib/model_fitter.h:29-79 with its device objects and
 * library handles as arguments (Renderer, Optimizer, Logbarrier_Initializer, the cuSOLVER and
 * cuBLAS handles, OptimizationVariables, Camera, RotationComputer, DerivativeComputer, Solver).
 * Here one FitContext takes their place: a tdi_batch holds the model, the frames (cameras, crops,
 * observations, landmarks and boxes) and the workspace in HBM, and `batched` picks the device-
 * resident batched fit or the frame-by-frame host-driven one.  The unknowns are host arrays
 * [n][K] in the OptimizationVariables order of tdi_api.h.  As in the reference, each function
 * returns whether the fit counts as successful: more than niters_success outer iterations
 * (model_fitter.cu:374, 660) for the shape fits, for every frame of the batch; a failed library
 * call also returns false (tdi_last_error() says why).  Per-frame outcomes are in `reports`.
 */
#ifndef TDI_MODEL_FITTER_H
#define TDI_MODEL_FITTER_H

#include "tdi_api.h"

namespace tdi {

struct FitContext {
    tdi_batch* batch;      // frames + workspace (tdi_batch_create)
    bool batched;          // true: tdi_fit_batched; false: tdi_fit_plain
};

/* fit_3DMM_shape_rigid (model_fitter.h:29-32, model_fitter.cu:121-375). */
bool fit_3DMM_shape_rigid(FitContext& ctx, float* theta, tdi_fit_report* reports);

/* fit_3DMM_epsilon_finetune (model_fitter.h:76-79, model_fitter.cu:397-661). */
bool fit_3DMM_epsilon_finetune(FitContext& ctx, float* theta, tdi_fit_report* reports);

/* fit_3DMM_lambdas (model_fitter.h:70-74, model_fitter.cu:981-1073): the light position of each
 * frame, shape and pose fixed at theta; lambdas [n][3] and Lintensity [n] in/out; the batch's
 * fixed light is updated to the result (the next shape fit reads it). */
bool fit_3DMM_lambdas(FitContext& ctx, const float* theta, float* lambdas, float* Lintensity,
                      tdi_fit_report* reports);

/* fit_3DMM_Lintensity (model_fitter.h:63-67, model_fitter.cu:1084-1180): the light intensity,
 * likewise. */
bool fit_3DMM_Lintensity(FitContext& ctx, const float* theta, float* lambdas, float* Lintensity,
                         tdi_fit_report* reports);

/* Logbarrier_Initializer::set_minimal_slack then fit_model with the slack variable
 * (logbarrier_initializer.cu:1041-1267), as fit_to_video_frame runs them on its landmark object
 * li_init (video_fitter.cu:479-486): theta [n][K] host, in/out (the pose is the start; a frame
 * that reaches s < -0.01 gets the fitted pose and first k_exp expression coefficients, the
 * others keep theta); returns fit_success for every frame (tdi_fit_landmarks_slack). */
bool fit_model_with_slack(FitContext& ctx, float* theta, int k_exp, tdi_lmk_report* reports);

} // namespace tdi

#endif
