"""Python binding of libtdi.so (the gfx950 fitting step) over its C ABI.

Device memory and streams come from PyTorch; every computation runs in the HIP library.
The library must be built (``python 3di_amd/build.py``); loading fails loudly otherwise —
there is no CPU fallback.

Mirrors the reference's fitting entry points for one frame and for many:
``Batch.fit_plain`` is fit_3DMM_shape_rigid (model_fitter.cu:121-375) run frame by frame,
``Batch.fit_batched`` runs the same algorithm for every frame of the batch on the device.
"""
from __future__ import annotations

import ctypes as C
import importlib
import os
from typing import List, Sequence, Tuple

import numpy as np

abi = importlib.import_module("3di_amd.abi")

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtdi.so")
_lib = None

_vp = C.c_void_p


class TdiError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TdiError(f"{LIB_PATH} is missing: build it with `python 3di_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        L.tdi_last_error.restype = C.c_char_p
        L.tdi_num_unknowns.restype = C.c_int32
        L.tdi_ctx_create.argtypes = [_vp, _vp, C.c_int, C.POINTER(_vp)]
        L.tdi_ctx_destroy.argtypes = [_vp]
        L.tdi_ctx_release_cached.argtypes = [_vp]
        L.tdi_batch_create.argtypes = [_vp, C.c_int32, _vp, C.POINTER(_vp)]
        L.tdi_batch_destroy.argtypes = [_vp]
        L.tdi_rotation.argtypes = [C.c_int32, _vp, _vp, _vp, _vp]
        L.tdi_ortho_init.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_synthesize.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.tdi_rasterize.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_render.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_normal_equations.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_normal_equations_mixed.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_joint_unknowns.argtypes = [_vp]
        L.tdi_joint_unknowns.restype = C.c_int64
        L.tdi_normal_equations_joint.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_fit_multiframe.argtypes = [_vp, _vp, _vp]
        L.tdi_cholesky_solve.argtypes = [C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32, _vp]
        L.tdi_fit_plain.argtypes = [_vp, _vp, _vp]
        L.tdi_fit_batched.argtypes = [_vp, _vp, _vp, _vp]
        L.tdi_fit_plain_finetune.argtypes = [_vp, _vp, _vp]
        L.tdi_fit_batched_finetune.argtypes = [_vp, _vp, _vp, _vp]
        L.tdi_blur_taps.argtypes = [C.c_float, _vp]
        L.tdi_stage_times.argtypes = [_vp, _vp, C.c_int32, C.c_int32, _vp]
        L.tdi_last_pixel_counts.argtypes = [_vp, _vp]
        L.tdi_workspace_frames.argtypes = [_vp]
        L.tdi_fit_lambdas.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.tdi_batch_set_light.argtypes = [_vp, _vp, _vp]
        L.tdi_batch_size.argtypes = [_vp]
        L.tdi_batch_unknowns.argtypes = [_vp]
        L.tdi_fit_lintensity.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.tdi_fit_buffers.argtypes = [_vp, _vp, C.c_int32, _vp, _vp, _vp, _vp]
        L.tdi_barrier_terms.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_fit_landmarks_slack.argtypes = [_vp, _vp, C.c_int32, _vp, _vp]
        L.tdi_render_loss_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_batch_autotune.argtypes = [_vp, _vp, C.c_int32, _vp, _vp]
        L.tdi_batch_launch_shape.argtypes = [_vp, _vp]
        L.tdi_landmark_slack_terms.argtypes = [_vp, _vp, C.c_int32, _vp, C.c_double, _vp, _vp, _vp, _vp]
        L.tdi_fit_landmarks_joint.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp]
        L.tdi_fit_landmarks_each.argtypes = [_vp, _vp, C.c_int32, C.c_int32, _vp]
        L.tdi_landmark_joint_terms.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_double, _vp, _vp,
                                               _vp]
        L.tdi_prepare_crops.argtypes = [C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp]
        L.tdi_prepare_crops_noalloc.argtypes = [C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp]
        L.tdi_paint_innermouth_black.argtypes = [C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int64, _vp, _vp, C.c_float,
                                                 _vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise TdiError(f"tdi error {rc}: {lib().tdi_last_error().decode()}")


def _torch():
    import torch
    return torch


def _dptr(t) -> int:
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise TdiError("expected a contiguous device tensor")
    return t.data_ptr()


def _stream():
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def blur_taps(sigma: float) -> np.ndarray:
    t = np.zeros(15, np.float32)
    _check(lib().tdi_blur_taps(C.c_float(sigma), t.ctypes.data))
    return t


class Context:
    """The model resident in HBM (tdi_ctx)."""

    def __init__(self, model, opts=None, device: int = 0):
        self.model = model
        self.opts = opts if opts is not None else abi.default_options()
        self.md = abi.ModelDesc(model)
        self.K = lib().tdi_num_unknowns(self.md.ref, C.byref(self.opts))
        self.device = device
        h = _vp()
        _check(lib().tdi_ctx_create(C.addressof(self.md.desc), C.addressof(self.opts), device, C.byref(h)))
        self.h = h

    def release_cached(self):
        """Free the device blocks the context keeps for later batches (tdi_ctx_release_cached)."""
        _check(lib().tdi_ctx_release_cached(self.h))

    def close(self):
        if self.h:
            lib().tdi_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Frames resident in HBM plus a workspace (tdi_batch)."""

    def __init__(self, ctx: Context, frames: Sequence):
        self.ctx = ctx
        self.frames = list(frames)
        self.fd = abi.FrameDescs(self.frames)
        h = _vp()
        _check(lib().tdi_batch_create(ctx.h, len(self.frames), C.addressof(self.fd.arr), C.byref(h)))
        self.h = h
        self.areas = np.array([f.width * f.height for f in self.frames], np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.areas)])

    @property
    def n(self):
        return len(self.frames)

    @property
    def K(self):
        return self.ctx.K

    def close(self):
        if self.h:
            lib().tdi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _theta(self, theta):
        torch = _torch()
        if isinstance(theta, np.ndarray):
            theta = torch.from_numpy(np.ascontiguousarray(theta, np.float32)).cuda()
        return theta.float().contiguous().reshape(self.n, self.K)

    # -------------------------------------------------------------- building blocks
    def synthesize(self, theta):
        torch = _torch()
        th = self._theta(theta)
        V = self.ctx.model.n_vertices
        xyz = torch.empty((self.n, V, 3), device="cuda", dtype=torch.float32)
        xyp = torch.empty((self.n, V, 2), device="cuda", dtype=torch.float32)
        _check(lib().tdi_synthesize(self.h, _dptr(th), _dptr(xyz), _dptr(xyp), _stream()))
        return xyz, xyp

    def rasterize(self, xyz, xyp):
        torch = _torch()
        keys = torch.empty(int(self.offsets[-1]), device="cuda", dtype=torch.int64)
        bary = torch.empty(int(self.offsets[-1]) * 2, device="cuda", dtype=torch.float32)
        _check(lib().tdi_rasterize(self.h, _dptr(xyz.contiguous()), _dptr(xyp.contiguous()), _dptr(keys), _dptr(bary),
                                   _stream()))
        return keys, bary

    def render(self, theta, blurred: bool = False):
        """(img, obj, npix), or (img, blurred image, obj, npix) with blurred=True."""
        torch = _torch()
        th = self._theta(theta)
        img = torch.empty(int(self.offsets[-1]), device="cuda", dtype=torch.float32)
        blr = torch.empty(int(self.offsets[-1]), device="cuda", dtype=torch.float32) if blurred else None
        obj = torch.empty(self.n, device="cuda", dtype=torch.float32)
        npix = torch.empty(self.n, device="cuda", dtype=torch.int32)
        _check(lib().tdi_render(self.h, _dptr(th), _dptr(img), _dptr(blr), _dptr(obj), _dptr(npix), _stream()))
        return (img, blr, obj, npix) if blurred else (img, obj, npix)

    def fit_buffers(self, theta, mixed_flags: int = 1):
        """The fit's projected vertices [n, V, 4], neighbour positions [sum W*H, 4] (frame i's
        rendered pixels at offsets[i] .. offsets[i] + npix[i]) and pixel counts (tdi_fit_buffers)."""
        torch = _torch()
        th = self._theta(theta)
        V = self.ctx.model.n_vertices
        vproj = torch.empty((self.n, V, 4), device="cuda", dtype=torch.float32)
        nbr = torch.zeros((int(self.offsets[-1]), 4), device="cuda", dtype=torch.int32)
        npix = torch.empty(self.n, device="cuda", dtype=torch.int32)
        _check(lib().tdi_fit_buffers(self.h, _dptr(th), int(mixed_flags), _dptr(vproj), _dptr(nbr), _dptr(npix),
                                     _stream()))
        return vproj, nbr, npix

    def barrier_terms(self, theta):
        """(H [n, K, K] float64, g [n, K] float64, obj [n] float64) of the log barrier alone."""
        torch = _torch()
        th = self._theta(theta)
        K = self.K
        H = torch.empty((self.n, K, K), device="cuda", dtype=torch.float64)
        g = torch.empty((self.n, K), device="cuda", dtype=torch.float64)
        obj = torch.empty(self.n, device="cuda", dtype=torch.float64)
        _check(lib().tdi_barrier_terms(self.h, _dptr(th), _dptr(H), _dptr(g), _dptr(obj), _stream()))
        return H, g, obj

    def autotune(self, theta, reps: int = 3) -> dict:
        """Pick the per-pixel Jacobian grid and the J^T J split count by timing (tdi_batch_autotune)."""
        th = self._theta(theta)
        choice = np.zeros(4, np.int32)
        rep = np.zeros(11, np.float32)
        _check(lib().tdi_batch_autotune(self.h, _dptr(th), int(reps), choice.ctypes.data, rep.ctypes.data))
        a0 = int(choice[2])
        areas = [max(1, a0 // 2), a0, 2 * a0, 4 * a0]
        return dict(jtj_splits=int(choice[0]), dI_area=int(choice[1]), mean_crop=a0, default_splits=int(choice[3]),
                    dI_ms={str(a): round(float(t), 4) for a, t in zip(areas, rep[:4])},
                    jtj_ms={str(1 << i): round(float(t), 4) for i, t in enumerate(rep[4:]) if t >= 0})

    def launch_shape(self):
        """(J^T J splits per frame, area the per-pixel Jacobian grid is sized for)."""
        sh = np.zeros(2, np.int32)
        _check(lib().tdi_batch_launch_shape(self.h, sh.ctypes.data))
        return int(sh[0]), int(sh[1])

    def render_loss(self, theta, dverts=None, gloss=None, grad: bool = True):
        """The render loss of every frame and, with grad, its adjoint (tdi_render_loss_grad).
        theta: [n, K]; dverts: per-vertex model-space offsets [n, V, 3] or None; gloss: upstream
        gradient [n] of each frame's loss (None = 1).  Returns (loss [n] float64, gtheta [n, K]
        float32 or None, gverts [n, V, 3] float32 or None)."""
        torch = _torch()
        th = self._theta(theta)
        V = self.ctx.model.n_vertices
        dv = None
        if dverts is not None:
            dv = dverts.float().contiguous().reshape(self.n, V, 3)
        gl = None
        if gloss is not None:
            gl = gloss.double().contiguous().reshape(self.n)
        loss = torch.empty(self.n, device="cuda", dtype=torch.float64)
        gth = torch.empty((self.n, self.K), device="cuda", dtype=torch.float32) if grad else None
        gv = torch.empty((self.n, V, 3), device="cuda", dtype=torch.float32) if grad else None
        _check(lib().tdi_render_loss_grad(self.h, _dptr(th), _dptr(dv) if dv is not None else None,
                                          _dptr(gl) if gl is not None else None, _dptr(loss),
                                          _dptr(gth) if gth is not None else None,
                                          _dptr(gv) if gv is not None else None, _stream()))
        return loss, gth, gv

    def fit_landmarks_slack(self, theta, k_exp: int):
        """The landmark-only fit with slack of every frame (tdi_fit_landmarks_slack).  theta:
        device tensor [n, K], updated in place (pose and the first k_exp expression coefficients of
        the frames that succeed).  Returns (theta, reports)."""
        th = self._theta(theta)
        reps = (abi.tdi_lmk_report * self.n)()
        _check(lib().tdi_fit_landmarks_slack(self.h, _dptr(th), int(k_exp), C.addressof(reps), _stream()))
        return th, list(reps)

    def landmark_slack_terms(self, theta, k_exp: int, z, tau: float):
        """(H [n, k_exp+7, k_exp+7], g [n, k_exp+7], obj [n]) float64 of that fit at z [n, k_exp+7]."""
        torch = _torch()
        th = self._theta(theta)
        nz = int(k_exp) + 7
        if isinstance(z, np.ndarray):
            z = torch.from_numpy(np.ascontiguousarray(z, np.float64)).cuda()
        z = z.double().contiguous().reshape(self.n, nz)
        H = torch.empty((self.n, nz, nz), device="cuda", dtype=torch.float64)
        g = torch.empty((self.n, nz), device="cuda", dtype=torch.float64)
        obj = torch.empty(self.n, device="cuda", dtype=torch.float64)
        _check(lib().tdi_landmark_slack_terms(self.h, _dptr(th), int(k_exp), _dptr(z), float(tau), _dptr(H), _dptr(g),
                                              _dptr(obj), _stream()))
        return H, g, obj

    def joint_landmark_unknowns(self, k_id: int, k_exp: int, slack: bool) -> int:
        return int(k_id) + (1 if slack else 0) + self.n * (int(k_exp) + 6)

    def fit_landmarks_joint(self, theta0, k_id: int, k_exp: int, slack: bool):
        """The joint landmark-only fit over the batch's frames (tdi_fit_landmarks_joint): theta0
        host [n, K]; returns (theta [n, K] float32, report)."""
        th = np.ascontiguousarray(np.array(theta0, np.float32).reshape(self.n, self.K))
        rep = abi.tdi_lmk_report()
        _check(lib().tdi_fit_landmarks_joint(self.h, th.ctypes.data, int(k_id), int(k_exp), int(bool(slack)),
                                             C.addressof(rep)))
        return th, rep

    def fit_landmarks_each(self, theta0, k_id: int, k_exp: int):
        """Every frame's landmark-only fit with the slack on its own, batched
        (tdi_fit_landmarks_each): theta0 host [n, K]; returns (theta [n, K] float32, reports)."""
        th = np.ascontiguousarray(np.array(theta0, np.float32).reshape(self.n, self.K))
        reps = (abi.tdi_lmk_report * self.n)()
        _check(lib().tdi_fit_landmarks_each(self.h, th.ctypes.data, int(k_id), int(k_exp), C.addressof(reps)))
        return th, list(reps)

    def landmark_joint_terms(self, theta, k_id: int, k_exp: int, slack: bool, z, tau: float):
        """(obj, g [N], H [N, N]) float64 of that fit at the joint unknowns z (host, FP64)."""
        th = np.ascontiguousarray(np.array(theta, np.float32).reshape(self.n, self.K))
        N = self.joint_landmark_unknowns(k_id, k_exp, slack)
        z = np.ascontiguousarray(np.array(z, np.float64).reshape(N))
        H = np.zeros((N, N))
        g = np.zeros(N)
        obj = C.c_double(0.0)
        _check(lib().tdi_landmark_joint_terms(self.h, th.ctypes.data, int(k_id), int(k_exp), int(bool(slack)),
                                              z.ctypes.data, float(tau), H.ctypes.data, g.ctypes.data,
                                              C.addressof(obj)))
        return obj.value, g, H

    def normal_equations(self, theta, mixed: bool = False):
        torch = _torch()
        th = self._theta(theta)
        K = self.K
        H = torch.empty((self.n, K, K), device="cuda", dtype=torch.float32)
        g = torch.empty((self.n, K), device="cuda", dtype=torch.float32)
        obj = torch.empty(self.n, device="cuda", dtype=torch.float32)
        fn = lib().tdi_normal_equations_mixed if mixed else lib().tdi_normal_equations
        _check(fn(self.h, _dptr(th), _dptr(H), _dptr(g), _dptr(obj), _stream()))
        return H, g, obj

    # -------------------------------------------------------------- identity stage
    def joint_unknowns(self) -> int:
        """Unknowns of the joint fit of the batch's frames: the shared albedo and identity block
        plus every frame's expression, pose and light."""
        return int(lib().tdi_joint_unknowns(self.h))

    def joint_normal_equations(self, theta):
        """(HJ [KJ, KJ] float64, gJ [KJ] float64, obj [1] float64) of the joint system at the
        frames' theta [n, K] (the shared block is frame 0's)."""
        torch = _torch()
        th = self._theta(theta)
        KJ = self.joint_unknowns()
        H = torch.empty((KJ, KJ), device="cuda", dtype=torch.float64)
        g = torch.empty(KJ, device="cuda", dtype=torch.float64)
        obj = torch.empty(1, device="cuda", dtype=torch.float64)
        _check(lib().tdi_normal_equations_joint(self.h, _dptr(th), _dptr(H), _dptr(g), _dptr(obj), _stream()))
        return H, g, obj

    def fit_multiframe(self, theta0):
        """The identity stage's joint Gauss-Newton fit (tdi_fit_multiframe): theta0 host [n, K];
        returns (theta [n, K] with the shared block on every frame, report)."""
        th = np.ascontiguousarray(np.array(theta0, np.float32).reshape(self.n, self.K))
        rep = abi.tdi_fit_report()
        _check(lib().tdi_fit_multiframe(self.h, th.ctypes.data, C.addressof(rep)))
        return th, rep

    def stage_times(self, theta, mixed_flags: int = 7, reps: int = 3) -> np.ndarray:
        """Mean ms per stage of one evaluation + solve (see tdi_stage_times)."""
        th = self._theta(theta)
        ms = np.zeros(7, np.float32)
        _check(lib().tdi_stage_times(self.h, _dptr(th), mixed_flags, reps, ms.ctypes.data))
        return ms

    def workspace_frames(self) -> int:
        return lib().tdi_workspace_frames(self.h)

    def pixel_counts(self) -> np.ndarray:
        out = np.zeros(min(self.n, self.workspace_frames()), np.int32)
        _check(lib().tdi_last_pixel_counts(self.h, out.ctypes.data))
        return out

    # -------------------------------------------------------------- fits
    def fit_plain(self, theta0, finetune: bool = False) -> Tuple[np.ndarray, List]:
        """fit_3DMM_shape_rigid one frame at a time (tdi_fit_plain), or with finetune=True
        fit_3DMM_epsilon_finetune (tdi_fit_plain_finetune).  theta0: host [n, K]."""
        th = np.ascontiguousarray(np.array(theta0, np.float32).reshape(self.n, self.K))
        reps = (abi.tdi_fit_report * self.n)()
        fn = lib().tdi_fit_plain_finetune if finetune else lib().tdi_fit_plain
        _check(fn(self.h, th.ctypes.data, C.addressof(reps)))
        return th, list(reps)

    def _light_fit(self, fn, theta, lam, Lint):
        th = np.ascontiguousarray(np.array(theta, np.float32).reshape(self.n, self.K))
        lam = np.ascontiguousarray(np.array(lam, np.float32).reshape(self.n, 3))
        Li = np.ascontiguousarray(np.array(Lint, np.float32).reshape(self.n))
        reps = (abi.tdi_fit_report * self.n)()
        _check(fn(self.h, th.ctypes.data, lam.ctypes.data, Li.ctypes.data, C.addressof(reps)))
        return lam, Li, list(reps)

    def set_light(self, lam, Lint):
        """Replace every frame's fixed light (used by shape fits when the light is not an unknown)."""
        la = np.ascontiguousarray(np.array(lam, np.float32).reshape(self.n, 3))
        Li = np.ascontiguousarray(np.array(Lint, np.float32).reshape(self.n))
        _check(lib().tdi_batch_set_light(self.h, la.ctypes.data, Li.ctypes.data))

    def fit_lambdas(self, theta, lam, Lint):
        """fit_3DMM_lambdas (model_fitter.cu:981-1073) on every frame: returns (lam, Lint, reports)."""
        return self._light_fit(lib().tdi_fit_lambdas, theta, lam, Lint)

    def fit_lintensity(self, theta, lam, Lint):
        """fit_3DMM_Lintensity (model_fitter.cu:1084-1180) on every frame."""
        return self._light_fit(lib().tdi_fit_lintensity, theta, lam, Lint)

    def fit_batched(self, theta, finetune: bool = False):
        """theta: device tensor [n, K], updated in place.  Returns the per-frame reports.  finetune:
        the expression finetune (tdi_fit_batched_finetune) instead of the shape fit."""
        th = self._theta(theta)
        reps = (abi.tdi_fit_report * self.n)()
        fn = lib().tdi_fit_batched_finetune if finetune else lib().tdi_fit_batched
        _check(fn(self.h, _dptr(th), C.addressof(reps), _stream()))
        return th, list(reps)


def cholesky_solve(H, g, mode: int = 0):
    torch = _torch()
    n, K = g.shape
    delta = torch.empty_like(g)
    status = torch.empty(n, device="cuda", dtype=torch.int32)
    _check(lib().tdi_cholesky_solve(n, K, _dptr(H.contiguous()), _dptr(g.contiguous()), _dptr(delta), _dptr(status),
                                    mode, _stream()))
    return delta, status


NLANDMARKS = 51   # TDI_NLANDMARKS


def ortho_init(ctx: Context, lmk, cams):
    """Orthographic landmark initialisation and pose seeding of n frames on the GPU
    (tdi_ortho_init).  lmk: [n, 2, 51] image-pixel landmarks (x row, y row); cams: [n, 4]
    (phix, phiy, cx, cy).  Returns device tensors x_orth [n, 7], pose [n, 6] (tau, u),
    angles [n, 3] (yaw, pitch, roll, degrees) and info [n, 4] (iterations, refit iterations,
    refit from the mirrored pose, status)."""
    torch = _torch()
    lmk = lmk.float().contiguous()
    cams = cams.float().contiguous()
    n = lmk.shape[0]
    if tuple(lmk.shape[1:]) != (2, NLANDMARKS) or tuple(cams.shape) != (n, 4):
        raise ValueError("ortho_init: lmk must be [n, 2, 51] and cams [n, 4]")
    dev = dict(device="cuda", dtype=torch.float32)
    x_orth = torch.empty((n, 7), **dev)
    pose = torch.empty((n, 6), **dev)
    angles = torch.empty((n, 3), **dev)
    info = torch.empty((n, 4), device="cuda", dtype=torch.int32)
    _check(lib().tdi_ortho_init(ctx.h, n, _dptr(lmk), _dptr(cams), _dptr(x_orth), _dptr(pose), _dptr(angles),
                                _dptr(info), _stream()))
    return dict(x_orth=x_orth, pose=pose, angles=angles, info=info)


def prepare_crops(frames, coef, crop, out=None, noalloc: bool = False):
    """Grey FP32 crops of resized 8-bit BGR frames on the GPU (tdi_prepare_crops).  frames: uint8
    device tensor [n, H, W, 3] (B, G, R); coef: [n] resize coefficients; crop: [n, 4] (x0, y0,
    width, height) in resized-frame pixels.  Returns a float32 device tensor with the crops
    concatenated, each column-major (frame i at the prefix sum of width*height); `out`, when
    given, is that tensor.  noalloc: call tdi_prepare_crops_noalloc, which allocates no device
    memory (at most TDI_PREP_NOALLOC_MAX_FRAMES frames)."""
    import numpy as np
    torch = _torch()
    if frames.dtype != torch.uint8 or frames.dim() != 4 or frames.shape[3] != 3 or not frames.is_cuda:
        raise ValueError("prepare_crops: frames must be a uint8 CUDA tensor [n, H, W, 3]")
    frames = frames.contiguous()
    n, H, W = (int(v) for v in frames.shape[:3])
    coef = np.ascontiguousarray(coef, np.float32).reshape(-1)
    crop = np.ascontiguousarray(crop, np.int32).reshape(-1, 4)
    if coef.shape[0] != n or crop.shape[0] != n:
        raise ValueError("prepare_crops: coef must be [n] and crop [n, 4]")
    total = int((crop[:, 2].astype(np.int64) * crop[:, 3]).sum())
    if out is None:
        out = torch.empty(total, device=frames.device, dtype=torch.float32)
    elif out.dtype != torch.float32 or not out.is_cuda or out.numel() < total:
        raise ValueError("prepare_crops: out must be a float32 CUDA tensor of at least the crops' size")
    fn = lib().tdi_prepare_crops_noalloc if noalloc else lib().tdi_prepare_crops
    _check(fn(n, _dptr(frames), W, H, H * W * 3, coef.ctypes.data, crop.ctypes.data, _dptr(out), _stream()))
    return out


def paint_innermouth_black(frames, lmk_x, lmk_y, ref_face_size: float = 86.0):
    """Inner-mouth masking of 8-bit BGR frames on the GPU, in place (tdi_paint_innermouth_black).
    frames: uint8 contiguous device tensor [n, H, W, 3]; lmk_x, lmk_y: [n, 51] frame pixels.
    Returns frames."""
    import numpy as np
    torch = _torch()
    if (frames.dtype != torch.uint8 or frames.dim() != 4 or frames.shape[3] != 3 or not frames.is_cuda
            or not frames.is_contiguous()):
        raise ValueError("paint_innermouth_black: frames must be a contiguous uint8 CUDA tensor [n, H, W, 3]")
    n, H, W = (int(v) for v in frames.shape[:3])
    lx = np.ascontiguousarray(lmk_x, np.float32).reshape(-1, 51)
    ly = np.ascontiguousarray(lmk_y, np.float32).reshape(-1, 51)
    if lx.shape[0] != n or ly.shape[0] != n:
        raise ValueError("paint_innermouth_black: landmarks must be [n, 51]")
    _check(lib().tdi_paint_innermouth_black(n, _dptr(frames), W, H, H * W * 3, lx.ctypes.data, ly.ctypes.data,
                                            float(ref_face_size), _stream()))
    return frames


def rotation(u):
    torch = _torch()
    u = u.float().contiguous()
    n = u.shape[0]
    R = torch.empty((n, 9), device="cuda", dtype=torch.float32)
    dR = torch.empty((n, 27), device="cuda", dtype=torch.float32)
    _check(lib().tdi_rotation(n, _dptr(u), _dptr(R), _dptr(dR), _stream()))
    return R, dR
