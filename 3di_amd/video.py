"""Python binding of the C++ host layer (libtdi_host.so, include/tdi_host.h).

Mirrors the reference's host-side pieces around the per-frame fit: ``VideoOutput``
(video_fitter.cu:2603-2828) with its .expressions / .poses / .illums writers,
``euler_angles`` (RotationComputer::compute_euler_angles, rotation_header.cu:421-449),
``read_config`` (config.cu:300-560, the keys of this path) and ``fit_video_frames``
(the per-frame sequence of VideoFitter::fit_to_video_frame, video_fitter.cu:554-615).
"""
from __future__ import annotations

import ctypes as C
import importlib
import os

import numpy as np

abi = importlib.import_module("3di_amd.abi")
tdi = importlib.import_module("3di_amd.tdi")

HOST_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtdi_host.so")
_lib = None
_vp = C.c_void_p

EXPRESSIONS, EXPRESSIONS_ALL, POSES, ILLUMS = 0, 1, 2, 3


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise tdi.TdiError(f"{HOST_LIB} is missing: build it with `python 3di_amd/build.py`")
        tdi.lib()                     # libtdi.so first (the host layer links against it)
        L = C.CDLL(HOST_LIB)
        L.tdi_vo_create.restype = _vp
        L.tdi_vo_create.argtypes = [C.c_int32, C.c_int32]
        L.tdi_vo_destroy.argtypes = [_vp]
        L.tdi_vo_add.argtypes = [_vp, C.c_int64, _vp, _vp, _vp]
        L.tdi_vo_save.argtypes = [_vp, C.c_int32, C.c_char_p, C.c_int32]
        L.tdi_vo_dump.argtypes = [_vp, C.c_char_p]
        L.tdi_vo_load.argtypes = [_vp, C.c_char_p]
        L.tdi_fit_video_long.argtypes = [_vp, _vp, _vp, FRAME_SOURCE, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp]
        L.tdi_euler_angles.argtypes = [_vp, _vp]
        L.tdi_read_config.argtypes = [C.c_char_p, _vp, _vp, _vp]
        L.tdi_fit_video_frames.argtypes = [_vp, C.c_int32, C.c_int32, _vp, _vp, _vp, _vp]
        L.tdi_initialize_with_orthographic.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_read_landmarks.argtypes = [C.c_char_p, C.c_int32, _vp, _vp, _vp]
        L.tdi_frame_geometry.argtypes = [C.c_int32, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_float, C.c_int32,
                                         C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_default_video_params.argtypes = [_vp]
        L.tdi_read_video_config.argtypes = [C.c_char_p, _vp, _vp]
        L.tdi_fit_video_frames_nres.argtypes = [_vp, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int64, _vp, _vp,
                                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.tdi_default_identity_params.argtypes = [_vp]
        L.tdi_bin_identity_frames.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp]
        L.tdi_select_identity_frames.argtypes = [_vp, _vp, C.c_int32, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp, _vp,
                                                 C.c_int32, _vp, _vp]
        L.tdi_learn_identity.argtypes = [_vp, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp,
                                         _vp, _vp]
        _lib = L
    return _lib


def _f32p(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(a, np.float32)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


# tdi_frame_source: (user, first, count) -> device pointer to count BGR8 frames
FRAME_SOURCE = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int32, C.c_int32)


class LongVideo(C.Structure):
    """tdi_long_video (include/tdi_host.h)."""
    _fields_ = [("n_frames", C.c_int32), ("src_w", C.c_int32), ("src_h", C.c_int32), ("chunk_frames", C.c_int32),
                ("max_chunks", C.c_int32), ("state_path", C.c_char_p), ("resume", C.c_int32)]


class VideoOutput:
    """Per-frame results of several fits and the reference's output files."""

    def __init__(self, k_exp: int, nres_coefs: int = 1):
        self.k_exp = k_exp
        self.h = lib().tdi_vo_create(k_exp, nres_coefs)
        if not self.h:
            raise ValueError("k_exp and nres_coefs must be positive")

    def add(self, frame_id: int, exp, pose=None, illum=None):
        e = _f32p(exp, self.k_exp)
        p = _f32p(pose, 6)
        il = _f32p(illum, 4)
        rc = lib().tdi_vo_add(self.h, int(frame_id), e.ctypes.data, None if p is None else p.ctypes.data,
                              None if il is None else il.ctypes.data)
        if rc:
            raise ValueError("bad frame")

    def dump(self, path: str):
        """Appends the fits added since the last dump to path (exact, hexadecimal)."""
        if lib().tdi_vo_dump(self.h, os.fsencode(path)):
            raise OSError(path)

    def load(self, path: str):
        """Replays the fits of a dump."""
        if lib().tdi_vo_load(self.h, os.fsencode(path)):
            raise ValueError(f"{path}: not a VideoOutput dump of {self.k_exp} expressions")

    def save(self, kind: int, path: str, flag: int = 0):
        if lib().tdi_vo_save(self.h, kind, os.fsencode(path), flag):
            raise ValueError("nothing to save or unknown kind")

    def save_expressions(self, path, ignore_first=False):
        self.save(EXPRESSIONS, path, int(ignore_first))

    def save_expressions_all(self, path):
        self.save(EXPRESSIONS_ALL, path)

    def save_poses(self, path, with_euler=True):
        self.save(POSES, path, int(with_euler))

    def save_illums(self, path):
        self.save(ILLUMS, path)

    def close(self):
        if self.h:
            lib().tdi_vo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VideoParams(C.Structure):
    """tdi_video_params (include/tdi_host.h)."""
    _fields_ = [("ref_face_size", C.c_float), ("nres_coefs", C.c_int32), ("twostage_illum", C.c_int32),
                ("finetune_expressions", C.c_int32), ("pad", C.c_int32), ("crop_w", C.c_int32), ("crop_h", C.c_int32),
                ("batched", C.c_int32), ("init_pose", C.c_int32), ("box_half_width", C.c_float),
                ("paint_innermouth_black", C.c_int32), ("lmk_k_exp", C.c_int32)]


def video_params(**kw) -> VideoParams:
    """The per-frame stage's parameters, the reference's defaults (config.cu) overridden by kw."""
    p = VideoParams()
    lib().tdi_default_video_params(C.addressof(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def read_video_config(path: str, opts=None, params=None):
    """Applies a 3DI config file to fit options and the per-frame stage's parameters."""
    o = opts if opts is not None else abi.default_options()
    p = params if params is not None else video_params()
    if lib().tdi_read_video_config(os.fsencode(path), C.addressof(o), C.addressof(p)):
        raise FileNotFoundError(path)
    return o, p


def fit_video_frames_nres(ctx, params, frames, lmk_x, lmk_y, cams, theta, lam, frame_ids=None, boxes=None,
                          vo=None, Lint=None):
    """The per-frame stage of a video (tdi_fit_video_frames_nres): frames uint8 CUDA tensor
    [n, H, W, 3] (B, G, R); lmk_x, lmk_y [n, 51], cams [n, 4] of the unresized frames; theta [n, K]
    and lam [n, 3] the start.  The frames are painted in place when params.paint_innermouth_black
    (the inner-mouth masking), so they must be contiguous.  Returns (theta, lam, Lint, status,
    reports)."""
    n, H, W = (int(v) for v in frames.shape[:3])
    if not frames.is_contiguous():
        raise ValueError("fit_video_frames_nres: frames must be a contiguous uint8 CUDA tensor [n, H, W, 3]")
    K = ctx.K
    th = np.ascontiguousarray(np.array(theta, np.float32).reshape(n, K))
    la = np.ascontiguousarray(np.array(lam, np.float32).reshape(n, 3))
    Li = np.ascontiguousarray(np.full(n, 0.005, np.float32) if Lint is None else np.array(Lint, np.float32))
    lx = np.ascontiguousarray(lmk_x, np.float32)
    ly = np.ascontiguousarray(lmk_y, np.float32)
    cm = np.ascontiguousarray(cams, np.float32)
    ids = np.ascontiguousarray(np.arange(n) if frame_ids is None else frame_ids, np.int64)
    bx = None if boxes is None else np.ascontiguousarray(boxes, np.float32)
    st = np.zeros(n, np.int32)
    reps = (abi.tdi_fit_report * n)()
    rc = lib().tdi_fit_video_frames_nres(ctx.h, C.addressof(params), n, frames.data_ptr(), W, H, H * W * 3,
                                         lx.ctypes.data, ly.ctypes.data, cm.ctypes.data,
                                         None if bx is None else bx.ctypes.data, ids.ctypes.data, th.ctypes.data,
                                         la.ctypes.data, Li.ctypes.data, None if vo is None else vo.h,
                                         st.ctypes.data, C.addressof(reps))
    if rc:
        raise tdi.TdiError(f"tdi_fit_video_frames_nres: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return th, la, Li, st, list(reps)


def fit_video_long(ctx, params, n_frames: int, width: int, height: int, source, lmk_x, lmk_y, cams, theta, lam,
                   vo, Lint: float = 0.005, chunk_frames: int = 0, max_chunks: int = 0, state_path=None,
                   resume: bool = False):
    """The per-frame stage of a long video, chunk by chunk (tdi_fit_video_long).  source(first,
    count) returns a contiguous uint8 CUDA tensor [count, height, width, 3] of those frames (kept
    alive until the next call; painted in place).  theta [K], lam [3]: the start.  Returns
    (theta, lam, Lint, status [n_frames], frames_done)."""
    K = vo.k_exp + 6
    keep = {}

    def _src(user, first, count):
        t = source(int(first), int(count))
        if (t is None or not t.is_cuda or not t.is_contiguous() or t.dtype != torch_uint8()
                or tuple(t.shape) != (count, height, width, 3)):
            return None
        keep["frames"] = t
        return t.data_ptr()

    cb = FRAME_SOURCE(_src)
    lv = LongVideo(n_frames=n_frames, src_w=width, src_h=height, chunk_frames=chunk_frames, max_chunks=max_chunks,
                   state_path=None if state_path is None else os.fsencode(state_path), resume=int(resume))
    lx = np.ascontiguousarray(lmk_x, np.float32).reshape(n_frames, 51)
    ly = np.ascontiguousarray(lmk_y, np.float32).reshape(n_frames, 51)
    cm = np.ascontiguousarray(cams, np.float32).reshape(n_frames, 4)
    th = np.ascontiguousarray(np.array(theta, np.float32).reshape(K))
    la = np.ascontiguousarray(np.array(lam, np.float32).reshape(3))
    Li = np.array([Lint], np.float32)
    st = np.zeros(n_frames, np.int32)
    done = np.zeros(1, np.int32)
    rc = lib().tdi_fit_video_long(ctx.h, C.addressof(params), C.addressof(lv), cb, None, lx.ctypes.data,
                                  ly.ctypes.data, cm.ctypes.data, th.ctypes.data, la.ctypes.data, Li.ctypes.data,
                                  vo.h, st.ctypes.data, done.ctypes.data)
    if rc:
        raise tdi.TdiError(f"tdi_fit_video_long: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return th, la, float(Li[0]), st, int(done[0])


def torch_uint8():
    import torch
    return torch.uint8


def euler_angles(u) -> np.ndarray:
    """(yaw, pitch, roll) of the rotation with axis-angle vector u."""
    uu = np.ascontiguousarray(u, np.float32)
    out = np.zeros(3, np.float32)
    lib().tdi_euler_angles(uu.ctypes.data, out.ctypes.data)
    return out


def read_config(path: str, opts=None):
    """Applies a 3DI config file to fit options; returns (opts, twostage_illum, nres_coefs)."""
    o = opts if opts is not None else abi.default_options()
    two = C.c_int32(0)
    nres = C.c_int32(1)
    if lib().tdi_read_config(os.fsencode(path), C.addressof(o), C.addressof(two), C.addressof(nres)):
        raise FileNotFoundError(path)
    return o, two.value, nres.value


def fit_video_frames(batch, theta, lam, twostage_illum: bool = False, batched: bool = False):
    """fit_to_video_frame's sequence for every frame of `batch` (a tdi.Batch whose light is not
    an unknown): L = 0.005, lambda fit, shape fit (+ two more rounds with twostage_illum).
    Returns (theta, lam, Lint, reports of the last shape fit)."""
    n, K = batch.n, batch.K
    th = np.ascontiguousarray(np.array(theta, np.float32).reshape(n, K))
    la = np.ascontiguousarray(np.array(lam, np.float32).reshape(n, 3))
    Li = np.zeros(n, np.float32)
    reps = (abi.tdi_fit_report * n)()
    rc = lib().tdi_fit_video_frames(batch.h, int(twostage_illum), int(batched), th.ctypes.data, la.ctypes.data,
                                    Li.ctypes.data, C.addressof(reps))
    if rc:
        raise tdi.TdiError(f"tdi_fit_video_frames: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return th, la, Li, list(reps)


def initialize_with_orthographic(ctx, lmk_x, lmk_y, cams):
    """initialize_with_orthographic_t for n frames (host arrays): lmk_x, lmk_y [n, 51], cams [n, 4]
    (phix, phiy, cx, cy).  Returns pose [n, 6] (tau, u), angles [n, 3] (degrees), info [n, 4]."""
    lx = np.ascontiguousarray(lmk_x, np.float32)
    ly = np.ascontiguousarray(lmk_y, np.float32)
    cm = np.ascontiguousarray(cams, np.float32)
    n = lx.shape[0]
    pose = np.zeros((n, 6), np.float32)
    ang = np.zeros((n, 3), np.float32)
    info = np.zeros((n, 4), np.int32)
    rc = lib().tdi_initialize_with_orthographic(ctx.h, n, lx.ctypes.data, ly.ctypes.data, cm.ctypes.data,
                                                pose.ctypes.data, ang.ctypes.data, info.ctypes.data)
    if rc:
        raise tdi.TdiError(f"tdi_initialize_with_orthographic: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return pose, ang, info


def read_landmarks(path: str):
    """Landmark text file -> (lmk_x, lmk_y) [n, 51] float32 (LandmarkData::init_from_txtfile)."""
    n = C.c_int32(0)
    rc = lib().tdi_read_landmarks(path.encode(), 0, None, None, C.byref(n))
    if rc:
        raise tdi.TdiError(f"tdi_read_landmarks: {rc}: cannot read {path}")
    lx = np.zeros((n.value, tdi.NLANDMARKS), np.float32)
    ly = np.zeros_like(lx)
    rc = lib().tdi_read_landmarks(path.encode(), n.value, lx.ctypes.data, ly.ctypes.data, C.byref(n))
    if rc:
        raise tdi.TdiError(f"tdi_read_landmarks: {rc}: cannot read {path}")
    return lx[:n.value], ly[:n.value]


def frame_geometry(lmk_x, lmk_y, cams, frame_w: int, frame_h: int, ref_face_size: float, pad: bool = True,
                   crop_size=(256, 256)):
    """Per-frame resize coefficient, resized landmarks and camera, crop and face size of n frames
    (tdi_frame_geometry).  Returns a dict of host arrays: coef [n], lmk_x, lmk_y [n, 51],
    cams [n, 4], crop [n, 4], face_size [n], status [n]."""
    lx = np.ascontiguousarray(lmk_x, np.float32)
    ly = np.ascontiguousarray(lmk_y, np.float32)
    cm = np.ascontiguousarray(cams, np.float32)
    n = lx.shape[0]
    if lx.shape != (n, tdi.NLANDMARKS) or ly.shape != lx.shape or cm.shape != (n, 4):
        raise ValueError("frame_geometry: lmk_x, lmk_y must be [n, 51] and cams [n, 4]")
    out = dict(coef=np.zeros(n, np.float32), lmk_x=np.zeros_like(lx), lmk_y=np.zeros_like(ly),
               cams=np.zeros_like(cm), crop=np.zeros((n, 4), np.int32), face_size=np.zeros(n, np.float32),
               status=np.zeros(n, np.int32))
    rc = lib().tdi_frame_geometry(n, lx.ctypes.data, ly.ctypes.data, cm.ctypes.data, int(frame_w), int(frame_h),
                                  float(ref_face_size), int(bool(pad)), int(crop_size[0]), int(crop_size[1]),
                                  out["coef"].ctypes.data, out["lmk_x"].ctypes.data, out["lmk_y"].ctypes.data,
                                  out["cams"].ctypes.data, out["crop"].ctypes.data, out["face_size"].ctypes.data,
                                  out["status"].ctypes.data)
    if rc:
        raise tdi.TdiError(f"tdi_frame_geometry: {rc}")
    return out


def prepare_frames(frames, lmk_x, lmk_y, cams, ref_face_size: float, pad: bool = True, crop_size=(256, 256)):
    """The frame preparation of the video loop for n frames: frame_geometry on the host, then the
    grey crops of the frames that passed (status 0) on the GPU (tdi.prepare_crops).  frames: uint8
    CUDA tensor [n, H, W, 3] (B, G, R).  Returns the geometry dict with "obs", the crops of the
    frames that passed, concatenated column-major in frame order, and "frames", their indices."""
    n, H, W = (int(v) for v in frames.shape[:3])
    g = frame_geometry(lmk_x, lmk_y, cams, W, H, ref_face_size, pad, crop_size)
    keep = np.flatnonzero(g["status"] == 0)
    g["frames"] = keep
    if keep.size:
        import torch
        sel = frames if keep.size == n else frames[torch.from_numpy(keep).to(frames.device)]
        g["obs"] = tdi.prepare_crops(sel, g["coef"][keep], g["crop"][keep])
    else:
        g["obs"] = None
    return g


# ------------------------------------------------------------------------ identity stage
class IdentityParams(C.Structure):
    """tdi_identity_params (include/tdi_host.h)."""
    _fields_ = [("nframes", C.c_int32), ("nframes_per_angle_bin", C.c_int32), ("nperms", C.c_int32),
                ("ntot_reconstrs", C.c_int32), ("nmax_frames", C.c_int32), ("max_vid_frames", C.c_int32),
                ("skip_first_n_secs", C.c_float), ("fps", C.c_float), ("ref_face_size", C.c_float),
                ("pad", C.c_int32), ("crop_w", C.c_int32), ("crop_h", C.c_int32),
                ("paint_innermouth_black", C.c_int32), ("lmk_k_exp", C.c_int32), ("box_half_width", C.c_float),
                ("lmk_k_id", C.c_int32)]


def identity_params(**kw) -> IdentityParams:
    """The identity stage's parameters, the reference's defaults (config.cu) overridden by kw."""
    p = IdentityParams()
    lib().tdi_default_identity_params(C.addressof(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def bin_identity_frames(p, yaw, pitch, cap=None):
    """tdi_bin_identity_frames: selected frame indices (groups of p.nframes) from per-frame yaw and
    pitch in radians (NaN: no angle)."""
    yaw = np.ascontiguousarray(yaw, np.float32)
    pitch = np.ascontiguousarray(pitch, np.float32)
    n = yaw.size
    cap = cap or max(1, p.ntot_reconstrs * p.nframes)
    out = np.zeros(cap, np.int32)
    ns = C.c_int32(0)
    rc = lib().tdi_bin_identity_frames(C.addressof(p), n, yaw.ctypes.data, pitch.ctypes.data, out.ctypes.data, cap,
                                       C.addressof(ns))
    if rc:
        raise tdi.TdiError(f"tdi_bin_identity_frames: {rc}")
    return out[:ns.value].copy()


def select_identity_frames(ctx, p, lmk_x, lmk_y, width, height, cam):
    """tdi_select_identity_frames: (selected frame indices, angles [n, 3] radians, resize coefficient)."""
    lx = np.ascontiguousarray(lmk_x, np.float32)
    ly = np.ascontiguousarray(lmk_y, np.float32)
    n = lx.shape[0]
    cm = np.ascontiguousarray(cam, np.float32).reshape(4)
    ang = np.zeros((n, 3), np.float32)
    cap = max(1, p.ntot_reconstrs * p.nframes)
    out = np.zeros(cap, np.int32)
    ns = C.c_int32(0)
    coef = C.c_float(0)
    rc = lib().tdi_select_identity_frames(ctx.h, C.addressof(p), n, lx.ctypes.data, ly.ctypes.data, int(width),
                                          int(height), cm.ctypes.data, ang.ctypes.data, out.ctypes.data, cap,
                                          C.addressof(ns), C.addressof(coef))
    if rc:
        raise tdi.TdiError(f"tdi_select_identity_frames: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return out[:ns.value].copy(), ang, coef.value


def learn_identity(ctx, p, frames, lmk_x, lmk_y, cam):
    """tdi_learn_identity over the selected frames (uint8 CUDA tensor [n, H, W, 3], contiguous, in
    selection order; masked in place): (alpha [k_id], beta [k_tex], groups used)."""
    n, H, W = (int(v) for v in frames.shape[:3])
    if not frames.is_contiguous():
        raise ValueError("learn_identity: frames must be a contiguous uint8 CUDA tensor [n, H, W, 3]")
    lx = np.ascontiguousarray(lmk_x, np.float32).reshape(n, 51)
    ly = np.ascontiguousarray(lmk_y, np.float32).reshape(n, 51)
    cm = np.ascontiguousarray(cam, np.float32).reshape(4)
    m = ctx.model
    alpha = np.zeros(max(1, m.k_id), np.float32)
    beta = np.zeros(max(1, m.k_tex), np.float32)
    used = C.c_int32(0)
    rc = lib().tdi_learn_identity(ctx.h, C.addressof(p), n, frames.data_ptr(), W, H, W * H * 3, lx.ctypes.data,
                                  ly.ctypes.data, cm.ctypes.data, alpha.ctypes.data, beta.ctypes.data,
                                  C.addressof(used))
    if rc:
        raise tdi.TdiError(f"tdi_learn_identity: {rc}: {tdi.lib().tdi_last_error().decode()}")
    return alpha[:m.k_id], beta[:m.k_tex], used.value
