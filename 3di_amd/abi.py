"""ctypes mirror of include/tdi_api.h (structs, status codes, descriptor builders).

Descriptors hold raw pointers into numpy arrays; ``ModelDesc``/``FrameDescs`` keep
references to those arrays so they stay alive as long as the descriptor does.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

TDI_OK = 0
TDI_ERR_ARG = 1
TDI_ERR_HIP = 2
TDI_ERR_NOMEM = 3
TDI_NOT_PD = 4
TDI_NAN = 5
TDI_OUT_OF_FRAME = 6
TDI_TOO_FEW_PIXELS = 7

STOP_MAXITER, STOP_DECREMENT, STOP_STEP, STOP_NAN, STOP_NOT_PD = 1, 2, 4, 8, 16

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class tdi_model_desc(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("n_triangles", C.c_int32),
                ("k_id", C.c_int32), ("k_tex", C.c_int32), ("k_exp", C.c_int32),
                ("mean_shape", _fp), ("id_basis", _fp), ("exp_basis", _fp),
                ("tex_mean", _fp), ("tex_basis", _fp), ("triangles", _ip), ("landmark_ids", _ip),
                ("id_lb", _fp), ("id_ub", _fp), ("tex_lb", _fp), ("tex_ub", _fp),
                ("exp_lb", _fp), ("exp_ub", _fp)]


class tdi_frame_desc(C.Structure):
    _fields_ = [("phix", C.c_float), ("phiy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("x0", C.c_int32), ("y0", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("face_size", C.c_float), ("obs", _fp), ("lmk_x", _fp), ("lmk_y", _fp),
                ("box_lx", _fp), ("box_ux", _fp), ("box_ly", _fp), ("box_uy", _fp),
                ("lambda_", C.c_float * 3), ("Lintensity", C.c_float)]


class tdi_fit_options(C.Structure):
    _fields_ = [("fit_id", C.c_int32), ("fit_tex", C.c_int32), ("fit_exp", C.c_int32), ("fit_light", C.c_int32),
                ("max_iters", C.c_int32), ("max_ls", C.c_int32), ("ntmp_cap", C.c_int32),
                ("use_landmarks", C.c_int32), ("niters_success", C.c_int32),
                ("data_weight", C.c_float), ("armijo_alpha", C.c_float), ("armijo_beta", C.c_float),
                ("t_min", C.c_float), ("decrement_tol", C.c_float), ("blur_sigma", C.c_float),
                ("ref_face_size", C.c_float), ("mixed_flags", C.c_int32), ("finetune_coef", C.c_float)]


class tdi_fit_report(C.Structure):
    _fields_ = [("status", C.c_int32), ("stop", C.c_int32), ("iters", C.c_int32), ("ls_evals", C.c_int32),
                ("n_pixels", C.c_int32), ("success", C.c_int32), ("objective", C.c_float), ("pad", C.c_float)]


class tdi_lmk_report(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("rounds", C.c_int32), ("accepted", C.c_int32),
                ("slack0", C.c_float), ("slack", C.c_float)]


TDI_LMK_INFEASIBLE = 8


def default_options(**kw) -> tdi_fit_options:
    """Reference defaults (config.cu:15-140, model_fitter.cu:137-345)."""
    o = tdi_fit_options(fit_id=0, fit_tex=0, fit_exp=1, fit_light=0, max_iters=100, max_ls=100,
                        ntmp_cap=0, use_landmarks=1, niters_success=6, data_weight=25.0,
                        armijo_alpha=0.4, armijo_beta=0.8, t_min=0.075, decrement_tol=3e-4,
                        blur_sigma=0.75, ref_face_size=0.0, mixed_flags=7, finetune_coef=1.5)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


def _ptr(a, t=_fp):
    return C.cast(a.ctypes.data, t) if a is not None else C.cast(None, t)


class ModelDesc:
    """Owns contiguous float32/int32 copies of a synth.Model and the C descriptor."""

    def __init__(self, m):
        self.keep = dict(
            mean_shape=_f32(m.mean_shape), id_basis=_f32(m.id_basis), exp_basis=_f32(m.exp_basis),
            tex_mean=_f32(m.tex_mean), tex_basis=_f32(m.tex_basis),
            triangles=np.ascontiguousarray(m.triangles, np.int32),
            landmark_ids=np.ascontiguousarray(m.landmark_ids, np.int32),
            id_lb=_f32(m.id_lb), id_ub=_f32(m.id_ub), tex_lb=_f32(m.tex_lb), tex_ub=_f32(m.tex_ub),
            exp_lb=_f32(m.exp_lb), exp_ub=_f32(m.exp_ub))
        k = self.keep
        self.desc = tdi_model_desc(
            n_vertices=m.n_vertices, n_triangles=m.n_triangles, k_id=m.k_id, k_tex=m.k_tex, k_exp=m.k_exp,
            mean_shape=_ptr(k["mean_shape"]), id_basis=_ptr(k["id_basis"]), exp_basis=_ptr(k["exp_basis"]),
            tex_mean=_ptr(k["tex_mean"]), tex_basis=_ptr(k["tex_basis"]),
            triangles=_ptr(k["triangles"], _ip), landmark_ids=_ptr(k["landmark_ids"], _ip),
            id_lb=_ptr(k["id_lb"]), id_ub=_ptr(k["id_ub"]), tex_lb=_ptr(k["tex_lb"]), tex_ub=_ptr(k["tex_ub"]),
            exp_lb=_ptr(k["exp_lb"]), exp_ub=_ptr(k["exp_ub"]))

    @property
    def ref(self):
        return C.byref(self.desc)


class FrameDescs:
    """Array of tdi_frame_desc for a list of synth.Frame, keeping the arrays alive."""

    def __init__(self, frames: Sequence):
        n = len(frames)
        self.arr = (tdi_frame_desc * n)()
        self.keep = []
        for i, f in enumerate(frames):
            arrs = [_f32(f.obs), _f32(f.lmk_x), _f32(f.lmk_y), _f32(f.box_lx), _f32(f.box_ux),
                    _f32(f.box_ly), _f32(f.box_uy)]
            self.keep.append(arrs)
            d = self.arr[i]
            d.phix, d.phiy, d.cx, d.cy = f.phix, f.phiy, f.cx, f.cy
            d.x0, d.y0, d.width, d.height = f.x0, f.y0, f.width, f.height
            d.face_size = f.face_size
            d.obs, d.lmk_x, d.lmk_y, d.box_lx, d.box_ux, d.box_ly, d.box_uy = [_ptr(a) for a in arrs]
            for j in range(3):
                d.lambda_[j] = float(f.lam[j])
            d.Lintensity = float(f.Lint)
        self.n = n

    def __len__(self):
        return self.n

    def ptr(self, i=0):
        return C.cast(C.byref(self.arr, i * C.sizeof(tdi_frame_desc)), C.POINTER(tdi_frame_desc))
