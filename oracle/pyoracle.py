"""ctypes loader of the FP64 OpenMP oracle (libtdi_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  Never imported by the product.
"""
from __future__ import annotations

import ctypes as C
import importlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libtdi_oracle.so")
abi = importlib.import_module("3di_amd.abi")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_num_unknowns.restype = C.c_int
        L.orc_blur_taps.argtypes = [C.c_double, _dp]
        L.orc_rodrigues.argtypes = [_dp, _dp, _dp]
        L.orc_blur.argtypes = [_dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_rasterize.argtypes = [C.c_int, _ip, _fp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _up, _fp]
        L.orc_geometry.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _up, _fp, _dp, _dp, _dp]
        L.orc_render.restype = C.c_int
        L.orc_normal_equations.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _dp, _dp, _dp, _ip,
                                           C.c_int]
        L.orc_normal_equations.restype = C.c_int
        L.orc_barrier.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _dp]
        L.orc_barrier.restype = C.c_double
        L.orc_cholesky_solve.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_render_loss.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _ip, C.c_int]
        L.orc_render_loss.restype = C.c_double
        L.orc_fit_landmarks_slack.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_void_p]
        L.orc_landmark_slack_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _dp, _dp, C.c_double,
                                               _dp, _dp]
        L.orc_landmark_slack_terms.restype = C.c_double
        L.orc_fit_landmarks_joint.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                              _dp, C.c_void_p]
        L.orc_fit_landmarks_joint.restype = C.c_int
        L.orc_landmark_joint_terms.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                               _dp, _dp, C.c_double, _dp, _dp]
        L.orc_landmark_joint_terms.restype = C.c_double
        L.orc_fit.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _dp, C.c_void_p, C.c_int, C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_fit_light.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, _dp, _dp, _dp, C.c_void_p, C.c_int,
                                    C.c_int]
        L.orc_fit_light.restype = C.c_int
        L.orc_ortho_residuals.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_rotvec_from_euler.argtypes = [_dp, _dp]
        L.orc_euler_angles.argtypes = [_dp, _dp]
        L.orc_ortho_init.argtypes = [C.c_int, _fp, _fp, _dp, _dp, _dp, _dp, _ip]
        L.orc_prepare_crops.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_longlong, _fp, _ip, _fp]
        L.orc_resized_size.argtypes = [C.c_int, C.c_float]
        L.orc_paint_innermouth_black.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_longlong, _fp, _fp,
                                                 C.c_float]
        L.orc_resized_size.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def blur_taps(sigma):
    t = np.zeros(15)
    rc = lib().orc_blur_taps(sigma, p(t))
    if rc:
        raise ValueError(sigma)
    return t


def rodrigues(u):
    R = np.zeros(9)
    dR = np.zeros(27)
    lib().orc_rodrigues(p(_d(u)), p(R), p(dR))
    return R.reshape(3, 3).T.copy(), dR.reshape(3, 3, 3).transpose(0, 2, 1).copy()


def blur(img, W, H, sigma):
    out = np.zeros(W * H)
    lib().orc_blur(p(_d(img)), p(out), W, H, p(blur_taps(sigma)))
    return out


def rasterize(tris, xp, yp, z, x0, y0, W, H, ntmp=0):
    keys = np.zeros(W * H, np.uint64)
    bary = np.zeros(2 * W * H, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    xp, yp, z = (np.ascontiguousarray(a, np.float32) for a in (xp, yp, z))
    lib().orc_rasterize(tris.shape[0], p(tris, _ip), p(xp, _fp), p(yp, _fp), p(z, _fp), x0, y0, W, H, ntmp,
                        p(keys, _up), p(bary, _fp))
    return keys, bary


class Scene:
    """A model + frames + options bound to the oracle."""

    def __init__(self, model, frames, opts):
        self.model = model
        self.md = abi.ModelDesc(model)
        self.fd = abi.FrameDescs(frames)
        self.frames = frames
        self.opts = opts
        self.K = lib().orc_num_unknowns(self.md.ref, C.byref(opts))

    def _f(self, i):
        return self.fd.ptr(i)

    def geometry(self, i, theta):
        V = self.model.n_vertices
        xyz = np.zeros(3 * V)
        xyp = np.zeros(2 * V)
        R = np.zeros(9)
        dR = np.zeros(27)
        lib().orc_geometry(self.md.ref, self._f(i), C.byref(self.opts), p(_d(theta)), p(xyz), p(xyp), p(R), p(dR))
        return xyz.reshape(V, 3), xyp.reshape(V, 2), R, dR

    def render(self, i, theta):
        f = self.frames[i]
        A = f.width * f.height
        keys = np.zeros(A, np.uint64)
        bary = np.zeros(2 * A, np.float32)
        raw = np.zeros(A)
        img = np.zeros(A)
        obj = np.zeros(3)
        n = lib().orc_render(self.md.ref, self._f(i), C.byref(self.opts), p(_d(theta)), p(keys, _up), p(bary, _fp),
                             p(raw), p(img), p(obj))
        return dict(keys=keys, bary=bary, raw=raw, img=img, obj=obj, npix=n)

    def normal_equations(self, i, theta, want_dI=False, cap=200000):
        K = self.K
        H = np.zeros(K * K)
        g = np.zeros(K)
        obj = np.zeros(3)
        dI = np.zeros(cap * K) if want_dI else None
        nb = np.zeros(cap * 4, np.int32) if want_dI else None
        n = lib().orc_normal_equations(self.md.ref, self._f(i), C.byref(self.opts), p(_d(theta)), p(H), p(g),
                                       p(obj), p(dI), p(nb, _ip), cap)
        out = dict(H=H.reshape(K, K), g=g, obj=obj, npix=n)
        if want_dI:
            out["dI"] = dI[:n * K].reshape(n, K)
            out["nbr"] = nb[:4 * n].reshape(n, 4)
        return out

    def barrier(self, i, theta):
        K = self.K
        g = np.zeros(K)
        H = np.zeros(K * K)
        obj = lib().orc_barrier(self.md.ref, self._f(i), C.byref(self.opts), p(_d(theta)), p(g), p(H))
        return obj, g, H.reshape(K, K)

    def render_loss(self, i, theta, dverts=None, tri=None, frozen=False):
        """The render loss of frame i (sum over visible pixels of |g^_r - g^_s|^2 plus the squared
        landmark reprojection errors) at theta with per-vertex offsets dverts [V, 3] (model space).
        tri: int32 [W*H] triangle per pixel: written (the rasteriser's) unless frozen, read when
        frozen (visibility held fixed, exact FP64 barycentrics).  Returns the loss."""
        dv = None if dverts is None else _d(dverts).reshape(-1)
        return lib().orc_render_loss(self.md.ref, self._f(i), C.byref(self.opts), p(_d(theta)), p(dv),
                                     p(tri, _ip), int(frozen))

    def fit_landmarks_slack(self, theta0, k_exp):
        """The landmark-only fit with slack of every frame: (theta [n, K] float64, reports)."""
        n = len(self.frames)
        th = np.ascontiguousarray(np.array(theta0, np.float64).reshape(n, self.K))
        reps = (abi.tdi_lmk_report * n)()
        rc = lib().orc_fit_landmarks_slack(self.md.ref, n, self.fd.ptr(0), C.byref(self.opts), int(k_exp), p(th), reps)
        if rc:
            raise ValueError(f"orc_fit_landmarks_slack: error {rc}")
        return th, list(reps)

    def landmark_slack_terms(self, i, theta, k_exp, z, tau, derivs=True):
        """(objective, g [k_exp+7], H [k_exp+7, k_exp+7]) of that fit at z for frame i."""
        nz = int(k_exp) + 7
        z = _d(z)
        assert z.size == nz
        g = np.zeros(nz) if derivs else None
        H = np.zeros(nz * nz) if derivs else None
        obj = lib().orc_landmark_slack_terms(self.md.ref, self._f(i), C.byref(self.opts), int(k_exp), p(_d(theta)),
                                             p(z), float(tau), p(g) if derivs else None, p(H) if derivs else None)
        return (obj, g, H.reshape(nz, nz)) if derivs else (obj, None, None)

    def joint_landmark_unknowns(self, k_id, k_exp, slack):
        return int(k_id) + (1 if slack else 0) + len(self.frames) * (int(k_exp) + 6)

    def fit_landmarks_joint(self, theta0, k_id, k_exp, slack):
        """The joint landmark-only fit over all frames of the scene (one group): (theta [n, K]
        float64, report)."""
        n = len(self.frames)
        th = np.ascontiguousarray(np.array(theta0, np.float64).reshape(n, self.K))
        rep = abi.tdi_lmk_report()
        rc = lib().orc_fit_landmarks_joint(self.md.ref, n, self.fd.ptr(0), C.byref(self.opts), int(k_id), int(k_exp),
                                           int(bool(slack)), p(th), C.byref(rep))
        if rc:
            raise ValueError(f"orc_fit_landmarks_joint: error {rc}")
        return th, rep

    def landmark_joint_terms(self, theta, k_id, k_exp, slack, z, tau, derivs=True):
        """(objective, g [N], H [N, N]) of the joint landmark fit at z (joint order: identity,
        slack, then each frame's expressions and pose); theta: frame 0's row supplies the fixed
        identity coefficients past k_id."""
        nz = self.joint_landmark_unknowns(k_id, k_exp, slack)
        z = _d(z)
        assert z.size == nz
        g = np.zeros(nz) if derivs else None
        H = np.zeros(nz * nz) if derivs else None
        obj = lib().orc_landmark_joint_terms(self.md.ref, len(self.frames), self.fd.ptr(0), C.byref(self.opts),
                                             int(k_id), int(k_exp), int(bool(slack)), p(_d(theta)), p(z), float(tau),
                                             p(g) if derivs else None, p(H) if derivs else None)
        return (obj, g, H.reshape(nz, nz)) if derivs else (obj, None, None)

    def fit(self, theta0, n_threads=0, frames=None, finetune=False):
        n = len(self.frames)
        th = np.ascontiguousarray(np.array(theta0, np.float64).reshape(n, self.K))
        reps = (abi.tdi_fit_report * n)()
        lib().orc_fit(self.md.ref, n, self.fd.ptr(0), C.byref(self.opts), p(th), reps, n_threads, int(finetune))
        return th, list(reps)


def _light_fit(scene, theta, lam, Lint, which, n_threads):
    n = len(scene.frames)
    th = np.ascontiguousarray(np.array(theta, np.float64).reshape(n, scene.K))
    lam = np.ascontiguousarray(np.array(lam, np.float64).reshape(n, 3))
    Li = np.ascontiguousarray(np.array(Lint, np.float64).reshape(n))
    reps = (abi.tdi_fit_report * n)()
    rc = lib().orc_fit_light(scene.md.ref, n, scene.fd.ptr(0), C.byref(scene.opts), p(th), p(lam), p(Li), reps, which,
                             n_threads)
    if rc:
        raise ValueError("light sub-fits need options with fit_light = 0")
    return lam, Li, list(reps)


def fit_lambdas(scene, theta, lam, Lint, n_threads=0):
    """fit_3DMM_lambdas on every frame of `scene` (light position, shape and pose fixed)."""
    return _light_fit(scene, theta, lam, Lint, 0, n_threads)


def fit_lintensity(scene, theta, lam, Lint, n_threads=0):
    """fit_3DMM_Lintensity on every frame of `scene`."""
    return _light_fit(scene, theta, lam, Lint, 1, n_threads)


def cholesky_solve(H, g):
    K = len(g)
    d = np.zeros(K)
    rc = lib().orc_cholesky_solve(K, p(_d(H)), p(_d(g)), p(d))
    return rc, d


def ortho_residuals(x, xlb, ylb, p0L, jacobian=True):
    """Residuals [102] and Jacobian [102, 7] of the orthographic landmark model at x."""
    e = np.zeros(102)
    J = np.zeros(102 * 7) if jacobian else None
    lib().orc_ortho_residuals(p(_d(x)), p(_d(xlb)), p(_d(ylb)), p(_d(p0L)), p(e), p(J))
    return (e, J.reshape(102, 7)) if jacobian else e


def rotvec_from_euler(e):
    """Rotation vector of Rz(e[2]) Ry(e[1]) Rx(e[0])."""
    r = np.zeros(3)
    lib().orc_rotvec_from_euler(p(_d(e)), p(r))
    return r


def euler_angles(R):
    """(yaw, pitch, roll) in radians of a rotation matrix R (3x3, as a numpy matrix)."""
    ypr = np.zeros(3)
    lib().orc_euler_angles(p(_d(np.asarray(R).T.reshape(-1))), p(ypr))
    return ypr


def ortho_init(lmk, cams, p0L):
    """Orthographic initialisation of n frames: lmk [n, 2, 51], cams [n, 4] (phix, phiy, cx, cy),
    p0L [51, 3].  Returns x_orth [n, 7], pose [n, 6], angles [n, 3] (degrees), info [n, 4]."""
    lmk = np.ascontiguousarray(lmk, np.float32)
    cams = np.ascontiguousarray(cams, np.float32)
    n = lmk.shape[0]
    xo = np.zeros((n, 7))
    pose = np.zeros((n, 6))
    ang = np.zeros((n, 3))
    info = np.zeros((n, 4), np.int32)
    lib().orc_ortho_init(n, p(lmk, _fp), p(cams, _fp), p(_d(p0L)), p(xo), p(pose), p(ang), p(info, _ip))
    return dict(x_orth=xo, pose=pose, angles=ang, info=info)


def prepare_crops(frames, coef, crop):
    """Grey crops of resized 8-bit BGR frames: frames [n, sh, sw, 3] uint8, coef [n], crop [n, 4]
    (x0, y0, width, height).  Returns the crops concatenated, each column-major (width*height)."""
    frames = np.ascontiguousarray(frames, np.uint8)
    coef = np.ascontiguousarray(coef, np.float32)
    crop = np.ascontiguousarray(crop, np.int32)
    n, sh, sw = frames.shape[:3]
    out = np.zeros(int((crop[:, 2].astype(np.int64) * crop[:, 3]).sum()), np.float32)
    lib().orc_prepare_crops(n, frames.ctypes.data, sw, sh, sh * sw * 3, p(coef, _fp), p(crop, _ip), p(out, _fp))
    return out


def resized_size(n, coef):
    """cv::resize's destination size for a source of n pixels at scale coef."""
    return lib().orc_resized_size(int(n), float(coef))


def paint_innermouth_black(frames, lmk_x, lmk_y, ref_face_size=86.0):
    """The inner-mouth masking of uint8 BGR frames [n, H, W, 3] (a painted copy is returned);
    lmk_x, lmk_y [n, 51] frame pixels."""
    fr = np.ascontiguousarray(np.array(frames, np.uint8))
    n, H, W = fr.shape[:3]
    lx = np.ascontiguousarray(lmk_x, np.float32)
    ly = np.ascontiguousarray(lmk_y, np.float32)
    lib().orc_paint_innermouth_black(n, fr.ctypes.data, W, H, H * W * 3, p(lx, _fp), p(ly, _fp), ref_face_size)
    return fr


class JointScene:
    """The identity stage's joint system on the oracle (the shape fit of fit_to_multi_images,
    model_fitter.cu:1252-1479: fit_3DMM_shape_rigid over T frames, model_fitter.cu:121-375, with
    the frames' products summed as get_obj_hessian_and_gradient_multiframe does, 21-116).  The
    frames share the albedo and identity coefficients, the first Ks unknowns of every frame's
    layout; x = [shared | frame 0's rest | frame 1's rest | ...].  Every frame's normal equations
    come from the oracle with the shared bounds left out (a copy of the model whose albedo and
    identity bounds are infinite, so that the barrier skips them); the shared bounds' terms are
    the oracle's barrier with and without them at frame 0, counted once."""

    def __init__(self, model, frames, opts):
        import dataclasses
        self.full = Scene(model, frames, opts)
        inf = lambda a: None if a is None else np.full(np.shape(a), np.inf, np.float32)
        free = dataclasses.replace(model, id_lb=None if model.id_lb is None else -inf(model.id_lb),
                                   id_ub=inf(model.id_ub), tex_lb=None if model.tex_lb is None else -inf(model.tex_lb),
                                   tex_ub=inf(model.tex_ub))
        self.free = Scene(free, frames, opts)
        self.n = len(frames)
        self.K = self.full.K
        self.Ks = (model.k_tex if opts.fit_tex else 0) + (model.k_id if opts.fit_id else 0)
        self.Kf = self.K - self.Ks
        self.KJ = self.Ks + self.n * self.Kf
        self.opts = opts

    def join(self, thetas):
        th = np.asarray(thetas, np.float64).reshape(self.n, self.K)
        return np.concatenate([th[0, :self.Ks]] + [th[t, self.Ks:] for t in range(self.n)])

    def split(self, x):
        x = np.asarray(x, np.float64)
        th = np.zeros((self.n, self.K))
        for t in range(self.n):
            th[t, :self.Ks] = x[:self.Ks]
            th[t, self.Ks:] = x[self.Ks + t * self.Kf:self.Ks + (t + 1) * self.Kf]
        return th

    def shared_bounds(self, th0):
        """(obj, g [Ks], H diagonal [Ks]) of the shared block's bounds at frame 0's theta."""
        of, gf, Hf = self.full.barrier(0, th0)
        oo, go, Ho = self.free.barrier(0, th0)
        Ks = self.Ks
        return of - oo, (gf - go)[:Ks], np.diag(Hf - Ho)[:Ks]

    def normal_equations(self, x, threads=1):
        """(HJ [KJ, KJ], gJ [KJ], objective) at x; threads > 1 evaluates frames concurrently."""
        th = self.split(x)
        Ks, Kf, KJ = self.Ks, self.Kf, self.KJ
        H = np.zeros((KJ, KJ))
        g = np.zeros(KJ)
        obj = 0.0
        if threads > 1:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(threads) as ex:
                nes = list(ex.map(lambda t: self.free.normal_equations(t, th[t]), range(self.n)))
        else:
            nes = [self.free.normal_equations(t, th[t]) for t in range(self.n)]
        for t in range(self.n):
            ne = nes[t]
            idx = np.concatenate([np.arange(Ks), Ks + t * Kf + np.arange(Kf)])
            H[np.ix_(idx, idx)] += ne["H"]
            g[idx] += ne["g"]
            obj += ne["obj"][2]
        ob, gb, hb = self.shared_bounds(th[0])
        g[:Ks] += gb
        H[np.arange(Ks), np.arange(Ks)] += hb
        return H, g, obj + ob

    def objective(self, x):
        th = self.split(x)
        obj = sum(self.free.render(t, th[t])["obj"][2] for t in range(self.n))
        return obj + self.shared_bounds(th[0])[0]

    def fit(self, x0, max_iters=None):
        """The joint Gauss-Newton fit from x0 (FP64, numpy Cholesky): Armijo line search with the
        options' constants and the reference's step threshold for several frames (0.001).
        Returns (x, iterations performed, stop flags)."""
        o = self.opts
        x = np.asarray(x0, np.float64).copy()
        t_min = 0.001 if self.n > 1 else o.t_min
        performed, stop, terminate = 0, 0, False
        iters = o.max_iters if max_iters is None else max_iters
        pose0 = self._pose_offset()
        for _ in range(iters):
            if terminate:
                break
            H, g, obj0 = self.normal_equations(x)
            performed += 1
            try:
                Lc = np.linalg.cholesky(H)
            except np.linalg.LinAlgError:
                stop |= 16
                break
            d = -np.linalg.solve(Lc.T, np.linalg.solve(Lc, g))
            dg = float(d @ g)
            t = 1.0
            inner = 0
            while inner < o.max_ls:
                if t < t_min:
                    terminate = True
                    stop |= 4
                    break
                xt = x + t * d
                if xt[pose0 + 2] < 0:
                    break
                objt = self.objective(xt)
                if not np.isfinite(objt):
                    t *= o.armijo_beta
                    continue
                inner += 1
                if -dg < o.decrement_tol:
                    terminate = True
                    stop |= 2
                    break
                if objt < obj0 + o.armijo_alpha * t * dg:
                    x = xt
                    break
                t *= o.armijo_beta
        if not terminate and not stop & 24:
            stop |= 1
        return x, performed, stop

    def _pose_offset(self):
        """Offset of frame 0's pose in x (the per-frame layout's pose offset)."""
        m, o = self.full.model, self.opts
        ntex = m.k_tex if o.fit_tex else 0
        nid = m.k_id if o.fit_id else 0
        nexp = m.k_exp if o.fit_exp else 0
        return ntex + nid + nexp
