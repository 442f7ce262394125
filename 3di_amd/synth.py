"""Seeded synthetic 3DMM scenes (the licensed BFM data the reference needs is not available).

The reference fits Basel-Face-Model bases read from text files (renderer.cu:22-442,
build/models/prepare_BFM.py:71-314).  These generators build stand-ins with the same
array layout and statistics family:

* mesh: an ``G x G`` vertex grid lifted onto an ellipsoid cap with semi-axes
  (80, 100, 60) mm, two triangles per grid cell, nose towards the camera (-z);
* bases: each column is seeded Gaussian noise over the grid, Gaussian-blurred, normalised
  to unit RMS and scaled by ``s0 * k**-0.7`` (s0 = 5 mm for shape, 0.1 for albedo), so
  coefficients are unit-variance.  Albedo columns are blurred with sigma = G/20 cells
  (5 cells on the 100x100 grid).  Shape (identity/expression) columns are blurred with
  sigma = G/4 cells: at 5 mm
  amplitude a 5-cell blur folds the surface (36% of the triangles of a 100x100 grid flip
  orientation in the image, and more at 231x231), whereas a face-scale blur leaves ~2%
  flipped, at the silhouette only;
* truths: coefficients ~ N(0, 1) clipped to +-3, bounds at +-3.5 so truths stay
  strictly feasible for the log barrier;
* 51 landmark vertices on a fixed face-like pattern; indices 0, 9 and 40 play the
  eye-corner / mouth roles the reference's face-size check reads (renderer.cu:836-857).

Arrays follow the C ABI conventions of ``include/tdi_api.h`` (float32, AoS vertices,
row-major [3V][K] bases, int32 triangles, column-major crops).  Seeds: base seed
0x3D1 + configuration index, per-frame seed base*1_000_003 + frame.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np
from scipy.ndimage import gaussian_filter

BASE_SEED = 0x3D1
CAP_AZ_DEG = 25.0
CAP_EL_DEG = 20.0
SEMI_AXES = (80.0, 100.0, 60.0)
# reference initial light (video_fitter.cu:115-116)
LAMBDA0 = np.array([-7.3627, 51.1364, 100.1784], np.float32)
L0 = 0.005
# start of a jointly fitted light: the truth perturbed by N(0, 2 mm) per position axis and
# N(0, 2%) in intensity, the same order of accuracy as the pose and coefficient starts
LIGHT_INIT_PERT = (2.0, 0.02)


@dataclasses.dataclass
class Model:
    grid: int
    mean_shape: np.ndarray          # [V,3] f32
    exp_basis: np.ndarray           # [3V,Ke] f32
    tex_mean: np.ndarray            # [V] f32
    triangles: np.ndarray           # [NT,3] i32
    landmark_ids: np.ndarray        # [51] i32
    exp_lb: np.ndarray
    exp_ub: np.ndarray
    id_basis: Optional[np.ndarray] = None   # [3V,Ki]
    tex_basis: Optional[np.ndarray] = None  # [V,Kt]
    id_lb: Optional[np.ndarray] = None
    id_ub: Optional[np.ndarray] = None
    tex_lb: Optional[np.ndarray] = None
    tex_ub: Optional[np.ndarray] = None

    @property
    def n_vertices(self):
        return self.mean_shape.shape[0]

    @property
    def n_triangles(self):
        return self.triangles.shape[0]

    @property
    def k_exp(self):
        return self.exp_basis.shape[1]

    @property
    def k_id(self):
        return 0 if self.id_basis is None else self.id_basis.shape[1]

    @property
    def k_tex(self):
        return 0 if self.tex_basis is None else self.tex_basis.shape[1]


@dataclasses.dataclass
class Frame:
    phix: float
    phiy: float
    cx: float
    cy: float
    x0: int
    y0: int
    width: int
    height: int
    face_size: float
    obs: np.ndarray                 # [W*H] f32 column-major
    lmk_x: np.ndarray
    lmk_y: np.ndarray
    box_lx: np.ndarray
    box_ux: np.ndarray
    box_ly: np.ndarray
    box_uy: np.ndarray
    lam: np.ndarray                 # fixed light position (3,)
    Lint: float


def _cap_vertices(G: int):
    az = np.deg2rad(np.linspace(-CAP_AZ_DEG, CAP_AZ_DEG, G))
    el = np.deg2rad(np.linspace(-CAP_EL_DEG, CAP_EL_DEG, G))
    A, E = np.meshgrid(az, el, indexing="ij")        # vertex v = i*G + j
    a, b, c = SEMI_AXES
    X = a * np.sin(A) * np.cos(E)
    Y = b * np.sin(E)
    Z = -c * np.cos(A) * np.cos(E)
    Z = Z - Z.min() - c * 0.5          # centre the cap depth around the origin
    Z = Z - Z.mean()
    return np.stack([X, Y, Z], -1).reshape(-1, 3), A, E


def _triangles(G: int, P: np.ndarray) -> np.ndarray:
    i, j = np.meshgrid(np.arange(G - 1), np.arange(G - 1), indexing="ij")
    v00 = (i * G + j).ravel()
    v10 = ((i + 1) * G + j).ravel()
    v11 = ((i + 1) * G + j + 1).ravel()
    v01 = (i * G + j + 1).ravel()
    t1 = np.stack([v00, v10, v11], 1)
    t2 = np.stack([v00, v11, v01], 1)
    tris = np.empty((2 * t1.shape[0], 3), np.int64)
    tris[0::2] = t1
    tris[1::2] = t2
    # orient so that n = (b-a) x (c-a) points towards the camera (-z), which makes the
    # reference shading term n.(light - p) positive for a light on the camera side
    b, c, a = P[tris[:, 0]], P[tris[:, 1]], P[tris[:, 2]]
    n = np.cross(b - a, c - a)
    flip = n[:, 2] > 0
    tris[flip] = tris[flip][:, [1, 0, 2]]
    return tris.astype(np.int32)


def _landmarks(G: int) -> np.ndarray:
    """51 grid vertices on a face-like pattern (brows, nose, eyes, mouth)."""
    pts = []
    # brows 0..9 (0 and 9 are the outer corners the face-size test reads)
    for u in np.linspace(-0.8, 0.8, 10):
        pts.append((u, 0.55 - 0.1 * u * u))
    # nose ridge 10..13 and base 14..18
    for v in np.linspace(0.35, 0.0, 4):
        pts.append((0.0, v))
    for u in np.linspace(-0.25, 0.25, 5):
        pts.append((u, -0.08))
    # eyes 19..30
    for cxe in (-0.45, 0.45):
        for t in np.linspace(0, 2 * np.pi, 6, endpoint=False):
            pts.append((cxe + 0.18 * np.cos(t), 0.3 + 0.07 * np.sin(t)))
    # mouth 31..50 (40 sits on the lower lip)
    for t in np.linspace(0, 2 * np.pi, 12, endpoint=False):
        pts.append((0.4 * np.cos(t), -0.45 + 0.15 * np.sin(t)))
    for t in np.linspace(0, 2 * np.pi, 8, endpoint=False):
        pts.append((0.25 * np.cos(t), -0.45 + 0.06 * np.sin(t)))
    pts = np.array(pts[:51])
    ij = np.clip(np.round((pts + 1) * 0.5 * (G - 1)).astype(int), 0, G - 1)
    return (ij[:, 0] * G + (G - 1 - ij[:, 1])).astype(np.int32)


def _smooth_fields(rng, G: int, n: int, sigma: float) -> np.ndarray:
    """n unit-RMS Gaussian-blurred noise fields over the G x G grid -> [G*G, n]."""
    out = np.empty((G * G, n), np.float64)
    for k in range(n):
        f = gaussian_filter(rng.standard_normal((G, G)), sigma, mode="reflect")
        f /= np.sqrt(np.mean(f * f))
        out[:, k] = f.ravel()
    return out


SHAPE_BLUR_FRACTION = 0.25   # blur sigma of shape basis columns, in units of the grid size
TEX_BLUR_FRACTION = 0.05     # blur sigma of albedo basis columns, in units of the grid size
ALBEDO_DETAIL = (40.0, 12.0)
SHAPE_AXIS_WEIGHT = (1.0, 1.0, 1.0)  # albedo mean structure at grid/40 and grid/12 cells


def _similarity_modes(P: np.ndarray) -> np.ndarray:
    """Orthonormal basis [3V, 7] of the similarity motions of the point set P: three
    translations, three infinitesimal rotations about the centroid and uniform scaling."""
    c = P - P.mean(0)
    V = P.shape[0]
    M = np.zeros((V, 3, 7))
    for a in range(3):
        M[:, a, a] = 1.0
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    M[:, 1, 3], M[:, 2, 3] = -z, y          # rotation about x
    M[:, 0, 4], M[:, 2, 4] = z, -x          # rotation about y
    M[:, 0, 5], M[:, 1, 5] = -y, x          # rotation about z
    M[:, :, 6] = c                          # scaling
    q, _ = np.linalg.qr(M.reshape(3 * V, 7))
    return q


def _shape_basis(rng, G: int, K: int, s0: float, P: np.ndarray) -> np.ndarray:
    """Smooth random deformation fields with the similarity motions of the mean shape P
    projected out, as in a PCA model built from Procrustes-aligned scans: a column that
    moved, turned or scaled the whole face would be indistinguishable from the pose."""
    V = G * G
    scale = s0 * np.arange(1, K + 1, dtype=np.float64) ** -0.7
    B = np.empty((V, 3, K), np.float64)
    for axis in range(3):
        B[:, axis, :] = _smooth_fields(rng, G, K, SHAPE_BLUR_FRACTION * G) * SHAPE_AXIS_WEIGHT[axis]
    B = B.reshape(3 * V, K)
    Q = _similarity_modes(P)
    B -= Q @ (Q.T @ B)
    B *= scale / np.sqrt(np.mean(B * B, axis=0))
    return B.astype(np.float32)


# Expression bounds shrink with the component like the reference's, which bounds each
# coefficient in proportion to its variance (prepare_BFM.py:47-50, logbarrier_initializer.cu:284-285):
# with unit-variance coefficients and column amplitudes s0 * k**-0.7 that is +-3.5 * k**-0.7.
# The tight bounds of the small components are what keeps their weak directions out of the
# Gauss-Newton steps.
EXP_BOUND_DECAY = 0.7


def exp_bounds(k_exp: int) -> np.ndarray:
    """Expression coefficient bounds +-3.5 * k**-EXP_BOUND_DECAY (k = 1..K_exp)."""
    return (3.5 * np.arange(1, k_exp + 1, dtype=np.float64) ** -EXP_BOUND_DECAY).astype(np.float32)


def make_model(grid: int, k_id: int, k_tex: int, k_exp: int, seed: int) -> Model:
    rng = np.random.default_rng(seed)
    P, A, E = _cap_vertices(grid)
    tris = _triangles(grid, P)
    V = grid * grid
    exp_basis = _shape_basis(rng, grid, k_exp, 5.0, P)
    id_basis = _shape_basis(rng, grid, k_id, 5.0, P) if k_id else None
    tex_basis = None
    if k_tex:
        tex_basis = (_smooth_fields(rng, grid, k_tex, TEX_BLUR_FRACTION * grid)
                     * (0.1 * np.arange(1, k_tex + 1, dtype=np.float64) ** -0.7)).astype(np.float32)
    # albedo mean: skin-like level with structure at 1/40 and 1/12 of the face, so that the
    # normalised-gradient objective has edges to follow at every crop size
    fine, coarse = ALBEDO_DETAIL
    tex = (0.55 + 0.18 * _smooth_fields(rng, grid, 1, grid / fine)[:, 0]
           + 0.08 * _smooth_fields(rng, grid, 1, grid / coarse)[:, 0])
    tex = np.clip(tex, 0.15, 0.95).astype(np.float32)
    bnd = np.float32(3.5)
    eb = exp_bounds(k_exp)
    m = Model(grid=grid, mean_shape=P.astype(np.float32), exp_basis=exp_basis, tex_mean=tex,
              triangles=tris, landmark_ids=_landmarks(grid),
              exp_lb=-eb, exp_ub=eb,
              id_basis=id_basis, tex_basis=tex_basis)
    if k_id:
        m.id_lb = np.full(k_id, -bnd, np.float32)
        m.id_ub = np.full(k_id, bnd, np.float32)
    if k_tex:
        m.tex_lb = np.full(k_tex, -bnd, np.float32)
        m.tex_ub = np.full(k_tex, bnd, np.float32)
    return m


def fold_identity(m: Model, alpha: np.ndarray, beta: np.ndarray) -> Model:
    """Per-frame-stage model: identity and albedo fixed and folded into the means
    (the per-frame app builds VideoFitter with K_alpha = K_beta = 0,
    video_from_saved_identity.cu:154-157)."""
    V = m.n_vertices
    mean = m.mean_shape.astype(np.float64).reshape(-1)
    if m.id_basis is not None:
        mean = mean + m.id_basis.astype(np.float64) @ alpha.astype(np.float64)
    tex = m.tex_mean.astype(np.float64)
    if m.tex_basis is not None:
        tex = tex + m.tex_basis.astype(np.float64) @ beta.astype(np.float64)
    return Model(grid=m.grid, mean_shape=mean.reshape(V, 3).astype(np.float32), exp_basis=m.exp_basis,
                 tex_mean=tex.astype(np.float32), triangles=m.triangles, landmark_ids=m.landmark_ids,
                 exp_lb=m.exp_lb, exp_ub=m.exp_ub)


def rotvec_from_euler(yaw, pitch, roll) -> np.ndarray:
    """Rotation vector of R = Rz(roll) Ry(yaw) Rx(pitch) (radians)."""
    cx, sx = math.cos(pitch), math.sin(pitch)
    cy, sy = math.cos(yaw), math.sin(yaw)
    cz, sz = math.cos(roll), math.sin(roll)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    R = Rz @ Ry @ Rx
    th = math.acos(max(-1.0, min(1.0, (np.trace(R) - 1) / 2)))
    if th < 1e-9:
        return np.array([1e-6, 0.0, 0.0])
    w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (2 * math.sin(th))
    return w * th


LIGHT_UNKNOWNS = {0: 0, 1: 4, 2: 9}   # fit_light: fixed, point light (lambda, L), + second-order terms
# second-order light terms of a drawn truth ~ N(0, SH2_SIGMA), start perturbed by N(0, SH2_INIT_PERT)
SH2_SIGMA = 0.05
SH2_INIT_PERT = 0.01


def light_unknowns(fit_light) -> int:
    return LIGHT_UNKNOWNS[int(fit_light)]


def layout(m: Model, fit_id: bool, fit_tex: bool, fit_exp: bool, fit_light):
    """Offsets of the unknown blocks (Optimizer.h:71-91 order).  fit_light: 0 (fixed), 1 (the
    reference's point light, 4 unknowns) or 2 (with the second-order terms, 9 unknowns)."""
    ntex = m.k_tex if fit_tex else 0
    nid = m.k_id if fit_id else 0
    nexp = m.k_exp if fit_exp else 0
    off = {"tex": 0, "id": ntex, "exp": ntex + nid, "pose": ntex + nid + nexp}
    off["light"] = off["pose"] + 6
    off["K"] = off["light"] + light_unknowns(fit_light)
    return off


@dataclasses.dataclass
class Truth:
    theta: np.ndarray       # full unknown vector of the truth (float64)
    init: np.ndarray        # starting point of the fit
    lam: np.ndarray
    Lint: float


def draw_truth(m: Model, rng, lay: dict, pose_deg: float, tz_range=(400.0, 600.0),
               fit_light: bool = False, exp_scale: float = 1.0,
               exp_coeffs: Optional[np.ndarray] = None, pose: Optional[np.ndarray] = None,
               coef_pert: float = 0.05) -> Truth:
    K = lay["K"]
    th = np.zeros(K)
    if lay["exp"] < lay["pose"]:
        ne = lay["pose"] - lay["exp"]
        # unit-variance draws clipped to +-3, shrunk with the component's bound so that the
        # truth stays strictly inside it
        th[lay["exp"]:lay["pose"]] = (exp_coeffs if exp_coeffs is not None
                                      else np.clip(rng.standard_normal(ne) * exp_scale, -3, 3)
                                      * (m.exp_ub[:ne].astype(np.float64) / 3.5))
    if lay["id"] < lay["exp"]:
        th[lay["id"]:lay["exp"]] = np.clip(rng.standard_normal(lay["exp"] - lay["id"]), -3, 3)
    if lay["tex"] < lay["id"]:
        th[lay["tex"]:lay["id"]] = np.clip(rng.standard_normal(lay["id"] - lay["tex"]), -3, 3)
    if pose is None:
        ang = np.deg2rad(rng.uniform(-pose_deg, pose_deg, 3))
        tz = rng.uniform(*tz_range)
        pose = np.concatenate([[rng.uniform(-10, 10), rng.uniform(-10, 10), tz], rotvec_from_euler(*ang)])
    th[lay["pose"]:lay["pose"] + 6] = pose
    lam = LAMBDA0.astype(np.float64) + rng.normal(0, 20.0, 3)
    Lint = L0 * rng.uniform(0.8, 1.2)
    nl = lay["K"] - lay["light"]
    if nl:
        # the point light (lambda, L); with nine light unknowns written as (a = L lambda, L)
        th[lay["light"]:lay["light"] + 3] = lam * (Lint if nl == 9 else 1.0)
        th[lay["light"] + 3] = Lint
    if nl == 9:
        th[lay["light"] + 4:lay["light"] + 9] = rng.normal(0, SH2_SIGMA, 5)
    init = th.copy()
    # start from a perturbed expression/identity, a perturbed pose and a perturbed light,
    # at roughly the one-pixel accuracy of the reference's landmark-driven initialisation
    # (video_fitter.cu:479-485); the caller makes it landmark-feasible with make_feasible().
    # Coefficients are perturbed in proportion to their bounds and kept inside them.
    nb = lay["pose"]
    cb = np.full(nb, 3.5)
    if lay["exp"] < lay["pose"]:
        cb[lay["exp"]:lay["pose"]] = m.exp_ub[:lay["pose"] - lay["exp"]]
    init[:nb] = np.clip(th[:nb] + rng.normal(0, coef_pert, nb) * cb / 3.5, -0.92 * cb, 0.92 * cb)
    p = lay["pose"]
    init[p:p + 3] += rng.normal(0, [0.3, 0.3, 1.0])
    init[p + 3:p + 6] += rng.normal(0, np.deg2rad(0.2), 3)
    if nl:
        lam_i = lam + rng.normal(0, LIGHT_INIT_PERT[0], 3)
        L_i = Lint * (1.0 + rng.normal(0, LIGHT_INIT_PERT[1]))
        init[lay["light"]:lay["light"] + 3] = lam_i * (L_i if nl == 9 else 1.0)
        init[lay["light"] + 3] = L_i
    if nl == 9:
        init[lay["light"] + 4:lay["light"] + 9] += rng.normal(0, SH2_INIT_PERT, 5)
    return Truth(theta=th, init=init, lam=lam, Lint=Lint)


def project_points(m: Model, th: np.ndarray, lay: dict, phix, phiy, cx, cy, ids) -> np.ndarray:
    """FP64 projection of selected vertices (for landmarks)."""
    p = m.mean_shape.astype(np.float64)[ids]
    rows = (3 * ids[:, None] + np.arange(3)[None, :])
    if lay["exp"] < lay["pose"]:
        p = p + (m.exp_basis[rows.ravel()].astype(np.float64) @ th[lay["exp"]:lay["pose"]]).reshape(-1, 3)
    if lay["id"] < lay["exp"]:
        p = p + (m.id_basis[rows.ravel()].astype(np.float64) @ th[lay["id"]:lay["exp"]]).reshape(-1, 3)
    u = th[lay["pose"] + 3:lay["pose"] + 6]
    t = np.linalg.norm(u)
    n = u / t
    N = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    R = np.eye(3) + math.sin(t) * N + (1 - math.cos(t)) * N @ N
    q = p @ R.T + th[lay["pose"]:lay["pose"] + 3]
    return np.stack([phix * q[:, 0] / q[:, 2] + cx, phiy * q[:, 1] / q[:, 2] + cy], 1)


def make_frame(m: Model, truth: Truth, lay: dict, rng, f: float, W: int, H: int,
               box: float = 0.15, lmk_noise: float = 1.0) -> Frame:
    """Frame geometry, landmarks and box; the observation is filled in by a renderer."""
    cx, cy = W / 2.0, H / 2.0
    lm = project_points(m, truth.theta, lay, f, f, cx, cy, m.landmark_ids.astype(np.int64))
    lm = lm + rng.normal(0, lmk_noise, lm.shape)
    lx = lm[:, 0].astype(np.float32)
    ly = lm[:, 1].astype(np.float32)
    fs = float(max(lx.max() - lx.min(), ly.max() - ly.min()))
    ones = np.ones(51, np.float32)
    return Frame(phix=f, phiy=f, cx=cx, cy=cy, x0=0, y0=0, width=W, height=H, face_size=fs,
                 obs=np.zeros(W * H, np.float32), lmk_x=lx, lmk_y=ly,
                 box_lx=-box * ones, box_ux=box * ones, box_ly=-box * ones, box_uy=box * ones,
                 lam=truth.lam.astype(np.float32), Lint=float(truth.Lint))


def landmark_slack(m: Model, th: np.ndarray, lay: dict, fr: "Frame") -> float:
    """Smallest landmark-box constraint value G (logbarrier_initializer.cu:501-507) at th."""
    pr = project_points(m, th, lay, fr.phix, fr.phiy, fr.cx, fr.cy, m.landmark_ids.astype(np.int64))
    fs = fr.face_size
    g = np.concatenate([pr[:, 0] - (fr.lmk_x + fr.box_lx * fs), pr[:, 1] - (fr.lmk_y + fr.box_ly * fs),
                        (fr.lmk_x + fr.box_ux * fs) - pr[:, 0], (fr.lmk_y + fr.box_uy * fs) - pr[:, 1]])
    return float(g.min())


def make_feasible(m: Model, truth: Truth, lay: dict, fr: "Frame", margin: float = 0.5) -> np.ndarray:
    """Pull the initial point towards the truth until every landmark-box constraint
    holds with `margin` pixels (the reference reaches a feasible start with its
    landmark-only fit, video_fitter.cu:479-485)."""
    init = truth.init.copy()
    for _ in range(30):
        if landmark_slack(m, init, lay, fr) > margin:
            break
        init = truth.theta + 0.7 * (init - truth.theta)
    truth.init = init
    return init


def add_noise(img: np.ndarray, rng, sigma: float = 0.02) -> np.ndarray:
    return (img + rng.normal(0, sigma, img.shape)).astype(np.float32)


def frame_seed(config_index: int, frame: int) -> int:
    return (BASE_SEED + config_index) * 1_000_003 + frame


# ----------------------------------------------------------------- configurations
CFG0 = dict(grid=100, k_id=40, k_tex=40, k_exp=20, crop=192, f=800.0, pose_deg=20.0, iters=20)
CFG1 = dict(grid=231, k_id=199, k_tex=199, k_exp=79)
CFG2 = dict(grid=231, k_exp=79, crop_min=96, crop_max=384, pose_deg=30.0, iters=12, frames=4096)


def cfg2_crop_sizes(n: int, seed: int = BASE_SEED + 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(CFG2["crop_min"], CFG2["crop_max"] + 1, n)


def cfg2_focal(S: int) -> float:
    """Focal length that keeps the small plain fit's face-to-crop proportion (f = 800 px at 192 px)."""
    return 800.0 * S / 192.0
