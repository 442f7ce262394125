"""Small synthetic scenes for the parity tests.  Observations are rendered by the FP64
oracle at the ground truth (plus noise), as the configurations prescribe."""
from __future__ import annotations

import dataclasses
import importlib

import numpy as np

import pyoracle as O

synth = importlib.import_module("3di_amd.synth")
abi = importlib.import_module("3di_amd.abi")


@dataclasses.dataclass
class Scene:
    model: object
    frames: list
    truths: list
    opts: object
    lay: dict

    @property
    def K(self):
        return self.lay["K"]

    def init(self):
        return np.stack([t.init for t in self.truths])

    def truth(self):
        return np.stack([t.theta for t in self.truths])

    def oracle(self, opts=None):
        return O.Scene(self.model, self.frames, opts or self.opts)


_model_cache = {}


def model(grid, k_id, k_tex, k_exp, seed):
    key = (grid, k_id, k_tex, k_exp, seed)
    if key not in _model_cache:
        _model_cache[key] = synth.make_model(grid, k_id, k_tex, k_exp, seed)
    return _model_cache[key]


def make_scene(grid=40, k_exp=10, k_id=0, k_tex=0, fit_id=False, fit_tex=False, fit_light=False,
               crops=(96, 96), pose_deg=15.0, noise=0.02, iters=12, seed=7, cfg=9, blur_sigma=0.75,
               ntmp_cap=0, use_landmarks=1):
    m = model(grid, k_id, k_tex, k_exp, seed)
    lay = synth.layout(m, fit_id, fit_tex, True, fit_light)
    opts = abi.default_options(max_iters=iters, fit_light=int(fit_light), fit_id=int(fit_id),
                               fit_tex=int(fit_tex), blur_sigma=blur_sigma, ntmp_cap=ntmp_cap,
                               use_landmarks=use_landmarks)
    frames, truths = [], []
    for i, S in enumerate(crops):
        rng = np.random.default_rng(synth.frame_seed(cfg, i))
        tr = synth.draw_truth(m, rng, lay, pose_deg, fit_light=fit_light)
        fr = synth.make_frame(m, tr, lay, rng, synth.cfg2_focal(S), S, S)
        sc = O.Scene(m, [fr], opts)
        raw = sc.render(0, tr.theta)["raw"].astype(np.float32)
        fr.obs = synth.add_noise(raw, rng, noise) if noise > 0 else raw
        synth.make_feasible(m, tr, lay, fr)
        frames.append(fr)
        truths.append(tr)
    return Scene(model=m, frames=frames, truths=truths, opts=opts, lay=lay)
