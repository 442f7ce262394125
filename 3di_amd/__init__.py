"""MI355X-native 3DMM per-frame fitting step (3DI's fit_3DMM_shape_rigid path on gfx950)."""
