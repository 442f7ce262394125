"""CPU tests of the C++ host layer (libtdi_host.so): the reference's output files
(VideoOutput, video_fitter.cu:2603-2828, written by write_2d_vector, funcs.h:112-124), Euler
angles (compute_euler_angles, rotation_header.cu:421-449) and the config keys of this path.

Known answers: the expected text is written out by hand for values whose %.7g form is
unambiguous; it follows the reference's rules (setprecision(7), one space after every value,
rows t < max(frame id) -- the last frame is not written -- except save_expressions_all, which
writes t <= max; missing frames are NaN rows; expressions average the NRES fits, poses all
fits, illumination keeps the first)."""
import ctypes as C
import importlib
import os
import re

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
video = importlib.import_module("3di_amd.video")
abi = importlib.import_module("3di_amd.abi")


def declared(header):
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", header)).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(tdi_[a-z0-9_]+)\s*\(", text)))


def test_host_library_exports_every_declared_symbol():
    L = video.lib()
    missing = [n for n in declared("tdi_host.h") if not hasattr(L, n)]
    assert not missing, missing


@pytest.fixture
def vo():
    v = video.VideoOutput(3, 2)
    v.add(0, [0.5, -1.25, 3.0], pose=[1, 2, 500, 0.1, 0, 0], illum=[-7.25, 51.5, 100, 0.005])
    v.add(0, [1.0, -0.25, 1.0], pose=[3, 4, 502, 0.1, 0, 0], illum=[9, 9, 9, 9])
    v.add(2, [0.125, 2.0, -3.5], pose=[0, 0, 450, 0, 0.2, 0], illum=[1, 2, 3, 0.25])
    v.add(3, [7, 8, 9], pose=[0, 0, 400, 0, 0, 0.3], illum=[4, 5, 6, 0.5])
    yield v
    v.close()


def read(path):
    return open(path).read()


def test_save_expressions(vo, tmp_path):
    p = str(tmp_path / "x.expressions")
    vo.save_expressions(p)
    assert read(p) == "0.75 -0.75 2 \nnan nan nan \n0.125 2 -3.5 \n"
    # ignore_first with a single fit averages nothing and leaves zeros (video_fitter.cu:2614-2634)
    vo.save_expressions(p, ignore_first=True)
    assert read(p) == "1 -0.25 1 \nnan nan nan \n0 0 0 \n"


def test_save_expressions_all(vo, tmp_path):
    p = str(tmp_path / "x.expressions_all")
    vo.save_expressions_all(p)
    assert read(p) == ("0.5 -1.25 3 1 -0.25 1 \n"
                       "nan nan nan nan nan nan \n"
                       "0.125 2 -3.5 nan nan nan \n"
                       "7 8 9 nan nan nan \n")


def test_save_poses(vo, tmp_path):
    p = str(tmp_path / "x.poses")
    vo.save_poses(p, with_euler=False)
    assert read(p) == "2 3 501 0.1 0 0 \nnan nan nan nan nan nan \n0 0 450 0 0.2 0 \n"
    vo.save_poses(p, with_euler=True)
    rows = [r.split() for r in read(p).splitlines()]
    assert len(rows) == 3 and all(len(r) == 9 for r in rows)
    assert rows[0][:6] == ["2", "3", "501", "0.1", "0", "0"]
    # a rotation of 0.1 rad about x is a pitch of 0.1, about y a yaw of 0.2
    assert float(rows[0][7]) == pytest.approx(0.1, abs=1e-6) and abs(float(rows[0][6])) < 1e-6
    assert float(rows[2][6]) == pytest.approx(0.2, abs=1e-6) and abs(float(rows[2][7])) < 1e-6
    assert rows[1] == ["nan"] * 9


def test_save_illums(vo, tmp_path):
    p = str(tmp_path / "x.illums")
    vo.save_illums(p)
    assert read(p) == "-7.25 51.5 100 0.005 \nnan nan nan nan \n1 2 3 0.25 \n"


def test_precision_is_seven_significant_digits(tmp_path):
    v = video.VideoOutput(2, 1)
    v.add(0, [1.0 / 3.0, 123456789.0])
    v.add(1, [0, 0])
    p = str(tmp_path / "x.expressions")
    v.save_expressions(p)
    assert read(p) == "0.3333333 1.234568e+08 \n"


@pytest.mark.parametrize("ypr", [(0.3, -0.2, 0.1), (-0.5, 0.4, -1.2), (0.05, 0.0, 0.0), (1.2, 0.3, 2.5)])
def test_euler_angles_against_scipy(ypr):
    """compute_euler_angles reads R as R = Rz(roll) Ry(yaw) Rx(pitch)."""
    yaw, pitch, roll = ypr
    u = Rotation.from_euler("ZYX", [roll, yaw, pitch]).as_rotvec()
    out = video.euler_angles(u)
    assert np.allclose(out, [yaw, pitch, roll], atol=2e-6)


def test_read_config():
    o, two, nres = video.read_config(os.path.join(ROOT, "tests", "golden", "config_path.txt"))
    assert o.blur_sigma == 2.0 and o.ref_face_size == 78.0 and o.niters_success == 4
    assert two == 1 and nres == 3
    d = abi.default_options()
    assert o.data_weight == d.data_weight and o.max_iters == d.max_iters
    with pytest.raises(FileNotFoundError):
        video.read_config(os.path.join(ROOT, "tests", "golden", "no_such_config.txt"))


def test_initialize_with_orthographic_rejects_bad_arguments():
    # argument checks come before any device call (TDI_ERR_ARG = 1)
    L = video.lib()
    buf = np.zeros(51 * 4, np.float32)
    p = buf.ctypes.data
    assert L.tdi_initialize_with_orthographic(None, 1, p, p, p, p, None, None) == 1
    assert L.tdi_initialize_with_orthographic(C.c_void_p(1), 0, p, p, p, p, None, None) == 1
    assert L.tdi_initialize_with_orthographic(C.c_void_p(1), 1, None, p, p, p, None, None) == 1


def test_reference_named_cpp_entry_points_link(tmp_path):
    """A C++ caller written against include/tdi_model_fitter.h (the reference's fitting functions
    by name, model_fitter.h:29-79) compiles and links against libtdi_host.so."""
    import shutil
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if not shutil.which("g++"):
        pytest.skip("no g++")
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "tdi_model_fitter.h"\n'
        "bool run(tdi_batch* b, float* th, float* lam, float* L, tdi_fit_report* r) {\n"
        "    tdi::FitContext c{b, true};\n"
        "    return tdi::fit_3DMM_lambdas(c, th, lam, L, r) && tdi::fit_3DMM_Lintensity(c, th, lam, L, r) &&\n"
        "           tdi::fit_3DMM_shape_rigid(c, th, r) && tdi::fit_3DMM_epsilon_finetune(c, th, r);\n"
        "}\n"
        "int main(int argc, char**) { return argc > 5 ? (int)run(nullptr, nullptr, nullptr, nullptr, nullptr) : 0; }\n")
    lib = os.path.join(root, "3di_amd")
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(root, "include"), str(src), "-L", lib,
                           "-ltdi_host", "-ltdi", "-Wl,-rpath," + lib, "-Wl,--allow-shlib-undefined",
                           "-o", str(tmp_path / "caller")])


def test_read_video_config(tmp_path):
    cfg = tmp_path / "c.txt"
    cfg.write_text("%YAML:1.0\nREF_FACE_SIZE: 90\nNRES_COEFS: 3\nTWOSTAGE_ILLUM: 1\nFINETUNE_EXPRESSIONS: 1\n"
                   "FINETUNE_COEF: 2.0\nPAD_RENDERING: 0\nKERNEL_SIGMA: 1.0\nPAINT_INNERMOUTH_BLACK: 0\n")
    o, p = video.read_video_config(str(cfg))
    assert (p.ref_face_size, p.nres_coefs, p.twostage_illum, p.finetune_expressions, p.pad) == (90.0, 3, 1, 1, 0)
    assert p.paint_innermouth_black == 0 and video.video_params().paint_innermouth_black == 1
    assert o.finetune_coef == 2.0 and o.blur_sigma == 1.0 and o.ref_face_size == 90.0
    d = video.video_params()
    assert (d.ref_face_size, d.nres_coefs, d.pad, d.crop_w, d.crop_h) == (86.0, 4, 1, 256, 256)


def test_video_output_dump_and_load(tmp_path):
    """VideoOutput's exact dump: replaying it gives byte-identical output files; a second dump
    appends only the fits added since the first."""
    rng = np.random.default_rng(8)
    vo = video.VideoOutput(5, 2)
    for fid in (0, 1, 1, 2, 3, 3):
        vo.add(fid, rng.normal(size=5).astype(np.float32), rng.normal(size=6).astype(np.float32),
               rng.normal(size=4).astype(np.float32))
    d = tmp_path / "fits.txt"
    vo.dump(str(d))
    vo.add(4, rng.normal(size=5).astype(np.float32), rng.normal(size=6).astype(np.float32))
    vo.dump(str(d))
    assert len(d.read_text().splitlines()) == 7
    vo2 = video.VideoOutput(5, 2)
    vo2.load(str(d))
    for v, tag in ((vo, "a"), (vo2, "b")):
        v.save_expressions(str(tmp_path / f"{tag}.e"))
        v.save_expressions_all(str(tmp_path / f"{tag}.ea"))
        v.save_poses(str(tmp_path / f"{tag}.p"))
        v.save_illums(str(tmp_path / f"{tag}.i"))
    for ext in ("e", "ea", "p", "i"):
        assert (tmp_path / f"a.{ext}").read_bytes() == (tmp_path / f"b.{ext}").read_bytes()
    with pytest.raises(ValueError):
        video.VideoOutput(6, 2).load(str(d))
