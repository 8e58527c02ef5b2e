"""C ABI checks that need no GPU: the library loads, exports every function
include/mpm_gpu.h declares, and its host-only scene code is deterministic."""
import ctypes
import os
import re

import numpy as np

import non_ficks_diffusion_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mpm_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(mpm_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(M.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole ABI
    assert set(names) <= set(M.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors agree with the C compiler's sizeof/offsetof of the header."""
    import subprocess
    src = tmp_path / "layout.c"
    checks = {"mpm_plane": M.Plane, "mpm_material": M.Material, "mpm_scene_desc": M.SceneDesc,
              "mpm_step_stats": M.StepStats, "mpm_launch_config": M.LaunchConfig,
              "mpm_serve_config": M.ServeConfig, "mpm_request_status": M.RequestStatus,
              "mpm_tick_stats": M.TickStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_gpu.h"', "int main(void){"]
    for cname, py in checks.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, py in checks.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_scene_generation_deterministic_and_sized():
    d, x, v, m = M.scene_generate("avalanche", 64, 0)
    d2, x2, v2, m2 = M.scene_generate("avalanche", 64, 0)
    assert np.array_equal(x, x2) and np.array_equal(m, m2)
    assert d.counts == 64 and d.dim == 3
    # particles inside [dx/2, 1-dx/2] and ppc 8 mass
    dx = 1.0 / 64
    assert x.min() >= dx / 2 and x.max() <= 1 - dx / 2
    assert np.allclose(m, 400 * dx ** 3 / 8)
    # slab volume 0.6 x 0.1 x 0.125 at 8 per cell
    assert abs(len(m) - 0.6 * 0.1 * 0.125 * 64 ** 3 * 8) / len(m) < 0.02
    _, n_full = M.scene_preset("avalanche", 256, 0)
    assert 0.99e6 < n_full < 1.02e6
    _, n0 = M.scene_preset("disk2d", 64, 0)
    assert 480 < n0 < 550
    _, xa, _, _ = M.scene_generate("blob", 64, 1)
    _, xb, _, _ = M.scene_generate("blob", 64, 2)
    assert xa.shape != xb.shape or not np.array_equal(xa, xb)


def test_presets_bound_dt_by_elastic_wave_speed():
    d, _ = M.scene_preset("avalanche", 256, 0)
    cp = np.sqrt((d.material.lambda0 + 2 * d.material.mu0) / 400.0)
    assert abs(d.max_dt - (1.0 / 256) / cp) / d.max_dt < 1e-5
    d, _ = M.scene_preset("ref3d_test18", 32, 0)
    assert abs(d.max_dt - 5e-3) < 1e-9  # soft reference material keeps the reference's clamp


def test_errors_are_status_codes():
    # invalid descriptors are rejected with a status and a message, no crash
    d, _ = M.scene_preset("disk2d", 64, 0)
    d.dim = 5
    h = ctypes.c_void_p()
    rc = M.lib().mpm_create(ctypes.byref(d), 10, ctypes.byref(h))
    assert rc == -1 and b"dim" in M.lib().mpm_last_error()
    rc = M.lib().mpm_scene_preset(b"no_such_scene", 0, 0, ctypes.byref(M.SceneDesc()), None)
    assert rc == -1
    # the copy-bandwidth probe checks its arguments before touching a device
    g = ctypes.c_double()
    for nbytes, reps in ((0, 1), (1000, 1), (1 << 20, 0)):
        assert M.lib().mpm_measure_copy_bandwidth(nbytes, reps, ctypes.byref(g)) == -1
    assert M.lib().mpm_measure_copy_bandwidth(1 << 20, 1, None) == -1


def test_reference_test_scenes_counts_mass_and_materials():
    """The Standard_Tests scenes: particle count, mass = mass density * shape
    area / count, velocity, planes and material table of each reference test
    (mpm_2d/Standard_Tests/Standard_Tests.h:55-283,
    mpm_3d/Standard_Tests/Standard_Tests.h:113-150)."""
    area = np.pi * 0.1 ** 2
    cases = {  # name: (count, mass of the first shape's particles, first velocity, planes)
        "ref2d_test1": (1, 2 * area, (0.5, 0.0), 0),
        "ref2d_test15": (2000, 2 * 0.04 / 2000, (-2.0, 0.1), 0),
        "ref2d_test16": (3000, 2 * area / 2000, (0.0, -1.5), 0),
        "ref2d_test17": (2000, 2 * area / 2000, (-2.0, 0.1), 2),
        "ref2d_test19": (2000, 2 * area / 2000, (0.0, -1.0), 4),
        "ref3d_test19": (20000, 2 * (4 / 3 * np.pi * 0.1 ** 3) / 20000, (0.0, -1.0, 0.0), 5),
    }
    for name, (count, m0, v0, planes) in cases.items():
        d, x, v, m = M.scene_generate(name)
        mat = M.scene_materials(name)
        assert len(m) == count and len(mat) == count, name
        assert abs(m[0] - m0) / m0 < 1e-6, name
        assert np.allclose(v[0], v0), name
        assert d.nplanes == planes, name
        assert d.gravity[1] == np.float32(-2.0) and d.cg_max_iterations == 10000, name
    # test 16: the ball (entry 0) then the layer (entry 1), mass density 1 over [0,1]x[0.1,0.125]
    d, x, v, m = M.scene_generate("ref2d_test16")
    mat = M.scene_materials("ref2d_test16")
    assert d.nmaterials == 2 and list(np.bincount(mat)) == [2000, 1000]
    layer = mat == 1
    assert np.allclose(m[layer], 1 * 0.025 / 1000) and np.all(v[layer] == 0)
    assert x[layer, 1].min() >= 0.1 and x[layer, 1].max() <= 0.125
    E, nu = 5.0, 0.4
    lay = d.materials()[1]
    assert np.isclose(lay.mu0, E / (2 * (1 + nu))) and np.isclose(lay.lambda0, E * nu / ((1 + nu) * (1 - 2 * nu)))
    assert np.isclose(lay.hardening_factor, 5.0) and np.isclose(lay.stretching_yield, 1.001)
    # test 19: EOS fluid with K=1, gamma=7, initial density 1
    for name in ("ref2d_test19", "ref3d_test19"):
        d, _ = M.scene_preset(name)
        assert d.material.eos == 1 and d.material.bulk_modulus == 1.0
        assert d.material.gamma == 7.0 and d.material.density == 1.0
    # a table entry outside [0, 8] is refused, so is an EOS material without a positive density
    d, _ = M.scene_preset("ref2d_test16")
    d.nmaterials = 9
    h = ctypes.c_void_p()
    assert M.lib().mpm_create(ctypes.byref(d), 10, ctypes.byref(h)) == -1
    d, _ = M.scene_preset("ref2d_test19")
    d.material.density = 0.0
    assert M.lib().mpm_create(ctypes.byref(d), 10, ctypes.byref(h)) == -1
    assert b"EOS" in M.lib().mpm_last_error()
