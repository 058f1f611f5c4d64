"""The C-ABI library loads and exports every declared entry point (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(INCLUDE, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(mp_\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("mp_create", "mp_destroy", "mp_step_aos", "mp_step", "mp_upload_field", "mp_last_error",
                 "mp_mesh_box", "mp_mesh_load_dat"):
        assert must in names


def test_library_exports_every_declared_symbol(mp):
    L = ctypes.CDLL(mp.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol(mp):
    missing = [n for n in declared_functions() if n not in mp.SIGNATURES]
    assert not missing, missing


def test_struct_layouts(mp):
    assert mp.TET_DTYPE.itemsize == 168 and mp.TET_DTYPE.fields["cell_con"][1] == 152
    assert mp.TET_DTYPE.fields["delta"][1] == 120 and mp.TET_DTYPE.fields["alpha"][1] == 24
    assert mp.PARTICLE_DTYPE.itemsize == 88 and mp.PARTICLE_DTYPE.fields["cell_index"][1] == 80
    assert mp.PARTICLE_DTYPE.fields["lc"][1] == 48
    assert mp.FACE_DTYPE.itemsize == 104
    assert ctypes.sizeof(mp.MpOptions) == 48 and ctypes.sizeof(mp.MpTuneTrial) == 56


def test_create_rejects_bad_tables(mp):
    m = mp.Mesh.box_mesh(2, 1.0)
    t = m.tets
    t["con"][3, 1] = m.n_nodes + 5
    with pytest.raises(mp.MpError) as e:
        mp.Mover(t, m.node_volume)
    assert e.value.code in (1, 5)


def test_create_without_gpu_fails_loudly(mp):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    m = mp.Mesh.box_mesh(2, 1.0)
    with pytest.raises(mp.MpError) as e:
        mp.Mover.from_mesh(m)
    assert e.value.code == 5


def test_mesh_api_errors(mp):
    with pytest.raises(mp.MpError):
        mp.Mesh.box_mesh(0, 1.0)
    with pytest.raises(mp.MpError):
        mp.Mesh.load_dat("/nonexistent/mesh.dat")


def test_synth_is_seeded_and_in_range(mp):
    a = mp.synth_uniform(5, 1000, [0, 0, 0], [1, 2, 3], 2.0, [7.0, 0, 0])
    b = mp.synth_uniform(5, 1000, [0, 0, 0], [1, 2, 3], 2.0, [7.0, 0, 0])
    c = mp.synth_uniform(6, 1000, [0, 0, 0], [1, 2, 3], 2.0, [7.0, 0, 0])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and not np.array_equal(a[0], c[0])
    p, v = a
    assert p.min() >= 0 and np.all(p.max(0) < [1, 2, 3])
    assert abs(v[:, 0].mean() - 7.0) < 0.3 and abs(v[:, 1].std() - 2.0) < 0.2
    q, _ = mp.synth_mixture(3, 2000, [0, 0, 0], [1, 1, 1], [0.5, 0.3, 0.2],
                            [[0.3, 0.3, 0.3], [0.7, 0.5, 0.5], [0.5, 0.8, 0.2]], 0.05, 3000.0, [0, 0, 0])
    assert q.min() >= 0 and q.max() < 1


def test_field_linear_potential_is_uniform(mp):
    """phi = a.x + b gives E = -a exactly up to rounding in every tet (computeNX/updateEf)."""
    m = mp.Mesh.box_mesh(3, 0.5)
    a = np.array([3.0, -2.0, 0.5])
    phi = m.node_pos @ a + 1.0
    ef = m.field_from_potential(phi)
    np.testing.assert_allclose(ef, np.broadcast_to(-a, ef.shape), rtol=1e-9, atol=1e-9)


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources absent (build container only)")
def test_dropin_header_binds_reference_types():
    """include/minipic_amd/move_particles.hpp instantiated with the reference's own Species,
    Volume, FESolver and Parameters (src/particles.h, meshes.h, FESolver.h, parameters.h),
    compiled with g++ -fsyntax-only where those headers lie (tests/cpp/dropin_reftypes.cpp)."""
    import subprocess

    src = os.path.join(ROOT, "tests", "cpp", "dropin_reftypes.cpp")
    r = subprocess.run(["g++", "-fsyntax-only", "-std=c++20", f"-I{REF_SRC}", f"-I{INCLUDE}", src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
