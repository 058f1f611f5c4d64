"""The host FE solver and the output writers against the reference's own code: src/FESolver.cpp,
src/meshes.cpp (OutputMesh) and src/particles.cpp (OutputParticles) compiled where they lie
into oracle/_ref (build container only; skipped where the reference sources are absent).

* computePhi with the NonLinear (Newton + Gauss-Seidel) and GaussSeidel methods: phi and E
  bit-identical to the reference's FESolver over successive solves (F1 accumulating as in the
  reference), on the coarse cylinder of the reference's example (its own inlet and wall
  surfaces) and on the 12^3 box of the coupled workload;
* MP_FE_DIRECT (the stand-in for the Lapack method's dsysv_, which needs an external LAPACK
  the reference harness does not link): the reference's own Lapack-method system J, G solved
  with LAPACK's dsysv (scipy), the same routine and triangle the reference calls, agrees to
  1e-10 of max |phi|;
* OutputMesh and OutputParticles: byte-identical files.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import REF_EXAMPLES

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF_EXAMPLES) and oracle.ref_available()),
                                reason="reference sources or examples absent (build container only)")

COARSE = os.path.join(REF_EXAMPLES, "one_stream", "coarse")


def write_volume_dat(mesh, path):
    """A SALOME .dat volume file of mesh (mm coordinates, 304 tet records) for both loaders."""
    pos = mesh.node_pos * 1000.0
    con = mesh.tets["con"]
    with open(path, "w") as f:
        f.write(f"{len(pos)} {len(con)}\n")
        for i, (x, y, z) in enumerate(pos):
            f.write(f"{i + 1} {float(x)!r} {float(y)!r} {float(z)!r}\n")
        for e, c in enumerate(con):
            f.write(f"{e + 1} 304 {c[0] + 1} {c[1] + 1} {c[2] + 1} {c[3] + 1}\n")


def load_both(mp, case, tmp_path):
    """(our mesh, ref) with the same volume, inlet (INLET) and wall (FIXED) files loaded by our
    loader and by the reference's LoadVolumeMesh / LoadSurfaceMesh."""
    import coupled as cp

    if case == "coarse":
        vol, inl, wal = (os.path.join(COARSE, f) for f in ("mesh.dat", "inlet.dat", "wall.dat"))
    else:
        box = mp.Mesh.box_mesh(12, 0.1)
        vol, inl, wal = (str(tmp_path / f) for f in ("mesh.dat", "inlet.dat", "wall.dat"))
        write_volume_dat(box, vol)
        cp.write_box_face(box, inl, "xmin")
        cp.write_box_face(box, wal, "xmax")
    mesh = mp.Mesh.load_dat(vol)
    mesh.load_surface(inl, mp.NODE_INLET)
    mesh.load_surface(wal, mp.NODE_FIXED)
    R = oracle.ref()
    assert R.ref_load_volume(vol.encode()) == 0
    assert R.ref_load_surface(inl.encode(), mp.NODE_INLET, 0) == 0
    assert R.ref_load_surface(wal.encode(), mp.NODE_FIXED, 0) == 0
    assert (mesh.node_type == 2).sum() > 0 and (mesh.node_type == 3).sum() > 0
    return mesh


# n0 1e10 m^-3, Te 2 eV, wall -3 V: the coupled workload's plasma, where the Newton iteration
# converges (the reference exits the process when it does not)
N0, KTE, WALL = 1e10, 8.617333262e-5 * 23209.036, 3.0


@pytest.mark.parametrize("method", [0, 1], ids=["newton", "gauss-seidel"])
@pytest.mark.parametrize("case", ["coarse", "box12"])
def test_fe_solver_bit_identical_to_reference(mp, case, method, tmp_path):
    mesh = load_both(mp, case, tmp_path)
    ref = oracle.RefFE(mesh.n_nodes, mesh.n_tets, N0, 0.0, KTE, WALL)
    ours = mp.FESolver(mesh, n0=N0, phi0=0.0, kTe=KTE, wall_potential=WALL)
    assert ours.neq == ref.neq
    rng = np.random.default_rng(11 + method)
    newton_its = 0
    for step in range(6):
        den = N0 * rng.uniform(0.5, 1.5, mesh.n_nodes)
        ref.compute_phi(den, method)
        st = ours.compute_phi(den, method)
        newton_its += st["nr_iterations"]
        ef = ours.update_ef()
        uh, ef_ref = ref.phi_ef()
        assert ours.phi.tobytes() == uh.tobytes(), f"step {step}: phi differs"
        assert ef.tobytes() == ef_ref.tobytes(), f"step {step}: E differs"
    assert np.abs(uh).max() > 1.0 and np.abs(ef).max() > 0
    if method == 0:
        assert newton_its > 6  # at least one solve took several Newton iterations


@pytest.mark.parametrize("case", ["coarse", "box12"])
def test_direct_method_matches_lapack_dsysv(mp, case, tmp_path):
    from scipy.linalg import lapack

    mesh = load_both(mp, case, tmp_path)
    ref = oracle.RefFE(mesh.n_nodes, mesh.n_tets, N0, 0.0, KTE, WALL)
    ours = mp.FESolver(mesh, n0=N0, phi0=0.0, kTe=KTE, wall_potential=WALL)
    den = N0 * np.random.default_rng(5).uniform(0.5, 1.5, mesh.n_nodes)
    J, G = ref.lapack_system(den)
    # dsysv_('L', ...) on the row-major copy reads J's upper triangle (column-major lower)
    A = np.triu(J) + np.triu(J, 1).T
    _, _, y, info = lapack.dsysv(A, G, lower=1)
    assert info == 0
    # phi = g + (d - y) with d = 0 before the first step: g = -wall on FIXED nodes, 0 on INLET
    types = mesh.node_type
    unknown = (types == 0) | (types == 1)
    want = np.where(types == 3, -WALL, 0.0)
    want[unknown] -= y
    ours.compute_phi(den, mp.FE_DIRECT)
    scale = np.abs(want).max()
    assert scale > 1.0
    np.testing.assert_allclose(ours.phi, want, rtol=0, atol=1e-10 * scale)


@pytest.mark.parametrize("case", ["coarse", "box12"])
def test_output_mesh_and_particles_byte_identical(mp, case, tmp_path, monkeypatch):
    mesh = load_both(mp, case, tmp_path)
    rng = np.random.default_rng(3)
    phi = rng.normal(0.0, 3.0, mesh.n_nodes)
    ef = rng.normal(0.0, 50.0, (mesh.n_tets, 3))
    den = rng.uniform(0.0, 2e10, mesh.n_nodes)
    n = 1000
    pos = mesh.node_pos[rng.integers(0, mesh.n_nodes, n)] + rng.normal(0, 1e-3, (n, 3))
    parts = mp.particles_from_arrays(pos, rng.normal(0, 1e3, (n, 3)), np.zeros(n, np.int32))
    refdir, ourdir = tmp_path / "ref", tmp_path / "ours"
    refdir.mkdir()
    ourdir.mkdir()
    monkeypatch.chdir(refdir)
    R = oracle.ref()
    R.ref_output_mesh(9, phi.ctypes.data, np.ascontiguousarray(ef).ctypes.data, den.ctypes.data)
    R.ref_output_particles(parts.ctypes.data, n)
    mp.write_mesh_vtu(mesh, str(ourdir / "mesh_0010.vtu"), phi, ef, den)
    mp.write_particles_vtp(str(ourdir / "particles.vtp"), parts)
    for f in ("mesh_0010.vtu", "particles.vtp"):
        a, b = (refdir / f).read_bytes(), (ourdir / f).read_bytes()
        assert len(a) > 1000 and a == b, f
