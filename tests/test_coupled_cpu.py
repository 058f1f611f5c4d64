"""The host side of mini-pic's coupled loop (no GPU): the random stream, InjectIons, the FE
Poisson-Boltzmann solver and the parameter file, each pinned by the reference's draws, closed
forms or the dense restatement of FESolver."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)


@pytest.fixture(scope="module")
def cp():
    import coupled

    return coupled


def test_rng_draws_match_reference(mp):
    """rnd() of src/maths.cpp:17-19: the draws the reference's own maths.cpp produced."""
    want = np.load(os.path.join(GOLDEN, "rnd_mt19937.npz"))["draws"]
    assert mp.Rng(0).uniform(len(want)).tobytes() == want.tobytes()


def test_inject_ions_known_answer(mp, cp, tmp_path):
    """InjectIons on the 8 inlet triangles of a 2^3 box, recomputed from the reference's draws:
    counts with the carried fraction, folded triangle sampling, the cold beam and the half-step
    velocity rewind in the start tet's field."""
    mesh = cp.box_case(2, 0.1, str(tmp_path))
    faces = mesh.inlet_faces
    assert len(faces) == 8 and np.all(faces["normal"] == [1.0, 0.0, 0.0])
    assert np.allclose(faces["area"], 0.05 ** 2 / 2, rtol=1e-15)
    ef = np.random.default_rng(3).normal(0.0, 50.0, (mesh.n_tets, 3))
    plasma_den, v0, dt, spwt = 1e8, 7000.0, 5e-7, 200.0  # 2.1875 macroparticles per face
    draws = np.load(os.path.join(GOLDEN, "rnd_mt19937.npz"))["draws"]
    parts, rem = mp.inject_ions(mesh, mp.Rng(0), plasma_den, v0, dt, spwt, QM, ef, 0.25)
    # the same arithmetic, face by face, from the recorded draws
    r, k, want_pos, want_vel, want_cell = 0.25, 0, [], [], []
    npos = mesh.node_pos
    for f in faces:
        fnum = plasma_den * v0 * f["area"] * dt / spwt + r
        num = int(fnum)
        r = fnum - num
        for _ in range(num):
            a, b = draws[k], draws[k + 1]
            k += 2
            if a + b > 1:
                a, b = 1 - a, 1 - b
            want_pos.append(a * f["u"] + b * f["v"] + npos[f["con"][0]])
            want_vel.append(f["normal"] * v0 - QM * ef[f["cell_con"]] * (0.5 * dt))
            want_cell.append(f["cell_con"])
    assert 0 < len(parts) == len(want_pos) and k <= len(draws)
    assert rem == r
    assert parts["pos"].tobytes() == np.array(want_pos).tobytes()
    assert parts["vel"].tobytes() == np.array(want_vel).tobytes()
    assert np.array_equal(parts["cell_index"], want_cell)
    # every sample lies on the x = 0 face, in its face's triangle, inside the box
    assert np.all(parts["pos"][:, 0] == 0.0) and np.all((parts["pos"] >= 0) & (parts["pos"] <= 0.1))


def test_thermal_injection_zero_spread_is_cold_beam(mp, cp, tmp_path):
    """mp_inject_ions_thermal with v_thermal = 0 injects the reference's cold beam bit for bit,
    and its position draws leave the mt19937 stream where the cold beam leaves it."""
    mesh = cp.box_case(4, 0.1, str(tmp_path))
    ef = np.random.default_rng(5).normal(0.0, 50.0, (mesh.n_tets, 3))
    r1, r2 = mp.Rng(7), mp.Rng(7)
    a, ra = mp.inject_ions(mesh, r1, 1e10, 7000.0, 5e-7, 1.0, QM, ef, 0.5)
    b, rb = mp.inject_ions(mesh, r2, 1e10, 7000.0, 5e-7, 1.0, QM, ef, 0.5, v_thermal=0.0)
    assert len(a) > 1000 and a.tobytes() == b.tobytes() and ra == rb
    assert r1.uniform(4).tobytes() == r2.uniform(4).tobytes()


def test_thermal_injection_maxwellian(mp, orc, cp, tmp_path):
    """The 1 eV Maxwellian of the coupled workload (sigma = sqrt(kT/m) = 2455.5 m/s for 16 amu):
    positions, cells and the carried fraction are the cold beam's; the velocity differs from the
    cold beam's by v_thermal * z with z the oracle's SplitMix64 Box-Muller normals of the k-th
    injected particle (continuing across calls); and those normals are standard normal: mean,
    variance, pairwise correlation within 5 standard errors and a Kolmogorov-Smirnov test."""
    from scipy import stats

    mesh = cp.box_case(12, 0.1, str(tmp_path))
    vth = mp.thermal_speed(1.0 / mp.KB, 16 * mp.AMU)
    assert abs(vth - 2455.5) < 0.5
    ef = np.random.default_rng(6).normal(0.0, 50.0, (mesh.n_tets, 3))
    plasma_den, v0, dt, spwt = 1e13, 7000.0, 5e-7, 200.0   # ~1.75M ions a call
    cold_rng, hot_rng = mp.Rng(3), mp.Rng(3)
    k0 = 0
    zs = []
    for call in range(2):
        cold, rc_ = mp.inject_ions(mesh, cold_rng, plasma_den, v0, dt, spwt, QM, ef, 0.0)
        hot, rh = mp.inject_ions(mesh, hot_rng, plasma_den, v0, dt, spwt, QM, ef, 0.0, v_thermal=vth)
        assert len(hot) == len(cold) > 100000 and rh == rc_
        assert hot["pos"].tobytes() == cold["pos"].tobytes()
        assert np.array_equal(hot["cell_index"], cold["cell_index"])
        z = orc.thermal_normals(3, k0, len(hot))
        k0 += len(hot)
        np.testing.assert_allclose(hot["vel"] - cold["vel"], vth * z, rtol=0, atol=1e-8)
        zs.append(z)
    z = np.concatenate(zs)
    n = len(z)
    assert np.all(np.abs(z.mean(axis=0)) < 5 / np.sqrt(n))
    assert np.all(np.abs(z.var(axis=0) - 1) < 5 * np.sqrt(2 / n))
    c = np.corrcoef(z.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 5 / np.sqrt(n))
    assert stats.kstest(z.ravel(), "norm").pvalue > 1e-4
    # the draws behind the normals are uniform on [0, 1): a plausible indexing slip (two items
    # sharing draws) would correlate consecutive particles
    assert abs(np.corrcoef(z[:-1, 0], z[1:, 0])[0, 1]) < 5 / np.sqrt(n)


def test_splitmix_draws_match_synthetic_generator(mp, orc):
    """The oracle's SplitMix64 restatement reproduces the library's seeded synthetic positions
    (draws 0..2 of particle p scaled into the box) bit for bit."""
    pos, _ = mp.synth_uniform(12345, 1000, [0, 0, 0], [1, 1, 1], 1.0, [0, 0, 0])
    p = np.arange(1000)
    want = np.stack([orc.splitmix_draws(12345, p, j) for j in range(3)], axis=1)
    assert pos.tobytes() == want.tobytes()


def test_fe_stiffness_closed_form(mp, cp, tmp_path):
    """K_PQ = sum_e 8 detJ_e (grad N_a . grad N_b) (2x2x2 unit weights over the element), with the
    gradients taken from the tet tables (grad lc_i = (-beta_i, gamma_i, -delta_i) / 6V) and
    detJ from numpy: an independent evaluation.  K is exactly symmetric."""
    mesh = cp.box_case(4, 0.1, str(tmp_path), wall=True)
    fe = mp.FESolver(mesh, n0=1e10, phi0=0.0, kTe=2.0, wall_potential=3.0)
    rp, ci, val = fe.stiffness()
    K = np.zeros((fe.neq, fe.neq))
    for i in range(fe.neq):
        K[i, ci[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    assert K.tobytes() == K.T.copy().tobytes()
    t = mesh.tets
    types = mesh.node_type
    ID = np.cumsum((types == 0) | (types == 1)) - 1
    ID[(types != 0) & (types != 1)] = -1
    pos = mesh.node_pos
    Kc = np.zeros_like(K)
    for e in range(mesh.n_tets):
        con = t["con"][e]
        grad = np.stack([-t["beta"][e], t["gamma"][e], -t["delta"][e]], axis=1) / (6 * t["volume"][e])
        x = pos[con]
        detJ = np.linalg.det((x[:3] - x[3]).T)
        ke = 8 * detJ * grad @ grad.T
        for a in range(4):
            for b in range(4):
                P, Q = ID[con[a]], ID[con[b]]
                if P >= 0 and Q >= 0:
                    Kc[P, Q] += ke[a, b]
    np.testing.assert_allclose(K, Kc, rtol=1e-11, atol=1e-13 * np.abs(Kc).max())


@pytest.mark.parametrize("method", [2, 0], ids=["direct", "nonlinear"])
def test_fe_laplace_reproduces_linear_potential(mp, cp, tmp_path, method):
    """No space charge (n0 = 0, no ions), phi = 0 on the x = 0 inlet and -V on the x = L wall,
    the other faces open: the exact solution phi = -V x / L is linear, so P1 elements reproduce
    it; E is uniform (V / L, 0, 0).  The direct solve holds it to rounding, Gauss-Seidel to its
    1e-4 residual tolerance."""
    V, L = 7.0, 0.1
    mesh = cp.box_case(5, L, str(tmp_path), wall=True)
    fe = mp.FESolver(mesh, n0=0.0, phi0=0.0, kTe=2.0, wall_potential=V)
    fe.compute_phi(np.zeros(mesh.n_nodes), method)
    exact = -V * mesh.node_pos[:, 0] / L
    ef = fe.update_ef()
    tol = 1e-9 if method == 2 else 2e-2
    np.testing.assert_allclose(fe.phi, exact, atol=tol * V)
    np.testing.assert_allclose(ef, np.tile([V / L, 0.0, 0.0], (mesh.n_tets, 1)), atol=tol * 50 * V / L)


@pytest.mark.parametrize("reset", [False, True], ids=["reference-F1", "reset-F1"])
@pytest.mark.parametrize("method", [0, 1], ids=["nonlinear", "gauss-seidel"])
def test_fe_matches_dense_restatement(mp, orc, cp, tmp_path, method, reset):
    """The compressed-row solver against the reference's dense loops (oracle), bit for bit over
    successive solves with a Boltzmann electron term and a fixed wall: phi and E, and the same
    Newton failure when the reference's accumulating F1 stops the iteration converging."""
    mesh = cp.box_case(5, 0.1, str(tmp_path), wall=True)
    args = dict(n0=1e10, phi0=0.0, kTe=2.0, wall_potential=5.0)
    fe = mp.FESolver(mesh, flags=mp.FE_RESET_F1 if reset else 0, **args)
    ref = orc.DenseFE(mesh.node_pos, mesh.node_type, mesh.tets, reset_f1=reset, **args)
    assert fe.neq == ref.neq
    rng = np.random.default_rng(11)
    for k in range(6):
        den = rng.uniform(0.5e10, 1.5e10, mesh.n_nodes)
        rc = ref.compute_phi(den, method)
        if rc != 0:
            with pytest.raises(mp.MpError, match="NR failed to converge"):
                fe.compute_phi(den, method)
            assert not reset
            return
        fe.compute_phi(den, method)
        uh, ef = ref.phi_ef()
        assert fe.phi.tobytes() == uh.tobytes(), f"solve {k}: phi differs"
        assert fe.update_ef().tobytes() == ef.tobytes(), f"solve {k}: E differs"


def test_parameter_file(cp):
    """Parameters::Parameters on the reference's coarse example file (its species spelling
    'Duterium' is what the reference's own pattern accepts)."""
    p = cp.Parameters.from_file(os.path.join(GOLDEN, "coarse_system.param"))
    assert (p.plasma_den, p.ion_velocity, p.electron_temperature) == (1e16, 1e8, 2e8)
    assert (p.wall_potential, p.max_iter, p.dt) == (10000.0, 250, 1e-12)
    assert p.plasma_species == "Deuterium" and p.fesolver_method == 2 and not p.invert_normals
    assert os.path.basename(p.mesh_files["inlet_mesh"]) == "inlet.dat"


def test_coupled_loop_on_host(mp, orc, cp, tmp_path):
    """The 12^3 box loop (ions at 7 km/s from the x = 0 face, n = 1e10, Te = 2 eV, dt = 5e-7 s)
    with the serial mover: 1750 ions a step (plasma_den v A dt / spwt = 1750 with the carried
    fraction), none leave before crossing the 0.1 m box (~28 steps), the Newton solve converges
    with F1 reset; with the reference's accumulating F1 it stops converging within 60 steps,
    as the reference would (it exits there)."""
    mesh = cp.box_case(12, 0.1, str(tmp_path))
    p = cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036, wall_potential=0.0,
                      dt=5e-7, max_iter=200)
    mover = orc.AosMover(mesh.tets, mesh.node_volume, QM, p.dt, 2e2)
    run = cp.CoupledRun(mesh, p, move=mover, fe_flags=mp.FE_RESET_F1)
    recs = [run.step() for _ in range(40)]
    added = np.array([r["added"] for r in recs])
    assert abs(added.sum() - 40 * 1750) <= 1 and set(added) <= {1749, 1750, 1751}
    assert all(r["removed"] == 0 for r in recs[:20]) and recs[-1]["removed"] > 0
    assert all(r["nr_iterations"] <= 3 for r in recs)
    run = cp.CoupledRun(mesh, p, move=mover)
    with pytest.raises(mp.MpError, match="NR failed to converge"):
        for _ in range(60):
            run.step()


def test_vtk_output_layout(mp, orc, cp, tmp_path):
    """OutputMesh / OutputParticles files of a short coupled run: the reference's VTK layout
    (ASCII UnstructuredGrid with node_index, node_type, phi, ion_den, ef, cell_volume; PolyData
    of positions), every number written as the C++ stream writes a double by default
    (printf %g: six significant digits)."""
    import xml.etree.ElementTree as ET

    mesh = cp.box_case(3, 0.1, str(tmp_path))
    p = cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036, wall_potential=0.0,
                      dt=5e-7, max_iter=10)
    out = tmp_path / "out"
    out.mkdir()
    run = cp.CoupledRun(mesh, p, move=orc.AosMover(mesh.tets, mesh.node_volume, QM, p.dt, 2e2),
                        fe_flags=mp.FE_RESET_F1, output_dir=str(out))
    for _ in range(10):
        run.step()
    run.finish()
    assert sorted(os.listdir(out)) == ["mesh_0010.vtu", "mesh_0011.vtu", "particles.vtp", "trace.csv"]
    rows = [line.split(",") for line in (out / "trace.csv").read_text().splitlines()]
    assert rows[0] == ["function_name", "total_time", "exclusive_time", "exclusive_proportion", "number_of_calls"]
    calls = {r[0]: int(r[4]) for r in rows[1:]}
    assert calls["InjectIons"] == calls["MoveParticles"] == calls["computePhi"] == calls["updateEf"] == 10
    assert calls["OutputMesh"] == 2 and calls["OutputParticles"] == 1
    # the solver's own methods (src/FESolver.cpp TRACE_ME): one Newton solve per step
    assert calls["FESolver"] == calls["computeNX"] == calls["preAssembly"] == 1
    assert calls["solve"] == calls["solveNonLinear"] == 10
    assert calls["buildF1Vector"] == calls["buildJmatrix"] == calls["solveLinear"] >= 10
    assert "solveLinearDirect" not in calls  # never called, so no row
    assert [r[0] for r in rows[-3:]] == ["traceEnter", "traceExit", "traceTotalTime"]
    assert len({r[0] for r in rows}) == len(rows)
    text = (out / "mesh_0011.vtu").read_text()
    root = ET.fromstring(text)
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == mesh.n_nodes and int(piece.get("NumberOfCells")) == mesh.n_tets
    arrays = {a.get("Name"): a.text.split() for a in root.iter("DataArray")}
    assert arrays["connectivity"] == [str(v) for v in mesh.tets["con"].ravel()]
    assert arrays["offsets"] == [str(4 * (e + 1)) for e in range(mesh.n_tets)]
    assert arrays["types"] == ["10"] * mesh.n_tets
    assert arrays["node_index"] == [str(n) for n in range(mesh.n_nodes)]
    assert arrays["node_type"] == [str(t) for t in mesh.node_type]
    assert arrays["phi"] == ["%g" % v for v in run.fe.phi]
    assert arrays["ion_den"] == ["%g" % v for v in run.den]
    assert arrays["ef"] == ["%g" % v for v in run.ef.ravel()]
    assert arrays["cell_volume"] == ["%g" % v for v in mesh.tets["volume"]]
    pts = root.find("UnstructuredGrid/Piece/Points/DataArray").text.split()
    assert pts == ["%g" % v for v in mesh.node_pos.ravel()]
    proot = ET.fromstring((out / "particles.vtp").read_text())
    assert int(proot.find("PolyData/Piece").get("NumberOfPoints")) == len(run.parts)
    assert proot.find("PolyData/Piece/Points/DataArray").text.split() == ["%g" % v for v in run.parts["pos"].ravel()]


def test_fe_profile_nests_like_the_reference_trace(mp, cp, tmp_path):
    """FESolver's trace rows: computePhi encloses solve, which encloses the Newton loop and its
    buildF1Vector / buildJmatrix / solveLinear calls, so each parent's callee time is the sum of
    its children's inclusive times and no row's exclusive time is negative."""
    mesh = cp.box_case(3, 0.1, str(tmp_path))
    fe = mp.FESolver(mesh, n0=1e10, phi0=0.0, kTe=mp.KB * 23209.036, wall_potential=0.0, flags=mp.FE_RESET_F1)
    den = np.full(mesh.n_nodes, 1e10)
    for method in (mp.FE_NONLINEAR, mp.FE_GAUSS_SEIDEL, mp.FE_DIRECT):
        fe.compute_phi(den, method)
    fe.update_ef()
    prof = fe.profile()
    assert len(prof) == mp.FE_PROFILE_ENTRIES
    n = {k: v[0] for k, v in prof.items()}
    tot = {k: v[1] for k, v in prof.items()}
    child = {k: v[2] for k, v in prof.items()}
    assert n["FESolver"] == n["computeNX"] == n["preAssembly"] == 1
    assert n["computePhi"] == n["solve"] == 3 and n["solveNonLinear"] == 1 and n["solveLinearDirect"] == 1
    assert n["buildF1Vector"] == n["buildJmatrix"] == n["solveLinear"] + 1 and n["updateEf"] == 1
    rel = 1e-9
    assert child["FESolver"] == pytest.approx(tot["computeNX"] + tot["preAssembly"], rel=rel)
    assert child["computePhi"] == pytest.approx(tot["solve"], rel=rel)
    assert child["solve"] == pytest.approx(tot["solveNonLinear"] + tot["solveLinearDirect"]
                                           + (tot["buildF1Vector"] + tot["buildJmatrix"] + tot["solveLinear"]
                                              - child["solveNonLinear"]), rel=rel)
    for k in ("computeNX", "preAssembly", "buildF1Vector", "buildJmatrix", "solveLinear", "solveLinearDirect",
              "updateEf"):
        assert child[k] == 0.0, k
    assert all(tot[k] >= child[k] >= 0 for k in prof)
    assert mp.lib().mp_fe_get_profile(fe._h, mp.FE_PROFILE_ENTRIES, None, None, None, None) == 1
    assert b"no such entry" in mp.lib().mp_fe_last_error()
