"""GPU parity of mini-pic's coupled loop: InjectIons -> MoveParticles -> computePhi -> updateEf
on the 12^3 box (ions at 7 km/s from the x = 0 face, n = 1e10 m^-3, Te = 2 eV, dt = 5e-7 s),
with the MI355X mover (drop-in mp_step_aos, deterministic scatter) against the serial
MoveParticles, both fed the same host FE field each step.

Bar: particles (pos, vel, lc, cell) and the normalised density bit-exact at every step; the
field then is identical on both sides by construction, so any divergence shows up at once."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)


@pytest.fixture(scope="module")
def cp():
    import coupled

    return coupled


ION_T = {"cold-beam": 0.0, "thermal-1eV": 11604.518121745585}  # K; 1 eV = 1 / KB kelvin
IN_FLIGHT = {"cold-beam": (45000, 52000), "thermal-1eV": (35000, 42000)}  # ions after 200 steps


def _params(cp, beam="cold-beam"):
    return cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036, wall_potential=0.0,
                         dt=5e-7, max_iter=200, ion_temperature=ION_T[beam])


def _setup(mp, orc, cp, tmp_path, fe_flags, beam="cold-beam"):
    mesh = cp.box_case(12, 0.1, str(tmp_path))
    p = _params(cp, beam)
    gpu = cp.GpuMover(mesh, QM, p.dt, 2e2, mp.SCATTER_DETERMINISTIC)
    run = cp.CoupledRun(mesh, p, move=gpu, fe_flags=fe_flags)
    cpu = orc.AosMover(mesh.tets, mesh.node_volume, QM, p.dt, 2e2)
    return run, gpu, cpu


def _lockstep(run, cpu, steps, done):
    """Advance the loop, moving a copy of the same particles with the serial mover too; appends
    each step whose field solve succeeded to done."""
    for ts in range(steps):
        run.inject()
        before = run.parts.copy()
        run.move_step()
        ref_parts, ref_den = cpu(before, run.ef)
        assert run.n_parts == len(ref_parts), f"step {ts}: particle count"
        assert run.parts.tobytes() == ref_parts.tobytes(), f"step {ts}: particles differ"
        assert run.den.tobytes() == ref_den.tobytes(), f"step {ts}: density differs"
        run.solve()
        done.append(ts)


@pytest.mark.parametrize("beam", list(ION_T))
def test_coupled_loop_200_steps_bit_exact(mp, orc, cp, tmp_path, beam):
    """200 steps (about 50k ions in flight once the cold beam crosses the box, about 38k with the
    1 eV Maxwellian, which loses ions through the side walls), F1 cleared per solve so the Newton
    iteration converges throughout."""
    run, gpu, cpu = _setup(mp, orc, cp, tmp_path, mp.FE_RESET_F1, beam)
    try:
        done = []
        _lockstep(run, cpu, 200, done)
        lo, hi = IN_FLIGHT[beam]
        assert len(done) == 200 and lo < run.n_parts < hi
    finally:
        gpu.close()


def test_coupled_loop_reference_solver_fails_alike(mp, orc, cp, tmp_path):
    """With the reference's accumulating F1 the Newton solve stops converging after some tens of
    steps (the reference exits there); the GPU-driven loop reaches that step with bit-identical
    state and fails the same way."""
    run, gpu, cpu = _setup(mp, orc, cp, tmp_path, 0)
    try:
        done = []
        with pytest.raises(mp.MpError, match="NR failed to converge"):
            _lockstep(run, cpu, 60, done)
        assert 10 < len(done) < 60
    finally:
        gpu.close()


@pytest.mark.parametrize("beam", list(ION_T))
def test_coupled_loop_resident_mover_200_steps(mp, orc, cp, tmp_path, beam):
    """The same loop with the ions resident in HBM (only the injected ions cross PCIe, cell
    sort every 10 steps): density bit-exact every step, and the ions themselves (ascending pid
    = the serial vector's order) bit-exact every 25 steps."""
    mesh = cp.box_case(12, 0.1, str(tmp_path))
    p = _params(cp, beam)
    gpu = cp.ResidentGpuMover(mesh, QM, p.dt, 2e2, mp.SCATTER_DETERMINISTIC, sort_interval=10)
    run = cp.CoupledRun(mesh, p, move=gpu, fe_flags=mp.FE_RESET_F1)
    cpu = orc.AosMover(mesh.tets, mesh.node_volume, QM, p.dt, 2e2)
    ref = np.zeros(0, mp.PARTICLE_DTYPE)
    try:
        for ts in range(200):
            run.inject()
            ref = np.concatenate([ref, run.parts])
            run.move_step()
            ref, ref_den = cpu(ref, run.ef)
            assert run.n_parts == len(ref), f"step {ts}: particle count"
            assert run.den.tobytes() == ref_den.tobytes(), f"step {ts}: density differs"
            if ts % 25 == 24:
                gp, gv, gc, gid = gpu.particles()
                assert np.array_equal(gc, ref["cell_index"]), f"step {ts}: cells differ"
                assert gp.tobytes() == ref["pos"].tobytes() and gv.tobytes() == ref["vel"].tobytes()
            run.solve()
        lo, hi = IN_FLIGHT[beam]
        assert lo < run.n_parts < hi
    finally:
        gpu.close()
