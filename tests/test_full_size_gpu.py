"""GPU parity at the workloads' own sizes, where index widths, sort digit widths, hash
overflow rates and fold grids are those the benchmark runs with.

* 20M particles on the 60^3 box (1,296,000 tets): mp_step_many (atomic, 16-step launches with
  the particles in registers) against per-step mp_step (deterministic) from the same state:
  particles bit-identical after each launch, every step's density within 1e-12;
* 2M-particle subsets checked bit for bit against the serial oracle: the 60^3 box (mp_step and
  mp_step_many), the 64-run ensemble on the 40^3 box (run * 68,921 + node keys, 23 bits;
  mp_batch_step and mp_batch_step_many), and the 100^3 clustered box with reflecting walls
  (1,030,301 nodes: 20-bit keys, long segments on hot nodes, sort every 8 steps);
* the adjoint at 2M particles, 30^3 and 50 steps against the oracle gradient.

Bar: cells, positions and velocities bit-exact; deterministic density bit-exact; atomic
density within 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)
SPWT = 200.0
ATOMIC_RTOL = 1e-12


def check_den(den, ref, mode):
    if mode == 1:
        assert den.tobytes() == ref.tobytes(), "deterministic density not bit-exact"
    else:
        np.testing.assert_allclose(den, ref, rtol=ATOMIC_RTOL, atol=0)


def check_particles(got, run):
    gp, gv, gc, gid = got
    assert np.array_equal(gid, run.pid), "particle set differs"
    assert np.array_equal(gc, run.cell), "cells differ"
    assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes(), "pos/vel not bit-exact"


def check_same(a, b):
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


@pytest.fixture(scope="module")
def box60(mp):
    mesh = mp.Mesh.box_mesh(60, 1.0)
    return mesh, mesh.field_sinusoid(10.0)


def test_step_many_20M_matches_per_step(mp, box60):
    """The headline workload: 20M particles, dt 2.75e-7 s (~30% cross a face per step), cell sort
    every 16 steps on cell >> 3, LDS node hash; 32 steps through mp_step_many = launches of 16
    steps.  The same 32 steps one mp_step at a time in deterministic mode from the same state:
    particles bit-identical after each launch, each step's atomic density within 1e-12."""
    mesh, ef = box60
    n, dt, k_launch = 20_000_000, 2.75e-7, 16
    pos, vel = mp.synth_uniform(20240601, n, [0, 0, 0], [1, 1, 1], 2000.0, [5000.0, 0, 0])
    cell = mesh.locate_box(pos)
    dens, snaps = [], []
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=k_launch, sort_shift=3, scatter_agg=2)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.set_timing(True)
        for _ in range(2):
            mv.step_many(QM, dt, SPWT, 0, k_launch)
            dens.append([mv.density_step(k) for k in range(k_launch)])
            snaps.append(mv.get_particles())
        _, launches = mv.timing()
        mv.set_timing(False)
        assert launches == 2, "two 16-step launches of k_push_steps"
        c_many = mv.counters()
        mv.set_particles(pos, vel, cell)
        for s in range(2):
            for k in range(k_launch):
                mv.step(QM, dt, SPWT, 1)
                np.testing.assert_allclose(dens[s][k], mv.density(), rtol=ATOMIC_RTOL, atol=0,
                                           err_msg=f"launch {s} step {k}")
            check_same(mv.get_particles(), snaps[s])
        c_one = mv.counters()
    for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits"):
        assert c_many[key] == c_one[key], key
    assert 0.25 < c_one["crossed"] / c_one["particle_steps"] < 0.40


@pytest.mark.parametrize("path", ["step", "step_many"])
def test_box60_2M_matches_oracle(mp, orc, box60, path):
    """2M particles on the 60^3 box against the serial oracle, 3 steps, both scatter modes (the
    multi-step path is atomic only: its deterministic mode steps one by one)."""
    mesh, ef = box60
    pos, vel = mp.synth_uniform(77, 2_000_000, [0, 0, 0], [1, 1, 1], 2000.0, [5000.0, 0, 0])
    cell = mesh.locate_box(pos)
    dt = 2.75e-7
    modes = (0, 1) if path == "step" else (0,)
    for mode in modes:
        run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
        with mp.Mover.from_mesh(mesh) as mv:
            mv.set_options(sort_interval=2, sort_shift=3)
            mv.upload_field(ef)
            mv.set_particles(pos, vel, cell)
            if path == "step":
                for _ in range(3):
                    mv.step(QM, dt, SPWT, mode)
                    check_den(mv.density(), run.step(QM, dt, SPWT), mode)
            else:
                mv.step_many(QM, dt, SPWT, mode, 3)
                for k in range(3):
                    check_den(mv.density_step(k), run.step(QM, dt, SPWT), mode)
            check_particles(mv.get_particles(), run)
            c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves"):
            assert c[key] == run.stats[key], key


def _ensemble40(mp, n_runs=64, seed=20240604, scale=1.0 / 40.0):
    """The ensemble's 64 runs on the 40^3 box (side 1 m) with per-run counts log-uniform in
    [1e5, 4e6] * scale (about 1.7M particles in all at 1/40), drift U[2, 8] km/s, sigma
    U[500, 3000] m/s, and per-run q/m, dt and spwt varied so a mix-up of runs shows."""
    mesh = mp.Mesh.box_mesh(40, 1.0)
    ef = mesh.field_sinusoid(10.0)
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(1e5 * scale), np.log(4e6 * scale), n_runs))).astype(np.int64)
    drifts = rng.uniform(2000.0, 8000.0, n_runs)
    sigmas = rng.uniform(500.0, 3000.0, n_runs)
    runs = []
    for r in range(n_runs):
        pos, vel = mp.synth_uniform(seed * 1000 + r, int(counts[r]), [0, 0, 0], [1, 1, 1], sigmas[r],
                                    [drifts[r], 0, 0])
        runs.append(dict(pos=pos, vel=vel, cell=mesh.locate_box(pos), qm=QM * (1 + 0.01 * (r % 7)),
                         dt=4e-7 * (1 + 0.02 * (r % 5)), spwt=SPWT * (1 + r % 3)))
    return mesh, ef, runs


@pytest.fixture(scope="module")
def ens40(mp):
    return _ensemble40(mp)


@pytest.mark.parametrize("order", [0, 1], ids=["cell", "morton"])
@pytest.mark.parametrize("how", ["step", "step_many"])
@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
def test_ensemble_64_runs_40cubed_matches_oracle(mp, orc, ens40, mode, how, order):
    """All 64 runs pushed together (the graph-replayed mp_batch_step, or mp_batch_step_many with
    a sort every 2 steps), 3 steps: every run's density each step and its particles against its
    own serial MoveParticles."""
    mesh, ef, runs = ens40
    assert sum(len(r["pos"]) for r in runs) > 1_000_000
    oracles = [orc.Run(mesh.tets, ef, mesh.node_volume, r["pos"], r["vel"], r["cell"]) for r in runs]
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=2, sort_shift=3, sort_order=order)
        mv.upload_field(ef)
        with mp.Batch(mv, len(runs)) as b:
            for k, r in enumerate(runs):
                b.set_run(k, r["pos"], r["vel"], r["cell"], r["qm"], r["dt"], r["spwt"])
            if how == "step_many":
                b.step_many(3, mode)
            for s in range(3):
                if how == "step":
                    b.step(mode)
                for k, r in enumerate(runs):
                    ref = oracles[k].step(r["qm"], r["dt"], r["spwt"])
                    got = b.density(k) if how == "step" else b.density_step(s, k)
                    check_den(got, ref, mode)
            for k in range(len(runs)):
                check_particles(b.get_particles(k), oracles[k])
            c = b.counters()
    assert c["particle_steps"] == sum(o.stats["particle_steps"] for o in oracles)
    assert c["crossed"] == sum(o.stats["crossed"] for o in oracles)


@pytest.fixture(scope="module")
def box100(mp):
    mesh = mp.Mesh.box_mesh(100, 1.0)
    return mesh, mesh.field_sinusoid(10.0)


@pytest.mark.parametrize("mode,order", [(1, 0), (1, 1), (0, 0)], ids=["deterministic-cell", "deterministic-morton",
                                                                    "atomic-cell"])
def test_clustered_100cubed_reflecting_matches_oracle(mp, orc, box100, mode, order):
    """The contention workload's mesh and distribution at 2M particles: 100^3 box (1,030,301
    nodes, so node keys take 20 bits and the record sort its two 10-bit digits), 3-Gaussian
    mixture (sigma 0.05 m, weights 0.5/0.3/0.2) putting more than 1024 records on the hottest
    nodes (the whole-wave fold), Maxwellian sigma 3000 m/s, reflecting walls, sort every 8 steps;
    10 steps so the sort runs inside the window."""
    mesh, ef = box100
    rng = np.random.default_rng(20240606)
    centres = rng.uniform(0.3, 0.7, (3, 3))
    pos, vel = mp.synth_mixture(20240606, 2_000_000, [0, 0, 0], [1, 1, 1], [0.5, 0.3, 0.2], centres, 0.05, 3000.0,
                                [0, 0, 0])
    cell = mesh.locate_box(pos)
    dt = 1.7e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=1)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=8, sort_shift=3, wall_mode=1, sort_order=order)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(10):
            mv.step(QM, dt, SPWT, mode)
            check_den(mv.density(), run.step(QM, dt, SPWT), mode)
        check_particles(mv.get_particles(), run)
        c = mv.counters()
    for key in ("particle_steps", "lost", "crossed", "moves", "reflections"):
        assert c[key] == run.stats[key], key
    assert c["lost"] == 0
    per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
    assert (per_node > 1024).sum() > 10, "the case must put segments longer than 1024 records on nodes"
    assert np.nonzero(per_node)[0].max() >= 1 << 19, "node ids must use the twentieth key bit"


def test_adjoint_2M_30cubed_50_steps_matches_oracle(mp, orc):
    """The adjoint workload: 2M particles on the 30^3 box (162,000 tets), 50 steps, checkpoint
    every 10, target density from a forward run at 1.05 x drift; loss, dL/dE (3 x 162,000) and
    dL/d(drift) against the oracle's FP64 gradient, in atomic and deterministic mode; the
    deterministic mode gives the same bits twice."""
    mesh = mp.Mesh.box_mesh(30, 1.0)
    ef = mesh.field_sinusoid(10.0)
    pos, u = mp.synth_uniform(20240605, 2_000_000, [0, 0, 0], [1, 1, 1], 2000.0, [0, 0, 0])
    cell = mesh.locate_box(pos)
    d = np.array([1.0, 0.0, 0.0])
    dt, steps, drift = 4.0e-7, 50, 5000.0
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos, u + 1.05 * drift * d, cell)
        for _ in range(steps):
            mv.step(QM, dt, SPWT, 1)
        tgt = mv.density()
        runs = [mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=10, mode=mode)
                for mode in (0, 1, 1)]
    Lr, gr, gdr = orc.adjoint(mesh.tets, ef, mesh.node_volume, pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d)
    scale = np.abs(gr).max()
    assert Lr > 0 and scale > 0 and np.count_nonzero(gr) > len(gr)
    for L, g, gd in runs:  # atomic, deterministic, deterministic again
        assert abs(L - Lr) <= 1e-12 * abs(Lr)
        np.testing.assert_allclose(g, gr, rtol=0, atol=1e-10 * scale)
        assert abs(gd - gdr) <= 1e-10 * abs(gdr)
    (L1, g1, d1), (L2, g2, d2) = runs[1], runs[2]
    assert L1 == L2 and d1 == d2 and g1.tobytes() == g2.tobytes(), "deterministic adjoint differs between runs"
