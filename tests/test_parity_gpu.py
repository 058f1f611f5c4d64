"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the golden
vectors written from the reference's own code.

Bar: cells, positions, velocities (and lc on the drop-in path) bit-exact; density
bit-exact in deterministic mode and within 1e-12 relative in atomic mode.
"""
import gzip
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)
SPWT = 200.0
ATOMIC_RTOL = 1e-12


def check_den(den, ref, mode, mp):
    if mode == mp.SCATTER_DETERMINISTIC:
        assert den.tobytes() == ref.tobytes(), "deterministic density not bit-exact"
    else:
        np.testing.assert_allclose(den, ref, rtol=ATOMIC_RTOL, atol=0)


def box_case(mp, n_hex=8, side=0.1, n=60000, seed=1, sigma=2000.0, drift=5000.0, amp=10.0):
    mesh = mp.Mesh.box_mesh(n_hex, side)
    ef = mesh.field_sinusoid(amp)
    pos, vel = mp.synth_uniform(seed, n, [0, 0, 0], [side] * 3, sigma, [drift, 0, 0])
    cell = mesh.locate_box(pos)
    return mesh, ef, pos, vel, cell


@pytest.fixture(scope="module")
def coarse(mp, tmp_path_factory):
    d = tmp_path_factory.mktemp("coarse")
    p = d / "mesh.dat"
    with gzip.open(os.path.join(GOLDEN, "coarse_mesh.dat.gz"), "rb") as f, open(p, "wb") as g:
        shutil.copyfileobj(f, g)
    return mp.Mesh.load_dat(str(p)), np.load(os.path.join(GOLDEN, "coarse_move.npz"))


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
def test_dropin_matches_reference_golden(mp, coarse, mode):
    """mp_step_aos (MoveParticles drop-in) vs the reference kernel body's own outputs."""
    mesh, g = coarse
    parts = mp.particles_from_arrays(g["pos0"], g["vel0"], g["cell0"])
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(g["ef"])
        for k in range(3):
            kept, den = mv.step_aos(parts, float(g["qm"]), float(g["dt"]), float(g["spwt"]), mode)
            parts = parts[:kept].copy()
            assert parts.tobytes() == g[f"parts{k}"].tobytes(), f"step {k}: particles differ"
            check_den(den, g[f"den{k}"], mode, mp)


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
def test_dropin_box_matches_oracle(mp, orc, mode):
    mesh, ef, pos, vel, cell = box_case(mp, n=40000, seed=3)
    dt = 3e-7
    parts = mp.particles_from_arrays(pos, vel, cell)
    ref = parts.copy()
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        for k in range(4):
            kept, den = mv.step_aos(parts, QM, dt, SPWT, mode)
            parts = parts[:kept].copy()
            raw, _ = orc.move_aos(mesh.tets, ef, ref, QM, dt, np.zeros(mesh.n_nodes))
            ref = orc.compact_aos(ref).copy()
            assert parts.tobytes() == ref.tobytes()
            check_den(den, orc.normalise(raw, mesh.node_volume, SPWT), mode, mp)


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
@pytest.mark.parametrize("sort_interval", [0, 2])
def test_resident_matches_oracle(mp, orc, mode, wall, sort_interval):
    mesh, ef, pos, vel, cell = box_case(mp, seed=5 + wall)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=wall)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=sort_interval, wall_mode=wall)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(5):
            mv.step(QM, dt, SPWT, mode)
            ref_den = run.step(QM, dt, SPWT)
            gp, gv, gc, gid = mv.get_particles()
            assert np.array_equal(gid, run.pid), f"step {k}: particle set differs"
            assert np.array_equal(gc, run.cell), f"step {k}: cells differ"
            assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
            check_den(mv.density(), ref_den, mode, mp)
        c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
            assert c[key] == run.stats[key], key
        if wall == 1:
            assert c["lost"] == 0 and c["reflections"] > 0


def test_walk_cap_matches_oracle(mp, orc):
    mesh, ef, pos, vel, cell = box_case(mp, seed=9, n=20000)
    dt = 2e-6  # several cells per step
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, walk_cap=2)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(walk_cap=2)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.step(QM, dt, SPWT, mp.SCATTER_DETERMINISTIC)
        run.step(QM, dt, SPWT)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert mv.counters()["cap_hits"] == run.stats["cap_hits"] > 0
        assert mv.density().tobytes() == run.den.tobytes()


def test_empty_and_all_lost(mp, orc):
    mesh, ef, pos, vel, cell = box_case(mp, n=500, seed=2)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        # empty drop-in step: nothing kept, zero density
        kept, den = mv.step_aos(np.zeros(0, mp.PARTICLE_DTYPE), QM, 1e-7, SPWT)
        assert kept == 0 and not den.any()
        # every particle flies out through the +x wall
        v = vel.copy()
        v[:, 0] = 1e9
        for mode in (0, 1):
            mv.set_particles(pos, v, cell)
            mv.step(QM, 1e-6, SPWT, mode)
            assert mv.num_active() == 0
            assert not mv.density().any()
            assert len(mv.get_particles()[0]) == 0
            mv.step(QM, 1e-6, SPWT, mode)  # a step on an empty set is fine
        mv.set_particles(pos[:0], vel[:0], cell[:0])
        mv.step(QM, 1e-6, SPWT, 0)
        assert not mv.density().any()


def test_injection_appends_pids(mp, orc):
    mesh, ef, pos, vel, cell = box_case(mp, n=30000, seed=4)
    dt = 3e-7
    half = len(pos) // 2
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half])
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos[:half], vel[:half], cell[:half])
        mv.step(QM, dt, SPWT, 1)
        run.step(QM, dt, SPWT)
        mv.add_particles(pos[half:], vel[half:], cell[half:])
        # the oracle appends the new particles after the survivors: pid order is preserved
        run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
        run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
        run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
        run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
        for _ in range(3):
            mv.step(QM, dt, SPWT, 1)
            run.step(QM, dt, SPWT)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes()
        assert mv.density().tobytes() == run.den.tobytes()


def test_full_size_properties(mp):
    """At the 20M-particle / 1.3M-tet scale: size-independent properties.

    * deterministic density is identical whatever the cell-sort interval;
    * atomic density equals it to 1e-12 relative;
    * the raw weights sum to the particles in the domain (spwt=1, volumes=1 via counters);
    * every particle that is still in the domain sits in its cell (sampled).
    """
    mesh = mp.Mesh.box_mesh(60, 1.0)
    ef = mesh.field_sinusoid(10.0)
    n = 20_000_000
    pos, vel = mp.synth_uniform(11, n, [0, 0, 0], [1, 1, 1], 2000.0, [5000.0, 0, 0])
    cell = mesh.locate_box(pos)
    dt = 2.5e-7
    dens = {}
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        for tag, mode, si in (("det0", 1, 0), ("det3", 1, 3), ("atomic", 0, 3)):
            mv.set_options(sort_interval=si)
            mv.set_particles(pos, vel, cell)
            for _ in range(3):
                mv.step(QM, dt, SPWT, mode)
            dens[tag] = mv.density()
        c = mv.counters()
        gp, gv, gc, gid = mv.get_particles()
    assert dens["det0"].tobytes() == dens["det3"].tobytes()
    np.testing.assert_allclose(dens["atomic"], dens["det0"], rtol=ATOMIC_RTOL, atol=0)
    nv = mesh.node_volume
    raw_sum = (dens["det0"] * nv / SPWT).sum()
    np.testing.assert_allclose(raw_sum, len(gid), rtol=1e-9)
    assert 0.2 < c["crossed"] / c["particle_steps"] < 0.45
    t = mesh.tets
    idx = np.random.default_rng(0).choice(len(gid), 20000, replace=False)
    x, cc = gp[idx], gc[idx]
    lc = ((1.0 / 6.0) * (t["alpha"][cc] - x[:, :1] * t["beta"][cc] + x[:, 1:2] * t["gamma"][cc]
                         - x[:, 2:3] * t["delta"][cc]) / t["volume"][cc][:, None])
    assert np.all(lc >= 0) and np.all(lc <= 1)


def test_cpp_dropin_program(mp):
    """A C++ host program written against the reference call surface (DeviceData +
    MoveParticles from include/minipic_amd/move_particles.hpp) matches the oracle."""
    import subprocess

    exe = os.path.join(os.path.dirname(__file__), "cpp", "dropin_check")
    assert os.path.exists(exe), "tests/cpp/dropin_check not built (run build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 2


@pytest.mark.parametrize("agg", [0, 2], ids=["plain", "lds-by-node"])
def test_atomic_aggregation_modes(mp, orc, agg):
    """Both atomic-scatter aggregations give the oracle's density to 1e-12 relative."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=13, n=80000)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(scatter_agg=agg, sort_interval=3)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for _ in range(4):
            mv.step(QM, dt, SPWT, mp.SCATTER_ATOMIC)
            ref = run.step(QM, dt, SPWT)
            check_den(mv.density(), ref, mp.SCATTER_ATOMIC, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and gp.tobytes() == run.pos.tobytes()


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
@pytest.mark.parametrize("ppt,block", [(2, 256), (4, 128), (8, 64)])
def test_particles_per_thread_walls_and_cap(mp, orc, wall, ppt, block):
    """Several rounds of particles per thread with the LDS hash in atomic mode: 6 steps without
    a sort, so lost slots stay in place between live ones, with a walk cap of 3 that the fast
    particles hit and a ragged last workgroup; cells, positions, velocities and counters
    bit-exact, every step's density to the atomic bar."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=97 + wall, n=50003)
    dt = 1.2e-6
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=wall, walk_cap=3)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(particles_per_thread=ppt, scatter_agg=2, block_size=block, sort_interval=0, wall_mode=wall,
                       walk_cap=3)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for _ in range(6):
            mv.step(QM, dt, SPWT, mp.SCATTER_ATOMIC)
            check_den(mv.density(), run.step(QM, dt, SPWT), mp.SCATTER_ATOMIC, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
        c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
            assert c[key] == run.stats[key], key
        assert c["cap_hits"] > 0 and c["lost"] > c["cap_hits"] * (1 - wall) and (wall == 0 or c["reflections"] > 0)


@pytest.mark.parametrize("ppt,agg,block", [(2, 2, 256), (4, 2, 128), (2, 0, 64), (4, 0, 256)])
def test_particles_per_thread_matches_oracle(mp, orc, ppt, agg, block):
    """Several particles per thread (with smaller blocks) change only the
    schedule: cells, positions, velocities bit-exact, both density modes to their bar."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=15, n=70001)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(particles_per_thread=ppt, scatter_agg=agg, block_size=block, sort_interval=2)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(4):
            mode = k % 2
            mv.step(QM, dt, SPWT, mode)
            ref = run.step(QM, dt, SPWT)
            check_den(mv.density(), ref, mode, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
        c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
            assert c[key] == run.stats[key], key


@pytest.mark.parametrize("sort_interval", [0, 3], ids=["unsorted", "sorted"])
@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_sorted_and_unsorted_alternating_modes(mp, orc, sort_interval, wall):
    """Cell-sorted particles (a few start tets per wave) and unsorted ones (nearly one per lane)
    with a ragged last wave, atomic and deterministic steps alternating: same bits, same
    counters, both density modes to their bar."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=27 + wall, n=90001)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=wall)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=sort_interval, wall_mode=wall)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(4):
            mode = k % 2
            mv.step(QM, dt, SPWT, mode)
            check_den(mv.density(), run.step(QM, dt, SPWT), mode, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
        c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
            assert c[key] == run.stats[key], key


@pytest.mark.parametrize("shift", [3, 7])
def test_sort_shift_matches_oracle(mp, orc, shift):
    """Sorting by coarser cell buckets (cell >> shift) changes only the memory order."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=17, n=50000)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_shift=shift, sort_interval=2)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(5):
            mode = k % 2
            mv.step(QM, dt, SPWT, mode)
            ref = run.step(QM, dt, SPWT)
            check_den(mv.density(), ref, mode, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
def test_clustered_reflecting_matches_oracle(mp, orc, mode):
    """The contention workload at small size: a 3-Gaussian mixture packs tens of thousands of
    scatter records onto single nodes (the deterministic fold's whole-wave path for segments
    longer than 1024 records), reflecting walls, sort every 3 steps."""
    side = 0.1
    mesh = mp.Mesh.box_mesh(10, side)
    ef = mesh.field_sinusoid(10.0)
    centres = np.array([[0.006, 0.05, 0.05], [0.07, 0.04, 0.06], [0.05, 0.07, 0.03]])  # one at a wall
    pos, vel = mp.synth_mixture(31, 200000, [0, 0, 0], [side] * 3, [0.5, 0.3, 0.2], centres, 0.003, 3000.0,
                                [0, 0, 0])
    cell = mesh.locate_box(pos)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=1)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=3, wall_mode=1)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(6):
            mv.step(QM, dt, SPWT, mode)
            ref = run.step(QM, dt, SPWT)
            check_den(mv.density(), ref, mode, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
    per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
    assert per_node.max() > 16 * 1024 and (per_node > 1024).sum() > 50, "the case must put long segments on nodes"
    assert run.stats["reflections"] > 10000


def _ensemble(mp, n_runs=5, seed=21, n_hex=6, side=0.1):
    mesh = mp.Mesh.box_mesh(n_hex, side)
    ef = mesh.field_sinusoid(10.0)
    rng = np.random.default_rng(seed)
    runs = []
    for r in range(n_runs):
        n = int(rng.integers(2000, 12000))
        drift, sigma = rng.uniform(2000, 8000), rng.uniform(500, 3000)
        pos, vel = mp.synth_uniform(seed * 100 + r, n, [0, 0, 0], [side] * 3, sigma, [drift, 0, 0])
        runs.append(dict(pos=pos, vel=vel, cell=mesh.locate_box(pos), qm=QM * (1 + 0.1 * r),
                         dt=4e-7 * (1 + 0.05 * r), spwt=SPWT * (r + 1)))
    return mesh, ef, runs


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "stream"])
def test_batch_matches_per_run_oracle(mp, orc, mode, graph):
    """Every run of the ensemble equals its own serial MoveParticles (per-run q/m, dt, spwt)."""
    mesh, ef, runs = _ensemble(mp)
    oracles = [orc.Run(mesh.tets, ef, mesh.node_volume, r["pos"], r["vel"], r["cell"]) for r in runs]
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=2, sort_shift=2)
        mv.upload_field(ef)
        with mp.Batch(mv, len(runs)) as b:
            b.set_graph(graph)
            for k, r in reversed(list(enumerate(runs))):  # set order does not matter
                b.set_run(k, r["pos"], r["vel"], r["cell"], r["qm"], r["dt"], r["spwt"])
            for step in range(5):
                b.step(mode)
                for k, r in enumerate(runs):
                    ref = oracles[k].step(r["qm"], r["dt"], r["spwt"])
                    check_den(b.density(k), ref, mode, mp)
            for k in range(len(runs)):
                gp, gv, gc, gid = b.get_particles(k)
                o = oracles[k]
                assert np.array_equal(gid, o.pid) and np.array_equal(gc, o.cell)
                assert gp.tobytes() == o.pos.tobytes() and gv.tobytes() == o.vel.tobytes()
            c = b.counters()
            assert c["particle_steps"] == sum(o.stats["particle_steps"] for o in oracles)
            assert c["lost"] == sum(o.stats["lost"] for o in oracles) > 0


def test_batch_graph_follows_option_changes(mp, orc):
    """The batch's captured step replays the launch it captured; an mp_set_options between graph
    steps (absorbing -> reflecting walls, then a walk cap of 2) must take effect at the next step:
    every run still equals its own oracle switched alike."""
    mesh, ef, runs = _ensemble(mp, n_runs=3, seed=37)
    oracles = [orc.Run(mesh.tets, ef, mesh.node_volume, r["pos"], r["vel"], r["cell"]) for r in runs]
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=0)
        mv.upload_field(ef)
        with mp.Batch(mv, len(runs)) as b:
            b.set_graph(True)
            for k, r in enumerate(runs):
                r = dict(r, dt=4 * r["dt"])  # fast enough to reach the walls and to walk several tets
                runs[k] = r
                b.set_run(k, r["pos"], r["vel"], r["cell"], r["qm"], r["dt"], r["spwt"])
            for wall, cap in ((0, 4096), (1, 4096), (1, 2), (0, 4096)):
                mv.set_options(wall_mode=wall, walk_cap=cap)
                for o in oracles:
                    o.wall_mode, o.walk_cap = wall, cap
                for _ in range(2):
                    b.step(mp.SCATTER_ATOMIC)
                    for k, r in enumerate(runs):
                        check_den(b.density(k), oracles[k].step(r["qm"], r["dt"], r["spwt"]), 0, mp)
            for k in range(len(runs)):
                gp, gv, gc, gid = b.get_particles(k)
                assert np.array_equal(gid, oracles[k].pid) and gp.tobytes() == oracles[k].pos.tobytes()
            c = b.counters()
            assert c["reflections"] == sum(o.stats["reflections"] for o in oracles) > 0
            assert c["cap_hits"] == sum(o.stats["cap_hits"] for o in oracles) > 0


def test_batch_sequential_equals_batched(mp):
    mesh, ef, runs = _ensemble(mp, n_runs=4, seed=5)
    out = {}
    for how in ("batched", "sequential"):
        with mp.Mover.from_mesh(mesh) as mv:
            mv.upload_field(ef)
            with mp.Batch(mv, len(runs)) as b:
                for k, r in enumerate(runs):
                    b.set_run(k, r["pos"], r["vel"], r["cell"], r["qm"], r["dt"], r["spwt"])
                for _ in range(4):
                    b.step() if how == "batched" else b.step_sequential()
                out[how] = [(b.density(k), b.get_particles(k)) for k in range(len(runs))]
    for (da, pa), (db, pb) in zip(out["batched"], out["sequential"]):
        np.testing.assert_allclose(da, db, rtol=ATOMIC_RTOL, atol=0)
        for x, y in zip(pa, pb):
            assert x.tobytes() == y.tobytes()


def test_autotune_leaves_state_and_keeps_parity(mp, orc):
    """mp_autotune times candidates from a saved state and restores it: particles, density and
    counters are unchanged, the chosen options are valid, and later steps still match the
    oracle bit for bit."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=23, n=120000)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for _ in range(2):
            mv.step(QM, dt, SPWT, mp.SCATTER_DETERMINISTIC)
            run.step(QM, dt, SPWT)
        before = mv.get_particles(), mv.density(), mv.counters()
        best, ms, trials = mv.autotune(QM, dt, SPWT, mp.SCATTER_ATOMIC, min_steps=3)
        after = mv.get_particles(), mv.density(), mv.counters()
        for a, b in zip(before[0], after[0]):
            assert a.tobytes() == b.tobytes()
        assert before[1].tobytes() == after[1].tobytes() and before[2] == after[2]
        assert mv.options == best
        assert best["particles_per_thread"] in (1, 2, 4, 8)
        assert best["block_size"] in (64, 128, 256) and best["scatter_agg"] in (0, 2)
        # the reported time is one of the chosen options' timings, within 1% of the fastest trial
        # or from the final re-check against the starting options
        assert len(trials) > 20 and ms > 0
        assert any(o == best and t == ms for o, t in trials)
        for k in range(3):
            mode = k % 2
            mv.step(QM, dt, SPWT, mode)
            check_den(mv.density(), run.step(QM, dt, SPWT), mode, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and gp.tobytes() == run.pos.tobytes()


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_keep_pid_order_matches_oracle(mp, orc, wall):
    """keep_pid_order: the periodic pass compacts the lost particles in pid order instead of
    sorting by cell, and deterministic records are written by slot.  Starts in cell-sorted
    order (so the first compaction restores pid order by a scatter on pid ranks), injects
    more particles midway, alternates the two density modes."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=31 + wall, n=60000)
    dt = 6e-7
    half = 40000
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half], wall_mode=wall)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(wall_mode=wall, sort_interval=2)
        mv.upload_field(ef)
        mv.set_particles(pos[:half], vel[:half], cell[:half])
        for k in range(7):
            if k == 2:
                mv.set_options(keep_pid_order=1)
            if k == 4:
                mv.add_particles(pos[half:], vel[half:], cell[half:])
                run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
                run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
                run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
                run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
            mode = (k + 1) % 2
            mv.step(QM, dt, SPWT, mode)
            check_den(mv.density(), run.step(QM, dt, SPWT), mode, mp)
            gp, gv, gc, gid = mv.get_particles()
            assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell), f"step {k}"
            assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
        if wall == 0:
            assert run.stats["lost"] > 0


def test_error_paths(mp):
    """Calls out of order or with bad arguments return their codes and a message, and leave
    the context usable."""
    mesh, ef, pos, vel, cell = box_case(mp, n=2000, seed=41)
    L = mp.lib()
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_particles(pos, vel, cell)
        with pytest.raises(mp.MpError) as e:  # no field yet
            mv.step(QM, 1e-7, SPWT, 0)
        assert e.value.code == 4
        with pytest.raises(mp.MpError) as e:
            mv.step_many(QM, 1e-7, SPWT, 0, 3)
        assert e.value.code == 4
        with pytest.raises(mp.MpError):  # no multi-step call yet
            mv.density_step(0)
        with pytest.raises(mp.MpError):
            mv.set_options(block_size=100)
        with pytest.raises(mp.MpError):
            mv.set_options(scatter_agg=1)
        with pytest.raises(mp.MpError):
            mv.step(QM, 1e-7, SPWT, 7)
        assert L.mp_set_sort_interval(mv._h, -1) == 1 and b"interval" in L.mp_last_error(mv._h)
        assert L.mp_set_sort_interval(mv._h, 3) == 0 and mv.options["sort_interval"] == 3
        bad = cell.copy()
        bad[5] = mesh.n_tets
        with pytest.raises(mp.MpError):
            mv.add_particles(pos, vel, bad)
        mv.upload_field(ef)
        mv.step(QM, 1e-7, SPWT, 0)
        assert mv.num_active() > 0
        with pytest.raises(mp.MpError):
            mv.step_many(QM, 1e-7, SPWT, 0, 0)
        with pytest.raises(mp.MpError):
            mv.step_many(QM, 1e-7, SPWT, 5, 2)
        with pytest.raises(mp.MpError):  # the autotuner rejects an unknown mode bit
            mv.autotune(QM, 1e-7, SPWT, 0x200)
        mv.step_many(QM, 1e-7, SPWT, 0, 2)
        assert mv.density_step(1).shape == (mesh.n_nodes,)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        with pytest.raises(mp.MpError) as e:  # nothing resident to tune on
            mv.autotune(QM, 1e-7, SPWT)
        assert e.value.code == 4


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
@pytest.mark.parametrize("agg", [2, 0], ids=["lds", "plain"])
@pytest.mark.parametrize("dt,sort_interval,n_steps", [(4e-7, 3, 7), (1.5e-6, 0, 5)],
                         ids=["sorted-3", "fast-unsorted"])
def test_step_many_matches_oracle(mp, orc, wall, agg, dt, sort_interval, n_steps):
    """mp_step_many: the steps between two sorts in one launch, the particles kept in
    registers (7 steps with a sort every 3 = launches of 2, 3 and 2 steps after the first
    step's carried-over interval; the fast unsorted case runs 5 steps in one launch and hits
    the walk cap).  Every step's density, the particles and the counters against the oracle
    stepped one step at a time."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=53 + wall, n=70001)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell, wall_mode=wall, walk_cap=6)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=sort_interval, wall_mode=wall, walk_cap=6, scatter_agg=agg)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.step(QM, dt, SPWT, 0)  # one ordinary step first: the multi-step call starts mid-interval
        check_den(mv.density(), run.step(QM, dt, SPWT), 0, mp)
        mv.step_many(QM, dt, SPWT, 0, n_steps)
        for k in range(n_steps):
            ref = run.step(QM, dt, SPWT)
            check_den(mv.density_step(k), ref, 0, mp)
        check_den(mv.density(), ref, 0, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
        assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
        c = mv.counters()
        for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
            assert c[key] == run.stats[key], key
        assert c["crossed"] > 0 and (wall == 0 or c["reflections"] > 0)
        with pytest.raises(mp.MpError):
            mv.density_step(n_steps)


def test_step_many_deterministic_and_timing(mp, orc):
    """Deterministic mode through mp_step_many (one step at a time): every step's density
    bit-exact; the push timing counts one launch per multi-step stretch in atomic mode."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=59, n=50001)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=4)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.step_many(QM, 4e-7, SPWT, 1, 3)
        for k in range(3):
            check_den(mv.density_step(k), run.step(QM, 4e-7, SPWT), 1, mp)
        mv.set_timing(True)
        mv.step_many(QM, 4e-7, SPWT, 0, 6)  # 1 step to the sort, then 4, then 1
        for k in range(6):
            check_den(mv.density_step(k), run.step(QM, 4e-7, SPWT), 0, mp)
        ms, launches = mv.timing()
        mv.set_timing(False)
        assert launches == 3 and ms > 0
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and gp.tobytes() == run.pos.tobytes()


@pytest.mark.parametrize("opts", [dict(scatter_agg=2), dict(scatter_agg=2, particles_per_thread=2)],
                         ids=["lds-hash", "lds-hash-2ppt"])
def test_node_hash_overflow_unsorted_large_mesh(mp, orc, opts):
    """Unsorted particles on a 20^3 box (9261 nodes): a workgroup's 256 particles bring up to
    1024 distinct nodes, more than the node hash holds within its probe limit, so contributions
    spill to direct HBM atomics.  Same bits and densities as the oracle."""
    mesh, ef, pos, vel, cell = box_case(mp, n_hex=20, side=0.1, n=40001, seed=61)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=0, **opts)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for _ in range(2):
            mv.step(QM, 4e-7, SPWT, 0)
            check_den(mv.density(), run.step(QM, 4e-7, SPWT), 0, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gc, run.cell) and gp.tobytes() == run.pos.tobytes()


@pytest.mark.parametrize("order", [0, 1], ids=["cell", "morton"])
@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
def test_batch_step_many_matches_per_run_oracle(mp, orc, mode, order):
    """mp_batch_step_many: every run's density at every step, its particles and the counters
    against its own serial MoveParticles; sort every 3 steps, 7 steps (launches of 3, 3, 1)."""
    mesh, ef, runs = _ensemble(mp)
    oracles = [orc.Run(mesh.tets, ef, mesh.node_volume, r["pos"], r["vel"], r["cell"]) for r in runs]
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=3, sort_shift=2, sort_order=order)
        mv.upload_field(ef)
        with mp.Batch(mv, len(runs)) as b:
            for k, r in enumerate(runs):
                b.set_run(k, r["pos"], r["vel"], r["cell"], r["qm"], r["dt"], r["spwt"])
            b.step_many(7, mode)
            for step in range(7):
                for k, r in enumerate(runs):
                    ref = oracles[k].step(r["qm"], r["dt"], r["spwt"])
                    check_den(b.density_step(step, k), ref, mode, mp)
                    if step == 6:
                        check_den(b.density(k), ref, mode, mp)
            for k in range(len(runs)):
                gp, gv, gc, gid = b.get_particles(k)
                o = oracles[k]
                assert np.array_equal(gid, o.pid) and np.array_equal(gc, o.cell)
                assert gp.tobytes() == o.pos.tobytes() and gv.tobytes() == o.vel.tobytes()
            c = b.counters()
            assert c["particle_steps"] == sum(o.stats["particle_steps"] for o in oracles)
            assert c["lost"] == sum(o.stats["lost"] for o in oracles) > 0
            with pytest.raises(mp.MpError):
                b.density_step(7, 0)


def test_autotune_multi_step_keeps_state_and_parity(mp, orc):
    """mp_autotune timing candidates through mp_step_many: the particle state is restored, the
    chosen options are among those the multi-step kernel takes, and mp_step_many afterwards
    still matches the oracle bit for bit (particles) and within 1e-12 (densities)."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=67, n=120000)
    dt = 4e-7
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.step_many(QM, dt, SPWT, 0, 2)
        for _ in range(2):
            run.step(QM, dt, SPWT)
        before = mv.get_particles()
        best, ms, trials = mv.autotune(QM, dt, SPWT, mp.SCATTER_ATOMIC, min_steps=4, multi_step=True)
        after = mv.get_particles()
        for a, b in zip(before, after):
            assert a.tobytes() == b.tobytes()
        assert mv.options == best and best["scatter_agg"] in (0, 2) and ms > 0
        assert len({t["sort_interval"] for t, _ in trials}) > 3
        mv.step_many(QM, dt, SPWT, 0, 5)
        for k in range(5):
            check_den(mv.density_step(k), run.step(QM, dt, SPWT), 0, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and gp.tobytes() == run.pos.tobytes()


@pytest.mark.parametrize("order", [0, 1], ids=["cell", "morton"])
@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_step_many_gather_sort(mp, orc, wall, order):
    """The gather sort (k_cell_rank writes only the permutation, the next k_push_steps launch
    moves the records) against the copy sort and the oracle: a sort every 3 steps, particles
    injected between two calls (sorted in before the next launch), every step's density, the
    particles and the counters; then steps that lose every particle, so that a later sort
    finds none in the domain."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=83 + wall, n=60001)
    dt = 4e-7
    half = len(pos) // 2
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half], wall_mode=wall,
                  walk_cap=6)
    movers = []
    try:
        for gather in (1, 0):
            mv = mp.Mover.from_mesh(mesh)
            movers.append(mv)
            mv.set_options(sort_interval=3, wall_mode=wall, walk_cap=6, gather_sort=gather, sort_order=order)
            assert mv.options["gather_sort"] == gather
            mv.upload_field(ef)
            mv.set_particles(pos[:half], vel[:half], cell[:half])
        for mv in movers:
            mv.step_many(QM, dt, SPWT, 0, 7)
        for k in range(7):
            ref = run.step(QM, dt, SPWT)
            for mv in movers:
                check_den(mv.density_step(k), ref, 0, mp)
        for mv in movers:
            mv.add_particles(pos[half:], vel[half:], cell[half:])
        run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
        run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
        run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
        run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
        for mv in movers:
            mv.step_many(QM, dt, SPWT, 0, 8)
        for k in range(8):
            ref = run.step(QM, dt, SPWT)
            for mv in movers:
                check_den(mv.density_step(k), ref, 0, mp)
        for mv in movers:
            gp, gv, gc, gid = mv.get_particles()
            assert np.array_equal(gid, run.pid) and np.array_equal(gc, run.cell)
            assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()
            c = mv.counters()
            for key in ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections"):
                assert c[key] == run.stats[key], key
        for mv in movers:
            mv.step_many(QM, 1e-3, SPWT, 0, 4)  # every particle is capped or leaves
            assert mv.num_active() == 0
            mv.step_many(QM, 1e-3, SPWT, 0, 4)  # sorts a set with no particle in the domain
            assert mv.num_active() == 0 and not mv.density_step(3).any()
            assert mv.get_particles()[3].size == 0
    finally:
        for mv in movers:
            mv.close()


def check_bucket_order(mv, shift, live_pids, rank=None):
    cell, pid = mv.slots()
    assert (cell >= 0).all(), "a cell sort drops the lost particles"
    pos = cell if rank is None else rank[cell]
    assert (np.diff(pos >> shift) >= 0).all(), "slots not in bucket order"
    assert np.array_equal(np.sort(pid), np.sort(live_pids))


def test_tet_order_is_a_compact_curve(mp):
    """sort_order 1's tet rank: a permutation of the tets, and consecutive ranks are spatial
    neighbours (a wrong centroid from the barycentric coefficients would scatter them)."""
    mesh = mp.Mesh.box_mesh(12, 0.1)
    with mp.Mover.from_mesh(mesh) as mv:
        rank = mv.tet_order(mesh.n_tets)
    assert np.array_equal(np.sort(rank), np.arange(mesh.n_tets))
    cent = mesh.node_pos[mesh.tets["con"]].mean(1)
    order = np.argsort(rank)
    step = np.linalg.norm(np.diff(cent[order], axis=0), axis=1)
    h = 0.1 / 12
    # tets of one hex sit 0.35-0.6 h apart; a wrong centroid would give a near-random order
    # (mean step ~8 h on this box)
    assert np.median(step) < 1.0 * h and np.mean(step) < 2.0 * h


@pytest.mark.parametrize("order", [0, 1], ids=["cell", "morton"])
@pytest.mark.parametrize("shift", [0, 3])
@pytest.mark.parametrize("gather", [0, 1], ids=["copy", "gather"])
def test_cell_sort_orders_buckets(mp, shift, gather, order):
    """The cell sort's result itself (parity holds for any permutation): from the host's random
    order, every live particle once, in non-decreasing cell >> shift; then a few steps later,
    from near-sorted order, with the lost particles dropped.  With the gather sort the sort
    is completed by the next mp_step_many launch, run here with dt = 0 so that no particle
    changes cell after it."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=91, n=300001)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=4, sort_shift=shift, gather_sort=gather, sort_order=order)
        rank = mv.tet_order(mesh.n_tets) if order else None
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        c0, p0 = mv.slots()
        assert np.array_equal(p0, np.arange(len(pos))) and np.array_equal(c0, cell)
        mv.sort_now()
        check_bucket_order(mv, shift, np.arange(len(pos)), rank)
        mv.step_many(QM, 4e-7, SPWT, 0, 3)  # no sort: 3 of the 4 steps of the interval
        live = mv.get_particles()[3]
        assert 0 < len(live) < len(pos)
        if gather:
            mv.step_many(QM, 0.0, SPWT, 0, 2)  # the 4th step, then the sort and one more step
        else:
            mv.sort_now()
        check_bucket_order(mv, shift, live, rank)


def test_step_many_long_and_empty(mp, orc):
    """mp_step_many over more steps than one launch takes (no periodic sort: 40 steps = launches
    of 32 and 8), and on an empty particle set (densities of zeros)."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=71, n=20000)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=0)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        mv.set_timing(True)
        mv.step_many(QM, 2e-7, SPWT, 0, 40)
        _, launches = mv.timing()
        mv.set_timing(False)
        assert launches == 2
        for k in range(40):
            check_den(mv.density_step(k), run.step(QM, 2e-7, SPWT), 0, mp)
        gp, gv, gc, gid = mv.get_particles()
        assert np.array_equal(gid, run.pid) and gp.tobytes() == run.pos.tobytes()
        assert mv.counters()["particle_steps"] == run.stats["particle_steps"]
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos[:0], vel[:0], cell[:0])
        mv.step_many(QM, 2e-7, SPWT, 0, 3)
        assert not mv.density_step(2).any() and mv.counters()["particle_steps"] == 0
