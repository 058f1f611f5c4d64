"""GPU parity of the deterministic scatters against the serial oracle.

det_algo 2 (the default) has the push place the records into slab segments sized from the
last step's counts: a cold node's one slab, or a hot node's pid slabs (a step that does not fit,
or with a slab over its limit, is rebuilt as with 1).  det_algo 1 rebuilds each
particle's records from the pushed state and groups them into node segments (dense nodes cut
into pid slabs).  Either way each segment is put in pid order on chip: a wave ranks records
within sub-buckets of the pid range (a register bitonic network where the pids bunch), and a
node with a slab over 256 records goes to a workgroup LSD radix sort.  det_algo 0 writes the
records in pid order and sorts them on the node bits with a device radix sort.  Every way the density must be the
ascending-pid fold of the serial MoveParticles, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)
SPWT = 200.0
DET = 1

# every register-count boundary of the wave sort (64 records a register), the last length
# it takes (512) and the first ones of the workgroup path
LENGTHS = [1, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 700, 3000]


def box_case(mp, n_hex=8, side=0.1, n=60000, seed=1):
    mesh = mp.Mesh.box_mesh(n_hex, side)
    ef = mesh.field_sinusoid(10.0)
    pos, vel = mp.synth_uniform(seed, n, [0, 0, 0], [side] * 3, 2000.0, [5000.0, 0, 0])
    return mesh, ef, pos, vel, mesh.locate_box(pos)


def check_state(mv, run, k):
    gp, gv, gc, gid = mv.get_particles()
    assert np.array_equal(gid, run.pid), f"step {k}: particle set differs"
    assert np.array_equal(gc, run.cell), f"step {k}: cells differ"
    assert gp.tobytes() == run.pos.tobytes() and gv.tobytes() == run.vel.tobytes()


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
@pytest.mark.parametrize("pid_order", [0, 1], ids=["cell-sorted", "pid-order"])
@pytest.mark.parametrize("algo", [0, 1, 2], ids=["radix", "segments", "placed"])
def test_det_algos_match_oracle(mp, orc, algo, pid_order, wall):
    """Every algorithm with cell-sorted slots and with pid-order compaction, 6 steps with a
    periodic pass every 2 steps and an injection midway (a third more particles: the placed
    records of det_algo 2 overflow their segments and the step is rebuilt); the algorithm is
    switched for one step in the middle."""
    mesh, ef, pos, vel, cell = box_case(mp, seed=71 + 2 * wall + pid_order)
    half = 45000
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half], wall_mode=wall)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=2, wall_mode=wall, keep_pid_order=pid_order, det_algo=algo)
        assert mv.options["det_algo"] == algo
        mv.upload_field(ef)
        mv.set_particles(pos[:half], vel[:half], cell[:half])
        for k in range(6):
            if k == 3:
                mv.add_particles(pos[half:], vel[half:], cell[half:])
                run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
                run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
                run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
                run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
            mv.set_options(det_algo=(algo + 1) % 3 if k == 4 else algo)
            mv.step(QM, 6e-7, SPWT, DET)
            ref = run.step(QM, 6e-7, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
            check_state(mv, run, k)
        assert run.stats["crossed"] > 0


def segment_case(mp, fillers, seed=3):
    """Particles at rest, with no field, in one tet of each of 15 cubes that share no nodes (every
    other cube along each axis): each node of the tet holding k particles folds exactly k
    records.  The pids are shuffled over the tets.  `fillers` particles first, each leaving the
    box in the first step, push the pids of the resting ones past 16 bits."""
    n_hex, side = 8, 0.1
    h = side / n_hex
    mesh = mp.Mesh.box_mesh(n_hex, side)
    rng = np.random.default_rng(seed)
    cubes = [(2 * a, 2 * b, 2 * c) for a in range(4) for b in range(4) for c in range(4)]
    pos = []
    for k, cube in zip(LENGTHS, cubes):
        p0 = (np.array(cube) + [0.3, 0.2, 0.1]) * h
        pos.append(p0 + rng.uniform(-0.01 * h, 0.01 * h, (k, 3)))
    pos = np.concatenate(pos)
    pos = pos[rng.permutation(len(pos))]
    vel = np.zeros_like(pos)
    if fillers:
        fp = rng.uniform(0.0, side, (fillers, 3))
        fv = np.tile([1e6, 0.0, 0.0], (fillers, 1))
        pos, vel = np.concatenate([fp, pos]), np.concatenate([fv, vel])
    return mesh, np.zeros((mesh.n_tets, 3)), pos, vel, mesh.locate_box(pos)


@pytest.mark.parametrize("algo", [1, 2], ids=["segments", "placed"])
@pytest.mark.parametrize("sort_interval", [0, 1], ids=["slots-in-pid-order", "cell-sorted"])
@pytest.mark.parametrize("fillers", [0, 70000], ids=["pids-13bit", "pids-17bit"])
def test_segment_lengths_at_every_boundary(mp, orc, fillers, sort_interval, algo):
    """Node segments of exactly 1, 63..65, 127..129, 255..257, 511..513, 700 and 3000 records, in
    slots left in pid order or cell-sorted; pids under 2^13 (two radix passes on the long
    segments) or from 70000 up (four)."""
    mesh, ef, pos, vel, cell = segment_case(mp, fillers)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=sort_interval, det_algo=algo)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(3):
            mv.step(QM, 1e-6, SPWT, DET)
            ref = run.step(QM, 1e-6, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
            check_state(mv, run, k)
    per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
    assert sorted(set(per_node.tolist()) - {0}) == LENGTHS
    assert run.pid.min() == fillers and len(run.pid) == sum(LENGTHS)


@pytest.mark.parametrize("algo", [1, 2], ids=["segments", "placed"])
def test_segments_with_hot_nodes_and_empty_steps(mp, orc, algo):
    """A dense cluster (dozens of nodes over 512 records, the hottest over 6 x 10^4) beside a
    smaller one, then every particle absorbed: the step after that has no records at all and
    must leave a zero density.  With det_algo 2 the first step is exact and the next ones place
    the hot nodes' records by pid slab."""
    side = 0.1
    mesh = mp.Mesh.box_mesh(10, side)
    ef = mesh.field_sinusoid(10.0)
    centres = np.array([[0.05, 0.05, 0.05], [0.02, 0.08, 0.03]])
    pos, vel = mp.synth_mixture(5, 150000, [0, 0, 0], [side] * 3, [0.7, 0.3], centres, 0.004, 1500.0, [0, 0, 0])
    cell = mesh.locate_box(pos)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=2, det_algo=algo)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        for k in range(4):
            mv.step(QM, 4e-7, SPWT, DET)
            ref = run.step(QM, 4e-7, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
        per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
        assert per_node.max() > 60000 and (per_node > 512).sum() > 50
        check_state(mv, run, 4)
        # everything out through the walls, then a step with nothing left
        mv.set_particles(pos[:1000], np.tile([1e6, 0.0, 0.0], (1000, 1)), cell[:1000])
        mv.step(QM, 1e-6, SPWT, DET)
        assert mv.num_active() == 0
        mv.step(QM, 1e-6, SPWT, DET)
        assert not mv.density().any()


@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_placed_records_match_oracle(mp, orc, wall):
    """det_algo 2 on a sparse 10^3 box (about 200 records per interior node): the first step is
    rebuilt (no segment layout yet), the next ones are placed by the push; an injection of a
    third more particles overflows the segments once (that step is rebuilt) and placement
    resumes.  Every step bit-exact against the serial fold."""
    mesh = mp.Mesh.box_mesh(10, 0.1)
    ef = mesh.field_sinusoid(10.0)
    # no drift against a reflecting wall: particles would pile up there into nodes over 512
    drift = [5000.0, 0, 0] if wall == 0 else [0.0, 0, 0]
    pos, vel = mp.synth_uniform(91 + wall, 52000, [0, 0, 0], [0.1] * 3, 2000.0, drift)
    cell = mesh.locate_box(pos)
    half = 39000
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half], wall_mode=wall)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=3, wall_mode=wall, det_algo=2)
        mv.upload_field(ef)
        mv.set_particles(pos[:half], vel[:half], cell[:half])
        for k in range(9):
            if k == 5:
                mv.add_particles(pos[half:], vel[half:], cell[half:])
                run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
                run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
                run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
                run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
            mv.step(QM, 6e-7, SPWT, DET)
            ref = run.step(QM, 6e-7, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
            check_state(mv, run, k)
        placed, rebuilt = mv.det_paths()
        per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
    assert per_node.max() <= 512
    assert placed >= 6 and rebuilt >= 2 and placed + rebuilt == 9, (placed, rebuilt)


@pytest.mark.parametrize("pid_bits", [0, 4], ids=["bucket-key", "pid-digit-key"])
@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_placed_hot_slabs_match_oracle(mp, orc, wall, pid_bits):
    """det_algo 2 where most records fall on hot nodes (a cluster with ~170 nodes over 512
    records): after the first (exact) step the push places every record into its slab, a hot
    node's slab picked by the pid digit; the cell sorts (with the pid digit in the key or
    without) change the slots but not the slabs; cold nodes that grow past 512 records are
    sorted by the workgroup fallback in their placed range (steps 4, 5, 8 and 9 here); the
    injection of a fifth more pids changes the digit's multiplier and overflows segments once
    (that step is rebuilt).  Every step bit-exact against the serial fold."""
    side = 0.1
    mesh = mp.Mesh.box_mesh(10, side)
    ef = mesh.field_sinusoid(10.0)
    centres = np.array([[0.05, 0.05, 0.05], [0.03, 0.07, 0.04]])
    pos, vel = mp.synth_mixture(17 + wall, 120000, [0, 0, 0], [side] * 3, [0.6, 0.4], centres, 0.012, 1500.0,
                                [0, 0, 0])
    cell = mesh.locate_box(pos)
    half = 100000
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos[:half], vel[:half], cell[:half], wall_mode=wall)
    hot_steps = 0
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=3, wall_mode=wall, det_algo=2, sort_pid_bits=pid_bits)
        mv.upload_field(ef)
        mv.set_particles(pos[:half], vel[:half], cell[:half])
        for k in range(10):
            if k == 6:
                mv.add_particles(pos[half:], vel[half:], cell[half:])
                run.pos = np.ascontiguousarray(np.concatenate([run.pos, pos[half:]]))
                run.vel = np.ascontiguousarray(np.concatenate([run.vel, vel[half:]]))
                run.cell = np.ascontiguousarray(np.concatenate([run.cell, cell[half:]]).astype(np.int32))
                run.pid = np.concatenate([run.pid, np.arange(half, len(pos))])
            mv.step(QM, 1e-7, SPWT, DET)
            ref = run.step(QM, 1e-7, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
            per_node = np.bincount(mesh.tets["con"][run.cell].ravel(), minlength=mesh.n_nodes)
            hot_steps += int((per_node > 512).sum() > 100)
        check_state(mv, run, 10)
        placed, rebuilt = mv.det_paths()
    assert hot_steps == 10
    # rebuilt: the first step (no layout yet) and the injection's
    assert placed >= 7 and rebuilt >= 2 and placed + rebuilt == 10, (placed, rebuilt)


def test_placed_slab_over_256_continues(mp, orc):
    """A motionless blob of 350 particles with consecutive new pids dropped into one tet of a
    uniform run: the injection step overflows the segments and is rebuilt, which cuts the
    blob's nodes into 8 pid slabs each, the top one holding the blob (~380 records, over the
    slab loop's 256); the placed steps after it stop each such node's chain at that slab and
    finish it in the continuation kernel.  Every step bit-exact against the serial fold."""
    side = 0.1
    mesh = mp.Mesh.box_mesh(10, side)
    ef = mesh.field_sinusoid(10.0)
    pos, vel = mp.synth_uniform(31, 100000, [0, 0, 0], [side] * 3, 2000.0, [0.0, 0, 0])
    cell = mesh.locate_box(pos)
    k_blob = 350
    bpos = np.tile([0.0552, 0.0517, 0.0473], (k_blob, 1))
    bvel = np.zeros((k_blob, 3))
    bcell = mesh.locate_box(bpos)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, vel, cell)
    n0 = len(pos)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=4, det_algo=2)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        top = 0
        for k in range(8):
            if k == 3:
                mv.add_particles(bpos, bvel, bcell)
                run.pos = np.ascontiguousarray(np.concatenate([run.pos, bpos]))
                run.vel = np.ascontiguousarray(np.concatenate([run.vel, bvel]))
                run.cell = np.ascontiguousarray(np.concatenate([run.cell, bcell]).astype(np.int32))
                run.pid = np.concatenate([run.pid, np.arange(n0, n0 + k_blob)])
            mv.step(QM, 4e-7, SPWT, DET)
            ref = run.step(QM, 4e-7, SPWT)
            assert mv.density().tobytes() == ref.tobytes(), f"step {k}: density not bit-exact"
            if k >= 4:
                # records of the blob's first node in the top eighth of the pid digit (its top
                # slab when it has 8): over 256
                nodes = mesh.tets["con"][run.cell]
                p_all = n0 + k_blob
                mul = (1 << 48) // p_all
                live = run.cell >= 0
                node = int(mesh.tets["con"][bcell[0]][0])
                hit = (nodes == node).any(axis=1) & live
                d = (run.pid[hit].astype(np.uint64) * np.uint64(mul)) >> np.uint64(32)
                top = max(top, int((d >> np.uint64(13) == 7).sum()))
        check_state(mv, run, 8)
        placed, rebuilt = mv.det_paths()
    assert 256 < top <= 512, top
    assert rebuilt >= 2 and placed >= 5 and placed + rebuilt == 8, (placed, rebuilt)
