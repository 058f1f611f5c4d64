"""GPU adjoint (mp_adjoint) against the CPU FP64 oracle gradient, central finite differences
through the forward path, and the torch.autograd.Function binding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)
SPWT = 200.0


def problem(mp, n_hex=5, side=0.1, n=6000, steps=12, seed=31, drift=5000.0):
    mesh = mp.Mesh.box_mesh(n_hex, side)
    ef = mesh.field_sinusoid(10.0)
    pos, u = mp.synth_uniform(seed, n, [0, 0, 0], [side] * 3, 2000.0, [0, 0, 0])
    cell = mesh.locate_box(pos)
    d = np.array([1.0, 0.0, 0.0])
    dt = 2e-7
    return mesh, ef, pos, u, cell, d, dt, steps, drift


def target(mp, mesh, ef, pos, u, cell, d, dt, steps, drift):
    """Target density: a forward run with the drift raised by 5 %."""
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        mv.set_particles(pos, u + 1.05 * drift * d, cell)
        for _ in range(steps):
            mv.step(QM, dt, SPWT, 1)
        return mv.density()


def test_adjoint_matches_oracle(mp, orc):
    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp)
    tgt = target(mp, mesh, ef, pos, u, cell, d, dt, steps, drift)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        L, g, gd = mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=5)
    Lr, gr, gdr = orc.adjoint(mesh.tets, ef, mesh.node_volume, pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d)
    assert abs(L - Lr) <= 1e-12 * abs(Lr)
    scale = np.abs(gr).max()
    assert scale > 0
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-10 * scale)
    assert abs(gd - gdr) <= 1e-10 * abs(gdr)


def test_adjoint_deterministic_same_bits_and_matches_oracle(mp, orc):
    """mp_adjoint_ex in deterministic mode: the same loss, dL/dE and dL/d(drift) bits on every
    run, also with a different cell-sort key (which reorders the particles and so the atomics);
    against the oracle gradient to the atomic mode's bar."""
    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp, n=20000)
    tgt = target(mp, mesh, ef, pos, u, cell, d, dt, steps, drift)
    outs = []
    for shift, order in ((0, 0), (3, 1), (0, 0)):
        with mp.Mover.from_mesh(mesh) as mv:
            mv.set_options(sort_shift=shift, sort_order=order)
            mv.upload_field(ef)
            outs.append(mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=5,
                                   mode=mp.SCATTER_DETERMINISTIC))
    for o in outs[1:]:
        assert o[0] == outs[0][0] and o[2] == outs[0][2]
        assert o[1].tobytes() == outs[0][1].tobytes()
    L, g, gd = outs[0]
    Lr, gr, gdr = orc.adjoint(mesh.tets, ef, mesh.node_volume, pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d)
    assert abs(L - Lr) <= 1e-12 * abs(Lr)
    scale = np.abs(gr).max()
    assert scale > 0 and np.count_nonzero(g) > 0
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-10 * scale)
    assert abs(gd - gdr) <= 1e-10 * abs(gdr)


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "deterministic"])
@pytest.mark.parametrize("ckpt", [1, 5])
def test_adjoint_reflecting_walls_matches_oracle(mp, orc, mode, ckpt):
    """Reflecting walls (a 2 cm box, so that most particles reach a wall in the 12 steps and
    some reflect twice): the loss, dL/dE and dL/d(drift) against the oracle, whose reverse
    sweep applies each step's reflections found by re-walking the step (pinned against central
    differences in the CPU suite); the deterministic mode gives the same bits on two runs."""
    side = 0.02
    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp, side=side)
    run = orc.Run(mesh.tets, ef, mesh.node_volume, pos, u + drift * d, cell, wall_mode=1)
    for _ in range(steps):
        run.step(QM, dt, SPWT)
    assert run.stats["reflections"] > len(pos) // 2 and run.stats["lost"] == 0
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(wall_mode=1)
        mv.upload_field(ef)
        mv.set_particles(pos, u + 1.05 * drift * d, cell)
        for _ in range(steps):
            mv.step(QM, dt, SPWT, 1)
        tgt = mv.density()
        outs = [mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=ckpt, mode=mode)
                for _ in range(2 if mode else 1)]
    L, g, gd = outs[0]
    if mode:
        assert outs[1][0] == L and outs[1][2] == gd and outs[1][1].tobytes() == g.tobytes()
    Lr, gr, gdr = orc.adjoint(mesh.tets, ef, mesh.node_volume, pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d,
                              wall_mode=1)
    assert abs(L - Lr) <= 1e-12 * abs(Lr)
    scale = np.abs(gr).max()
    assert scale > 0 and np.count_nonzero(g) > 0
    np.testing.assert_allclose(g, gr, rtol=0, atol=1e-10 * scale)
    assert abs(gd - gdr) <= 1e-10 * abs(gdr)


@pytest.mark.parametrize("ckpt", [1, 4, 12])
def test_adjoint_checkpointing_invariant(mp, ckpt):
    """The gradient does not depend on the checkpoint interval (replay is exact)."""
    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp, n=3000, steps=8)
    tgt = np.full(mesh.n_nodes, 1e7)
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        ref = mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=steps)
        got = mv.adjoint(pos, u + drift * d, cell, QM, dt, SPWT, steps, tgt, d, checkpoint_every=ckpt)
    assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
    np.testing.assert_allclose(got[1], ref[1], rtol=0, atol=1e-10 * np.abs(ref[1]).max())
    assert abs(got[2] - ref[2]) <= 1e-10 * abs(ref[2])


def test_adjoint_vs_central_differences(mp):
    """32 random E components and the drift: central differences of the GPU forward
    (deterministic scatter) with steps small enough that no particle changes its tet path.
    On a fixed path the loss is quadratic in E and in the drift, so central differences
    have no truncation error."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from torch_adjoint import FrozenFieldProblem

    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp, n=8000, steps=10)
    tgt = target(mp, mesh, ef, pos, u, cell, d, dt, steps, drift)
    with mp.Mover.from_mesh(mesh) as mv:
        pb = FrozenFieldProblem(mp, mv, pos, u, cell, d, QM, dt, SPWT, steps, tgt, checkpoint_every=4)
        L0, path0, _ = pb.forward(ef, drift)
        L, g, gd = pb.adjoint(ef, drift)
        assert abs(L - L0) <= 1e-12 * L0
        rng = np.random.default_rng(7)
        cand = np.flatnonzero(np.abs(g).max(1) > 1e-3 * np.abs(g).max())
        picks = [(int(e), int(rng.integers(3))) for e in rng.choice(cand, 32, replace=False)]
        scale = np.abs(g).max()
        h = 1e-4 * np.abs(ef).max()
        for e, k in picks:
            ep, em = ef.copy(), ef.copy()
            ep[e, k] += h
            em[e, k] -= h
            Lp, pp, _ = pb.forward(ep, drift)
            Lm, pm, _ = pb.forward(em, drift)
            assert all(np.array_equal(a, b) for a, b in zip(pp, path0)), "a tet path changed"
            assert all(np.array_equal(a, b) for a, b in zip(pm, path0)), "a tet path changed"
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - g[e, k]) <= 1e-6 * scale, (e, k, fd, g[e, k])
        hd = 1e-3
        Lp, pp, _ = pb.forward(ef, drift + hd)
        Lm, pm, _ = pb.forward(ef, drift - hd)
        assert all(np.array_equal(a, b) for a, b in zip(pp, path0))
        fd = (Lp - Lm) / (2 * hd)
        assert abs(fd - gd) <= 1e-6 * abs(gd)


def test_torch_autograd_binding(mp):
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    import torch
    from torch_adjoint import AdjointLoss, FrozenFieldProblem

    mesh, ef, pos, u, cell, d, dt, steps, drift = problem(mp, n=3000, steps=6)
    tgt = target(mp, mesh, ef, pos, u, cell, d, dt, steps, drift)
    with mp.Mover.from_mesh(mesh) as mv:
        pb = FrozenFieldProblem(mp, mv, pos, u, cell, d, QM, dt, SPWT, steps, tgt, checkpoint_every=3)
        E = torch.tensor(ef, dtype=torch.float64, requires_grad=True)
        dr = torch.tensor(drift, dtype=torch.float64, requires_grad=True)
        loss = AdjointLoss.apply(E, dr, pb)
        (3.0 * loss).backward()
        L, g, gd = pb.adjoint(ef, drift)
    # two adjoint calls differ only in the order of their FP64 atomic sums
    assert abs(loss.item() - L) <= 1e-12 * L
    np.testing.assert_allclose(E.grad.numpy(), 3.0 * g, rtol=0, atol=1e-10 * np.abs(3.0 * g).max())
    assert abs(dr.grad.item() - 3.0 * gd) <= 1e-10 * abs(3.0 * gd)
