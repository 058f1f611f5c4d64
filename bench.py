"""Benchmark of the MI355X particle mover: FP64 particle-steps/s of push + tet walk + scatter.

Workload (one GPU): a 60x60x60 six-tet box of side 1 m (1,296,000 tets, 226,981 nodes),
20,000,000 FP64 particles uniform in the box, Maxwellian velocities (sigma 2000 m/s) plus a
5000 m/s +x drift, frozen field from phi = 10 V sin(pi x) sin(pi y) sin(pi z), O+ ions
(16 amu, +1 e), dt = 2.75e-7 s (about 30% of particles cross a face per step), absorbing
walls, cell sort every 10 steps.  Particles stay resident in HBM between steps.

Prints ONE JSON line.  value = particle-steps/s over the timed steps (whole step: push,
walk, scatter, normalisation, periodic sort) of the frozen-field run through mp_step_many,
which pushes the steps between two periodic sorts in one launch with the particles held in
registers; per_step = the same workload one mp_step (one MoveParticles) per step.  roofline
= the push kernel's algorithmic HBM bytes (104 B per active particle-step + 192 B per tet
and 16 B per node per step) per launch over its average launch time from HIP events on the
context stream; traffic = the HBM bytes rocprofv3 measured per launch.  cpu_baseline = the
oracle's OpenMP port of the reference kernel body on the host cores, on a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mini-fem-pic-main_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import minipic_amd as mp  # noqa: E402

HBM_PEAK_GBS = 8000.0
N_HEX, SIDE, N_PART = 60, 1.0, 20_000_000
SIGMA, DRIFT, AMP = 2000.0, 5000.0, 10.0
DT = 2.75e-7
SPWT = 200.0
QM = mp.QE / (16 * mp.AMU)
SEED = 20240601


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measured_traffic(name="k_push_soa_box60_20M_traffic.txt", key="traffic_bytes"):
    """HBM bytes from the committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs of
    this workload, calibrated on a copy kernel of known size): per push launch for the per-step
    kernel, per step for the multi-step kernel; None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.split() and line.split()[0] == key:
                return float(line.split()[1])
    except OSError:
        pass
    return None


def traffic_fields(traffic, launch_s, profile):
    """roofline fields from the committed profile's measured HBM bytes per launch: the bytes and
    the HBM bandwidth they mean at this run's launch time, as a fraction of the peak."""
    if traffic is None:
        return {"traffic": None, "traffic_source": None, "measured_gbs": None, "measured_frac": None}
    gbs = traffic / launch_s / 1e9
    return {"traffic": traffic,
            "traffic_source": f"profiles/{profile}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, "
                              "calibrated on a copy kernel of known size (not measured in this run)",
            "measured_gbs": gbs, "measured_frac": gbs / HBM_PEAK_GBS}


def push_bytes(active_particle_steps, launches, n_tets, n_nodes):
    return 104.0 * active_particle_steps + launches * (192.0 * n_tets + 16.0 * n_nodes)


def run_gpu_many(mesh, ef, pos, vel, cell, steps, warmup, options, tune=True):
    """The workload through mp_step_many (frozen field): each stretch of steps between two
    periodic sorts is one launch with the particles held in registers (256-thread workgroups,
    one particle per thread).  With tune, mp_autotune times its candidates through
    mp_step_many first (untimed; the particle state is restored)."""
    tuned = None
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(**options)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        if tune:
            t = time.perf_counter()
            best, best_ms, trials = mv.autotune(QM, DT, SPWT, mp.SCATTER_ATOMIC, min_steps=10, multi_step=True)
            tuned = {"options": best, "ms_per_step": best_ms, "trials": len(trials),
                     "tuner_s": time.perf_counter() - t}
            log(f"autotune multi-step ({len(trials)} trials, {tuned['tuner_s']:.1f} s): {best} at {best_ms:.3f} ms/step")
        mv.step_many(QM, DT, SPWT, mp.SCATTER_ATOMIC, warmup)
        mv.sync()
        mv.reset_counters()
        mv.set_timing(True)
        t0 = time.perf_counter()
        mv.step_many(QM, DT, SPWT, mp.SCATTER_ATOMIC, steps)
        mv.sync()
        t1 = time.perf_counter()
        push_ms, launches = mv.timing()
        mv.set_timing(False)
        c = mv.counters()
        opts = mv.options
    wall = t1 - t0
    return {"value": c["particle_steps"] / wall, "ms_per_step": 1e3 * wall / steps, "steps": steps,
            "push_ms": push_ms, "push_ms_per_step": push_ms / steps, "launches": launches,
            "steps_per_launch": steps / max(launches, 1), "particle_steps": c["particle_steps"],
            "counters": c, "options": opts, "tuned": tuned}


def run_gpu(mesh, ef, pos, vel, cell, steps, warmup, mode, sort_interval, agg=2, shift=0, block=256, ppt=1,
            tune=False, pid_order=0):
    # Synchronisation goes through the C ABI (hipStreamSynchronize on the context stream,
    # the only stream the mover uses); torch is not loaded: it bundles a second HIP runtime.
    tuned = None
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=sort_interval, scatter_agg=agg, sort_shift=shift, block_size=block,
                       particles_per_thread=ppt, keep_pid_order=pid_order)
        mv.upload_field(ef)
        mv.set_particles(pos, vel, cell)
        if tune:
            # untimed: the autotuner restores the particle state it started from
            t = time.perf_counter()
            best, best_ms, trials = mv.autotune(QM, DT, SPWT, mode, min_steps=10)
            tuned = {"options": best, "ms_per_step": best_ms, "trials": len(trials),
                     "tuner_s": time.perf_counter() - t}
            log(f"autotune ({len(trials)} trials, {tuned['tuner_s']:.1f} s): {best} at {best_ms:.3f} ms/step")
        for _ in range(warmup):
            mv.step(QM, DT, SPWT, mode)
        mv.sync()
        mv.reset_counters()
        mv.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            mv.step(QM, DT, SPWT, mode)
        mv.sync()
        t1 = time.perf_counter()
        push_ms, launches = mv.timing()
        mv.set_timing(False)
        c = mv.counters()
        opts = mv.options
    wall = t1 - t0
    return {"wall_s": wall, "particle_steps": c["particle_steps"], "rate": c["particle_steps"] / wall,
            "push_ms": push_ms, "launches": launches, "counters": c, "options": opts, "tuned": tuned}


def run_ensemble(steps, warmup, n_runs=64, seed=SEED + 3):
    """Batched ensemble: n_runs independent runs on a 40^3 six-tet box (side 1 m) under the
    frozen sinusoidal field; per-run counts log-uniform in [1e5, 4e6], drift U[2, 8] km/s,
    thermal sigma U[500, 3000] m/s.  One graph-captured launch per step for all runs versus
    one launch sequence per run."""
    mesh = mp.Mesh.box_mesh(40, 1.0)
    ef = mesh.field_sinusoid(AMP)
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(1e5), np.log(4e6), n_runs))).astype(np.int64)
    drifts = rng.uniform(2000.0, 8000.0, n_runs)
    sigmas = rng.uniform(500.0, 3000.0, n_runs)
    dt = 4.0e-7
    runs = []
    for r in range(n_runs):
        pos, vel = mp.synth_uniform(seed * 1000 + r, int(counts[r]), [0, 0, 0], [1, 1, 1], sigmas[r],
                                    [drifts[r], 0, 0])
        runs.append((pos, vel, mesh.locate_box(pos)))
    out = {}
    with mp.Mover.from_mesh(mesh) as mv:
        # buckets of 16 tets along the Morton curve of the tet centroids: at a few particles per
        # tet a workgroup's particles then share nodes (31.4 against 29 G/s in cell-id order for
        # the multi-step line on 8-tet buckets every 10 steps); sorting every 12 steps on 16-tet
        # buckets: 32.4-32.6 G/s over 100 steps, the batched and sequential lines unchanged
        mv.set_options(sort_interval=12, sort_shift=4, sort_order=1)
        mv.upload_field(ef)
        for how in ("batched", "sequential", "multi_step"):
            with mp.Batch(mv, n_runs) as b:
                for r, (pos, vel, cell) in enumerate(runs):
                    b.set_run(r, pos, vel, cell, QM, dt, SPWT)
                if how == "multi_step":
                    b.step_many(warmup)
                else:
                    go = b.step if how == "batched" else b.step_sequential
                    for _ in range(warmup):
                        go()
                b.sync()
                b.reset_counters()
                t0 = time.perf_counter()
                if how == "multi_step":
                    b.step_many(steps)
                else:
                    for _ in range(steps):
                        go()
                b.sync()
                wall = time.perf_counter() - t0
                c = b.counters()
            out[how] = {"value": c["particle_steps"] / wall, "ms_per_step": 1e3 * wall / steps}
            log(f"ensemble {how}: {out[how]['value'] / 1e9:.3f} G particle-steps/s, {out[how]['ms_per_step']:.3f} ms/step")
    return {"value": out["batched"]["value"], "unit": "particle-steps/s",
            "ms_per_step": out["batched"]["ms_per_step"],
            "sequential_value": out["sequential"]["value"],
            "sequential_ms_per_step": out["sequential"]["ms_per_step"],
            "speedup_vs_sequential": out["batched"]["value"] / out["sequential"]["value"],
            "multi_step_value": out["multi_step"]["value"],
            "multi_step_ms_per_step": out["multi_step"]["ms_per_step"],
            "multi_step_speedup_vs_sequential": out["multi_step"]["value"] / out["sequential"]["value"],
            "steps": steps, "runs": n_runs, "particles": int(counts.sum()),
            "workload": "64 runs on a 40^3 six-tet box (384,000 tets), counts log-uniform 1e5..4e6, "
                        "drift U[2,8] km/s, sigma U[500,3000] m/s, frozen sinusoidal field, dt 4e-7 s, "
                        "cell sort every 12 steps on the Morton rank of the tet >> 4; "
                        "one hipGraph launch per step vs 64 per-run launch sequences; multi_step: "
                        "mp_batch_step_many, one launch per sort interval (12 steps) with the particles in "
                        "registers"}


def run_contention(steps, n=200_000_000, warmup=2, seed=SEED + 5, shift=3, order=0, pid_bits=8, det_shift=5,
                   det_order=1):
    """Contention stress test: a 100^3 six-tet box (side 1 m, 6,000,000 tets, 1,030,301 nodes),
    n particles from a 3-Gaussian mixture (sigma 0.05 m, weights 0.5/0.3/0.2, seeded centres in
    [0.3, 0.7]^3) so a few thousand nodes take most of the scatter, Maxwellian velocities
    (sigma 3000 m/s, no drift), frozen sinusoidal field, reflecting walls, cell sort every 8
    steps.  The same initial state is run in deterministic mode (the configuration's mode, its
    cell sort keyed by bucket and the top pid_bits of a pid digit, so that the dense nodes' pid
    slabs stay few per tile; det_algo 2: after an exact first step the push places each record
    into its pid slab's segment, sized from the last step's counts; its cell sort on buckets of
    2^det_shift tets along det_order, wider buckets cheapening the sort more than they slow the
    push) and in atomic mode (buckets of 2^shift tets along order); the difference of the two
    step times is the cost of the sorted reduce.  A third run keeps the particles in pid order (compaction
    instead of the cell sort) with the device radix sort of pid-ordered records (det_algo 0)."""
    t = time.perf_counter()
    mesh = mp.Mesh.box_mesh(100, 1.0)
    ef = mesh.field_sinusoid(AMP)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.3, 0.7, (3, 3))
    pos, vel = mp.synth_mixture(seed, n, [0, 0, 0], [1, 1, 1], [0.5, 0.3, 0.2], centres, 0.05, 3000.0,
                                [0, 0, 0])
    cell = mesh.locate_box(pos)
    log(f"contention setup {time.perf_counter() - t:.1f}s: {mesh.n_tets} tets, {n} particles")
    dt = 1.7e-7
    res = {}
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=8, wall_mode=mp.WALL_REFLECT, sort_shift=shift, sort_order=order)
        mv.upload_field(ef)
        for name, mode, pid_order in (("deterministic", mp.SCATTER_DETERMINISTIC, 0), ("atomic", mp.SCATTER_ATOMIC, 0),
                                      ("deterministic_pid_order", mp.SCATTER_DETERMINISTIC, 1)):
            cell_det = mode == mp.SCATTER_DETERMINISTIC and not pid_order
            mv.set_options(keep_pid_order=pid_order, det_algo=0 if pid_order else 2,
                           sort_pid_bits=pid_bits if cell_det else 0, sort_shift=det_shift if cell_det else shift,
                           sort_order=det_order if cell_det else order)
            mv.set_particles(pos, vel, cell)
            for _ in range(warmup):
                mv.step(QM, dt, SPWT, mode)
            mv.sync()
            mv.reset_counters()
            mv.set_timing(True)
            t0 = time.perf_counter()
            for _ in range(steps):
                mv.step(QM, dt, SPWT, mode)
            mv.sync()
            wall = time.perf_counter() - t0
            push_ms, launches = mv.timing()
            mv.set_timing(False)
            c = mv.counters()
            # push-kernel algorithmic bytes; the deterministic push also writes 4 (u32 node,
            # f64 weight) records per particle
            per_part = 104.0 + (48.0 if mode == mp.SCATTER_DETERMINISTIC else 0.0)
            avg_s = push_ms / 1e3 / max(launches, 1)
            bpl = (per_part * c["particle_steps"] + launches * (192.0 * mesh.n_tets + 16.0 * mesh.n_nodes)) / max(launches, 1)
            res[name] = {"value": c["particle_steps"] / wall, "ms_per_step": 1e3 * wall / steps,
                         "push_ms_per_launch": avg_s * 1e3, "push_gbs": bpl / avg_s / 1e9,
                         "crossing_fraction": c["crossed"] / max(c["particle_steps"], 1),
                         "reflections_per_step": c["reflections"] / steps}
            if mode == mp.SCATTER_DETERMINISTIC:
                placed, rebuilt = mv.det_paths()
                res[name]["steps_placed"], res[name]["steps_rebuilt"] = placed, rebuilt
            log(f"contention {name}: {res[name]['value'] / 1e9:.3f} G particle-steps/s, "
                f"{res[name]['ms_per_step']:.2f} ms/step, push {avg_s * 1e3:.3f} ms ({res[name]['push_gbs']:.0f} GB/s)")
    del pos, vel, cell
    d, a = res["deterministic"], res["atomic"]
    # both deterministic layouts give the same bits: the faster one is the line's value
    best = "deterministic" if d["value"] >= res["deterministic_pid_order"]["value"] else "deterministic_pid_order"
    b = res[best]
    return {"value": b["value"], "unit": "particle-steps/s", "mode": "deterministic", "layout": best, "steps": steps,
            "ms_per_step": b["ms_per_step"], "push_gbs": b["push_gbs"], "atomic": a,
            "deterministic": d, "sorted_reduce_ms_per_step": b["ms_per_step"] - a["ms_per_step"],
            "deterministic_pid_order": res["deterministic_pid_order"],
            "workload": f"100^3 six-tet box (6,000,000 tets, 1,030,301 nodes), {n} particles from a "
                        "3-Gaussian mixture (sigma 0.05 m, weights 0.5/0.3/0.2), Maxwellian sigma 3000 m/s, "
                        f"no drift, reflecting walls, sort every 8 steps (shift {shift}, order {order}; "
                        f"deterministic: shift {det_shift}, order {det_order}, sort_pid_bits {pid_bits}), "
                        f"dt {dt:g} s"}


ION_TEMPERATURE_1EV = 11604.518121745585  # kelvin: 1 eV / k_B


def coupled_loop(mesh, p, steps, resident):
    """One coupled run of `steps` steps with the ions moved by the drop-in GPU mover (host
    particle array) or kept resident on the device; times per phase."""
    import coupled as cp

    mover = (cp.ResidentGpuMover if resident else cp.GpuMover)(mesh, QM, p.dt, 2e2, mp.SCATTER_DETERMINISTIC)
    try:
        run = cp.CoupledRun(mesh, p, move=mover, fe_flags=mp.FE_RESET_F1)
        t_inj = t_move = t_fe = 0.0
        psteps = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            run.inject()
            b = time.perf_counter()
            psteps += run.n_parts
            run.move_step()
            c = time.perf_counter()
            run.solve()
            t_inj, t_move, t_fe = t_inj + b - a, t_move + c - b, t_fe + time.perf_counter() - c
        wall = time.perf_counter() - t0
        return {"ms_per_step": 1e3 * wall / steps, "inject_ms_per_step": 1e3 * t_inj / steps,
                "move_ms_per_step": 1e3 * t_move / steps, "fe_ms_per_step": 1e3 * t_fe / steps,
                "move_particle_steps_per_s": psteps / t_move, "particles_final": run.n_parts}
    finally:
        mover.close()


def run_coupled(steps=200):
    """mini-pic's coupled loop on a 12^3 six-tet box of side 0.1 m (10,368 tets, 2,197 nodes):
    O+ ions injected at 7 km/s through the x = 0 face (n = 1e10 m^-3, spwt 200: 1750 a step),
    Boltzmann electrons at Te = 2 eV, dt = 5e-7 s; each step InjectIons (host) -> MoveParticles
    (GPU, deterministic scatter) -> FE Poisson-Boltzmann (host, Newton + Gauss-Seidel, F1
    cleared per solve) -> updateEf.  Run with the reference's cold beam (the top-level fields)
    and with the workload's 1 eV Maxwellian spread on the drift ("thermal_1eV"), each with the
    drop-in mover and the device-resident one.  Also runs the reference's own solver setting
    (F1 accumulating) until its Newton iteration stops converging, as the reference's would."""
    import tempfile

    import coupled as cp

    with tempfile.TemporaryDirectory() as d:
        mesh = cp.box_case(12, 0.1, d)
    out = {"steps": steps}
    for name, temp in (("cold_beam", 0.0), ("thermal_1eV", ION_TEMPERATURE_1EV)):
        p = cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036, wall_potential=0.0,
                          dt=5e-7, max_iter=steps, ion_temperature=temp)
        r = coupled_loop(mesh, p, steps, resident=False)
        rr = coupled_loop(mesh, p, steps, resident=True)
        r["resident_ms_per_step"] = rr["ms_per_step"]
        r["resident_move_ms_per_step"] = rr["move_ms_per_step"]
        r["ion_temperature_K"] = temp
        if name == "cold_beam":
            out.update(r)
        else:
            out[name] = r
        log(f"coupled ({name}): {r['ms_per_step']:.2f} ms/step (move {r['move_ms_per_step']:.2f}, FE "
            f"{r['fe_ms_per_step']:.2f}), resident {r['resident_ms_per_step']:.2f} ms/step (move "
            f"{r['resident_move_ms_per_step']:.2f}), {r['particles_final']} ions at the end")
    p = cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036, wall_potential=0.0,
                      dt=5e-7, max_iter=steps)
    gpu = cp.GpuMover(mesh, QM, p.dt, 2e2, mp.SCATTER_DETERMINISTIC)
    try:
        run = cp.CoupledRun(mesh, p, move=gpu)
        fail = None
        for ts in range(steps):
            try:
                run.step()
            except mp.MpError:
                fail = ts
                break
        out["reference_solver_newton_fails_at_step"] = fail
    finally:
        gpu.close()
    log(f"coupled: reference-F1 Newton fails at step {out['reference_solver_newton_fails_at_step']}")
    out["workload"] = ("12^3 six-tet box, side 0.1 m, O+ beam 7 km/s from x = 0 (cold, and with a 1 eV Maxwellian "
                       "spread), n 1e10 m^-3, Te 2 eV, dt 5e-7 s, GPU MoveParticles (deterministic) + host FE "
                       "Newton/Gauss-Seidel, F1 cleared per solve")
    return out


def run_adjoint(steps=50, n=2_000_000, seed=SEED + 4, reps=3):
    """Adjoint: 2M particles on a 30^3 six-tet box (side 1 m), frozen sinusoidal field, K = 50
    steps, checkpoint every 10; L = |den_K - den_target|^2 with the target from a forward run
    at 1.05 x drift.  Compares mp_adjoint (forward with checkpoints + replayed reverse sweep)
    with the plain K-step forward (upload + K atomic steps) on the same particles."""
    mesh = mp.Mesh.box_mesh(30, 1.0)
    ef = mesh.field_sinusoid(AMP)
    pos, u = mp.synth_uniform(seed, n, [0, 0, 0], [1, 1, 1], SIGMA, [0, 0, 0])
    cell = mesh.locate_box(pos)
    d = np.array([1.0, 0.0, 0.0])
    dt = 4.0e-7
    with mp.Mover.from_mesh(mesh) as mv:
        mv.set_options(sort_interval=steps, sort_shift=3)
        mv.upload_field(ef)

        def forward(drift):
            mv.set_particles(pos, u + drift * d, cell)
            for _ in range(steps):
                mv.step(QM, dt, SPWT, mp.SCATTER_ATOMIC)
            mv.sync()

        forward(1.05 * DRIFT)
        tgt = mv.density()
        fwd, adj, adj_det = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            forward(DRIFT)
            fwd.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            loss, g, gd = mv.adjoint(pos, u + DRIFT * d, cell, QM, dt, SPWT, steps, tgt, d, 10)
            adj.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            mv.adjoint(pos, u + DRIFT * d, cell, QM, dt, SPWT, steps, tgt, d, 10, mode=mp.SCATTER_DETERMINISTIC)
            adj_det.append(time.perf_counter() - t0)
        # the same with reflecting walls: the forward and the adjoint (which re-walks every step
        # from a stored start state to find the reflections)
        mv.set_options(wall_mode=1)
        fwd_r, adj_r = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            forward(DRIFT)
            fwd_r.append(time.perf_counter() - t0)
            refl = mv.counters()["reflections"]  # of the forward's 50 steps
            t0 = time.perf_counter()
            mv.adjoint(pos, u + DRIFT * d, cell, QM, dt, SPWT, steps, tgt, d, 10)
            adj_r.append(time.perf_counter() - t0)
    f, a, ad, fr, ar = min(fwd), min(adj), min(adj_det), min(fwd_r), min(adj_r)
    log(f"adjoint: forward {1e3 * f:.1f} ms, adjoint {1e3 * a:.1f} ms, ratio {a / f:.2f}; deterministic "
        f"{1e3 * ad:.1f} ms, ratio {ad / f:.2f}; reflecting walls: forward {1e3 * fr:.1f} ms, adjoint "
        f"{1e3 * ar:.1f} ms, ratio {ar / fr:.2f}")
    return {"forward_ms": 1e3 * f, "adjoint_ms": 1e3 * a, "ratio": a / f, "deterministic_adjoint_ms": 1e3 * ad,
            "deterministic_ratio": ad / f, "reflecting_forward_ms": 1e3 * fr, "reflecting_adjoint_ms": 1e3 * ar,
            "reflecting_ratio": ar / fr, "reflections_counted": int(refl), "steps": steps, "particles": n,
            "loss": loss, "grad_norm": float(np.linalg.norm(g)), "dL_ddrift": gd,
            "workload": "2M particles on a 30^3 six-tet box (162,000 tets), 50 steps, checkpoint every 10; both "
                        "timings include the particle upload (best of 3)"}


def run_cpu_baseline(mesh, ef, pos, vel, cell, budget_s, max_steps=30, sample=2_000_000, coupled_steps=0):
    """The oracle's restatement of the reference kernel body on the host: OpenMP over all
    threads (atomic density adds) on the full workload until the budget is spent; one thread
    on a sample; one thread running the whole MoveParticles over 88 B Particle records (copy
    in, kernel body, stable erase of the lost, normalisation, copy out) on the sample; and the
    12^3 coupled loop with that serial MoveParticles and the same host FE solver."""
    import oracle

    tets = mesh.tets
    p, v, c = pos.copy(), vel.copy(), cell.copy()
    den = np.zeros(mesh.n_nodes)
    done, steps, t0 = 0, 0, time.perf_counter()
    threads = 0
    while steps < max_steps:
        den[:] = 0
        _, n, threads = oracle.move_omp(tets, ef, p, v, c, QM, DT, mesh.n_nodes, den=den)
        done += n
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    out = {"value": done / wall, "steps": steps, "threads": threads, "wall_s": wall}
    # one thread, resident arrays
    p, v, c = pos[:sample].copy(), vel[:sample].copy(), cell[:sample].copy()
    t0 = time.perf_counter()
    n1 = 0
    for _ in range(2):
        n1 += int((c >= 0).sum())
        oracle.move(tets, ef, p, v, c, QM, DT, mesh.n_nodes)
    out["single_thread_value"] = n1 / (time.perf_counter() - t0)
    # one thread, the whole MoveParticles over Particle records
    mover = oracle.AosMover(tets, mesh.node_volume, QM, DT, SPWT)
    parts = mp.particles_from_arrays(pos[:sample], vel[:sample], cell[:sample])
    t0 = time.perf_counter()
    n2 = 0
    for _ in range(2):
        n2 += len(parts)
        dev = parts.copy()                     # the host -> device copy
        dev, den = mover(dev, ef)
        parts = dev.copy()                     # the device -> host copy
    out["full_moveparticles_value"] = n2 / (time.perf_counter() - t0)
    out["sample"] = sample
    if coupled_steps > 0:
        # the 12^3 coupled loop with the serial MoveParticles restatement on the host
        import tempfile

        import coupled as cp

        with tempfile.TemporaryDirectory() as d:
            cmesh = cp.box_case(12, 0.1, d)
        for key, temp in (("coupled", 0.0), ("coupled_thermal_1eV", ION_TEMPERATURE_1EV)):
            prm = cp.Parameters(plasma_den=1e10, ion_velocity=7000.0, electron_temperature=23209.036,
                                wall_potential=0.0, dt=5e-7, max_iter=coupled_steps, ion_temperature=temp)
            run = cp.CoupledRun(cmesh, prm, move=oracle.AosMover(cmesh.tets, cmesh.node_volume, QM, prm.dt, 2e2),
                                fe_flags=mp.FE_RESET_F1)
            t_move = 0.0
            t0 = time.perf_counter()
            for _ in range(coupled_steps):
                run.inject()
                t_move += run.move_step()[1]
                run.solve()
            out[f"{key}_ms_per_step"] = 1e3 * (time.perf_counter() - t0) / coupled_steps
            out[f"{key}_move_ms_per_step"] = 1e3 * t_move / coupled_steps
        out["coupled_steps"] = coupled_steps
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", choices=["atomic", "deterministic"], default="atomic")
    ap.add_argument("--sort-interval", type=int, default=10)
    ap.add_argument("--many-steps", type=int, default=None,
                    help="timed steps of the multi-step (mp_step_many) line (default: --steps; 0: skip)")
    ap.add_argument("--many-sort-interval", type=int, default=16,
                    help="multi-step line: steps between two sorts (one launch each), and the tuner's start")
    ap.add_argument("--headline", choices=["multi", "per-step"], default="multi",
                    help="which atomic-mode line is value/roofline (the other is reported beside it)")
    ap.add_argument("--det-steps", type=int, default=20, help="timed steps of the deterministic-mode line (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--particles", type=int, default=N_PART)
    ap.add_argument("--agg", type=int, default=2, help="atomic scatter aggregation: 0 one atomic per weight, 2 LDS hash per workgroup")
    ap.add_argument("--sort-shift", type=int, default=3, help="cell sort key = cell >> shift")
    ap.add_argument("--ensemble-steps", type=int, default=100, help="timed steps of the batched ensemble (0: skip)")
    ap.add_argument("--block", type=int, default=256, help="push workgroup size (64, 128, 256)")
    ap.add_argument("--ppt", type=int, default=1, help="particles per thread (1, 2, 4)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (use the options given)")
    ap.add_argument("--det-pid-order", type=int, default=0, help="deterministic line: keep pid order (1) or cell sort (0)")
    ap.add_argument("--gather-sort", type=int, default=0,
                    help="multi-step line: the sort writes a permutation that the next launch gathers through (1), "
                         "or copies the records (0)")
    ap.add_argument("--adjoint-steps", type=int, default=50, help="steps of the adjoint comparison (0: skip)")
    ap.add_argument("--contention-steps", type=int, default=50, help="timed steps of the 200M contention test (0: skip)")
    ap.add_argument("--contention-particles", type=int, default=200_000_000)
    ap.add_argument("--contention-pid-bits", type=int, default=8,
                    help="contention line, deterministic run: cell sort refined by this many pid-digit bits")
    ap.add_argument("--contention-sort-order", type=int, default=0,
                    help="contention line, atomic run: cell sort by cell id (0) or along the Morton curve (1)")
    ap.add_argument("--contention-det-shift", type=int, default=5,
                    help="contention line, deterministic run: cell sort buckets of 2^shift tets")
    ap.add_argument("--contention-det-order", type=int, default=1,
                    help="contention line, deterministic run: cell sort by cell id (0) or along the Morton curve (1)")
    ap.add_argument("--coupled-steps", type=int, default=200, help="steps of the coupled PIC loop (0: skip)")
    args = ap.parse_args()
    if args.many_steps is None:
        args.many_steps = args.steps
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the mover runs on a single "
              "MI355X; several GPUs would each take a slab of cell-sort buckets and exchange the particles "
              "that cross slab faces once per step)", file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    mesh = mp.Mesh.box_mesh(N_HEX, SIDE)
    ef = mesh.field_sinusoid(AMP)
    n = args.particles
    pos, vel = mp.synth_uniform(SEED, n, [0, 0, 0], [SIDE] * 3, SIGMA, [DRIFT, 0, 0])
    cell = mesh.locate_box(pos)
    log(f"setup {time.perf_counter() - t:.1f}s: {mesh.n_tets} tets, {mesh.n_nodes} nodes, {n} particles")

    mode = mp.SCATTER_ATOMIC if args.mode == "atomic" else mp.SCATTER_DETERMINISTIC
    r = run_gpu(mesh, ef, pos, vel, cell, args.steps, args.warmup, mode, args.sort_interval, args.agg,
                args.sort_shift, args.block, args.ppt, tune=not args.no_tune)
    log(f"gpu {args.mode}: {r['rate'] / 1e9:.3f} G particle-steps/s, push {r['push_ms'] / max(r['launches'], 1):.3f} ms/launch")
    avg_push_s = r["push_ms"] / 1e3 / max(r["launches"], 1)
    bytes_per_launch = push_bytes(r["particle_steps"], r["launches"], mesh.n_tets, mesh.n_nodes) / max(r["launches"], 1)
    achieved = bytes_per_launch / avg_push_s / 1e9
    per_step = {
        "value": r["rate"], "ms_per_step": 1e3 * r["wall_s"] / args.steps, "steps": args.steps,
        "path": "mp_step: one MoveParticles step per call (k_push_soa)", "options": r["options"],
        "autotune": r["tuned"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     **traffic_fields(measured_traffic(), avg_push_s, "k_push_soa_box60_20M_traffic.txt"),
                     "kernel": "k_push_soa", "avg_launch_ms": avg_push_s * 1e3,
                     "bytes_per_launch": bytes_per_launch}}
    head, roof, path, head_steps = r, per_step["roofline"], per_step["path"], args.steps
    extra = {}
    if args.many_steps > 0 and mode == mp.SCATTER_ATOMIC:
        m = run_gpu_many(mesh, ef, pos, vel, cell, args.many_steps, args.warmup,
                         dict(sort_interval=args.many_sort_interval, sort_shift=args.sort_shift, scatter_agg=args.agg,
                              gather_sort=args.gather_sort),
                         tune=not args.no_tune)
        log(f"gpu atomic multi-step: {m['value'] / 1e9:.3f} G particle-steps/s, push {m['push_ms_per_step']:.3f} "
            f"ms/step, {m['steps_per_launch']:.1f} steps per launch")
        m_launch_s = m["push_ms"] / 1e3 / max(m["launches"], 1)
        m_bytes = push_bytes(m["particle_steps"], m["steps"], mesh.n_tets, mesh.n_nodes) / max(m["launches"], 1)
        m_ach = m_bytes / m_launch_s / 1e9
        tps = measured_traffic("k_push_steps_box60_20M_traffic.txt", "traffic_bytes_per_step")
        multi = {
            "value": m["value"], "ms_per_step": m["ms_per_step"], "steps": m["steps"],
            "path": "mp_step_many: the steps between two periodic sorts in one launch (k_push_steps), particles and "
                    "their tet records held in registers, lanes stepping independently (frozen field)",
            "options": m["options"], "autotune": m["tuned"], "steps_per_launch": m["steps_per_launch"],
            "roofline": {"bound": "hbm", "achieved": m_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": m_ach / HBM_PEAK_GBS,
                         **traffic_fields(None if tps is None else tps * m["steps_per_launch"], m_launch_s,
                                          "k_push_steps_box60_20M_traffic.txt"),
                         "kernel": "k_push_steps", "avg_launch_ms": m_launch_s * 1e3,
                         "bytes_per_launch": m_bytes,
                         "note": "achieved and frac count the algorithmic bytes of every step the launch takes "
                                 "(an effective bandwidth: the state is read and written once per launch); "
                                 "measured_gbs / measured_frac are the HBM bytes the launch really moves"}}
        if args.headline == "multi":
            head = {"rate": m["value"], "wall_s": m["ms_per_step"] * m["steps"] / 1e3, "counters": m["counters"],
                    "options": m["options"], "tuned": m["tuned"]}
            roof, path, head_steps = multi["roofline"], multi["path"], m["steps"]
            extra["per_step"] = per_step
        else:
            extra["multi_step"] = multi
    if args.det_steps > 0 and mode == mp.SCATTER_ATOMIC:
        d = run_gpu(mesh, ef, pos, vel, cell, args.det_steps, 2, mp.SCATTER_DETERMINISTIC, args.sort_interval,
                    shift=args.sort_shift, tune=not args.no_tune, pid_order=args.det_pid_order)
        extra["deterministic"] = {"value": d["rate"], "ms_per_step": 1e3 * d["wall_s"] / args.det_steps,
                                  "steps": args.det_steps, "options": d["options"]}
        log(f"gpu deterministic: {d['rate'] / 1e9:.3f} G particle-steps/s")
    if args.ensemble_steps > 0:
        extra["batched"] = run_ensemble(args.ensemble_steps, 3)
    if args.adjoint_steps > 0:
        extra["adjoint"] = run_adjoint(args.adjoint_steps)
    if args.contention_steps > 0:
        extra["contention"] = run_contention(args.contention_steps, args.contention_particles,
                                             order=args.contention_sort_order, pid_bits=args.contention_pid_bits,
                                             det_shift=args.contention_det_shift,
                                             det_order=args.contention_det_order)
    if args.coupled_steps > 0:
        extra["coupled"] = run_coupled(args.coupled_steps)
    cpu = None
    if args.cpu_seconds > 0:
        cb = run_cpu_baseline(mesh, ef, pos, vel, cell, args.cpu_seconds, coupled_steps=args.coupled_steps)
        cpu = {"value": cb["value"], "unit": "particle-steps/s", "cores": cb["threads"], "kind": "port",
               "sample": f"the same 20M particles and mesh, {cb['steps']} steps from the same initial state "
                         f"({cb['wall_s']:.1f} s), OpenMP port of the reference kernel body with atomic "
                         f"density adds on {cb['threads']} threads of {cpu_model()}",
               "single_thread_value": cb["single_thread_value"],
               "full_moveparticles_1thread_value": cb["full_moveparticles_value"],
               "single_thread_sample": f"first {cb['sample']} particles, 2 steps; the full MoveParticles line adds "
                                       "the 2 x 88 B record copies, the stable erase and the normalisation"}
        if "coupled_ms_per_step" in cb:
            for k in ("coupled_ms_per_step", "coupled_move_ms_per_step", "coupled_thermal_1eV_ms_per_step",
                      "coupled_thermal_1eV_move_ms_per_step"):
                cpu[k] = cb[k]
            cpu["coupled_sample"] = (f"the 12^3 coupled loop, {cb['coupled_steps']} steps, with the oracle's serial "
                                     "MoveParticles (one thread) and the same host injection and FE solver, cold beam "
                                     "and 1 eV thermal injection")
            log(f"cpu coupled loop: {cb['coupled_ms_per_step']:.2f} ms/step (move {cb['coupled_move_ms_per_step']:.2f})")
        log(f"cpu baseline: {cb['value'] / 1e6:.2f} M particle-steps/s on {cb['threads']} threads")
    c = head["counters"]
    out = {
        "metric": "FP64 particle-steps/s (push+tet-walk+scatter)",
        "value": head["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": head_steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["wall_s"] / head_steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 + Box-Muller on the host)",
        "config": {"workload": "push-only scale test: 60^3 six-tet box, side 1 m (1,296,000 tets, 226,981 nodes), "
                               "20M FP64 particles, frozen sinusoidal field, absorbing walls",
                   "particles": n, "tets": mesh.n_tets, "nodes": mesh.n_nodes, "dt": DT,
                   "scatter": args.mode, "path": path, "options": head["options"], "autotune": head["tuned"],
                   "crossing_fraction": c["crossed"] / max(c["particle_steps"], 1),
                   "lost": c["lost"]},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
