"""mini-pic's coupled PIC loop around the MI355X mover.

Each step follows src/mini-pic.cpp:124-167:
  InjectIons (host, src/mini-pic.cpp:208-265) -> MoveParticles (the GPU mover, drop-in
  mp_step_aos with the field uploaded every step) -> FESolver::computePhi (host FE
  Poisson-Boltzmann) -> FESolver::updateEf.
Set-up follows src/mini-pic.cpp:62-118: the parameter file (src/parameters.cpp), the volume
and surface meshes, phi0 = 0, n0 = plasma_den, kTe = Kb * electron_temperature, and ions of
charge +1 e, mass 16 amu, specific weight 200.

The mover is pluggable (`move(parts, ef) -> (kept parts, den)`), so the same loop drives
the GPU path or any other implementation of MoveParticles on identical inputs.
"""
import os
import re
import time
from dataclasses import dataclass, field

import numpy as np

import minipic_amd as mp


@dataclass
class Parameters:
    """Parameters (src/parameters.h:18-38) with the reference's defaults, plus ion_temperature
    (kelvin; an extension, key `ion_temperature`): 0 injects the reference's cold beam, T > 0
    adds a Maxwellian of per-component deviation sqrt(kT/m) to it."""
    plasma_den: float = 1e20
    ion_velocity: float = 7000.0
    electron_temperature: float = 24000.0
    max_iter: int = 250
    wall_potential: float = 100.0
    dt: float = 1e-7
    mesh_files: dict = field(default_factory=dict)
    invert_normals: bool = False
    plasma_species: str = "Deuterium"
    fesolver_method: int = mp.FE_NONLINEAR
    ion_temperature: float = 0.0

    @classmethod
    def from_file(cls, path):
        """Parameters::Parameters (src/parameters.cpp:16-93): `key = value` lines, `#` comments."""
        p = cls()
        base = os.path.dirname(os.path.abspath(path))
        with open(path) as f:
            for line in f:
                line = re.sub(r"\s*(#.*)?$", "", line.rstrip("\r\n"))
                if not line:
                    continue
                m = re.fullmatch(r"(\w+)\s*=\s*(.*)", line)
                if not m:
                    raise ValueError(f"Invalid input line: '{line}'")
                key, value = m.group(1), m.group(2)
                if key in ("plasma_den", "ion_velocity", "electron_temperature", "wall_potential", "dt",
                           "ion_temperature"):
                    setattr(p, key, float(value))
                elif key == "max_iter":
                    p.max_iter = int(value)
                elif key in ("global_mesh", "inlet_mesh", "wall_mesh"):
                    p.mesh_files[key] = value if os.path.isabs(value) else os.path.join(base, value)
                elif key == "invert_normals":
                    if re.fullmatch(r"t(rue)?", value, re.I) or value == "1":
                        p.invert_normals = True
                    elif re.fullmatch(r"f(alse)?", value, re.I) or value == "0":
                        p.invert_normals = False
                elif key == "plasma_species":
                    if re.fullmatch(r"D(uterium)?", value):
                        p.plasma_species = "Deuterium"
                    elif re.fullmatch(r"O(xygen)?", value):
                        p.plasma_species = "Oxygen"
                    else:
                        raise ValueError(f"Invalid plasma species: '{value}'")
                elif key == "fesolver_method":
                    methods = {"nonlinear": mp.FE_NONLINEAR, "gaussseidel": mp.FE_GAUSS_SEIDEL,
                               "lapack": mp.FE_DIRECT}
                    if value.lower() not in methods:
                        raise ValueError(f"Invalid fesolver_method: '{value}'")
                    p.fesolver_method = methods[value.lower()]
                else:
                    raise ValueError(f"Unrecognised parameter: '{key}'")
        return p


def write_box_face(mesh, path, side="xmin"):
    """A surface mesh (.dat, as LoadSurfaceMesh reads it) of the x = 0 ("xmin") or x = Lx
    ("xmax") face of a box mesh: the 2 * ny * nz triangles of the six-tet split.  On the x = 0
    face they are wound so that the inlet normal -(u x v) points into the box (+x)."""
    nx, ny, nz = mesh.box[:3]
    i = 0 if side == "xmin" else nx
    nid = lambda j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    nodes = sorted({nid(j, k) for j in range(ny + 1) for k in range(nz + 1)})
    pos = mesh.node_pos
    tris = []
    for k in range(nz):
        for j in range(ny):
            c0, cz, cy, cyz = nid(j, k), nid(j, k + 1), nid(j + 1, k), nid(j + 1, k + 1)
            # the square's diagonal runs c0 -> cyz (the tets share the hex's main diagonal)
            tris.append((c0, cz, cyz))
            tris.append((c0, cyz, cy))
    with open(path, "w") as f:
        f.write(f"{len(nodes)} {len(tris)}\n")
        for n in nodes:
            x, y, z = pos[n] * 1000.0
            f.write(f"{n + 1} {x:.15e} {y:.15e} {z:.15e}\n")
        for e, (a, b, c) in enumerate(tris):
            f.write(f"{e + 1} 203 {a + 1} {b + 1} {c + 1}\n")


def box_case(n_hex, side, workdir, wall=False):
    """A box mesh with its x = 0 face as the inlet (and, with wall=True, its x = Lx face as the
    FIXED wall); the other boundary nodes stay OPEN."""
    mesh = mp.Mesh.box_mesh(n_hex, side)
    inlet = os.path.join(workdir, "inlet.dat")
    write_box_face(mesh, inlet, "xmin")
    mesh.load_surface(inlet, mp.NODE_INLET)
    if wall:
        w = os.path.join(workdir, "wall.dat")
        write_box_face(mesh, w, "xmax")
        mesh.load_surface(w, mp.NODE_FIXED)
    return mesh


class Trace:
    """Per-phase wall-time profile in the reference's trace.csv format (src/trace.cpp:36-78):
    function_name, total_time, exclusive_time, exclusive_proportion, number_of_calls, then the
    traceEnter / traceExit / traceTotalTime rows; numbers as the C++ stream writes them.  The
    reference keeps rows in a std::unordered_map (src/trace.h:44), so its row order is
    unspecified; here they come in first-call order."""

    def __init__(self):
        self.start = time.perf_counter()
        self.funcs = {}
        self.stack = []
        self.entry_time = self.exit_time = 0.0

    def enter(self, name):
        t = time.perf_counter()
        f = self.funcs.setdefault(name, {"calls": 0, "total": 0.0, "child": 0.0, "start": 0.0})
        f["calls"] += 1
        f["start"] = t
        self.stack.append(name)
        self.entry_time += time.perf_counter() - t

    def exit(self, name):
        t = time.perf_counter()
        f = self.funcs[name]
        dt = t - f["start"]
        f["total"] += dt
        f["start"] = 0.0
        self.stack.pop()
        if self.stack:
            self.funcs[self.stack[-1]]["child"] += dt
        self.exit_time += time.perf_counter() - t

    def phase(self, name):
        tr = self

        class _Phase:
            def __enter__(self):
                tr.enter(name)

            def __exit__(self, *exc):
                tr.exit(name)

        return _Phase()

    def write_profile(self, path, extra=None):
        """trace.csv with the phases timed here plus `extra` rows {name: (calls, total, child)}
        timed elsewhere (the FE solver's methods), rows with no calls left out."""
        calc = time.perf_counter() - self.start
        g = lambda v: "%g" % v  # noqa: E731
        rows = {name: (e["calls"], e["total"], e["child"]) for name, e in self.funcs.items() if not e["start"]}
        for name, r in (extra or {}).items():
            if r[0]:
                rows[name] = r
        with open(path, "w") as f:
            f.write("function_name,total_time,exclusive_time,exclusive_proportion,number_of_calls\n")
            for name, (calls, total, child) in rows.items():
                excl = total - child
                f.write(f"{name},{g(total)},{g(excl)},{g(excl / calc)},{calls}\n")
            f.write(f"traceEnter,{g(self.entry_time)},{g(self.entry_time)},{g(self.entry_time / calc)},1\n")
            f.write(f"traceExit,{g(self.exit_time)},{g(self.exit_time)},{g(self.entry_time / calc)},1\n")
            f.write(f"traceTotalTime,{g(calc)},0,0,1\n")


class GpuMover:
    """MoveParticles on the MI355X (drop-in mp_step_aos: field copied every step, survivors
    compacted in order, den normalised)."""

    def __init__(self, mesh, q_over_m, dt, spwt, mode=mp.SCATTER_DETERMINISTIC):
        self.mv = mp.Mover.from_mesh(mesh)
        self.q_over_m, self.dt, self.spwt, self.mode = q_over_m, dt, spwt, mode

    def __call__(self, parts, ef):
        self.mv.upload_field(ef)
        kept, den = self.mv.step_aos(parts, self.q_over_m, self.dt, self.spwt, self.mode)
        return parts[:kept], den

    def close(self):
        self.mv.close()


class ResidentGpuMover:
    """MoveParticles with the ions kept in HBM between steps: each step only the newly
    injected ions cross PCIe (mp_add_particles, which gives them the next pids, so the
    deterministic density keeps the serial order), the field is uploaded and the density
    read back."""

    resident = True

    def __init__(self, mesh, q_over_m, dt, spwt, mode=mp.SCATTER_DETERMINISTIC, sort_interval=10):
        self.mv = mp.Mover.from_mesh(mesh)
        self.mv.set_options(sort_interval=sort_interval)
        self.mv.set_particles(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
        self.q_over_m, self.dt, self.spwt, self.mode = q_over_m, dt, spwt, mode

    def __call__(self, new_parts, ef):
        if len(new_parts):
            self.mv.add_particles(new_parts["pos"], new_parts["vel"], new_parts["cell_index"])
        self.mv.upload_field(ef)
        self.mv.step(self.q_over_m, self.dt, self.spwt, self.mode)
        return self.mv.num_active(), self.mv.density()

    def particles(self):
        """(pos, vel, cell, pid) of the ions in the domain, ascending pid."""
        return self.mv.get_particles()

    def close(self):
        self.mv.close()


class CoupledRun:
    """The main loop's state: particles, carried injection fraction, field and solver.  With a
    resident mover the particles live on the device and `parts` holds only the ions injected
    this step."""

    def __init__(self, mesh, params, move=None, mode=mp.SCATTER_DETERMINISTIC, seed=0, spwt=2e2,
                 charge=mp.QE, mass=16 * mp.AMU, fe_flags=0, output_dir=None):
        self.mesh, self.params = mesh, params
        self.output_dir = output_dir
        self.trace = Trace()
        self.spwt = spwt
        self.q_over_m = charge / mass
        self.v_thermal = mp.thermal_speed(params.ion_temperature, mass) if params.ion_temperature > 0 else 0.0
        self.fe = mp.FESolver(mesh, n0=params.plasma_den, phi0=0.0, kTe=mp.KB * params.electron_temperature,
                              wall_potential=params.wall_potential, flags=fe_flags)
        self.rng = mp.Rng(seed)
        self.move = move if move is not None else GpuMover(mesh, self.q_over_m, params.dt, spwt, mode)
        self.resident = getattr(self.move, "resident", False)
        self.parts = np.zeros(0, mp.PARTICLE_DTYPE)
        self._buf = None  # host particle storage with spare capacity (a std::vector's growth)
        self.n_parts = 0
        self.rem = 0.0
        self.ef = np.zeros((mesh.n_tets, 3))  # zero until the first solve
        self.den = np.zeros(mesh.n_nodes)
        self.ts = 0

    def inject(self):
        p = self.params
        new, self.rem = mp.inject_ions(self.mesh, self.rng, p.plasma_den, p.ion_velocity, p.dt, self.spwt,
                                       self.q_over_m, self.ef, self.rem, self.v_thermal)
        if self.resident:
            self.parts = new
        else:
            # append in amortised O(1), as the reference's vector insert does: the survivors are
            # a prefix view of the buffer after the move (it compacts in place)
            n, k = len(self.parts), len(new)
            buf = self._buf
            if buf is None or n + k > len(buf):
                buf = np.empty(max(2 * (n + k), 1024), mp.PARTICLE_DTYPE)
                buf[:n] = self.parts
                self._buf = buf
            elif n and self.parts.ctypes.data != buf.ctypes.data:
                buf[:n] = self.parts
            buf[n:n + k] = new
            self.parts = buf[:n + k]
        self.n_parts += len(new)
        return len(new)

    def move_step(self):
        old = self.n_parts
        t0 = time.perf_counter()
        if self.resident:
            self.n_parts, self.den = self.move(self.parts, self.ef)
        else:
            self.parts, self.den = self.move(self.parts, self.ef)
            self.n_parts = len(self.parts)
        return old, time.perf_counter() - t0

    def solve(self, den=None):
        st = self.fe.compute_phi(self.den if den is None else den, self.params.fesolver_method)
        self.ef = self.fe.update_ef()
        return st

    def step(self):
        """One time step; returns the reference's per-step report as a dict."""
        with self.trace.phase("InjectIons"):
            added = self.inject()
        with self.trace.phase("MoveParticles"):
            old, duration = self.move_step()
        # computePhi and updateEf, with the solver's inner methods, are timed by the solver itself
        st = self.fe.compute_phi(self.den, self.params.fesolver_method)
        self.ef = self.fe.update_ef()
        phi = self.fe.phi
        rec = {"ts": self.ts, "np": self.n_parts, "added": added, "removed": old - self.n_parts,
               "max_den": float(max(self.den.max(initial=0.0), 0.0)), "max_phi": float(np.abs(phi).max()),
               "move_s": duration, "nr_iterations": st["nr_iterations"], "gs_sweeps": st["gs_sweeps"]}
        if self.output_dir and (self.ts + 1) % 10 == 0:
            with self.trace.phase("OutputMesh"):
                self.write_mesh(self.ts)
        self.ts += 1
        return rec

    def write_mesh(self, ts):
        """OutputMesh(ts, ...): mesh_{ts+1:04d}.vtu in the output directory."""
        path = os.path.join(self.output_dir, f"mesh_{ts + 1:04d}.vtu")
        mp.write_mesh_vtu(self.mesh, path, self.fe.phi, self.ef, self.den)
        return path

    def finish(self):
        """The end of main (src/mini-pic.cpp:160-164): the last mesh file, particles.vtp and
        the phase profile trace.csv."""
        if not self.output_dir:
            return
        with self.trace.phase("OutputMesh"):
            self.write_mesh(self.ts)
        with self.trace.phase("OutputParticles"):
            pos = self.move.particles()[0] if self.resident else self.parts
            mp.write_particles_vtp(os.path.join(self.output_dir, "particles.vtp"), pos)
        self.trace.write_profile(os.path.join(self.output_dir, "trace.csv"), self.fe.profile())
