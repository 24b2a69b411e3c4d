"""MI355X RKDG shallow-water stage engine: Python host side.

Mirrors the reference's driver surface (Serial::Simulation<SWE::RKDG::Problem>,
source/simulation/serial/simulation.hpp:14-98): a Simulation is built from a mesh and the
problem inputs, steps with SSP-RK (Run/step_serial), exposes the modal state and reports the
L2 residual (ComputeL2Residual).  All numerics run in librkdg.so (HIP kernels + C++ host);
this module only marshals arrays through the C-ABI of include/rkdg.h.
"""
import ctypes as C
import math

import numpy as np

from . import _native as N
from ._native import (BC_FLOW, BC_FUNCTION, BC_LAND, BC_LEVEE, BC_TIDE, FRICTION_CHEZY, FRICTION_MANNING,
                      FRICTION_NONE, PROBLEM_DAM_BREAK, PROBLEM_MANUFACTURED, PROBLEM_NONE,
                      PROBLEM_PARABOLIC_BOWL, RKDG_FP32_BULK, RKDG_FP64_FAST, RKDG_FP64_PLAIN, RkdgError)

__all__ = [
    "Mesh", "quad_mesh", "read_adcirc", "write_adcirc", "read_meta", "write_meta", "partition_rcb", "submesh", "renumber", "PartitionedSimulation", "make_params", "Simulation", "Batch", "MlpTrainer", "RkdgError", "ndof_of",
    "BC_LAND", "BC_TIDE", "BC_FLOW", "BC_FUNCTION", "BC_LEVEE",
    "PROBLEM_NONE", "PROBLEM_MANUFACTURED", "PROBLEM_PARABOLIC_BOWL", "PROBLEM_DAM_BREAK",
    "FRICTION_NONE", "FRICTION_CHEZY", "FRICTION_MANNING", "RKDG_FP64_PLAIN", "RKDG_FP32_BULK", "RKDG_FP64_FAST",
]


def ndof_of(p):
    return (p + 1) * (p + 2) // 2


class Mesh:
    """ADCIRC-style mesh: nodes (x, y, depth), triangles, typed boundary edges, forcing."""

    def __init__(self, node_xyz, elt_nodes, bound_edges, bound_types, frequency=(), forcing_fact=(),
                 equilib_arg=(), node_amplitude=None, node_phase=None, node_manning_n=None, levee_nodes=None,
                 levee_param=None, _owner=None):
        self.node_xyz = np.ascontiguousarray(node_xyz, dtype=np.float64).reshape(-1, 3)
        self.elt_nodes = np.ascontiguousarray(elt_nodes, dtype=np.int32).reshape(-1, 3)
        self.bound_edges = np.ascontiguousarray(bound_edges, dtype=np.int32).reshape(-1, 2)
        self.bound_types = np.ascontiguousarray(bound_types, dtype=np.uint8).reshape(-1)
        self.frequency = np.ascontiguousarray(frequency, dtype=np.float64)
        self.forcing_fact = np.ascontiguousarray(forcing_fact, dtype=np.float64)
        self.equilib_arg = np.ascontiguousarray(equilib_arg, dtype=np.float64)
        nc = len(self.frequency)
        self.node_amplitude = None if node_amplitude is None else np.ascontiguousarray(node_amplitude, np.float64)
        self.node_phase = None if node_phase is None else np.ascontiguousarray(node_phase, np.float64)
        self.node_manning_n = None if node_manning_n is None else np.ascontiguousarray(node_manning_n, np.float64)
        if nc and (self.node_amplitude is None or self.node_phase is None):
            raise ValueError("constituents given without nodal amplitude/phase")
        self.levee_nodes = np.zeros((0, 2), np.int32) if levee_nodes is None else \
            np.ascontiguousarray(levee_nodes, dtype=np.int32).reshape(-1, 2)
        self.levee_param = np.zeros((0, 3), np.float64) if levee_param is None else \
            np.ascontiguousarray(levee_param, dtype=np.float64).reshape(-1, 3)
        if len(self.levee_nodes) != len(self.levee_param):
            raise ValueError("levee_nodes and levee_param differ in length")
        self._owner = _owner
        self.desc = N.MeshDesc()
        d = self.desc
        d.n_nodes, d.n_elements = len(self.node_xyz), len(self.elt_nodes)
        d.node_xyz, d.elt_nodes = N.dptr(self.node_xyz), N.iptr(self.elt_nodes)
        d.n_bound_edges = len(self.bound_types)
        d.bound_edge_nodes, d.bound_edge_type = N.iptr(self.bound_edges), N.uptr(self.bound_types)
        d.n_constituents = nc
        d.frequency = N.dptr(self.frequency) if nc else None
        d.forcing_fact = N.dptr(self.forcing_fact) if nc else None
        d.equilib_arg = N.dptr(self.equilib_arg) if nc else None
        d.node_amplitude = N.dptr(self.node_amplitude)
        d.node_phase = N.dptr(self.node_phase)
        d.node_manning_n = N.dptr(self.node_manning_n)
        d.n_levee_pairs = len(self.levee_nodes)
        d.levee_nodes = N.iptr(self.levee_nodes) if len(self.levee_nodes) else None
        d.levee_param = N.dptr(self.levee_param) if len(self.levee_nodes) else None

    @property
    def n_elements(self):
        return len(self.elt_nodes)

    def with_manning(self, node_manning_n):
        """The same mesh with nodal Manning n (one value per node)."""
        n = np.ascontiguousarray(node_manning_n, np.float64).reshape(-1)
        if len(n) != len(self.node_xyz):
            raise ValueError("one Manning n per node")
        return Mesh(self.node_xyz, self.elt_nodes, self.bound_edges, self.bound_types, self.frequency,
                    self.forcing_fact, self.equilib_arg, self.node_amplitude, self.node_phase, n,
                    self.levee_nodes, self.levee_param)

    def min_edge_length(self):
        x, y = self.node_xyz[:, 0], self.node_xyz[:, 1]
        e = self.elt_nodes
        lens = [np.hypot(x[e[:, a]] - x[e[:, b]], y[e[:, a]] - y[e[:, b]]) for a, b in ((0, 1), (1, 2), (2, 0))]
        return float(min(l.min() for l in lens))


def quad_mesh(x1, y1, x2, y2, m, n, bc=(BC_LAND,) * 4, pattern=0, bathy="constant", b0=1.0, alpha=1e-7,
              bmin=5.0, bmax=4000.0, n_modes=16, jitter=0.0, seed=0, step=(0.0, 1.0, 1.0), frequency=0.0,
              forcing_fact=1.0, equilib_arg=0.0, amplitude=1.0, phase=0.0, manning_n=0.0):
    """Quad-split triangle mesh from the library's generator (quad_mesh_generator semantics).

    bc: codes of the sides y=y1, x=x2, y=y2, x=x1 (reference order bottom, right, top, left).
    bathy: "constant" (b0), "bowl" (-alpha (x^2+y^2)), "fourier" (seeded modes clamped to
    [bmin, bmax]) or "step" (step=(x_step, depth_left, depth_right)).
    """
    lib = N.load_library()
    sp = N.QuadMeshSpec()
    sp.x1, sp.y1, sp.x2, sp.y2 = x1, y1, x2, y2
    sp.m, sp.n, sp.pattern = m, n, pattern
    for i in range(4):
        sp.bc[i] = bc[i]
    sp.bathy_kind = {"constant": 0, "bowl": 1, "fourier": 2, "step": 3}[bathy]
    sp.b0, sp.alpha, sp.bmin, sp.bmax, sp.n_modes = b0, alpha, bmin, bmax, n_modes
    sp.jitter, sp.seed = jitter, seed
    sp.step_x, sp.step_b_left, sp.step_b_right = step
    sp.frequency, sp.forcing_fact, sp.equilib_arg = frequency, forcing_fact, equilib_arg
    sp.amplitude, sp.phase, sp.manning_n = amplitude, phase, manning_n
    handle = C.c_void_p()
    desc = N.MeshDesc()
    st = lib.rkdg_mesh_generate(C.byref(sp), C.byref(handle), C.byref(desc))
    if st != N.RKDG_OK:
        raise RkdgError(st, "mesh generation failed")
    try:
        return _mesh_from_desc(desc)
    finally:
        lib.rkdg_mesh_free(handle)


def _mesh_from_desc(desc):
    """Copies a library-owned mesh description into a Mesh."""
    nn, ne = desc.n_nodes, desc.n_elements
    xyz = np.ctypeslib.as_array(desc.node_xyz, shape=(nn * 3,)).copy()
    elt = np.ctypeslib.as_array(desc.elt_nodes, shape=(ne * 3,)).copy()
    nb = desc.n_bound_edges
    be = np.ctypeslib.as_array(desc.bound_edge_nodes, shape=(nb * 2,)).copy() if nb else np.zeros(0, np.int32)
    bt = np.ctypeslib.as_array(desc.bound_edge_type, shape=(nb,)).copy() if nb else np.zeros(0, np.uint8)
    nc = desc.n_constituents
    kw = {}
    if nc:
        kw = dict(frequency=np.ctypeslib.as_array(desc.frequency, shape=(nc,)).copy(),
                  forcing_fact=np.ctypeslib.as_array(desc.forcing_fact, shape=(nc,)).copy(),
                  equilib_arg=np.ctypeslib.as_array(desc.equilib_arg, shape=(nc,)).copy(),
                  node_amplitude=np.ctypeslib.as_array(desc.node_amplitude, shape=(nn * nc,)).copy(),
                  node_phase=np.ctypeslib.as_array(desc.node_phase, shape=(nn * nc,)).copy())
    if desc.node_manning_n:
        kw["node_manning_n"] = np.ctypeslib.as_array(desc.node_manning_n, shape=(nn,)).copy()
    nl = desc.n_levee_pairs
    if nl:
        kw["levee_nodes"] = np.ctypeslib.as_array(desc.levee_nodes, shape=(nl * 2,)).copy()
        kw["levee_param"] = np.ctypeslib.as_array(desc.levee_param, shape=(nl * 3,)).copy()
    return Mesh(xyz, elt, be, bt, **kw)


def read_adcirc(fort14, bcis=None):
    """Reads an ADCIRC .14 mesh and its .bcis boundary data through the library's reader
    (AdcircFormat + SWE::Inputs::read_bcis of the reference)."""
    lib = N.load_library()
    handle = C.c_void_p()
    desc = N.MeshDesc()
    st = lib.rkdg_mesh_read_adcirc(str(fort14).encode(), None if bcis is None else str(bcis).encode(),
                                   C.byref(handle), C.byref(desc))
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())
    try:
        return _mesh_from_desc(desc)
    finally:
        lib.rkdg_mesh_free(handle)


def write_adcirc(mesh, fort14, bcis=None, name="mesh"):
    """Writes a Mesh as an ADCIRC .14 (and .bcis) through the library's writer; read_adcirc of the
    files returns the same mesh (boundary edges as a set, in the reader's order)."""
    lib = N.load_library()
    st = lib.rkdg_mesh_write_adcirc(C.byref(mesh.desc), name.encode(), str(fort14).encode(),
                                    None if bcis is None else str(bcis).encode())
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())


def write_meta(mesh, path, name="mesh"):
    """Writes a Mesh as a MeshMetaData (.meta) file (the reference partitioner's submesh format)."""
    lib = N.load_library()
    st = lib.rkdg_mesh_write_meta(C.byref(mesh.desc), name.encode(), str(path).encode())
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())


def read_meta(path, bcis=None):
    """Reads a MeshMetaData (.meta) file and, optionally, its .bcis boundary data."""
    lib = N.load_library()
    handle = C.c_void_p()
    desc = N.MeshDesc()
    st = lib.rkdg_mesh_read_meta(str(path).encode(), None if bcis is None else str(bcis).encode(),
                                 C.byref(handle), C.byref(desc))
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())
    try:
        return _mesh_from_desc(desc)
    finally:
        lib.rkdg_mesh_free(handle)


def partition_rcb(mesh, nparts):
    """Element -> part (recursive coordinate bisection of the barycentres, equal part sizes)."""
    lib = N.load_library()
    part = np.empty(mesh.n_elements, np.int32)
    st = lib.rkdg_partition_rcb(C.byref(mesh.desc), nparts, N.iptr(part))
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())
    return part


def submesh(mesh, part, p):
    """Part p's submesh (its elements + ghost neighbours, global order) and the global index of
    each local element."""
    lib = N.load_library()
    part = np.ascontiguousarray(part, dtype=np.int32)
    if part.shape != (mesh.n_elements,):
        raise ValueError(f"part has shape {part.shape}, expected ({mesh.n_elements},)")
    handle, desc, gid = C.c_void_p(), N.MeshDesc(), C.POINTER(C.c_int32)()
    st = lib.rkdg_submesh(C.byref(mesh.desc), N.iptr(part), p, C.byref(handle), C.byref(desc), C.byref(gid))
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())
    try:
        return _mesh_from_desc(desc), np.ctypeslib.as_array(gid, shape=(desc.n_elements,)).copy()
    finally:
        lib.rkdg_mesh_free(handle)


def renumber(mesh):
    """The mesh with elements in Hilbert-curve order of their barycentres (nodes in order of first
    use) and, per new element, its index in `mesh`."""
    lib = N.load_library()
    handle, desc, old = C.c_void_p(), N.MeshDesc(), C.POINTER(C.c_int32)()
    st = lib.rkdg_mesh_renumber(C.byref(mesh.desc), C.byref(handle), C.byref(desc), C.byref(old))
    if st != N.RKDG_OK:
        raise RkdgError(st, lib.rkdg_create_error().decode())
    try:
        return _mesh_from_desc(desc), np.ctypeslib.as_array(old, shape=(desc.n_elements,)).copy()
    finally:
        lib.rkdg_mesh_free(handle)


class PartitionedSimulation:
    """A mesh split into parts (one context each, all on this GPU) stepped with a halo exchange
    after every stage: the distributed-boundary path of the reference's MPI run, with the
    exchange as device-side pack / unpack kernels (on several GPUs the buffers would travel
    between devices, e.g. by RCCL send / receive).  Equals the whole-mesh run bit for bit."""

    def __init__(self, mesh, params, precision=RKDG_FP64_PLAIN, nparts=2, part=None):
        self.mesh, self.params, self.nparts = mesh, params, nparts
        self.part = partition_rcb(mesh, nparts) if part is None else np.ascontiguousarray(part, np.int32)
        self.sims, self.gids = [], []
        try:
            for p in range(nparts):
                sub, gid = submesh(mesh, self.part, p)
                self.sims.append(Simulation(sub, params, precision))
                self.gids.append(gid)
            # exchange plan: part q's ghosts owned by p, in global order, as local indices on both
            # sides; slot q on p packs, slot 32 + p on q unpacks
            if nparts > 32:
                raise ValueError("at most 32 parts (64 halo slots per context)")
            self.links = []
            loc = [{int(g): i for i, g in enumerate(gid)} for gid in self.gids]
            for q in range(nparts):
                gq = self.gids[q]
                own_q = self.part[gq]
                for p in range(nparts):
                    if p == q:
                        continue
                    g = gq[own_q == p]
                    if g.size == 0:
                        continue
                    self.sims[p].halo_define(q, [loc[p][int(x)] for x in g])
                    self.sims[q].halo_define(32 + p, [loc[q][int(x)] for x in g])
                    self.links.append((p, q))
            self.own = [self.part[gid] == p for p, gid in enumerate(self.gids)]
            self.exchange()
        except Exception:
            self.close()
            raise

    def exchange(self):
        for s in self.sims:
            s.sync()
        for p, q in self.links:
            self.sims[p].halo_pack(q)
        for s in self.sims:
            s.sync()
        for p, q in self.links:
            self.sims[q].halo_unpack(32 + p, *self.sims[p].halo_buffer(q))
        for s in self.sims:
            s.sync()

    def stage(self):
        if self.params.slope_limiting:  # the limiter reads the neighbours' new state
            for s in self.sims:
                s.stage_begin()
            self.exchange()
            for s in self.sims:
                s.stage_finish()
        else:
            for s in self.sims:
                s.stage()
        self.exchange()

    def step(self, nsteps=1):
        for _ in range(nsteps * self.params.nstages):
            self.stage()

    def set_state(self, q, wet=None):
        """Set the whole mesh's state [3, ndof, ne] (and wet flags): each part takes its own elements
        and its ghosts."""
        q = np.asarray(q, np.float64)
        for s, gid in zip(self.sims, self.gids):
            s.set_state(np.ascontiguousarray(q[:, :, gid]), None if wet is None else np.asarray(wet)[gid])

    def state(self):
        """The whole mesh's state (q, wet) gathered from the parts' own elements."""
        ndof = (self.params.p + 1) * (self.params.p + 2) // 2
        q = np.zeros((3, ndof, self.mesh.n_elements))
        wet = np.zeros(self.mesh.n_elements, np.uint8)
        for s, gid, own in zip(self.sims, self.gids, self.own):
            st = s.state()
            q[:, :, gid[own]] = st["q"][:, :, own]
            wet[gid[own]] = st["wet"][own]
        return {"q": q, "wet": wet}

    def close(self):
        for s in self.sims:
            s.close()
        self.sims = []


def make_params(p=1, nstages=2, order=2, dt=1.0, g=9.81, ramp_duration=0.0, problem=PROBLEM_NONE,
                ic_constant=True, ze0=0.0, qx0=0.0, qy0=0.0, function_source=False, friction=FRICTION_NONE,
                friction_cf=0.0, wetting_drying=False, h_o=0.1, slope_limiting=False, sl_M=1.0e-8, sl_nu=1.5,
                problem_param=(0.0, 0.0, 0.0, 0.0), spherical_mesh=False, sp_correction=False, sph_R=6378200.0,
                sph_lon_o=0.0, sph_lat_o=0.0):
    """rkdg_params with the reference's input defaults (problem_input/swe_inputs.hpp:90-155)."""
    prm = N.Params()
    prm.p, prm.nstages, prm.order, prm.dt, prm.g, prm.ramp_duration = p, nstages, order, dt, g, ramp_duration
    prm.problem, prm.ic_constant = problem, int(bool(ic_constant))
    prm.ze0, prm.qx0, prm.qy0 = ze0, qx0, qy0
    prm.function_source, prm.friction, prm.friction_cf = int(bool(function_source)), friction, friction_cf
    prm.wetting_drying, prm.h_o = int(bool(wetting_drying)), h_o
    prm.slope_limiting, prm.sl_M, prm.sl_nu = int(bool(slope_limiting)), sl_M, sl_nu
    for i in range(4):
        prm.problem_param[i] = problem_param[i]
    prm.spherical_mesh, prm.sp_correction = int(bool(spherical_mesh)), int(bool(sp_correction))
    prm.sph_R, prm.sph_lon_o, prm.sph_lat_o = sph_R, sph_lon_o, sph_lat_o
    return prm


class Simulation:
    """One mesh + problem on the GPU (the reference's Serial::Simulation)."""

    def __init__(self, mesh, params, precision=RKDG_FP64_PLAIN):
        self.lib = N.load_library()
        self.mesh, self.params, self.precision = mesh, params, precision
        self.ne = mesh.n_elements
        self.ndof = ndof_of(params.p)
        h = C.c_void_p()
        st = self.lib.rkdg_create(C.byref(mesh.desc), C.byref(params), precision, C.byref(h))
        if st != N.RKDG_OK:
            raise RkdgError(st, self.lib.rkdg_create_error().decode())
        self.ctx = h

    def _check(self, st):
        if st != N.RKDG_OK:
            raise RkdgError(st, self.lib.rkdg_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.rkdg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, nsteps=1):
        self._check(self.lib.rkdg_step(self.ctx, nsteps))

    def stage(self):
        self._check(self.lib.rkdg_stage(self.ctx))

    def sync(self):
        self._check(self.lib.rkdg_sync(self.ctx))

    def stage_begin(self):
        """The stage kernel of one stage (rkdg_stage_begin); stage_finish runs its limiter."""
        self._check(self.lib.rkdg_stage_begin(self.ctx))

    def stage_finish(self):
        self._check(self.lib.rkdg_stage_finish(self.ctx))

    # halo exchange of a partitioned mesh (rkdg_halo_*)
    def halo_define(self, slot, local_elements):
        idx = np.ascontiguousarray(local_elements, dtype=np.int32)
        self._check(self.lib.rkdg_halo_define(self.ctx, slot, N.iptr(idx) if idx.size else None, idx.size))

    def halo_buffer(self, slot):
        buf, nbytes = C.c_void_p(), C.c_int64()
        self._check(self.lib.rkdg_halo_buffer(self.ctx, slot, C.byref(buf), C.byref(nbytes)))
        return buf.value, nbytes.value

    def halo_pack(self, slot):
        self._check(self.lib.rkdg_halo_pack(self.ctx, slot))

    def halo_unpack(self, slot, device_buffer, nbytes):
        self._check(self.lib.rkdg_halo_unpack(self.ctx, slot, C.c_void_p(device_buffer), nbytes))

    def state(self):
        """Modal state [3, ndof, ne] plus wet / limited [3, ne] / troubled flags."""
        q = np.empty(3 * self.ndof * self.ne)
        wet = np.empty(self.ne, np.uint8)
        lim = np.empty(3 * self.ne, np.uint8)
        tr = np.empty(self.ne, np.uint8)
        self._check(self.lib.rkdg_get_state(self.ctx, N.dptr(q), N.uptr(wet), N.uptr(lim), N.uptr(tr)))
        return {"q": q.reshape(3, self.ndof, self.ne), "wet": wet, "limited": lim.reshape(3, self.ne),
                "troubled": tr}

    def set_state(self, q, wet=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        w = None if wet is None else np.ascontiguousarray(wet, dtype=np.uint8)
        self._check(self.lib.rkdg_set_state(self.ctx, N.dptr(q), N.uptr(w)))

    def time(self):
        t = C.c_double()
        step, stage = C.c_int32(), C.c_int32()
        self._check(self.lib.rkdg_get_time(self.ctx, C.byref(t), C.byref(step), C.byref(stage)))
        return t.value, step.value, stage.value

    def set_time(self, t, step):
        self._check(self.lib.rkdg_set_time(self.ctx, t, step))

    def nan_info(self):
        e, s, k = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.rkdg_nan_info(self.ctx, C.byref(e), C.byref(s), C.byref(k)))
        return e.value, s.value, k.value

    def l2_residual(self):
        v = C.c_double()
        self._check(self.lib.rkdg_l2_residual(self.ctx, C.byref(v)))
        return v.value

    def reset_timers(self):
        self._check(self.lib.rkdg_reset_timers(self.ctx))

    def set_timing(self, per_kernel=True):
        """Starts a timing window; per_kernel=False keeps only the events around each step call."""
        self._check(self.lib.rkdg_set_timing(self.ctx, 1 if per_kernel else 0))

    def report(self):
        r = N.PerfReport()
        self._check(self.lib.rkdg_report(self.ctx, C.byref(r)))
        return {f: getattr(r, f) for f, _ in N.PerfReport._fields_}

    def run(self, run_time):
        """Simulation::Run: ceil(run_time / dt) full steps (simulation.hpp:47, :68-82)."""
        n = int(math.ceil(run_time / self.params.dt))
        self.step(n)
        return n

    # ---- output (the reference's writers; prefix = output path + mesh name)
    def write_modal(self, prefix):
        self._check(self.lib.rkdg_write_modal(self.ctx, str(prefix).encode()))

    def write_vtu(self, prefix):
        self._check(self.lib.rkdg_write_vtu(self.ctx, str(prefix).encode()))

    def write_vtk(self, prefix):
        self._check(self.lib.rkdg_write_vtk(self.ctx, str(prefix).encode()))

    # ---- learned troubled-cell indicator
    def set_indicator(self, flat=None, threshold=0.0, judge=False):
        """Gate the limiter with the MLP of flat parameters (w1 b1 w2 b2 w3 b3 w4 b4); None removes it."""
        if flat is None:
            self._check(self.lib.rkdg_set_indicator(self.ctx, None, 0))
            return
        w = N.weights_from_flat(flat, threshold)
        self._check(self.lib.rkdg_set_indicator(self.ctx, C.byref(w), 1 if judge else 0))

    def indicator_stats(self, reset=False):
        st = N.IndicatorStats()
        self._check(self.lib.rkdg_get_indicator_stats(self.ctx, C.byref(st), 1 if reset else 0))
        return {f: getattr(st, f) for f, _ in N.IndicatorStats._fields_}

    def capture(self, on=True):
        """Record features / labels (/ logits) of each limiter stage's state before limiting."""
        self._check(self.lib.rkdg_capture(self.ctx, 1 if on else 0))

    def get_capture(self, logits=False):
        X = np.empty((N.MLP_IN, self.ne), np.float32)
        lab = np.empty(self.ne, np.uint8)
        ok = np.empty(self.ne, np.uint8)
        lg = np.empty(self.ne, np.float32) if logits else None
        ch = np.empty(self.ne, np.float32)
        self._check(self.lib.rkdg_get_capture(self.ctx, N.fptr(X), N.uptr(lab), N.uptr(ok), N.fptr(lg), N.fptr(ch)))
        out = {"features": X, "label": lab, "may_act": ok, "change": ch}
        if logits:
            out["logit"] = lg
        return out


class MlpTrainer:
    """The indicator's network trained on the GPU (include/rkdg.h rkdg_mlp_*)."""

    def __init__(self, flat, capacity, threshold=0.0):
        self.lib = N.load_library()
        w = N.weights_from_flat(flat, threshold)
        h = C.c_void_p()
        st = self.lib.rkdg_mlp_trainer_create(C.byref(w), int(capacity), C.byref(h))
        if st != N.RKDG_OK:
            raise RkdgError(st, "trainer create failed")
        self.h, self.capacity = h, int(capacity)

    def _check(self, st):
        if st != N.RKDG_OK:
            raise RkdgError(st, self.lib.rkdg_mlp_trainer_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.rkdg_mlp_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32).reshape(-1)
        if X.shape != (N.MLP_IN, y.size):
            raise ValueError("X must be [36][n] with n = len(y)")
        self._check(self.lib.rkdg_mlp_set_data(self.h, N.fptr(X), N.fptr(y), y.size))

    def train(self, epochs, batch=1024, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8, pos_weight=1.0, seed=5):
        prm = N.TrainParams(batch, epochs, lr, beta1, beta2, eps, pos_weight, seed)
        losses = np.empty(max(1, epochs), np.float32)
        self._check(self.lib.rkdg_mlp_train(self.h, C.byref(prm), N.fptr(losses)))
        return losses[:epochs]

    def weights(self):
        w = N.MlpWeights()
        self._check(self.lib.rkdg_mlp_get_weights(self.h, C.byref(w)))
        return N.weights_to_flat(w)

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(X.shape[1], np.float32)
        self._check(self.lib.rkdg_mlp_forward(self.h, N.fptr(X), X.shape[1], N.fptr(out)))
        return out

    def gradient(self, X, y, pos_weight=1.0):
        """(loss, flat gradient) of one batch under the current weights."""
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32).reshape(-1)
        loss = C.c_float()
        g = N.MlpWeights()
        self._check(self.lib.rkdg_mlp_gradient(self.h, N.fptr(X), N.fptr(y), y.size, pos_weight, C.byref(loss),
                                               C.byref(g)))
        return loss.value, N.weights_to_flat(g)


class Batch:
    """An ensemble of independent meshes stepped together on one GPU (rkdg_create_batch).
    Instances of equal p share one merged element range and one stage launch per stage."""

    def __init__(self, meshes, params, precision=RKDG_FP32_BULK):
        if len(meshes) != len(params) or not meshes:
            raise ValueError("one params entry per mesh, at least one mesh")
        self.lib = N.load_library()
        self.meshes, self.params, self.precision = list(meshes), list(params), precision
        n = len(self.meshes)
        descs = (N.MeshDesc * n)(*[m.desc for m in self.meshes])
        prms = (N.Params * n)(*self.params)
        h = C.c_void_p()
        st = self.lib.rkdg_create_batch(n, descs, prms, precision, C.byref(h))
        if st != N.RKDG_OK:
            raise RkdgError(st, self.lib.rkdg_create_error().decode())
        self.ctx = h

    def __len__(self):
        return len(self.meshes)

    def _check(self, st):
        if st != N.RKDG_OK:
            raise RkdgError(st, self.lib.rkdg_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.rkdg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1):
        self._check(self.lib.rkdg_step(self.ctx, nsteps))

    def sync(self):
        self._check(self.lib.rkdg_sync(self.ctx))

    def state(self, i):
        m, p = self.meshes[i], self.params[i]
        ne, nd = m.n_elements, ndof_of(p.p)
        q = np.empty(3 * nd * ne)
        w = np.empty(ne, np.uint8)
        lim = np.empty(3 * ne, np.uint8)
        tr = np.empty(ne, np.uint8)
        self._check(self.lib.rkdg_batch_get_state(self.ctx, i, N.dptr(q), N.uptr(w), N.uptr(lim), N.uptr(tr)))
        return {"q": q.reshape(3, nd, ne), "wet": w, "limited": lim.reshape(3, ne), "troubled": tr}

    def set_state(self, i, q, wet=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        w = None if wet is None else np.ascontiguousarray(wet, dtype=np.uint8)
        self._check(self.lib.rkdg_batch_set_state(self.ctx, i, N.dptr(q), N.uptr(w)))

    def time(self, i):
        t = C.c_double()
        step, stage = C.c_int32(), C.c_int32()
        self._check(self.lib.rkdg_batch_get_time(self.ctx, i, C.byref(t), C.byref(step), C.byref(stage)))
        return t.value, step.value, stage.value

    def nan_info(self):
        v = [C.c_int32() for _ in range(4)]
        self._check(self.lib.rkdg_batch_nan_info(self.ctx, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @property
    def schedule(self):
        sc = N.Schedule()
        self._check(self.lib.rkdg_batch_get_schedule(self.ctx, C.byref(sc)))
        return {"block": sc.block, "groups_per_xcd": sc.groups_per_xcd, "graph": sc.graph}

    def set_schedule(self, block=128, groups_per_xcd=0, graph=0):
        sc = N.Schedule(block, groups_per_xcd, graph)
        self._check(self.lib.rkdg_batch_set_schedule(self.ctx, C.byref(sc)))

    def autotune(self, steps_per_trial=3):
        sc = N.Schedule()
        ms = C.c_double()
        self._check(self.lib.rkdg_batch_autotune(self.ctx, steps_per_trial, C.byref(sc), C.byref(ms)))
        return {"block": sc.block, "groups_per_xcd": sc.groups_per_xcd, "graph": sc.graph}, ms.value

    def reset_timers(self):
        self._check(self.lib.rkdg_reset_timers(self.ctx))

    def set_timing(self, per_kernel=True):
        """Starts a timing window; per_kernel=False keeps only the events around each step call."""
        self._check(self.lib.rkdg_set_timing(self.ctx, 1 if per_kernel else 0))

    def report(self):
        r = N.PerfReport()
        self._check(self.lib.rkdg_report(self.ctx, C.byref(r)))
        return {f: getattr(r, f) for f, _ in N.PerfReport._fields_}
