"""The engine's named workloads (meshes + problem inputs), at full size or scaled down.

  manufactured  - examples/swe_manufactured_solution: [-1,1]^2, 64x64 quads x 2 triangles,
                  depth 2, g = 1, p = 1, SSP-RK2, dt = 1e-3, function BCs + function source.
  bowl          - examples/swe_parabolic_bowl: [-4000,4000]^2, 128x128 quads x 2, depth
                  -1e-7 (x^2+y^2), g = 9.81, p = 1, SSP-RK2, dt = 1 s, wetting-drying
                  (h_o = 0.01) and the Cockburn-Shu limiter (M = 1e-8, nu = 1.5).
  tidal         - 708x708 quads x 2 (1,002,528 triangles) over 708 km, interior vertices
                  jittered +-20 % (seed 3), depth 5-4000 m from 16 seeded Fourier modes (seed 3),
                  M2 tide (omega = 1.405189e-4 rad/s, 1 m) on the x = x1 side, land elsewhere,
                  Manning n = 0.025, p = 2, SSP-RK3, dt = 0.1 h_min / (sqrt(g b_max) (2p + 1)).
  dambreak      - 500x500 quads x 2 over 5 km x 5 km, p = 1, SSP-RK2: a 10 m deep reservoir
                  (x < 2 km) behind a dam, 1 m of water downstream, and a dry floodplain strip
                  (x > 4 km) whose bed sits 0.5 m above the downstream level (dry at t = 0); wetting-drying and
                  limiter on.
Every builder takes a scale factor (quads per side are multiplied by it) so tests can run the
same recipe small.
"""
import math

import numpy as np

from . import (BC_FUNCTION, BC_LAND, BC_TIDE, FRICTION_MANNING, PROBLEM_DAM_BREAK, PROBLEM_MANUFACTURED, Mesh,
               PROBLEM_NONE, PROBLEM_PARABOLIC_BOWL, make_params, quad_mesh)

M2_OMEGA = 1.405189e-4


def _n(base, scale):
    return max(2, int(round(base * scale)))


def manufactured(scale=1.0, p=1, nstages=2, order=2, dt=1.0e-3):
    m = _n(64, scale)
    mesh = quad_mesh(-1, -1, 1, 1, m, m, bc=(BC_FUNCTION,) * 4, bathy="constant", b0=2.0)
    prm = make_params(p=p, nstages=nstages, order=order, dt=dt, g=1.0, problem=PROBLEM_MANUFACTURED,
                      ic_constant=False, function_source=True)
    return mesh, prm


def bowl(scale=1.0, p=1, wetting_drying=True, slope_limiting=True):
    m = _n(128, scale)
    mesh = quad_mesh(-4000, -4000, 4000, 4000, m, m, bc=(BC_LAND,) * 4, bathy="bowl", alpha=1.0e-7)
    prm = make_params(p=p, nstages=2, order=2, dt=1.0, g=9.81, problem=PROBLEM_PARABOLIC_BOWL, ic_constant=False,
                      wetting_drying=wetting_drying, h_o=0.01, slope_limiting=slope_limiting, sl_M=1.0e-8,
                      sl_nu=1.5)
    return mesh, prm


def tidal(scale=1.0, p=2, nstages=3, order=3, seed=3):
    m = _n(708, scale)
    L = 708000.0 * (m / 708.0)
    mesh = quad_mesh(0.0, 0.0, L, L, m, m, bc=(BC_LAND, BC_LAND, BC_LAND, BC_TIDE), bathy="fourier", bmin=5.0,
                     bmax=4000.0, n_modes=16, jitter=0.2, seed=seed, frequency=M2_OMEGA, forcing_fact=1.0,
                     equilib_arg=0.0, amplitude=1.0, phase=0.0, manning_n=0.025)
    g = 9.81
    dt = 0.1 * mesh.min_edge_length() / (math.sqrt(g * 4000.0) * (2 * p + 1))
    prm = make_params(p=p, nstages=nstages, order=order, dt=dt, g=g, problem=PROBLEM_NONE, ic_constant=True,
                      friction=FRICTION_MANNING, friction_cf=0.0)
    return mesh, prm


def dambreak(scale=1.0, p=1):
    m = _n(500, scale)
    L = 5000.0
    # depth (positive down): 10 m up to the floodplain, 8.5 m beyond x = 4 km (bed 0.5 m above the
    # downstream water level)
    mesh = quad_mesh(0.0, 0.0, L, L, m, m, bc=(BC_LAND,) * 4, bathy="step", step=(4000.0, 10.0, 8.5))
    # free surface: 0 m behind the dam (x < 2 km), -9 m downstream (1 m of water over the 10 m bed)
    dt = 0.2 * (L / m) / (math.sqrt(9.81 * 10.0) * (2 * p + 1))
    prm = make_params(p=p, nstages=2, order=2, dt=dt, g=9.81, problem=PROBLEM_DAM_BREAK, ic_constant=False, wetting_drying=True, h_o=0.01,
                      slope_limiting=True, problem_param=(2000.0, 0.0, -9.0, 0.0))
    return mesh, prm


def ensemble(n=64, scale=1.0, seed=4, ne_min=20000, ne_max=400000, nstages=3, order=3):
    """Ensemble of n tidal instances of the tidal recipe (1 km quads, jitter 20 %, Fourier depth
    5-4000 m, M2 tide on one side, Manning 0.025): triangle counts log-uniform in [ne_min, ne_max]
    (square grids of m = floor(sqrt(Ne / 2)) quads, Ne scaled by `scale`), p in {1, 2} with
    probability 1/2, tide amplitude uniform in [0.1, 1.5] m, SSP-RK3, each with its own CFL dt.
    Returns (meshes, params, info)."""
    rng = np.random.default_rng(seed)
    meshes, params, info = [], [], []
    g = 9.81
    for i in range(n):
        ne = math.exp(rng.uniform(math.log(ne_min), math.log(ne_max))) * scale
        p = int(rng.integers(1, 3))
        amp = float(rng.uniform(0.1, 1.5))
        m = max(2, int(math.sqrt(ne / 2)))
        L = 1000.0 * m
        mesh = quad_mesh(0.0, 0.0, L, L, m, m, bc=(BC_LAND, BC_LAND, BC_LAND, BC_TIDE), bathy="fourier", bmin=5.0,
                         bmax=4000.0, n_modes=16, jitter=0.2, seed=seed * 1000 + i, frequency=M2_OMEGA,
                         forcing_fact=1.0, equilib_arg=0.0, amplitude=amp, phase=0.0, manning_n=0.025)
        dt = 0.1 * mesh.min_edge_length() / (math.sqrt(g * 4000.0) * (2 * p + 1))
        prm = make_params(p=p, nstages=nstages, order=order, dt=dt, g=g, problem=PROBLEM_NONE, ic_constant=True,
                          friction=FRICTION_MANNING, friction_cf=0.0)
        meshes.append(mesh)
        params.append(prm)
        info.append({"triangles": mesh.n_elements, "p": p, "amplitude": amp, "dt": dt})
    return meshes, params, info


WORKLOADS = {"manufactured": manufactured, "bowl": bowl, "tidal": tidal, "dambreak": dambreak}


def shuffled(mesh, seed=1):
    """The same mesh with its elements and nodes in random order (a mesh file with no locality),
    and, per new element, its index in `mesh`."""
    rng = np.random.default_rng(seed)
    pe = rng.permutation(mesh.n_elements)
    pn = rng.permutation(len(mesh.node_xyz))  # new -> old
    inv = np.empty_like(pn)
    inv[pn] = np.arange(pn.size)
    kw = {}
    if len(mesh.frequency):
        nc = len(mesh.frequency)
        kw = dict(frequency=mesh.frequency, forcing_fact=mesh.forcing_fact, equilib_arg=mesh.equilib_arg,
                  node_amplitude=mesh.node_amplitude.reshape(-1, nc)[pn].reshape(-1),
                  node_phase=mesh.node_phase.reshape(-1, nc)[pn].reshape(-1))
    if mesh.node_manning_n is not None:
        kw["node_manning_n"] = mesh.node_manning_n[pn]
    if len(mesh.levee_nodes):
        kw["levee_nodes"], kw["levee_param"] = inv[mesh.levee_nodes], mesh.levee_param
    return Mesh(mesh.node_xyz[pn], inv[mesh.elt_nodes[pe]], inv[mesh.bound_edges], mesh.bound_types, **kw), pe

