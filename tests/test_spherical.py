"""Spherical coordinates: node projection (preprocess_mesh_data, swe_pre_preprocess_mesh.hpp:5-27)
and the spherical correction sp = cos(lat_o) / cos(y / R) of the volume x-fluxes and the LLF flux
(swe_pre_init_data.hpp:31-48, 80-160; rkdg_swe_proc_volume.hpp:25-30; rkdg_swe_LLF_flux.hpp:14-45).

Pins of the oracle's spherical path that do not retype its formulas: a mesh given in degrees with
the projection on equals the same mesh given in projected metres (the projection path), the
correction with sp identically 1 (lat_o = 0, huge radius) equals the Cartesian run bit for bit,
and the correction changes a run by the size the sp factor predicts.  The GPU plain form then
matches the oracle bit for bit and the fast forms within their gates.
"""
import math

import numpy as np
import pytest

from dgswemv2_amd import (BC_LAND, BC_TIDE, RKDG_FP32_BULK, RKDG_FP64_FAST, RKDG_FP64_PLAIN, Mesh, Simulation,
                          make_params, quad_mesh)
from oracle.oracle_py import OracleSim

PI = 3.14159265359  # general_definitions.hpp:227
LON_O, LAT_O, R = -79.5, 35.0, 6378200.0


def degree_mesh(m=10):
    """A 0.5 x 0.5 degree box around (LON_O, LAT_O) in longitude / latitude, tide on the west side."""
    return quad_mesh(LON_O - 0.25, LAT_O - 0.25, LON_O + 0.25, LAT_O + 0.25, m, m,
                     bc=(BC_LAND, BC_LAND, BC_LAND, BC_TIDE), pattern=1, bathy="fourier", bmin=10.0, bmax=60.0,
                     n_modes=4, jitter=0.15, seed=7, frequency=1.405189e-4, forcing_fact=1.0, equilib_arg=0.0,
                     amplitude=0.5, phase=0.0)


def params(**kw):
    base = dict(p=2, nstages=3, order=3, dt=4.0, g=9.81, ic_constant=True, spherical_mesh=True, sp_correction=True,
                sph_R=R, sph_lon_o=LON_O, sph_lat_o=LAT_O)
    base.update(kw)
    return make_params(**base)


def projected(mesh):
    xyz = mesh.node_xyz.copy()
    R_o = R * math.cos(LAT_O * PI / 180.0)
    xyz[:, 0] = R_o * (xyz[:, 0] - LON_O) * PI / 180.0
    xyz[:, 1] = R * xyz[:, 1] * PI / 180.0
    return Mesh(xyz, mesh.elt_nodes, mesh.bound_edges, mesh.bound_types, frequency=mesh.frequency,
                forcing_fact=mesh.forcing_fact, equilib_arg=mesh.equilib_arg, node_amplitude=mesh.node_amplitude,
                node_phase=mesh.node_phase)


def run_oracle(mesh, prm, steps):
    o = OracleSim(mesh, prm)
    o.step(steps)
    q = o.state()["q"]
    o.close()
    return q


def test_projection_path():
    m = degree_mesh()
    a = run_oracle(m, params(), 6)
    b = run_oracle(projected(m), params(spherical_mesh=False), 6)
    assert np.abs(a).max() > 0 and np.array_equal(a, b)


def test_unit_correction_equals_cartesian():
    m = projected(degree_mesh())
    # lat_o = 0 and a radius of 1e30: cos(lat_o) = 1 and cos(y / R) = 1, so sp = 1 exactly
    a = run_oracle(m, params(spherical_mesh=False, sph_lat_o=0.0, sph_R=1e30), 6)
    b = run_oracle(m, params(spherical_mesh=False, sp_correction=False), 6)
    assert np.array_equal(a, b)


def test_correction_size():
    """sp - 1 is O(1e-3) over a 0.5 degree box centred on lat_o; the correction changes the run by
    that order relative to the change the run itself makes, not by rounding and not by O(1)."""
    m = degree_mesh()
    q_sp = run_oracle(m, params(), 6)
    q_no = run_oracle(m, params(sp_correction=False), 6)
    q0 = run_oracle(m, params(sp_correction=False), 0)
    y = np.radians(np.array([LAT_O - 0.25, LAT_O + 0.25]))
    spmax = float(np.max(np.abs(math.cos(math.radians(LAT_O)) / np.cos(y) - 1)))
    rel = np.linalg.norm(q_sp - q_no) / np.linalg.norm(q_no - q0)
    assert 1e-2 * spmax < rel < 10 * spmax, (rel, spmax)


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2])
def test_gpu_spherical_plain_vs_oracle(p):
    m = degree_mesh()
    prm = params(p=p, nstages=3 if p == 2 else 2, order=3 if p == 2 else 2)
    o = OracleSim(m, prm)
    s = Simulation(m, prm, RKDG_FP64_PLAIN)
    for k in range(12):
        o.stage(), s.stage()
        assert np.array_equal(s.state()["q"], o.state()["q"]), k
    o.close(), s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(RKDG_FP64_FAST, 1e-12), (RKDG_FP32_BULK, 1e-5)])
def test_gpu_spherical_fast_forms(precision, tol):
    m = degree_mesh()
    prm = params()
    a = Simulation(m, prm, RKDG_FP64_PLAIN)
    b = Simulation(m, prm, precision)
    a.step(10), b.step(10)
    qa, qb = a.state()["q"], b.state()["q"]
    for k in range(3):
        assert np.linalg.norm(qb[k] - qa[k]) <= tol * np.linalg.norm(qa[k]), k
    a.close(), b.close()
