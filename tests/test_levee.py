"""ADCIRC ingest and the levee (internal barrier) interface, on the reference's own weir fixture.

Fixture: test/files_for_testing/weir/{weir.14, weir.bcis, dgswemv2_input.15} of the reference,
copied as data into tests/golden/weir/.  The mesh is 18 x 26 km of 1 km quads split into
triangles with one row removed: a barrier joins the node string at y = 15 km (front) to the one
at y = 16 km (back), crest 2.0 m for x <= 7 km and 0.1 m for x >= 8 km, weir coefficients 1.
Flow forcing (two constituents) enters through y = 0.

Pins for the oracle's levee flux (ISP::Levee::ComputeFlux, rkdg_swe_is_levee.hpp:62-174), all
independent of its code: the broad-crested (free) weir law q = C (2h/3) sqrt(g 2h/3) and the
submerged weir law q = C h_d sqrt(2 g (h_u - h_d)), each entering through an LLF flux whose mass
component with zero interior momentum is q/2; a still level below both crests is a lake at rest.
"""
import numpy as np
import pytest

from conftest import relerr
from dgswemv2_amd import BC_LAND, BC_FLOW, BC_LEVEE, Mesh, RkdgError, make_params, read_adcirc
from oracle.oracle_py import OracleSim

from pathlib import Path

WEIR = Path(__file__).resolve().parent / "golden" / "weir"


def weir_mesh(flow_as_land=False):
    m = read_adcirc(WEIR / "weir.14", WEIR / "weir.bcis")
    if not flow_as_land:
        return m
    bt = m.bound_types.copy()
    bt[bt == BC_FLOW] = BC_LAND
    return Mesh(m.node_xyz, m.elt_nodes, m.bound_edges, bt, levee_nodes=m.levee_nodes, levee_param=m.levee_param)


def weir_params(p=1, nstages=2, order=2, dt=3.0, slope_limiting=True):
    # dgswemv2_input.15 of the fixture: p = 1, SSP-RK2, dt = 3 s, g = 9.81, Cockburn-Shu M = 1e-8, nu = 1.5
    return make_params(p=p, nstages=nstages, order=order, dt=dt, g=9.81, slope_limiting=slope_limiting, sl_M=1e-8,
                       sl_nu=1.5)


def test_reader_weir_fixture():
    m = read_adcirc(WEIR / "weir.14", WEIR / "weir.bcis")
    assert m.n_elements == 900 and len(m.node_xyz) == 513
    types, counts = np.unique(m.bound_types, return_counts=True)
    assert dict(zip(types.tolist(), counts.tolist())) == {BC_LAND: 68, BC_FLOW: 18, BC_LEVEE: 36}
    # element IDs in the file have gaps (0..935); they are relabelled in order
    assert m.elt_nodes[0].tolist() == [0, 20, 19] and m.elt_nodes[-1].tolist() == [492, 493, 512]
    # two flow constituents from the .bcis, forcing factor -1 (inflow)
    assert np.allclose(m.frequency, [0.0, 0.0001454441]) and np.allclose(m.forcing_fact, [-1.0, -1.0])
    amp = m.node_amplitude.reshape(-1, 2)
    assert np.count_nonzero(amp[:, 0]) == 19 and np.allclose(amp[:19], 0.2)
    ph = m.node_phase.reshape(-1, 2)
    assert np.allclose(ph[:19, 0], 0.0) and np.allclose(ph[:19, 1], 1.5707963268)
    # 19 barrier pairs front y = 15 km -> back y = 16 km, crest 2.0 then 0.1
    assert m.levee_nodes.shape == (19, 2)
    assert np.all(m.node_xyz[m.levee_nodes[:, 0], 1] == 15000.0)
    assert np.all(m.node_xyz[m.levee_nodes[:, 1], 1] == 16000.0)
    assert np.allclose(m.levee_param[:8], [2.0, 1.0, 1.0]) and np.allclose(m.levee_param[8:], [0.1, 1.0, 1.0])


def test_reader_errors(tmp_path):
    with pytest.raises(RkdgError, match="not found"):
        read_adcirc(tmp_path / "missing.14")
    bad = tmp_path / "bad.14"
    # one triangle whose land string covers only the edge (0, 1)
    bad.write_text("name\n1 3\n0 0 0 1\n1 1 0 1\n2 0 1 1\n0 3 0 1 2\n0\n0\n1\n2\n2 0\n0\n1\n")
    with pytest.raises(RkdgError, match="boundary not found"):
        read_adcirc(bad)
    ok = tmp_path / "ok.14"
    ok.write_text("name\n1 3\n0 0 0 1\n1 1 0 1\n2 0 1 1\n0 3 0 1 2\n0\n0\n1\n4\n4 0\n0\n1\n2\n0\n")
    m = read_adcirc(ok)
    assert m.n_elements == 1 and np.all(m.bound_types == BC_LAND) and len(m.bound_types) == 3


def _mode0_scale(mesh, p):
    # modal coefficient of a unit constant (the mean = q0 * scale)
    return OracleSim(mesh, make_params(p=p, ic_constant=True, ze0=1.0)).state()["q"][0, 0, 0]


def _areas(mesh):
    x, y = mesh.node_xyz[:, 0], mesh.node_xyz[:, 1]
    e = mesh.elt_nodes
    return 0.5 * np.abs((x[e[:, 1]] - x[e[:, 0]]) * (y[e[:, 2]] - y[e[:, 0]]) -
                        (x[e[:, 2]] - x[e[:, 0]]) * (y[e[:, 1]] - y[e[:, 0]]))


@pytest.mark.parametrize("p", [1, 2])
def test_levee_lake_at_rest_below_crest(p):
    mesh = weir_mesh(flow_as_land=True)
    o = OracleSim(mesh, weir_params(p=p, slope_limiting=False))
    c = _mode0_scale(mesh, p)
    q = np.zeros_like(o.state()["q"])
    q[0, 0, :] = 0.05 * c  # 5 cm above datum, below the 0.1 m crest by more than the 1 cm tolerance
    o.set_state(q)
    o.step(3)
    s = o.state()["q"]
    assert np.max(np.abs(s[1:])) < 1e-12 and np.max(np.abs(s[0] - q[0])) < 1e-14


def _weir_flux_expected(mesh, up, down, p):
    """Total mass flux (m^3/s) across the barrier at t = 0 with still water at ze = up below and
    ze = down above, from the weir laws: each barrier edge contributes int q_weir/2 ds over
    Gauss-Legendre points with the crest interpolated linearly between its two node pairs."""
    xyz, ln, lp = mesh.node_xyz, mesh.levee_nodes, mesh.levee_param
    z, w = np.polynomial.legendre.leggauss(p + 1)
    g = 9.81
    total = 0.0
    for i in range(len(ln) - 1):  # consecutive front nodes form the barrier edges
        L = abs(xyz[ln[i + 1, 0], 0] - xyz[ln[i, 0], 0])
        for zi, wi in zip(z, w):
            H = lp[i, 0] * (1 - zi) / 2 + lp[i + 1, 0] * (1 + zi) / 2
            Cs = lp[i, 1] * (1 - zi) / 2 + lp[i + 1, 1] * (1 + zi) / 2
            Cp = lp[i, 2] * (1 - zi) / 2 + lp[i + 1, 2] * (1 + zi) / 2
            hu, hd = up - H, down - H
            if (hu <= 0.01 and hd <= 0.01) or abs(hu - hd) <= 0.01:
                q = 0.0
            elif hd > 2 * hu / 3:  # submerged
                q = Cs * hd * np.sqrt(2 * g * (hu - hd))
            else:  # free (broad-crested)
                q = Cp * (2 * hu / 3) * np.sqrt(g * 2 * hu / 3)
            total += wi * L / 2 * q / 2
    return total


@pytest.mark.parametrize("p,down", [(1, 0.0), (1, 0.9), (2, 0.0), (2, 0.9)])
def test_levee_weir_discharge(p, down):
    """One forward-Euler stage from rest: the mass leaving the upstream side equals the weir
    discharge (free weir for down = 0, submerged for down = 0.9) through the LLF flux."""
    mesh = weir_mesh(flow_as_land=True)
    dt = 0.01
    o = OracleSim(mesh, weir_params(p=p, nstages=1, order=1, dt=dt, slope_limiting=False))
    c = _mode0_scale(mesh, p)
    yc = mesh.node_xyz[mesh.elt_nodes, 1].mean(axis=1)
    upside = yc < 15000
    up = 1.1
    q = np.zeros_like(o.state()["q"])
    q[0, 0, :] = np.where(upside, up, down) * c
    o.set_state(q)
    area = _areas(mesh)
    m0_up = np.sum(q[0, 0, upside] / c * area[upside])
    m0_dn = np.sum(q[0, 0, ~upside] / c * area[~upside])
    o.step(1)
    s = o.state()["q"]
    m1_up = np.sum(s[0, 0, upside] / c * area[upside])
    m1_dn = np.sum(s[0, 0, ~upside] / c * area[~upside])
    expected = _weir_flux_expected(mesh, up, down, p) * dt
    assert expected > 0
    assert abs((m0_up - m1_up) - expected) <= 1e-9 * expected
    assert abs((m1_dn - m0_dn) - expected) <= 1e-9 * expected


def test_levee_weir_run_serial_vs_openmp():
    """The OpenMP oracle (coloured interfaces, levees included) agrees with the serial one."""
    mesh = weir_mesh()
    prm = weir_params(p=1)
    a, b = OracleSim(mesh, prm), OracleSim(mesh, prm, threads=4)
    c = _mode0_scale(mesh, 1)
    yc = mesh.node_xyz[mesh.elt_nodes, 1].mean(axis=1)
    q = np.zeros_like(a.state()["q"])
    q[0, 0, :] = np.where(yc < 15000, 0.5, 0.0) * c
    a.set_state(q), b.set_state(q)
    a.step(10), b.step(10)
    assert max(relerr(b.state()["q"], a.state()["q"])) < 1e-10
