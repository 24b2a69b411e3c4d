"""Partitions and the halo exchange (the distributed-boundary path).

The reference's MPI run partitions the mesh (partitioner/partition.cpp) and exchanges
distributed-boundary data every stage (rkdg_swe_proc_ompi_step.hpp:38-102); its serial and
parallel runs are compared on L2 errors (scripts/correctness/compare_l2_errors.py:34-38).  Here a
partitioned run must equal the whole-mesh run bit for bit: on the CPU with the oracle stepping
each submesh and the ghosts refreshed from their owners after every stage, and on the GPU through
rkdg_halo_pack / rkdg_halo_unpack in every precision.
"""
import numpy as np
import pytest

from dgswemv2_amd import (BC_LAND, RKDG_FP32_BULK, RKDG_FP64_FAST, RKDG_FP64_PLAIN, PartitionedSimulation, RkdgError,
                          Simulation, partition_rcb, submesh)
from dgswemv2_amd import workloads as W
from oracle.oracle_py import OracleSim
from conftest import relerr


def _adjacency(mesh):
    edges = {}
    for e, nodes in enumerate(mesh.elt_nodes.tolist()):
        for k in range(3):
            edges.setdefault(frozenset((nodes[(k + 1) % 3], nodes[(k + 2) % 3])), []).append(e)
    nb = {e: set() for e in range(mesh.n_elements)}
    for els in edges.values():
        if len(els) == 2:
            nb[els[0]].add(els[1]), nb[els[1]].add(els[0])
    return edges, nb


@pytest.mark.parametrize("nparts", [1, 3, 8])
def test_partition_and_submeshes(nparts):
    mesh, _ = W.tidal(scale=0.05)
    part = partition_rcb(mesh, nparts)
    counts = np.bincount(part, minlength=nparts)
    assert counts.sum() == mesh.n_elements and counts.max() - counts.min() <= 1
    edges, nb = _adjacency(mesh)
    gbt = {frozenset(map(int, be)): int(t) for be, t in zip(mesh.bound_edges, mesh.bound_types)}
    for p in range(nparts):
        sub, gid = submesh(mesh, part, p)
        own = set(np.flatnonzero(part == p).tolist())
        ghosts = set().union(*(nb[e] for e in own)) - own
        assert np.all(np.diff(gid) > 0) and set(gid.tolist()) == own | ghosts
        # same triangles (through the submesh's node numbering), nodes in global order
        xy = sub.node_xyz[sub.elt_nodes]
        assert np.array_equal(xy, mesh.node_xyz[mesh.elt_nodes[gid]])
        # boundary faces: the global ones keep their codes; ghost faces out of the submesh are land
        sbt = {}
        for be, t in zip(sub.bound_edges, sub.bound_types):
            key = frozenset(tuple(sub.node_xyz[v, :2]) for v in be)
            sbt[key] = int(t)
        for le, e in enumerate(gid.tolist()):
            nodes = mesh.elt_nodes[e].tolist()
            for k in range(3):
                ge = frozenset((nodes[(k + 1) % 3], nodes[(k + 2) % 3]))
                key = frozenset(tuple(mesh.node_xyz[v, :2]) for v in ge)
                if len(edges[ge]) == 1:
                    assert sbt[key] == gbt[ge]
                elif not set(edges[ge]) <= set(gid.tolist()):
                    assert e in ghosts and sbt[key] == BC_LAND
                else:
                    assert key not in sbt


def test_partition_errors():
    from dgswemv2_amd import Mesh
    from test_levee import weir_mesh
    mesh, _ = W.tidal(scale=0.03)
    for bad in (0, mesh.n_elements + 1):
        with pytest.raises(RkdgError, match="bad arguments"):
            partition_rcb(mesh, bad)
    part = partition_rcb(mesh, 3)
    with pytest.raises(RkdgError, match="part 3 has no elements"):
        submesh(mesh, part, 3)
    with pytest.raises(ValueError, match="shape"):
        submesh(mesh, part[:-1], 0)
    # a barrier edge whose nodes have no barrier partners
    m = weir_mesh()
    cut = Mesh(m.node_xyz, m.elt_nodes, m.bound_edges, m.bound_types, levee_nodes=m.levee_nodes[:-3],
               levee_param=m.levee_param[:-3])
    with pytest.raises(RkdgError, match="levee edge without barrier node pairs"):
        submesh(cut, np.zeros(m.n_elements, np.int32), 0)


def _oracle_partitioned(mesh, prm, nparts, stages, part=None, q0=None):
    """The oracle stepping each submesh, ghosts refreshed from their owners after every stage."""
    part = partition_rcb(mesh, nparts) if part is None else part
    subs = [submesh(mesh, part, p) for p in range(nparts)]
    sims = [OracleSim(s, prm) for s, _ in subs]
    if q0 is not None:
        for (_, gid), s in zip(subs, sims):
            s.set_state(np.ascontiguousarray(q0[:, :, gid]))
    for _ in range(stages):
        for s in sims:
            s.stage()
        st = [s.state() for s in sims]
        gq = np.zeros((3, st[0]["q"].shape[1], mesh.n_elements))
        gw = np.zeros(mesh.n_elements, np.uint8)
        for p, ((_, gid), S) in enumerate(zip(subs, st)):
            own = part[gid] == p
            gq[:, :, gid[own]] = S["q"][:, :, own]
            gw[gid[own]] = S["wet"][own]
        for (_, gid), s in zip(subs, sims):
            s.set_state(np.ascontiguousarray(gq[:, :, gid]), np.ascontiguousarray(gw[gid]))
    for s in sims:
        s.close()
    return gq, gw


@pytest.mark.parametrize("name,kw,nparts,stages", [
    ("tidal", dict(scale=0.04), 4, 9),
    ("bowl", dict(scale=0.1, slope_limiting=False), 3, 12),
    ("manufactured", dict(scale=0.2), 5, 6),
])
def test_oracle_partitioned_equals_whole(name, kw, nparts, stages):
    mesh, prm = getattr(W, name)(**kw)
    whole = OracleSim(mesh, prm)
    for _ in range(stages):
        whole.stage()
    ref = whole.state()
    whole.close()
    q, wet = _oracle_partitioned(mesh, prm, nparts, stages)
    assert np.array_equal(q, ref["q"]) and np.array_equal(wet, ref["wet"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP64_FAST, RKDG_FP32_BULK])
@pytest.mark.parametrize("name,kw,nparts", [
    ("tidal", dict(scale=0.05), 4),
    ("bowl", dict(scale=0.1, slope_limiting=False), 3),
    ("manufactured", dict(scale=0.25), 8),
    ("bowl", dict(scale=0.1), 4),         # wetting-drying + slope limiter
    ("dambreak", dict(scale=0.06), 5),    # wetting-drying + slope limiter, a moving front
])
def test_gpu_partitioned_equals_whole(precision, name, kw, nparts):
    mesh, prm = getattr(W, name)(**kw)
    a = Simulation(mesh, prm, precision)
    b = PartitionedSimulation(mesh, prm, precision, nparts=nparts)
    for k in range(3 * prm.nstages):
        a.stage()
        b.stage()
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["wet"], sb["wet"]), k
    a.close(), b.close()
    # against the oracle stepping the same submeshes with its own ghost exchange (the serial role
    # of the reference's partitioned-vs-serial check, scripts/correctness/compare_l2_errors.py:34-38);
    # the oracle's partitioned run exchanges only between stages, so the cases with the limiter
    # (which needs the exchange between stage and limiter) compare with the whole-mesh oracle
    stages = 3 * prm.nstages
    if prm.slope_limiting:
        o = OracleSim(mesh, prm)
        for _ in range(stages):
            o.stage()
        q_ref, w_ref = o.state()["q"], o.state()["wet"]
        o.close()
    else:
        q_ref, w_ref = _oracle_partitioned(mesh, prm, nparts, stages, part=b.part)
    e = relerr(sb["q"], q_ref)
    if precision != RKDG_FP32_BULK:
        assert max(e) <= 1e-12, e
    elif not prm.wetting_drying:
        assert max(e) <= 1e-5, e
    elif not prm.slope_limiting:
        # wetting-drying workloads start from still water: after a few stages the momentum is
        # tiny, and FP32 storage of the state alone moves it by about 2e-3 of its norm (the FP64
        # oracle with its state rounded to FP32 each stage, tests/test_fp32_floor.py); the surface
        # stays within 1e-5
        mom = float(np.linalg.norm(sb["q"][1:] - q_ref[1:]) / np.linalg.norm(q_ref[1:]))
        assert e[0] <= 1e-5 and mom <= 1e-2, (e, mom)
    # FP32 with wetting-drying and the limiter: the wet/dry and limiting decisions taken on FP32
    # states differ from the FP64 ones at a few front elements of these coarse meshes, which moves
    # the state by far more than rounding; there the partitioned run is checked only against the
    # whole-mesh FP32 run (bit for bit, above)
    if precision == RKDG_FP64_PLAIN:
        assert np.array_equal(sb["wet"], w_ref)


@pytest.mark.gpu
def test_gpu_halo_errors():
    from test_levee import weir_mesh, weir_params
    s = Simulation(weir_mesh(), weir_params(), RKDG_FP64_PLAIN)
    with pytest.raises(RkdgError, match="without stage_begin"):
        s.stage_finish()
    s.stage_begin()
    with pytest.raises(RkdgError, match="twice"):
        s.stage_begin()
    s.stage_finish()
    s.close()
    mesh, prm = W.tidal(scale=0.03)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    with pytest.raises(RkdgError, match="out of range"):
        s.halo_define(0, [0, mesh.n_elements])
    with pytest.raises(RkdgError, match="not defined"):
        s.halo_pack(5)
    s.halo_define(1, [3, 4])
    with pytest.raises(RkdgError, match="already defined"):
        s.halo_define(1, [5])
    buf, nbytes = s.halo_buffer(1)
    assert buf and nbytes == 2 * (3 * 6 + 1) * 8
    with pytest.raises(RkdgError, match="smaller than"):
        s.halo_unpack(1, buf, nbytes - 8)
    s.halo_pack(1)
    s.halo_unpack(1, buf, nbytes)  # its own data back in place
    s.close()


# ---- levees across parts (the reference's distributed levee, rkdg_swe_dbc_distributed_levee.hpp)
def _weir_cut():
    """The weir fixture cut into 4 parts along the barrier (y = 15000 / 16000) and across it
    (x = 9000): every barrier edge joins elements of different parts."""
    from test_levee import weir_mesh
    mesh = weir_mesh()
    c = mesh.node_xyz[mesh.elt_nodes, :2].mean(axis=1)
    return mesh, ((c[:, 1] > 15500) + 2 * (c[:, 0] > 9000)).astype(np.int32)


def _weir_state(mesh, p):
    """Still water at ze = 1.1 below the barrier and 0.2 above: the weir discharges from the start."""
    from test_levee import _mode0_scale
    nd = (p + 1) * (p + 2) // 2
    up = mesh.node_xyz[mesh.elt_nodes, 1].mean(axis=1) < 15000
    q = np.zeros((3, nd, mesh.n_elements))
    q[0, 0] = np.where(up, 1.1, 0.2) * _mode0_scale(mesh, p)
    return q


def _levee_partners(mesh):
    """(element, face) -> element across the barrier, from the barrier node pairs."""
    edges, _ = _adjacency(mesh)
    partner = {}
    for f, b in mesh.levee_nodes.tolist():
        partner[f], partner[b] = b, f
    out = {}
    for e, nodes in enumerate(mesh.elt_nodes.tolist()):
        for k in range(3):
            a, b = nodes[(k + 1) % 3], nodes[(k + 2) % 3]
            if len(edges[frozenset((a, b))]) == 1 and a in partner and b in partner:
                pe = edges.get(frozenset((partner[a], partner[b])))
                if pe is not None:
                    out[(e, k)] = pe[0]
    return out


def test_levee_submeshes():
    mesh, part = _weir_cut()
    lp = _levee_partners(mesh)
    assert len(lp) == 36 and all(part[e] != part[x] for (e, _), x in lp.items())
    _, nb = _adjacency(mesh)
    for p in range(4):
        sub, gid = submesh(mesh, part, p)
        own = set(np.flatnonzero(part == p).tolist())
        across = {x for (e, _), x in lp.items() if e in own}
        assert across and set(gid.tolist()) == own | set().union(*(nb[e] for e in own)) | across
        # a levee face stays a levee exactly when the element across it is in the submesh
        slp = _levee_partners(sub)
        kept = {(int(gid[e]), k): int(gid[x]) for (e, k), x in slp.items()}
        want = {(e, k): x for (e, k), x in lp.items() if e in set(gid.tolist()) and x in set(gid.tolist())}
        assert kept == want and all((e, k) in kept for (e, k) in lp if e in own)
        ltype = {frozenset(map(int, be)): int(t) for be, t in zip(sub.bound_edges, sub.bound_types)}
        for (e, k), x in _levee_partners(mesh).items():
            if e in set(gid.tolist()) and x not in set(gid.tolist()):  # a ghost's barrier face
                le = int(np.flatnonzero(gid == e)[0])
                nodes = sub.elt_nodes[le].tolist()
                assert part[e] != p and ltype[frozenset((nodes[(k + 1) % 3], nodes[(k + 2) % 3]))] == BC_LAND


@pytest.mark.parametrize("cut", ["barrier", "rcb3"])
def test_oracle_partitioned_levees_equal_whole(cut):
    from test_levee import weir_params
    mesh, part = _weir_cut()
    if cut == "rcb3":
        part = partition_rcb(mesh, 3)
    prm = weir_params(p=1, dt=1.0, slope_limiting=False)
    q0 = _weir_state(mesh, 1)
    whole = OracleSim(mesh, prm)
    whole.set_state(q0)
    for _ in range(12):
        whole.stage()
    ref = whole.state()
    whole.close()
    assert np.abs(ref["q"][1:]).max() > 1e-3  # water moved over the barrier
    q, _ = _oracle_partitioned(mesh, prm, int(part.max()) + 1, 12, part=part, q0=q0)
    assert np.array_equal(q, ref["q"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP64_FAST, RKDG_FP32_BULK])
@pytest.mark.parametrize("cut,p,limiter", [("barrier", 1, False), ("barrier", 1, True), ("rcb3", 2, True)])
def test_gpu_partitioned_levees_equal_whole(precision, cut, p, limiter):
    from test_levee import weir_params
    mesh, part = _weir_cut()
    if cut == "rcb3":
        part = partition_rcb(mesh, 3)
    prm = weir_params(p=p, dt=1.0, slope_limiting=limiter)
    q0 = _weir_state(mesh, p)
    a = Simulation(mesh, prm, precision)
    b = PartitionedSimulation(mesh, prm, precision, nparts=int(part.max()) + 1, part=part)
    a.set_state(q0)
    b.set_state(q0)
    for k in range(6 * prm.nstages):
        a.stage()
        b.stage()
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["wet"], sb["wet"]), k
    assert np.abs(sa["q"][1:]).max() > 1e-4
    a.close(), b.close()


# ---- locality renumbering (rkdg_mesh_renumber)
def _close(mesh, window=256):
    """Share of neighbour pairs less than `window` elements apart (one work-group's tile)."""
    _, nb = _adjacency(mesh)
    d = np.array([abs(e - n) for e, s in nb.items() for n in s])
    return float(np.mean(d < window))


def test_renumber_locality_and_same_mesh():
    from dgswemv2_amd import renumber
    mesh, prm = W.tidal(scale=0.05)
    sh, pe = W.shuffled(mesh)
    rn, old = renumber(sh)
    assert sorted(old.tolist()) == list(range(mesh.n_elements))
    # the same triangles (by coordinates), and neighbours close in index again
    assert np.array_equal(rn.node_xyz[rn.elt_nodes], sh.node_xyz[sh.elt_nodes[old]])
    assert _close(rn) > 0.85 and _close(sh) < 0.25, (_close(rn), _close(sh))
    # the same discretisation: the oracle's runs agree to rounding, element by element
    a = OracleSim(mesh, prm)
    b = OracleSim(rn, prm)
    a.step(4), b.step(4)
    qa, qb = a.state()["q"], b.state()["q"]
    qa_in_rn = qa[:, :, pe[old]]
    for k in range(3):
        assert np.linalg.norm(qb[k] - qa_in_rn[k]) <= 1e-12 * np.linalg.norm(qa[k]), k
    a.close(), b.close()


def test_renumber_weir_levees():
    from dgswemv2_amd import renumber
    from test_levee import weir_mesh, weir_params
    mesh = weir_mesh()
    rn, old = renumber(mesh)
    prm = weir_params(p=1, slope_limiting=False)
    prm.ze0 = 0.5
    a, b = OracleSim(mesh, prm), OracleSim(rn, prm)
    a.step(5), b.step(5)
    qa, qb = a.state()["q"][:, :, old], b.state()["q"]
    assert np.abs(qa[1:]).max() > 0
    # rounding-level differences against the size of the state (the momenta are small here)
    assert np.linalg.norm(qb - qa) <= 1e-13 * np.linalg.norm(qa)
    a.close(), b.close()
