"""Terrain zones (ic_topo_props = 10) against the reference's own code.

g_zones_map holds what the reference's Topo_main and Which_Topo_Zone
("2 SPH_FEM_topo_2019.f90":516-576, 2635-2808) give on the 81 x 41 plane:
the zone (basal types 1 and 2) at 20,909 points of a 2 m lattice that
overhangs the terrain by 2 m, and topol(1), topol(15) after Get_Eros_Pts
(2:909-1110) for erodible-point files of type_Eros 1, -1, -2, -3.  The
library's Get_Eros_Pts (host code, gfsph_terrain_eros_pts) must reproduce
topol(1) and topol(15) bit for bit, and with its node lists the oracle's
zone lookup must give the reference's zone at every point -- geometries 1
(a range of the terrain Z), 2, 4, 5 (turned by +30 and -35 degrees) and 101.

The state fixtures g_zones_friction, g_zones_erosion, g_zones_erosion_main
and g_zones_pwp (reference runs, checked against the oracle with the
other state fixtures in test_oracle_golden.py) pin the zones' effect on the
flow; the last test here shows that each of their zones changes the
reference's state, so a zone the oracle ignored or mislaid would fail.
"""
import copy

import numpy as np
import pytest

import goldens
from gfsph.dem import DEM, TerrainZones, Zone
from oracle import oracle


def _map():
    f = goldens.load("g_zones_map")
    zones = [Zone(int(t), int(l), tuple(p)) for t, l, p in zip(f["zone_type"], f["zone_law"], f["zone_props"])]
    return f, zones


@pytest.mark.parametrize("tag,typ", [("t1", 1), ("tm1", -1), ("tm2", -2), ("tm3", -3)])
def test_erodible_points_match_reference(tag, typ):
    from gfsph import api
    f, zones = _map()
    dem = DEM(Z=f["dem_Z"], xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    Z, avail, listed, box = api.terrain_eros_pts(dem, f[f"{tag}_pts"], typ, 6.0)
    np.testing.assert_array_equal(avail, f[f"{tag}_avail"])
    np.testing.assert_array_equal(Z.ravel(), f[f"{tag}_z"])
    assert np.count_nonzero(listed) > 50
    if typ == 1:
        np.testing.assert_array_equal(Z, f["dem_Z"])
    else:  # the terrain changed where the points reached
        assert not np.array_equal(Z, f["dem_Z"])


@pytest.mark.parametrize("tag", ["t1", "tm2"])
def test_zone_lookup_matches_reference(tag):
    from gfsph import api
    f, zones = _map()
    dem = DEM(Z=f["dem_Z"], xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    Z, avail, listed, box = api.terrain_eros_pts(dem, f[f"{tag}_pts"], -2 if tag == "tm2" else 1, 6.0)
    dem = DEM(Z=Z, xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    k101 = [k for k, z in enumerate(zones) if int(z.props[0]) == 101][0]
    dem.zones = TerrainZones(zones, nodes=np.where(listed, 1 << k101, 0).astype(np.int32), avail=avail)
    P = f["points"]
    inside = (P[:, 0] >= 0) & (P[:, 0] <= 400) & (P[:, 1] >= 0) & (P[:, 1] <= 200)
    for btype, key in ((1, "zone1"), (2, "zone2")):
        ref = f[f"{tag}_{key}"]
        ref = np.where(ref < 0, -1, ref - 1)
        mine = oracle.which_zone(dem, P, btype)
        # off the terrain the reference's Z lookup of geometry 1 reads outside
        # topol (Get_Z_Topo only warns, 2:2421-2435): those points are not
        # compared for the type with a Z-range zone
        sel = inside if btype == 1 else np.ones_like(inside)
        assert np.array_equal(mine[sel], ref[sel]), (btype, np.count_nonzero(mine[sel] != ref[sel]))
        # every zone of the type is hit somewhere, and some points are in none
        want = sorted(k for k, z in enumerate(zones) if z.type == btype)
        assert sorted(set(ref.tolist()) - {-1}) == want and np.any(ref == -1)


def _without(case, k):
    """The case with zone k switched off: removed when it only shadows later
    zones (a friction zone with tan(phi) < 0, an erosion zone with law 0) or
    is a pore-pressure zone, else its tan(phi) or erosion law changed."""
    c2 = copy.deepcopy(case)
    zs = list(c2.dem.zones.zones)
    z = zs[k]
    if (z.type == 1 and z.props[6] < 0) or (z.type == 2 and z.law == 0) or z.type == 3:
        del zs[k]
        nodes = c2.dem.zones.nodes
        if nodes is not None:  # node-list bits follow their zones
            low = nodes & ((1 << k) - 1)
            c2.dem.zones.nodes = (low | ((nodes >> (k + 1)) << k)).astype(np.int32)
    elif z.type == 1:
        p = list(z.props)
        p[6] = -1.0
        zs[k] = Zone(1, 0, tuple(p))
    else:
        zs[k] = Zone(2, 0, z.props)
    c2.dem.zones.zones = zs
    return c2


@pytest.mark.parametrize("name", goldens.ZONE_CASES)
def test_every_zone_changes_the_reference_state(name):
    case, f = goldens.case_of(name)
    s = int(f["dumps"][-1])
    for k in range(len(case.dem.zones.zones)):
        c2 = _without(case, k)
        st = oracle.init(c2)
        oracle.run(c2, st, s)
        err = goldens.rel_err(st.v, f[f"s{s}_v"])
        if "s100_mass" in f:
            err = max(err, goldens.rel_err(st.mass, f[f"s{s}_mass"]))
        assert err > 1e-4, (name, k, err)


def test_terrain_file_zones_read_as_the_reference(tmp_path):
    """The terrain file's zone block and <terrain>.eros.pts through the
    library's reader (gfsph_inputs_read, an explicit grid: host code only):
    the zones as written, topol(15) and the zone lookup equal to the
    reference's Topo_main on the same files (g_zones_map, type_Eros 1)."""
    from gfsph import cases as C
    from gfsph.inputs import Inputs
    from oracle import refio
    f, zones = _map()
    dem = DEM(Z=f["dem_Z"], xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    dem.zones = TerrainZones(zones, eros_pts={"pts": f["t1_pts"], "type": 1, "dx_s": 6.0})
    case = C.Case("zones", C.Params(integrator=4, dt=0.05, h_inf=1e-3, const=C.voellmy(0.2, 500.0)), dem,
                  C.lattice_block(60.0, 80.0, 30.0, 20.0, 15, 10, 5.0, 1.5), 0)
    refio.write_inputs(str(tmp_path), case)
    inp = Inputs(str(tmp_path), "case", "topo")
    tz = inp.zones()
    assert len(tz.zones) == len(zones)
    for a, b in zip(tz.zones, zones):
        assert a.type == b.type and a.law == b.law
        if int(b.props[0]) != 101:  # 101 zones take their points' box as props 2-5 (2:989-992)
            assert a.props == b.props
    np.testing.assert_array_equal(tz.avail, f["t1_avail"])
    d = DEM(Z=f["dem_Z"], xmin=0.0, ymin=0.0, dx=5.0, dy=5.0)
    d.zones = tz
    P = f["points"]
    ref = f["t1_zone2"]
    assert np.array_equal(oracle.which_zone(d, P, 2), np.where(ref < 0, -1, ref - 1))
    inp.close()
