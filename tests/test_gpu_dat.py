"""GPU parity of the Planar-DAT truncation passes (predictor pass, per-color
passes, failed-surface restore): the reference's analytic division-plane
tests (PlanarDATTests.swift) on the HIP backend through the C ABI, each
checked against the closed form and against the oracle (positions within
1e-6 m, counters equal), plus a fired-sheets no-pass-through run and a
per-color parity run on a settled two-sheet stack."""
import numpy as np
import pytest

import oracle
from avbd_amd import GPUSolver, RuntimeFailure
from conftest import flush_print
from dat_fixtures import (RELAX, bilateral, edge_edge, fraction, place, sheets_fired, small_triangles, strictest,
                          two_triangles)

pytestmark = pytest.mark.gpu


def _case(name):
    flush_print(f"[gpu case] {name}")


def _pair(scene, rq=0.0, pairCap=0, softCap=0, setup=None, **kw):
    g = GPUSolver(scene, planarDATPairCapacityForTesting=pairCap, planarDATSoftCapacityForTesting=softCap, **kw)
    o = oracle.OracleSolver(scene)
    if rq > 0:
        g.setDATQueryRadiusForTesting(rq)
    o.setDatTestHooks(queryRadius=rq, pairCapacity=pairCap, softCapacity=softCap)
    if setup:
        setup(g)
        setup(o)
    return g, o


def _step_both(g, o, expect_failure=None):
    o.step()
    if expect_failure is None:
        g.step()
        g.synchronize()
        return None
    with pytest.raises(RuntimeFailure) as e:
        g.step()
        g.synchronize()
    assert e.value.kind == expect_failure, e.value
    assert e.value.frame == 1
    return e.value


def _agree(g, o, atol=1e-6):
    pg, po = g.positions(), o.positions()
    err = float(np.abs(pg - po).max())
    flush_print(f"  max|x_gpu - x_oracle| = {err:.3e}")
    assert err <= atol
    st, d = g.stats(), o.datStats()
    assert st.dat_truncations == d["dat_truncations"]
    assert st.dat_invalid_anchors == d["dat_invalid_anchors"]
    assert st.dat_vt_degeneracies == d["dat_vt_degeneracies"]
    assert st.dat_ee_degeneracies == d["dat_ee_degeneracies"]
    assert st.dat_nonfinite == d["dat_nonfinite"]
    cs = o.clothStats()
    assert st.dat_pairs == cs["dat_pairs"] and st.dat_vt_pairs == cs["dat_vt_pairs"]
    return pg


def test_vt_predictor_analytic_plane():
    _case("Planar-DAT: V-T predictor plane (PlanarDATTests.swift:272)")
    s, _, src = two_triangles(0.2, (0, 0, -0.4))
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2 - 0.4 * fraction((0.2 - 0.01) / 0.4), atol=2e-5)
    assert g.stats().dat_vt_pairs >= 3 and g.stats().dat_truncations >= 3


def test_bilateral_and_exact_hit_and_strictest_plane():
    _case("Planar-DAT: bilateral plane, exact hit, strictest plane (:768-850)")
    s, tgt, src = bilateral()
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2 - 0.3 * 0.5 * RELAX, atol=2e-5)
    np.testing.assert_allclose(p[tgt, 2], 0.1 * 0.5 * RELAX, atol=2e-5)
    s, _, src = two_triangles(0.2, (0, 0, -0.19))
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2 - 0.19 * RELAX, atol=2e-5)
    assert np.all(p[src, 2] > 0.01)
    s, src = strictest()
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2 - 0.4 * fraction((0.2 - 0.105) / 0.4), atol=2e-5)


def test_tangential_motion_and_translation_invariance():
    _case("Planar-DAT: tangential progress, translated scene (:852-914)")
    s, _, src = two_triangles(0.05, (0.1, 0, 0))
    starts = np.array([s.bodies[i]["position"] for i in src], np.float32)
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o)
    assert np.all(p[src, 0] - starts[:, 0] > 0.095)
    np.testing.assert_allclose(p[src, 2], 0.05, atol=1e-5)
    off = (10000.0, -10000.0, 10000.0)
    s, _, src = two_triangles(0.2, (0, 0, -0.4), offset=off)
    g, o = _pair(s)
    _step_both(g, o)
    p = _agree(g, o, atol=1e-3)
    np.testing.assert_allclose(p[src, 2] - off[2], 0.2 - 0.4 * fraction(0.95 * 0.2 / 0.4), atol=0.002)


@pytest.mark.parametrize("vz", [-0.4, -0.19])
def test_edge_edge_analytic_plane(vz):
    _case(f"Planar-DAT: E-E predictor plane, vz={vz} (:916-975)")
    s, src = edge_edge(velocity_z=vz)
    g, o = _pair(s, rq=0.5)
    _step_both(g, o)
    p = _agree(g, o)
    assert g.stats().dat_vt_pairs == 0 and g.stats().dat_ee_pairs > 0
    want = 0.2 - 0.4 * fraction(0.95 * 0.2 / 0.4) if vz == -0.4 else 0.2 - 0.19 * RELAX
    np.testing.assert_allclose(p[src[:2], 2], want, atol=2e-5)


def test_intersecting_edges_fallback_normal():
    _case("Planar-DAT: crossing edge anchors take the fallback normal (:977-1001)")
    s, src = edge_edge(height=0.0, velocity_z=0.0)
    g, o = _pair(s, rq=0.3)
    _step_both(g, o)
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.0, atol=1e-6)
    st = g.stats()
    assert st.dat_ee_degeneracies > 0 and st.dat_nonfinite == 0
    assert g.runtimeFailure is None


def test_failures_freeze_or_restore_the_surface():
    _case("Planar-DAT: non-finite prediction, grid span, pair and soft overflow (:652, 1312-1432)")
    s, _, src = two_triangles(0.2, (0, 0, 0))
    g, o = _pair(s, setup=lambda x: place(x, s, src, velocity=(0, 0, np.inf)))
    f = _step_both(g, o, "planarDATInvalidAnchor")
    assert "non-finite" in f.message
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2, atol=1e-6)

    s, tgt, src = two_triangles(0.2, (0, 0, 0))
    g, o = _pair(s, setup=lambda x: place(x, s, [tgt[0]], x=-100.0))
    f = _step_both(g, o, "planarDATElementGridSpan")
    assert "element AABBs" in f.message
    p = _agree(g, o)
    np.testing.assert_allclose(p[tgt[0], 0], -100.0, atol=1e-6)
    np.testing.assert_allclose(p[src, 2], 0.2, atol=1e-6)

    s, _, src = two_triangles(0.2, (0, 0, -0.4), iterations=2)
    g, o = _pair(s, pairCap=1)
    f = _step_both(g, o, "planarDATPairCapacity")
    assert f.required == g.stats().dat_pairs and f.capacity == 1
    p = _agree(g, o)
    np.testing.assert_allclose(p[src, 2], 0.2, atol=1e-6)

    s, _, src = two_triangles(0.04, (0.1, 0, 0), iterations=2)
    starts = np.array([s.bodies[i]["position"] for i in src], np.float32)
    g, o = _pair(s, softCap=1)
    f = _step_both(g, o, "softContactCapacity")
    assert f.required > 1 and f.capacity == 1
    np.testing.assert_allclose(g.positions()[src], starts, atol=1e-6)


def test_accepted_pose_queries():
    _case("Planar-DAT: accepted-pose query drops / finds pairs (:474-563)")
    s, _, src = two_triangles(0.5, (0, 0, 0))
    g, o = _pair(s)
    rq = o.clothInfo()["dat_query_radius"]
    assert g.clothInfo()["dat_query_radius"] == rq
    for x in (g, o):
        place(x, s, src, z=0.8 * rq, velocity=(0, 0, 2 * rq))
    _step_both(g, o)
    p = _agree(g, o)
    assert g.stats().dat_pairs > 0 and len(g.datPairs()) == 0
    np.testing.assert_allclose(p[src, 2], 0.8 * rq + 0.5 * RELAX * rq, atol=2e-5)

    s, _, src = small_triangles()
    g, o = _pair(s)
    rq = o.clothInfo()["dat_query_radius"]
    for x in (g, o):
        place(x, s, src, z=1.01 * rq, velocity=(0, 0, -2 * rq))
    _step_both(g, o)
    p = _agree(g, o)
    assert g.stats().num_soft >= 3
    np.testing.assert_allclose(p[src, 2], 1.01 * rq - 0.5 * RELAX * rq, atol=2e-5)
    assert np.array_equal(g.datPairs(), o.datPairs())


def test_fired_sheets_per_color_passes_match_oracle():
    _case("Planar-DAT: fired sheets, 6 steps x 10 iterations of per-color passes vs oracle")
    s, (lower, upper) = sheets_fired()
    g = GPUSolver(s, lanesPerBody=8, sweepMode=1)
    o = oracle.OracleSolver(s)
    worst, truncs = 0.0, []
    for k in range(6):
        g.step()
        o.step()
        pg, po = g.positions(), o.positions()
        worst = max(worst, float(np.abs(pg - po).max()))
        assert pg[upper, 2].min() > pg[lower, 2].max() - 1e-4, k
        assert np.array_equal(g.datPairs(), o.datPairs()), k
        st, d = g.stats(), o.datStats()
        assert st.dat_truncations == d["dat_truncations"], k
        assert st.dat_invalid_anchors == 0, k
        truncs.append(st.dat_truncations)
    flush_print(f"  truncations per step {truncs}, worst |dx| {worst:.3e} m")
    assert truncs[0] > 0      # the first step's predictor and sweeps are truncated
    assert worst <= 1e-5
