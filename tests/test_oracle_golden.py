"""The oracle against the golden fixtures written by the compiled reference.

Every comparison is bit-exact (numpy.array_equal on the raw float bits): the
oracle restates the reference's arithmetic operation for operation, so any
difference means a dropped term, a reordered sum or a wrong narrowing point.
"""
import numpy as np
import pytest

from conftest import SCENES, golden


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", SCENES)
def test_sam_matches_reference(oracle_lib, name):
    g = golden(name)
    o = oracle_lib
    assert same_bits(o.library_norms(g["library"]), g["norms"])
    score, target = o.sam_best(g["cube"], g["library"], g["norms"])
    assert same_bits(score, g["sam_score"])
    assert same_bits(target, g["sam_target"])
    assert same_bits(o.sam_all(g["cube"], g["library"], g["norms"]), g["sam_all"])


@pytest.mark.parametrize("name", SCENES)
def test_moments_and_statistics_match_reference(oracle_lib, name):
    g = golden(name)
    o = oracle_lib
    count, s, so = o.moments(g["cube"])
    assert count == g["m_count"]
    assert same_bits(s, g["m_sum"])
    assert same_bits(so, g["m_outer"])
    st = o.build_statistics(count, s, so, g["library"], float(g["loading"]))
    assert st.valid, st.error
    assert same_bits(st.mean, g["st_mean"])
    assert same_bits(st.whitener, g["st_whitener"])
    assert same_bits(st.target_white, g["st_target_white"])
    assert same_bits(st.target_norm2, g["st_target_norm2"])
    assert same_bits(st.cem_weight, g["st_cem_weight"])


@pytest.mark.parametrize("name", SCENES)
def test_second_order_detectors_match_reference(oracle_lib, name):
    g = golden(name)
    o = oracle_lib
    st = o.statistics_of_cube(g["cube"], g["library"], float(g["loading"]))
    assert same_bits(o.rx(g["cube"], st), g["rx_score"])
    ace, acet, acerx = o.ace_best(g["cube"], st)
    assert same_bits(ace, g["ace_score"])
    assert same_bits(acet, g["ace_target"])
    assert same_bits(acerx, g["ace_rx"])
    cem, cemt = o.cem_best(g["cube"], st)
    assert same_bits(cem, g["cem_score"])
    assert same_bits(cemt, g["cem_target"])


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("det", ["sam", "rx", "ace", "cem"])
@pytest.mark.parametrize("nms", [0, 2])
def test_detection_lists_match_reference(oracle_lib, name, det, nms):
    g = golden(name)
    _, h, w = g["cube"].shape
    score = g[f"{det}_score"]
    target = None if det == "rx" else g[f"{det}_target"]
    higher = det != "sam"
    dets, found = oracle_lib.detect(score, target, h, w, float(g[f"{det}_threshold"]), nms, higher)
    want_xyt = g[f"{det}_det{nms}_xyt"]
    assert found == want_xyt.shape[0]
    got_xyt = np.stack([dets["x"], dets["y"], dets["target"]], axis=1) if found else want_xyt
    assert np.array_equal(got_xyt, want_xyt)
    assert same_bits(dets["score"].astype(np.float32), g[f"{det}_det{nms}_score"])


def test_edge_case_sam(oracle_lib):
    g = golden("edge_cases")
    score, target = oracle_lib.sam_best(g["edge_cube"], g["edge_library"], g["edge_norms"])
    assert same_bits(score, g["edge_sam_score"])
    assert same_bits(target, g["edge_sam_target"])
    assert same_bits(oracle_lib.sam_all(g["edge_cube"], g["edge_library"], g["edge_norms"]),
                     g["edge_sam_all"])


@pytest.mark.parametrize("case", ["blob", "plateau", "pol_low", "pol_high", "extremum",
                                  "ties_low", "ties_high"])
def test_edge_case_detect(oracle_lib, case):
    g = golden("edge_cases")
    mp = g[f"det_{case}_map"]
    thr, nms, higher = g[f"det_{case}_args"]
    h, w = mp.shape
    tg = (np.arange(h * w) % 3).astype(np.int32)
    dets, found = oracle_lib.detect(mp, tg, h, w, float(thr), int(nms), bool(higher))
    want = g[f"det_{case}_xyt"]
    assert found == want.shape[0]
    assert np.array_equal(np.stack([dets["x"], dets["y"], dets["target"]], axis=1), want)
    assert same_bits(dets["score"].astype(np.float32), g[f"det_{case}_score"])


@pytest.mark.parametrize("n", [1, 2, 5, 17, 40])
def test_linear_algebra_matches_reference(oracle_lib, n):
    g = golden("edge_cases")
    ok, l = oracle_lib.cholesky(g[f"la{n}_a"])
    assert ok
    assert same_bits(l, g[f"la{n}_l"])
    assert same_bits(oracle_lib.triangular_inverse_lower(l), g[f"la{n}_inv"])
    assert same_bits(oracle_lib.cholesky_solve(l, g[f"la{n}_rhs"]), g[f"la{n}_x"])


def test_singular_and_chi2_match_reference(oracle_lib):
    g = golden("edge_cases")
    ok, _ = oracle_lib.cholesky(np.array([[1.0, 1.0], [1.0, 1.0]]))
    assert int(ok) == int(g["la_singular_ok"]) == 0
    for i, b in enumerate(g["chi2_bands"]):
        for j, r in enumerate(g["chi2_rates"]):
            assert oracle_lib.rx_threshold(int(b), float(r)) == g["chi2_threshold"][i, j]
    cube = g["sing_cube"]
    for key, loading in (("sing_l0", 0.0), ("sing_l3", 1e-3)):
        st = oracle_lib.statistics_of_cube(cube, None, loading)
        assert int(st.valid) == int(g[f"{key}_ok"])
        want_err = bytes(g[f"{key}_error"]).decode()
        assert st.error == want_err
        if st.valid:
            assert same_bits(st.whitener, g[f"{key}_whitener"])
            assert same_bits(st.mean, g[f"{key}_mean"])
        else:
            assert "loading" in st.error


def test_mean_spectrum_matches_reference(oracle_lib):
    g = golden("edge_cases")
    for i, mask in enumerate(g["ms_masks"]):
        if g["ms_ok"][i]:
            assert same_bits(oracle_lib.mean_spectrum(g["ms_cube"], mask), g["ms_mean"][i])
        else:
            with pytest.raises(ValueError, match="no pixels"):
                oracle_lib.mean_spectrum(g["ms_cube"], mask)
    assert list(g["ms_ok"]) == [1, 1, 1, 0]


def test_moments_from_shifted_matches_reference(oracle_lib):
    g = golden("edge_cases")
    s, so = oracle_lib.moments_from_shifted(g["mfs_s1"], g["mfs_s2"], 200.0, g["mfs_ref"])
    assert same_bits(s, g["mfs_sum"])
    assert same_bits(so, g["mfs_outer"])
    z, zo = oracle_lib.moments_from_shifted(g["mfs_s1"], g["mfs_s2"], 0.0, g["mfs_ref"])
    assert same_bits(z, g["mfs_zero_sum"]) and same_bits(zo, g["mfs_zero_outer"])
    assert not z.any() and not zo.any()


@pytest.mark.parametrize("n", [1, 3, 12, 33])
def test_cholesky_inverse_and_quadratic_form_match_reference(oracle_lib, n):
    g = golden("edge_cases")
    ok, l = oracle_lib.cholesky(g[f"ci{n}_a"])
    assert ok and same_bits(l, g[f"ci{n}_l"])
    assert same_bits(oracle_lib.cholesky_inverse(l), g[f"ci{n}_inv"])
    q = oracle_lib.quadratic_form(g[f"ci{n}_a"], g[f"ci{n}_y"])
    assert same_bits(np.array([q], np.float64), np.array([g[f"ci{n}_q"]], np.float64))


def test_near_singular_scene_matches_reference(oracle_lib):
    """The reference factorises this covariance although its smallest pivot is
    a rounding-level fraction of its diagonal; the oracle must agree bit for
    bit, and the pivot must lie in (0, 4*256*eps] of its diagonal entry - the
    band the device rejects where the reference does not."""
    g = golden("edge_cases")
    cube = g["nsing_cube"]
    for key, loading in (("nsing_l0", 0.0), ("nsing_l3", 1e-3)):
        st = oracle_lib.statistics_of_cube(cube, None, loading)
        assert st.valid, st.error
        assert same_bits(st.whitener, g[f"{key}_whitener"])
        assert same_bits(st.mean, g[f"{key}_mean"])
    count, s, so = oracle_lib.moments(cube)
    cov = oracle_lib.covariance(count, s, so, 0.0)
    ok, l = oracle_lib.cholesky(cov)
    rel = np.diag(l) ** 2 / np.diag(cov)
    tol = 4.0 * 256 * np.finfo(np.float64).eps
    assert ok and 0.0 < rel.min() <= tol
