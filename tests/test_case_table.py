"""MASTER.dat case tables (Driver_Input, "8 SPH_FEM_Driver_2019.f90":475-1236)
expanded by the host reader into OC_mat, checked against the closed forms of
the reference's construction: ic_cases_mat 1 and 2 (linear and logarithmic
value grids, the statistic lists, the mixed-radix order with the first
variable fastest), 20 (the FOSM sets); the sampled sets 3, 4, 31, 32 through
the distribution function of every law at the uniforms that drew each value
(the reference draws from a clock-seeded generator, so its values are not
reproducible: parity of the draws themselves is unpinned by construction).
No device work."""
import math

import numpy as np
import pytest

import goldens
from gfsph import GfsphError, Inputs
from oracle import refio


def _inputs(tmp_path, **kw):
    case, _ = goldens.case_of("g_voellmy_plane")
    refio.write_inputs(str(tmp_path), case, name="case", topo="topo")
    refio.write_master(str(tmp_path), "case", case.params.dt, **kw)
    return Inputs(str(tmp_path), "case", "topo", build=False)


def test_full_factorial_table(tmp_path):
    rows = [[9, 0.2, 3, 0.1, 0.3, 1],            # mu: linear, 3 values
            [3, 500.0, 2, 200.0, 800.0, 2],      # xi: logarithmic, 2 values
            [501, 1.0, 2, 0.0, 0.0, 100]]        # depth factor: statistic list
    lists = [([501, 0.8, 1.25], [501, 0.5, 0.5])]
    with _inputs(tmp_path, cases=2, case_rows=rows, lists=lists, stop=(0.05, 60.0, 1e-4)) as inp:
        t = inp.cases()
        assert (t["mode"], t["ncases"]) == (2, 3)
        assert (t["vstop"], t["timestop"], t["dtimestop"]) == (0.05, 60.0, 1e-4)
        T = t["table"]
        assert T.shape == (12, 3 + 3 + 2)
        mu = [0.1 + (0.3 - 0.1) / 2 * k for k in range(3)]
        lx = [math.log(200.0), math.log(800.0)]
        xi = [math.exp(lx[0] + (lx[1] - lx[0]) / 1 * k) for k in range(2)]
        fac = [0.8, 1.25]
        p = 0
        for c in range(2):
            for b in range(2):
                for a in range(3):  # variable 1 fastest
                    assert list(T[p, 3:6]) == [mu[a], xi[b], fac[c]], p
                    p += 1


def test_one_at_a_time_table(tmp_path):
    rows = [[9, 0.2, 3, 0.1, 0.3, 1, 0, 0], [521, 0.0, 2, -5.0, 5.0, 1, 2.0, 4.0]]
    with _inputs(tmp_path, cases=1, case_rows=rows) as inp:
        T = inp.cases()["table"]
        assert T.shape == (5, 3 + 2 + 2)
        np.testing.assert_array_equal(T[:, 0], [1, 1, 1, 2, 2])
        np.testing.assert_array_equal(T[:, 1], [1, 2, 3, 1, 2])
        np.testing.assert_array_equal(T[:, 2], [9, 9, 9, 521, 521])
        np.testing.assert_array_equal(T[:3, 3], [0.1 + (0.3 - 0.1) / 2 * k for k in range(3)])
        np.testing.assert_array_equal(T[3:, 3], [0.2, 0.2])   # the other variable at its default
        np.testing.assert_array_equal(T[3:, 4], [-5.0, 5.0])
        np.testing.assert_array_equal(T[:3, 4], [0.0, 0.0, 0.0])


def _uniforms(seed, k, n):
    """The sampled sets' uniforms of variable k (the generator the library
    documents: SplitMix64 from state seed + 0x9E3779B97F4A7C15 (k + 1), output
    u giving ((u >> 11) + 0.5) 2^-53)."""
    M, G = 1 << 64, 0x9E3779B97F4A7C15
    s = (seed + G * (k + 1)) % M
    out = []
    for _ in range(n):
        s = (s + G) % M
        z = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) % M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % M
        z ^= z >> 31
        out.append(((z >> 11) + 0.5) * 2.0 ** -53)
    return np.array(out)


def _phi(t):
    """The normal distribution function as the reference's sampler
    approximates it (the inverse it takes, 8:1323-1339): (1 + sign(t)
    sqrt(1 - exp(-5 t^2 / 8))) / 2."""
    return 0.5 * (1.0 + np.sign(t) * np.sqrt(1.0 - np.exp(-0.625 * t * t)))


def _tri_cdf(x, a, b, c):
    return np.where(x <= b, (x - a) ** 2 / ((b - a) * (c - a)), 1.0 - (c - x) ** 2 / ((c - b) * (c - a)))


LAWS = [(0, [0.21]),                        # constant
        (1, [0.1, 0.15, 0.35]),             # triangular (min, mode, max)
        (2, [600.0, 150.0]),                # normal (mean, sigma)
        (3, [400.0, 200.0]),                # log-normal (mean, sigma)
        (100, [0.8, 1.0, 1.3], [1.0, 3.0])]  # histogram: nodes, counts


def _check_draws(T, col, law, y):
    """The values of one sampled column against its law's distribution
    function at the uniforms that drew them."""
    x = T[:, col]
    if law[0] == 0:
        assert np.all(x == law[1][0])
    elif law[0] == 1:
        np.testing.assert_allclose(_tri_cdf(x, *law[1]), y, rtol=0, atol=1e-12)
    elif law[0] in (2, 3):
        m, s = law[1]
        if law[0] == 2:
            t = (x - m) / s
        else:  # log x is normal with the mean and variance that give X the mean m and sigma s
            v = np.log(1.0 + (s / m) ** 2)
            t = (np.log(x) - (np.log(m) - 0.5 * v)) / np.sqrt(v)
        np.testing.assert_allclose(_phi(t), np.clip(y, 1e-5, 0.99999), rtol=0, atol=1e-12)
    else:
        xh, cnt = np.array(law[1]), np.array(law[2])
        F = np.concatenate([[0.0], np.cumsum(cnt)]) / cnt.sum()
        np.testing.assert_allclose(np.interp(x, xh, F), y, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", [4, 5, 31])
def test_sampled_values(tmp_path, mode):
    """ic_cases_mat 4, 5 (one geological zone) and 31 (8:693-857,
    8:1053-1136): every variable's
    n_MCarlo_problems values inverted through its law's distribution
    function give back the uniforms that drew them; another seed draws
    others, the same seed the same."""
    rows = [[9, 0.2, 1, 0, 0, 1], [3, 500, 1, 0, 0, 1], [3, 500, 1, 0, 0, 1], [3, 500, 1, 0, 0, 1],
            [501, 1.0, 1, 0, 0, 1]]
    with _inputs(tmp_path, cases=mode, case_rows=rows, stop=(0.05, 60.0, 0.0), nmc=300, laws=LAWS) as inp:
        t = inp.cases()
        assert (t["mode"], t["ncases"]) == (mode, 5) and t["table"].shape == (300, 3 + 5 + 2)
        T = t["table"]
        for k, law in enumerate(LAWS):
            _check_draws(T, 3 + k, law, _uniforms(1, k, 300))
        inp.resample(77)
        T2 = inp.cases()["table"]
        for k, law in enumerate(LAWS):
            _check_draws(T2, 3 + k, law, _uniforms(77, k, 300))
        assert not np.array_equal(T2[:, 4], T[:, 4])
        inp.resample(1)
        np.testing.assert_array_equal(inp.cases()["table"], T)


def test_sampled_set_32(tmp_path):
    """ic_cases_mat 32 (8:1138-1234): the records with a law > 0 are the
    variables, in order; continuous laws."""
    recs = [(9, 1, [0.1, 0.2, 0.3]), (7, 0, [1.0]), (3, 3, [400.0, 100.0]), (501, 2, [1.0, 0.1])]
    with _inputs(tmp_path, cases=32, stop=(0.05, 60.0, 0.0), nmc=64, records=recs) as inp:
        t = inp.cases()
        assert (t["mode"], t["ncases"]) == (32, 3)
        T = t["table"]
        for k, rec in enumerate([r for r in recs if r[1] > 0]):
            _check_draws(T, 3 + k, (rec[1], rec[2]), _uniforms(1, k, 64))


def test_sampled_rows_of_the_table(tmp_path):
    """ic_cases_mat 3 (8:589-634): n_MCarlo_problems rows of the full
    factorial table at indices nint(y * nproblems) kept in 1 .. nproblems."""
    rows = [[9, 0.2, 3, 0.1, 0.3, 1], [3, 500.0, 4, 200.0, 800.0, 1]]
    with _inputs(tmp_path, cases=2, case_rows=rows) as inp:
        full = inp.cases()["table"]
    d = tmp_path / "s"
    d.mkdir()
    with _inputs(d, cases=3, case_rows=rows, nmc=10) as inp:
        t = inp.cases()
        T = t["table"]
    assert t["mode"] == 3 and T.shape == (10, full.shape[1])
    y = _uniforms(1, 0, 10)
    idx = np.clip(np.floor(y * 12 + 0.5).astype(int), 1, 12)
    np.testing.assert_array_equal(T, full[idx - 1])
    d2 = tmp_path / "s2"
    d2.mkdir()
    with _inputs(d2, cases=3, case_rows=rows, nmc=40) as inp:  # at most nproblems draws (8:593-596)
        assert inp.cases()["table"].shape == (12, full.shape[1])


def test_fosm_sets(tmp_path):
    """ic_cases_mat 20 (8:974-1051): 2 ncases + 1 problems -- the means, then
    for each variable c two rows with every column set to the mean of c (the
    reference fills the other variables with variable c's mean, 8:1039-1045)
    and column c at mean +- sigma/10.  The reference's own main program runs
    the same problems (test_oracle_case_driver.py, table t20)."""
    rows = [[9, 0.2, 1, 0, 0, 1], [3, 500.0, 1, 0, 0, 1]]
    with _inputs(tmp_path, cases=20, case_rows=rows, fosm=((0.2, 500.0), (0.5, 2000.0))) as inp:
        T = inp.cases()["table"][:, 3:5]
    np.testing.assert_allclose(T, [[0.2, 500.0], [0.25, 0.2], [0.15, 0.2], [500.0, 700.0], [500.0, 300.0]], rtol=1e-15)


def test_rejected_case_sets(tmp_path):
    with pytest.raises(GfsphError) as e:
        _inputs(tmp_path, cases=6, case_rows=[[9, 0.2, 3, 0.1, 0.3, 1]])
    assert e.value.code == -2 and "ic_cases_mat 6" in str(e.value)
    d = tmp_path / "b"
    d.mkdir()
    with pytest.raises(GfsphError) as e:
        _inputs(d, cases=4, case_rows=[[9, 0.2, 1, 0, 0, 1]], nmc=4, laws=[(7, [0.1])])
    assert e.value.code == -2 and "probability law 7" in str(e.value)
