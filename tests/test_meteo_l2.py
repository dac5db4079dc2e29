"""Forcing on the meteo grid (L2) mapped to the L1 cells.

meteo_forcings_wrapper (src/meteo/mo_meteo_helper.f90:119-143) maps each
forcing chunk from the meteo grid to L1: block averages over the valid L2
cells when L1 is coarser (spatial_aggregation_3d,
src/meteo/mo_meteo_spatial_tools.f90:94-200), the parent L2 cell or nodata
when L1 is finer (spatial_disaggregation_3d, 313-377), a copy for equal
cell sizes, then pack(L1, mask1).

CPU tests pin the oracle (oracle/mhm_oracle.c, or_meteo_l2_to_l1) with
closed forms on integer-valued fields, whose block sums are exact: block
means, masked means, the sum invariant, empty blocks (0/0), non-integer
cell-size ratios against exact rational ceilings, and the out-of-grid
error.  The GPU tests call mhmhip_set_forcing_l2 through the C-ABI and read
the mapped forcing back through a daily step, whose temperature and
precipitation outputs are the day's forcing itself; they must equal the
oracle bit for bit (NaN and nodata included).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O

NODATA = -9999.0


def _ints(rng, shape):
    return rng.integers(-500, 1500, size=shape).astype(np.float64)


def test_aggregation_block_means_are_exact():
    rng = np.random.default_rng(1)
    nd, f, nr1, nc1 = 3, 4, 5, 7
    data2 = _ints(rng, (nd, nr1 * f, nc1 * f))
    m2 = np.ones((nr1 * f, nc1 * f), bool)
    m1 = np.ones((nr1, nc1), bool)
    got = O.meteo_l2_to_l1(data2, m2, 250.0, m1, 1000.0)
    blocks = data2.reshape(nd, nr1, f, nc1, f).sum(axis=(2, 4)) / (f * f)
    want = np.stack([blocks[t].T.ravel() for t in range(nd)])  # pack: column-major
    assert np.array_equal(got, want)


def test_aggregation_masked_mean_sum_invariant_and_empty_blocks():
    rng = np.random.default_rng(2)
    nd, f, nr1, nc1 = 2, 2, 6, 4
    nr2, nc2 = nr1 * f, nc1 * f
    data2 = _ints(rng, (nd, nr2, nc2))
    m2 = rng.random((nr2, nc2)) > 0.3
    m2[0:2, 0:2] = False  # L1 cell (0, 0): no valid L2 cell -> 0 / 0
    m1 = np.ones((nr1, nc1), bool)
    m1[5, 3] = False      # outside the L1 mask: not packed
    got = O.meteo_l2_to_l1(data2, m2, 1.0, m1, 2.0)
    cnt = m2.reshape(nr1, f, nc1, f).sum(axis=(1, 3))
    for t in range(nd):
        s = (data2[t] * m2).reshape(nr1, f, nc1, f).sum(axis=(1, 3))
        with np.errstate(invalid="ignore"):
            want = (s / cnt).T.ravel()[m1.T.ravel()]
        assert np.array_equal(got[t], want, equal_nan=True)
        assert math.isnan(got[t][0])
        # sum invariant: count-weighted means give back the valid L2 total
        c = cnt.T.ravel()[m1.T.ravel()]
        ok = c > 0
        total = (data2[t] * m2)[:, :].reshape(nr1, f, nc1, f).sum(axis=(1, 3)).T.ravel()
        total = total[m1.T.ravel()]
        assert np.array_equal(got[t][ok] * c[ok], total[ok])


def test_aggregation_non_integer_ratio():
    """cellsize1 / cellsize2 = 1.5: blocks of 1 and 2 rows/columns by
    ceiling(i / 1.5); checked against exact rational ceilings."""
    rng = np.random.default_rng(3)
    cs1, cs2 = 3.0, 2.0
    nr2, nc2 = 9, 6
    nr1 = math.ceil(Fraction(nr2) * Fraction(2, 3))
    nc1 = math.ceil(Fraction(nc2) * Fraction(2, 3))
    data2 = _ints(rng, (2, nr2, nc2))
    m2 = np.ones((nr2, nc2), bool)
    m1 = np.ones((nr1, nc1), bool)
    got = O.meteo_l2_to_l1(data2, m2, cs2, m1, cs1)
    row = [math.ceil(Fraction(i) * Fraction(2, 3)) - 1 for i in range(1, nr2 + 1)]
    col = [math.ceil(Fraction(j) * Fraction(2, 3)) - 1 for j in range(1, nc2 + 1)]
    for t in range(2):
        s = np.zeros((nr1, nc1))
        n = np.zeros((nr1, nc1))
        for j in range(nc2):
            for i in range(nr2):
                s[row[i], col[j]] += data2[t, i, j]
                n[row[i], col[j]] += 1
        assert np.array_equal(got[t], (s / n).T.ravel())


def test_disaggregation_and_equal_grids():
    rng = np.random.default_rng(4)
    nr2, nc2, f = 4, 3, 3
    data2 = rng.normal(size=(2, nr2, nc2))
    m2 = np.ones((nr2, nc2), bool)
    m2[1, 2] = False
    m1 = np.ones((nr2 * f, nc2 * f), bool)
    got = O.meteo_l2_to_l1(data2, m2, 3000.0, m1, 1000.0)
    for t in range(2):
        up = np.where(m2, data2[t], NODATA).repeat(f, axis=0).repeat(f, axis=1)
        assert np.array_equal(got[t], up.T.ravel())
    # equal cell sizes: the grid itself, packed by mask1 (mask2 plays no part)
    m1 = rng.random((nr2, nc2)) > 0.4
    got = O.meteo_l2_to_l1(data2, m2, 500.0, m1, 500.0)
    for t in range(2):
        assert np.array_equal(got[t], data2[t].T.ravel()[m1.T.ravel()])


def test_cells_outside_the_other_grid_are_an_error():
    data2 = np.ones((1, 8, 8))
    # aggregation: L2 rows 5..8 map to L1 row 2 of a 1-row L1 grid
    assert O.meteo_l2_to_l1(data2, np.ones((8, 8), bool), 1.0, np.ones((1, 2), bool), 4.0) is None
    # disaggregation: L1 cells past the L2 grid
    assert O.meteo_l2_to_l1(np.ones((1, 2, 2)), np.ones((2, 2), bool), 2.0,
                            np.ones((6, 4), bool), 1.0) is None


# ---- GPU: mhmhip_set_forcing_l2 against the oracle ----------------------------------

def _daily_model(nr, nc):
    from mhm_amd import synth
    from pairs import make_pair
    dom = synth.make_domain(nr, nc, seed=21)
    model, _, steps, _ = make_pair([dom], timestep=24, n_steps=4, route=False, seed=2)
    return model, steps, dom


def _forcing_of_day(model, steps, k):
    """TEMP and PREC of daily step k: the day's tavg and pre themselves."""
    from mhm_amd import _native as N
    model.run(steps[k:k + 1], mode=N.RUN_PLAIN, want_gauges=False)
    model.sync()
    return model.get("TEMP_CALC"), model.get("PREC_CALC")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["aggregate4", "aggregate1.5", "disaggregate2", "equal"])
def test_device_mapping_equals_the_oracle(case):
    print(f"case: meteo l2 {case}", flush=True)
    nr1, nc1 = 12, 9
    model, steps, dom = _daily_model(nr1, nc1)
    m1 = np.ones((nr1, nc1), bool)
    rng = np.random.default_rng(7)
    if case == "aggregate4":
        cs1, cs2, nr2, nc2 = 1000.0, 250.0, nr1 * 4, nc1 * 4
    elif case == "aggregate1.5":
        cs1, cs2, nr2, nc2 = 3.0, 2.0, 18, 13
    elif case == "disaggregate2":
        cs1, cs2, nr2, nc2 = 1000.0, 2000.0, 6, 5
    else:
        cs1, cs2, nr2, nc2 = 1000.0, 1000.0, nr1, nc1
    nd = model.n_forcing_days
    tavg = rng.normal(8.0, 6.0, size=(nd, nr2, nc2))
    pre = rng.gamma(0.6, 4.0, size=(nd, nr2, nc2))
    m2 = rng.random((nr2, nc2)) > 0.25
    if case.startswith("aggregate"):
        m2[0:4, 0:4] = False  # empty blocks: NaN, as the reference's 0 / 0
    want_t = O.meteo_l2_to_l1(tavg, m2, cs2, m1, cs1)
    want_p = O.meteo_l2_to_l1(pre, m2, cs2, m1, cs1)
    assert want_t is not None and want_p is not None
    model.set_forcing_l2("TAVG", 0, tavg, m2, cs2, m1, cs1)
    model.set_forcing_l2("PRE", 0, pre, m2, cs2, m1, cs1)
    for k in range(2):
        day = int(steps[k].iMeteoTS) - 1
        t, p = _forcing_of_day(model, steps, k)
        assert np.array_equal(t, want_t[day], equal_nan=True), (case, k)
        assert np.array_equal(p, want_p[day], equal_nan=True), (case, k)
    model.close()


@pytest.mark.gpu
def test_device_mapping_rejects_cells_outside_the_grid():
    print("case: meteo l2 errors", flush=True)
    from mhm_amd import _native as N
    model, _, _ = _daily_model(4, 4)
    with pytest.raises(N.MhmHipError):
        model.set_forcing_l2("TAVG", 0, np.ones((model.n_forcing_days, 16, 4)),
                             np.ones((16, 4), bool), 1.0, np.ones((4, 4), bool), 2.0)
    with pytest.raises(N.MhmHipError):  # mask1 with the wrong number of cells
        model.set_forcing_l2("TAVG", 0, np.ones((model.n_forcing_days, 8, 8)),
                             np.ones((8, 8), bool), 1.0, np.ones((4, 3), bool), 2.0)
    model.close()
