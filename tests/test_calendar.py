"""The step table's calendar against the reference's step clock.

mhm_amd.calendar builds month, hour and day of year per model step from
Python's proleptic Gregorian datetime.  The reference composes its clock
from FORCES mo_julian (julday, caldat, dec2date; oracle/julian.py restates
the copy the reference ships, pre-proc/mtCLIM/mo_julian.f90) through
datetimeinfo_increment and time = newTime - 0.5.  These tests pin the
restatement (published Julian day numbers, the reference's own example,
Python's day ordinal over two centuries) and then require the product's
table to equal the reference's clock step for step.
"""
import datetime as dt
import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import julian as J  # noqa: E402

from mhm_amd import calendar  # noqa: E402

# Python's date ordinal 1 is 0001-01-01 (proleptic Gregorian), Julian day 1721426
JDN_OF_ORDINAL_0 = 1721425


def test_julian_day_numbers():
    # the reference's example (pre-proc/mtCLIM/mo_julian.f90:539): 2415021 is 01.01.1900
    assert J.caldat(2415021) == (1, 1, 1900)
    assert J.julday(1, 1, 1900) == 2415021
    # and (mo_julian.f90:655) 2415021.0 is 01.01.1900 12:00
    assert J.dec2date(2415021.0) == (1, 1, 1900, 12, 0, 0)
    # published: J2000 (2000-01-01), the Gregorian reform (15.10.1582 follows 4.10.1582)
    assert J.julday(1, 1, 2000) == 2451545
    assert J.julday(15, 10, 1582) == 2299161
    assert J.julday(4, 10, 1582) == 2299160
    assert J.caldat(2299160) == (4, 10, 1582)


def test_julday_and_caldat_against_the_day_ordinal():
    d = dt.date(1900, 1, 1)
    while d.year < 2101:
        jd = J.julday(d.day, d.month, d.year)
        assert jd == d.toordinal() + JDN_OF_ORDINAL_0, d
        assert J.caldat(jd) == (d.day, d.month, d.year), d
        d += dt.timedelta(days=1)


def test_dec2date_hours_and_the_seconds_carry():
    """Every hour of 1999-2001 read back from its fractional day; without the
    rounding of seconds some hours would come out one lower."""
    short = 0
    d = dt.datetime(1999, 1, 1)
    while d.year < 2002:
        jd = J.julday(d.day, d.month, d.year)
        t = jd + d.hour / 24.0 - 0.5
        assert J.dec2date(t)[:4] == (d.day, d.month, d.year, d.hour), d
        # the floor alone (no seconds carry)
        frac = t + 0.5 - float(int(t + 0.5))
        short += min(max(math.floor(frac * 24.0), 0), 23) != d.hour
        d += dt.timedelta(hours=1)
    assert short > 0


@pytest.mark.parametrize("start,n_steps,timestep", [
    (dt.datetime(1990, 1, 1), 3 * 8760, 1),
    (dt.datetime(2000, 2, 27), 24 * 10, 1),     # leap day
    (dt.datetime(2099, 12, 30), 24 * 800, 1),   # 2100 is not a leap year
    (dt.datetime(1990, 1, 1), 3650, 24),
    (dt.datetime(2003, 6, 30), 8 * 400, 3),
    (dt.datetime(1996, 12, 31), 4 * 800, 6),
])
def test_step_table_equals_the_reference_clock(start, n_steps, timestep):
    want = J.step_clock(start.day, start.month, start.year, n_steps, timestep)
    recs = calendar.step_records(start, n_steps, timestep)
    got = list(zip(recs["month"].tolist(), recs["hour"].tolist(), recs["doy"].tolist()))
    assert got == want


def test_a_shifted_clock_fails():
    """The checker sees a clock that skips time = newTime - 0.5."""
    start = dt.datetime(1990, 1, 1)
    want = J.step_clock(1, 1, 1990, 48, 1)
    shifted = []
    for k in range(48):
        dd, mm, yy, hh, _, _ = J.dec2date(J.julday(1, 1, 1990) + k / 24.0)
        shifted.append((mm, hh, J.julday(dd, mm, yy) - J.julday(1, 1, yy) + 1))
    assert shifted != want
    recs = calendar.step_records(start, 48, 1)
    assert list(zip(recs["month"].tolist(), recs["hour"].tolist(), recs["doy"].tolist())) == want
