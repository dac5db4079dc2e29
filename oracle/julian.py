"""Test infrastructure only: the reference's step clock restated in Python.

Checker for the step table of the time loop (month, hour and day of year of
every model step), never imported by the product.  It follows the
reference's own composition:

* the clock: datetimeinfo_init / datetimeinfo_increment
  (src/common/mo_common_datetime_type.f90:71-133): newTime starts at the
  simulation's first Julian day; each step adds `timestep` to the hour,
  newTime = julday(day, month, year) + hour / 24, then hour = mod(hour, 24)
  and the date comes from caldat(int(newTime));
* the getters and mHM() receive time = newTime - 0.5
  (src/mHM/mo_mhm_interface_run.f90:360,397) and take month, day, hour from
  dec2date(time) (src/mHM/mo_mhm.f90:440;
  src/meteo/mo_meteo_handler.f90:962,1149,1230);
* the day of year is that date's day count from 1 January
  (mo_meteo_handler.f90:963-964, FORCES datetime%doy).

julday, caldat and dec2date live in FORCES mo_julian (v0.6.0, not in the
tree).  They are restated from the copy of that module the reference ships
for its mtCLIM pre-processor (pre-proc/mtCLIM/mo_julian.f90: juldayJulian
966-1002, caldatJulian 555-591, dec2dateJulian 811-891, the standard
calendar), including its floating-point steps: the noon-based fractional
day, floor to the hour and the rounding of seconds that carries a
59:59.99... reading into the next hour.
"""
import math

import numpy as np

IGREG = 2299161
# juldayJulian multiplies by the default-real literal 30.6001 (single precision)
_F30 = float(np.float32(30.6001))


def julday(dd, mm, yy):
    """pre-proc/mtCLIM/mo_julian.f90:966-1002 (juldayJulian)."""
    if mm > 2:
        jm, jy = mm, yy
    else:
        jm, jy = mm + 12, yy - 1
    if dd + 31 * (mm + 12 * yy) >= 15 + 31 * (10 + 12 * 1582):
        a = int(jy / 100)  # Fortran integer division truncates toward zero
        b = 2 - a + int(a / 4)
    else:
        b = 0
    return int(365.25 * float(jy + 4716) + float(int(_F30 * float(jm + 1)))
               + float(dd + b) - 1524.5 + 0.5)


def _civil(a_base):
    """The day/month/year steps shared by caldatJulian and dec2dateJulian."""
    b = a_base + 1524
    c = int((float(b) - 122.1) / 365.25)
    d = int(365.25 * float(c))
    e = int(float(b - d) / 30.6001)
    day = b - d - int(30.6001 * float(e))
    month = e - 1 if e < 14 else e - 13
    year = c - 4716 if month > 2 else c - 4715
    return day, month, year


def _gregorian_a(z):
    if z < IGREG:
        return z
    g = int((float(z) - 1867216.25) / 36524.25)
    return z + 1 + g - int(g / 4)


def caldat(julian):
    """pre-proc/mtCLIM/mo_julian.f90:555-591 (caldatJulian): (dd, mm, yy)."""
    return _civil(_gregorian_a(julian))


def dec2date(julian):
    """pre-proc/mtCLIM/mo_julian.f90:811-891 (dec2dateJulian): (dd, mm, yy, hh, nn, ss)."""
    z = int(julian + 0.5)
    day, month, year = _civil(_gregorian_a(z))
    fraction = julian + 0.5 - float(z)
    hour = min(max(math.floor(fraction * 24.0), 0), 23)
    fraction = fraction - float(hour) / 24.0
    minute = min(max(math.floor(fraction * 1440.0), 0), 59)
    second = max(_nint((fraction - float(minute) / 1440.0) * 86400.0), 0)
    if second == 60:
        second = 0
        minute += 1
        if minute == 60:
            minute = 0
            hour += 1
            if hour == 24:
                hour = 0
                day, month, year = caldat(julday(day, month, year) + 1)
    return day, month, year, hour, minute, second


def _nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def step_clock(d0, m0, y0, n_steps, timestep):
    """[(month, hour, doy)] of steps tt = 1..n_steps as mHM() sees them
    (time = newTime - 0.5 through dec2date), the clock started at 00:00 of
    d0.m0.y0 and advanced by timestep hours per step."""
    new_time = float(julday(d0, m0, y0))
    day, month, year = caldat(int(new_time))
    hour = 0
    out = []
    for _ in range(n_steps):
        dd, mm, yy, hh, _, _ = dec2date(new_time - 0.5)
        doy = julday(dd, mm, yy) - julday(1, 1, yy) + 1
        out.append((mm, hh, doy))
        # datetimeinfo_increment
        hour = hour + timestep
        new_time = julday(day, month, year) + float(hour) / 24.0
        hour = hour % 24
        day, month, year = caldat(int(new_time))
    return out
