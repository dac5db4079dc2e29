"""Time-step driver, as the reference's `mhm.run` module
(pybind/src/wrapper.f90:172-244, src/mHM/mo_mhm_interface_run.f90)."""
from . import session as _s


def prepare():
    _s.prepare()


def get_ndomains():
    return len(_s.S.setup.domains)


def prepare_domain(domain=1):
    # all domains of a Setup are packed into one device context and stepped together
    _s.S.domain = domain


def finished():
    return _s.finished()


def do_time_step():
    _s.do_time_step()


def write_output():
    """NetCDF output is outside the accelerated path; states stay on the device
    and are read on demand through `get`."""


def finalize_domain():
    pass


def finalize():
    _s.finalize_run()


def current_time():
    t = _s.current_datetime()
    return t.year, t.month, t.day, t.hour
