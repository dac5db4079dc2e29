"""Lifecycle of the model, as the reference's `mhm.model` module
(pybind/src/wrapper.f90:13-170)."""
from . import session as _s


def config_coupling(couple_case=0, meteo_timestep=24, meteo_expect_pre=False,
                    meteo_expect_temp=False, meteo_expect_pet=False, meteo_expect_tmin=False,
                    meteo_expect_tmax=False, **_unused):
    """Forcing is pushed per time step through `set_meteo` (wrapper.f90:20-82)."""
    if couple_case == 0:
        _s.S.coupling = {}
        return
    if meteo_timestep != 24:
        raise ValueError("only daily coupled forcing is supported")
    _s.S.coupling = {"PRE": meteo_expect_pre, "TAVG": meteo_expect_temp, "PET": meteo_expect_pet,
                     "TMIN": meteo_expect_tmin, "TMAX": meteo_expect_tmax}


def init(setup):
    """mhm_interface_init (src/mHM/mo_mhm_interface.F90:37-228) with a Setup in
    place of the namelists in cwd."""
    _s.S.setup = setup
    _s.S.prepared = False


def set_verbosity(level=1):
    _s.S.verbosity = int(level)


def disable_output():
    _s.S.output = False


def run():
    """mhm_interface_run: the whole time loop."""
    from . import run as _run
    _run.prepare()
    _run.prepare_domain()
    while not _run.finished():
        _run.do_time_step()
    _run.finalize_domain()


def finalize():
    _s.finalize_run()
    _s.S.setup = None
    _s.S.coupling = {}
    _s.S.coupled = {}
