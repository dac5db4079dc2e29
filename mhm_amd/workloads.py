"""Builders for the benchmark workloads (synthetic, seeded).

  plain_domain   one basin, one parameter set   (4096 cells x 8760 h; 1024^2 cells x 2160 h)
  ensemble       one basin, many parameter sets on shared forcing (256 x 4096 cells, 3650 d)
  multi_basin    uneven basins packed in one context (64 basins, 500..200k cells, 720 h)
"""
import datetime as dt
import math

import numpy as np

from . import calendar, engine, parameters as PM, synth

START = dt.datetime(1990, 1, 1)


def _params_for(domains, gammas):
    parts = [synth.cell_params(d, g) for d, g in zip(domains, gammas)]
    return {k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}


def plain_domain(nrows, ncols, n_steps, timestep=1, seed=1, n_sets=1, forcing_seed=None,
                 soil=1, pet=1, start=START, device_params=False, tile_height=0):
    """Model + step table for one basin (n_sets parameter sets drawn within the
    bounds).  device_params: each set's parameter rows built on the device
    from its gamma vector (Model.mprlite_params, one launch per set) instead
    of in numpy and uploaded; the two agree bit for bit but for the root
    fractions' power.  tile_height: levels per routing tile (0: 16)."""
    dom = synth.make_domain(nrows, ncols, seed=seed)
    n_days = calendar.n_forcing_days_needed(n_steps)
    synth.make_forcing(dom, n_days, seed + 1000 if forcing_seed is None else forcing_seed,
                       doy0=start.timetuple().tm_yday)
    gam = PM.draw(np.random.default_rng(seed + 2000), n_sets)
    model, steps = model_on(dom, n_steps, gam, timestep, soil, pet, start, device_params,
                            tile_height)
    return model, steps, dom, gam


def model_on(dom, n_steps, gam, timestep=1, soil=1, pet=1, start=START, device_params=False,
             tile_height=0):
    """Model + step table on a built basin with its forcing, one parameter set
    per row of gam."""
    n_sets = len(gam)
    n_days = calendar.n_forcing_days_needed(n_steps)
    model = engine.Model([dom], n_sets=n_sets, timestep=timestep, process_soil=soil,
                         process_pet=pet, route=True, n_forcing_days=n_days)
    if tile_height:
        model.set_tile_height(tile_height)
    model.bind_network()
    if device_params:
        model.set_mprlite_inputs()
    for s in range(n_sets):
        if device_params:
            model.mprlite_params(gam[s], s)
        else:
            model.set_params(_params_for([dom], [gam[s]]), s)
        model.set_routing_params(PM.routing_gamma(gam[s]), s)
    model.default_states()
    model.set_domain_forcing(n_days=n_days)
    steps = calendar.step_table(start, n_steps, timestep)
    return model, steps


def multi_basin(n_basins, lo, hi, n_steps, timestep=1, seed=7, start=START):
    """Basins with log-uniform cell counts in [lo, hi], each with its own DEM,
    network, forcing and parameter set, packed into one context."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), n_basins)).astype(int)
    doms = []
    n_days = calendar.n_forcing_days_needed(n_steps)
    for b, n in enumerate(sizes):
        nr = max(2, int(round(math.sqrt(n) * rng.uniform(0.7, 1.3))))
        nc = max(2, int(round(n / nr)))
        d = synth.make_domain(nr, nc, seed=seed * 1000 + b)
        synth.make_forcing(d, n_days, seed * 1000 + 500 + b, doy0=start.timetuple().tm_yday)
        doms.append(d)
    model = engine.Model(doms, n_sets=1, timestep=timestep, route=True, n_forcing_days=n_days)
    model.bind_network()
    gam = PM.draw(np.random.default_rng(seed + 1), n_basins)
    model.set_params(_params_for(doms, gam), 0)
    for b in range(n_basins):
        model.set_routing_params(PM.routing_gamma(gam[b]), 0, b)
    model.default_states()
    model.set_domain_forcing()
    steps = calendar.step_table(start, n_steps, timestep)
    return model, steps, doms, gam
