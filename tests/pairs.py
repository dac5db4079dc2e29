"""Builds the HIP model and the oracle on identical seeded inputs."""
import datetime as dt

import numpy as np

from mhm_amd import calendar, engine, parameters as PM, synth
from oracle_sim import OracleSim, STATE_NAMES, FLUX_NAMES, rel_close

START = dt.datetime(1990, 1, 1)


def make_pair(domains, n_sets=1, timestep=1, n_steps=24, soil=1, pet=1, seed=0, route=True,
              forcing_seed=100, pet_forcing=False, start=START, inflow=None,
              forcing_columns=None, weights=False, tile_height=0):
    """HIP model and oracle on the same inputs.  inflow: {"domain", "node" (1-based),
    "headwater", "q" (n_steps, n_inflow)} inflow gauges, given to both.
    forcing_columns: the model gets that many forcing columns and no forcing
    (the caller streams it); the oracle gets all of it.
    pet 2, 3: the radiation-based PET cases get their daily inputs and
    monthly parameters; weights: meteo weights instead of day/night factors.
    tile_height: levels per routing tile of the model's partition (0: 16)."""
    n_days = calendar.n_forcing_days_needed(n_steps)
    doy0 = start.timetuple().tm_yday
    for i, d in enumerate(domains):
        if not d.forcing or d.n_days < n_days:
            synth.make_forcing(d, n_days, forcing_seed + i, doy0=doy0)
        if pet >= 2 and "NETRAD" not in d.forcing:
            synth.make_pet_forcing(d, forcing_seed + 70 + i, doy0=doy0)
        if pet_forcing and "PET" not in d.forcing:
            rng = np.random.default_rng(forcing_seed + 50 + i)
            d.forcing["PET"] = rng.gamma(2.0, 1.2, size=(d.n_days, d.n_cells))
    model = engine.Model(domains, n_sets=n_sets, timestep=timestep, process_soil=soil,
                         process_pet=pet, route=route,
                         n_forcing_days=forcing_columns or n_days)
    sim = OracleSim(domains, n_sets=n_sets, timestep=timestep, process_soil=soil, process_pet=pet,
                    n_forcing_days=n_days, route=route)
    if route:
        if tile_height:
            model.set_tile_height(tile_height)
        model.bind_network(inflow=inflow)
        if inflow is not None:
            model.set_inflow(inflow["q"])
            for i, d in enumerate(domains):
                sel = np.asarray(inflow["domain"]) == i
                if sel.any():
                    d._oracle_inflow = {"node": np.asarray(inflow["node"])[sel],
                                        "headwater": np.asarray(inflow["headwater"])[sel],
                                        "q": np.asarray(inflow["q"])[:, sel], "tt0": 0}
    rng = np.random.default_rng(seed)
    gam = PM.draw(rng, n_sets * len(domains)).reshape(n_sets, len(domains), -1)
    for s in range(n_sets):
        parts = [synth.cell_params(d, gam[s, i], pet_case=pet) for i, d in enumerate(domains)]
        params = {k: np.concatenate([p[k] for p in parts], axis=0) for k in parts[0]}
        model.set_params(params, s)
        sim.set_params(params, s)
        if route:
            for i in range(len(domains)):
                g5 = PM.routing_gamma(gam[s, i])
                model.set_routing_params(g5, s, i)
                sim.set_routing_params(g5, s, i)
    model.default_states()
    sim.default_states()
    if forcing_columns is None:
        model.set_domain_forcing()
    if pet_forcing:
        model.set_forcing("PET", np.concatenate([d.forcing["PET"] for d in domains], axis=1))
    sim.set_domain_forcing()
    if weights:
        w = synth.make_meteo_weights(sum(d.n_cells for d in domains), seed + 900)
        model.set_meteo_weights(*w)
        sim.set_meteo_weights(*w)
    steps = calendar.step_table(start, n_steps, timestep)
    recs = calendar.step_records(start, n_steps, timestep)
    return model, sim, steps, recs


def compare_cells(model, sim, rtol=1e-10, atol=1e-12, names=None):
    """Every state and flux of every set; returns list of (name, set, worst) failures."""
    bad = []
    for s in range(model.n_sets):
        for name in names or (STATE_NAMES + FLUX_NAMES):
            ok, worst = rel_close(model.get(name, s), sim.get(name, s), rtol, atol)
            if not ok:
                bad.append((name, s, worst))
    return bad


def compare_routing(model, sim, rtol=1e-10, atol=1e-12):
    bad = []
    for s in range(model.n_sets):
        a = model.routing_state(s)
        b = sim.routing_state(s)
        for k in ("qOUT", "qTIN", "qTR", "Qmod", "C1", "C2"):
            ok, worst = rel_close(a[k][:b[k].shape[0]], b[k], rtol, atol)
            if not ok:
                bad.append((k, s, worst))
    return bad
