"""Host-side owner of one device context (include/mhmhip.h).

A `Model` packs one or more domains (basins) and one or more parameter sets
into one context: cells of all domains are concatenated like the reference's
module arrays, parameter sets run side by side on the shared forcing.
"""
import ctypes as C

import numpy as np

from . import _native as N
from . import synth


def _fortran_cols(a):
    """(n,) or (n, k) array -> C-contiguous (k, n) block = Fortran X(n, k)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return np.ascontiguousarray(a[None, :]), 1
    return np.ascontiguousarray(a.T), a.shape[1]


class Model:
    def __init__(self, domains, n_sets=1, timestep=1, process_soil=1, process_pet=1,
                 route=True, n_forcing_days=1, hourly_forcing=False, device=0,
                 evap_coeff=synth.EVAP_COEFF, fnight_prec=synth.FNIGHT_PREC,
                 fnight_pet=synth.FNIGHT_PET, fnight_temp=synth.FNIGHT_TEMP):
        if isinstance(domains, synth.Domain):
            domains = [domains]
        self.domains = list(domains)
        nh = {d.horizon_depth.shape[0] for d in self.domains}
        if len(nh) != 1:
            raise ValueError("all packed domains need the same number of horizons")
        self.n_horizons = nh.pop()
        self.n_sets = int(n_sets)
        self.timestep = int(timestep)
        self.cell_start = np.zeros(len(self.domains) + 1, np.int64)
        self.cell_start[1:] = np.cumsum([d.n_cells for d in self.domains])
        self.n_cells = int(self.cell_start[-1])
        cfg = N.Config()
        cfg.n_cells = self.n_cells
        cfg.n_domains = len(self.domains)
        cfg.n_horizons = self.n_horizons
        cfg.n_sets = self.n_sets
        cfg.timestep = self.timestep
        cfg.process_soil = process_soil
        cfg.process_pet = process_pet
        cfg.process_route = 1 if route else 0
        cfg.hourly_forcing = 1 if hourly_forcing else 0
        cfg.n_forcing_days = int(n_forcing_days)
        cfg.device = device
        cfg.domain_cell_start = N.lptr(self.cell_start)
        tables = {"evap_coeff": evap_coeff, "fnight_prec": fnight_prec,
                  "fnight_pet": fnight_pet, "fnight_temp": fnight_temp}
        for name, vals in tables.items():
            arr = getattr(cfg, name)
            for m in range(12):
                arr[m] = float(vals[m])
        self.n_forcing_days = int(n_forcing_days)
        self.route = bool(route)
        self._ctx = C.c_void_p()
        N.check(N.lib().mhmhip_create(C.byref(cfg), C.byref(self._ctx)), None)
        self.n_gauges = 0

    # -- lifecycle -------------------------------------------------------------
    def close(self):
        if self._ctx:
            N.lib().mhmhip_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return N.check(rc, self._ctx)

    @property
    def device_bytes(self):
        return int(N.lib().mhmhip_device_bytes(self._ctx))

    def sync(self):
        self._chk(N.lib().mhmhip_sync(self._ctx))

    # -- cell data -------------------------------------------------------------
    def set_var(self, name, values, set_=0):
        """One parameter or state field over all packed cells ((n,), (n, nH) or (n, 12))."""
        var = N.P[name] if name in N.P else N.S[name]
        blk, ncols = _fortran_cols(values)
        if blk.shape[1] != self.n_cells:
            raise ValueError(f"{name}: expected {self.n_cells} cells, got {blk.shape[1]}")
        self._chk(N.lib().mhmhip_set_cell_var(self._ctx, var, set_, N.dptr(blk), self.n_cells,
                                              ncols))

    def set_params(self, params, set_=0):
        for k, v in params.items():
            self.set_var(k, v, set_)

    def get(self, name, set_=0):
        var = N.S[name] if name in N.S else N.X[name]
        ncols = self.n_horizons if name in N.HORIZON_VARS else 1
        out = np.zeros((ncols, self.n_cells))
        self._chk(N.lib().mhmhip_get_cell_var(self._ctx, var, set_, N.dptr(out), self.n_cells,
                                              ncols))
        return out[0] if ncols == 1 else out.T.copy()

    def default_states(self):
        hd = np.ascontiguousarray(self.domains[0].horizon_depth, dtype=np.float64)
        self._chk(N.lib().mhmhip_default_states(self._ctx, N.dptr(hd)))

    def set_forcing(self, name, data, day0=1):
        """data: (n_days, n_cells) = Fortran L1_x(n_cells, n_days) of all packed cells."""
        data = np.ascontiguousarray(data, dtype=np.float64)
        self._chk(N.lib().mhmhip_set_forcing(self._ctx, N.F[name], day0, data.shape[0],
                                             N.dptr(data), data.shape[1]))

    def stream_forcing(self, data, day0):
        """Days day0.. of the variables in `data` ({name: (n_days, n_cells)},
        Fortran L1_x(n_cells, n_days) of all packed cells) uploaded
        asynchronously into the device's forcing columns (day d in column
        (d - 1) mod n_forcing_days); returns once the arrays are staged."""
        ptrs = (N._dp * len(N.F_NAMES))()
        keep = []
        nd = None
        for name, arr in data.items():
            a = np.ascontiguousarray(arr, dtype=np.float64)
            if a.shape[1] != self.n_cells or (nd is not None and a.shape[0] != nd):
                raise ValueError(f"{name}: expected (n_days, {self.n_cells}) like the others")
            nd = a.shape[0]
            keep.append(a)
            ptrs[N.F[name]] = N.dptr(a)
        self._chk(N.lib().mhmhip_stream_forcing(self._ctx, int(day0), int(nd), ptrs,
                                                self.n_cells))

    def set_forcing_l2(self, name, domain, data2, mask2, cellsize2, mask1, cellsize1, day0=1):
        """One domain's forcing on its meteo grid (L2), mapped to its L1 cells on
        the device as meteo_forcings_wrapper does (mo_meteo_helper.f90:119-143).
        data2: (n_days, nrows2, ncols2); mask2 (nrows2, ncols2), mask1 (nrows1,
        ncols1) booleans, row i / column j as the reference's (i, j)."""
        data2 = np.asarray(data2, dtype=np.float64)
        nd, nr2, nc2 = data2.shape
        d2 = np.ascontiguousarray(np.transpose(data2, (0, 2, 1)))  # X(nrows2, ncols2, ndays)
        m2 = np.ascontiguousarray(np.asarray(mask2, dtype=bool).T.astype(np.uint8))
        m1b = np.asarray(mask1, dtype=bool)
        nr1, nc1 = m1b.shape
        m1 = np.ascontiguousarray(m1b.T.astype(np.uint8))
        self._chk(N.lib().mhmhip_set_forcing_l2(
            self._ctx, N.F[name], int(domain), int(day0), nd, N.dptr(d2), nr2, nc2,
            float(cellsize2), m2.ctypes.data, nr1, nc1, float(cellsize1), m1.ctypes.data))

    def set_mprlite_inputs(self):
        """The packed domains' MPR-lite inputs (cover fractions, soil texture,
        aspect, latitude; the horizon depths, which the packed domains share;
        the cover classes' monthly LAI) for mprlite_params."""
        from . import parameters as PM
        d = self.domains
        if any(not np.array_equal(x.horizon_depth, d[0].horizon_depth) for x in d[1:]):
            raise ValueError("MPR-lite on the device takes one set of horizon depths: the "
                             "packed domains differ")
        cols = [np.concatenate([x.fractions[k] for x in d]) for k in range(3)]
        cols += [np.concatenate([getattr(x, f) for x in d]) for f in ("texture", "aspect", "lat")]
        cells = np.ascontiguousarray(np.stack(cols), dtype=np.float64)  # X(n_cells, 6)
        hd = np.ascontiguousarray(d[0].horizon_depth, dtype=np.float64)
        lai = np.ascontiguousarray(PM.LAI, dtype=np.float64)
        self._chk(N.lib().mhmhip_mprlite_inputs(self._ctx, N.dptr(cells), self.n_cells,
                                                N.dptr(hd), N.dptr(lai)))

    def mprlite_params(self, gamma, set_=0):
        """Every parameter row of set_ built on the device from a gamma vector
        (parameters.GAMMA order; the routing entries are not read)."""
        g = np.ascontiguousarray(np.asarray(gamma, dtype=np.float64)[:N.MPRLITE_NGAMMA])
        if g.shape[0] != N.MPRLITE_NGAMMA:
            raise ValueError(f"gamma needs {N.MPRLITE_NGAMMA} entries")
        self._chk(N.lib().mhmhip_mprlite_params(self._ctx, set_, N.dptr(g)))

    def set_meteo_weights(self, pre_w=None, temp_w=None, pet_w=None):
        """Disaggregation by weights (read_meteo_weights): each (n_cells, 12, 24)
        = L1_x_weights(cell, month, hour + 1) of all packed cells; no
        arguments: back to the day/night factors."""
        if pre_w is None:
            self._chk(N.lib().mhmhip_set_meteo_weights(self._ctx, None, None, None, self.n_cells))
            return
        # Fortran X(ld, 12, 24): cell fastest, then month, then hour
        ws = [np.asfortranarray(np.asarray(w, dtype=np.float64).reshape(self.n_cells, 12, 24))
              for w in (pre_w, temp_w, pet_w)]
        self._chk(N.lib().mhmhip_set_meteo_weights(
            self._ctx, *[w.ctypes.data_as(N._dp) for w in ws], self.n_cells))

    def set_domain_forcing(self, day0=1, n_days=None):
        """Upload every domain's own forcing (synth.make_forcing) for all packed
        cells: all of its days, or its first n_days."""
        for name in ("PRE", "TAVG", "TMIN", "TMAX", "NETRAD", "ABSVAPPRESS", "WINDSPEED"):
            if all(name in d.forcing for d in self.domains):
                data = np.concatenate([d.forcing[name][:n_days] for d in self.domains], axis=1)
                self.set_forcing(name, data, day0)

    # -- network -----------------------------------------------------------------
    def set_tile_height(self, height):
        """Levels per routing tile of the partition bind_network builds (0 = 16;
        32 for callers that route one step per call).  Before bind_network."""
        self._chk(N.lib().mhmhip_set_tile_height(self._ctx, int(height)))

    def bind_network(self, inflow=None):
        doms = self.domains
        cat = lambda f, dt: np.ascontiguousarray(np.concatenate([f(d) for d in doms]).astype(dt))
        self.node_start = np.zeros(len(doms) + 1, dtype=np.int64)
        self.node_start[1:] = np.cumsum([d.n_nodes for d in doms])
        self.n_nodes = int(self.node_start[-1])
        self._net_arrays = arrs = dict(
            start=np.ascontiguousarray(self.node_start),
            nout=np.array([d.n_outlets for d in doms], np.int32),
            netPerm=cat(lambda d: d.netPerm, np.int32), fromN=cat(lambda d: d.fromN, np.int32),
            toN=cat(lambda d: d.toN, np.int32), length=cat(lambda d: d.length, np.float64),
            slope=cat(lambda d: d.slope, np.float64), fFPimp=cat(lambda d: d.fFPimp, np.float64),
            area=cat(lambda d: d.area_km2, np.float64),
            l1l11=cat(lambda d: np.arange(1, d.n_cells + 1) if d.l1_l11 is None else d.l1_l11,
                      np.int32),
            gdom=cat(lambda d: np.full(d.gauge_nodes.shape[0], doms.index(d)), np.int32),
            gnode=cat(lambda d: d.gauge_nodes, np.int32))
        net = N.Network()
        net.n_nodes = self.n_nodes
        net.domain_node_start = N.lptr(arrs["start"])
        net.n_outlets = N.iptr(arrs["nout"])
        net.netPerm = N.iptr(arrs["netPerm"])
        net.fromN = N.iptr(arrs["fromN"])
        net.toN = N.iptr(arrs["toN"])
        net.length = N.dptr(arrs["length"])
        net.slope = N.dptr(arrs["slope"])
        net.fFPimp = N.dptr(arrs["fFPimp"])
        net.L1_area = N.dptr(arrs["area"])
        net.L1_L11_Id = N.iptr(arrs["l1l11"])
        net.n_gauges = arrs["gnode"].shape[0]
        net.gauge_domain = N.iptr(arrs["gdom"])
        net.gauge_node = N.iptr(arrs["gnode"])
        if inflow:
            arrs["idom"] = np.ascontiguousarray(inflow["domain"], np.int32)
            arrs["inode"] = np.ascontiguousarray(inflow["node"], np.int32)
            arrs["ihw"] = np.ascontiguousarray(inflow["headwater"], np.int32)
            net.n_inflow = arrs["inode"].shape[0]
            net.inflow_domain = N.iptr(arrs["idom"])
            net.inflow_node = N.iptr(arrs["inode"])
            net.inflow_headwater = N.iptr(arrs["ihw"])
        self._chk(N.lib().mhmhip_bind_network(self._ctx, C.byref(net)))
        self.n_gauges = net.n_gauges
        self.n_inflow = net.n_inflow

    def set_routing_params(self, gamma5, set_=0, domain=-1):
        g = np.ascontiguousarray(gamma5, dtype=np.float64)
        self._chk(N.lib().mhmhip_set_routing_params(self._ctx, set_, domain, N.dptr(g)))

    def set_routing_celerity(self, celerity, set_=0, domain=-1):
        """Adaptive-timestep routing, constant celerity [m/s] (processCase(8) = 2)."""
        self._chk(N.lib().mhmhip_set_routing_celerity(self._ctx, set_, domain, float(celerity)))

    def set_routing_celerity_links(self, celerity, set_=0, domain=-1):
        """Adaptive-timestep routing with a celerity per link [m/s] (processCase(8) = 3);
        celerity over the domain's links (or every packed link for domain < 0)."""
        c = np.ascontiguousarray(celerity, dtype=np.float64)
        self._chk(N.lib().mhmhip_set_routing_celerity_links(self._ctx, set_, domain, N.dptr(c)))

    def set_inflow(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._chk(N.lib().mhmhip_set_inflow(self._ctx, q.shape[0], N.dptr(q)))

    def routing_state(self, set_=0):
        n = self.n_nodes
        nl = self._net_arrays["netPerm"].shape[0]
        out = {k: np.zeros(n) for k in ("qOUT", "qTIN", "qTR", "Qmod")}
        out["C1"] = np.zeros(max(nl, 1))
        out["C2"] = np.zeros(max(nl, 1))
        self._chk(N.lib().mhmhip_get_routing_state(self._ctx, set_, *[N.dptr(out[k]) for k in
                                                   ("qOUT", "qTIN", "qTR", "Qmod", "C1", "C2")]))
        return out

    def set_routing_state(self, qTIN, qTR, set_=0, C1=None, C2=None):
        """Routing states of one set as a restart holds them: qTIN, qTR per node,
        C1, C2 per link (optional; kept while steps run with read_states = 1)."""
        arr = lambda v: None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        a, b, c1, c2 = arr(qTIN), arr(qTR), arr(C1), arr(C2)
        self._chk(N.lib().mhmhip_set_routing_state(self._ctx, set_, N.dptr(a), N.dptr(b),
                                                   N.dptr(c1), N.dptr(c2)))

    def qmod_accumulate(self, on=True):
        """Start (sums zeroed) or stop summing every node's Qmod over the routed
        steps on the device (the output aggregation of mo_nc_output.f90:140-175)."""
        self._chk(N.lib().mhmhip_qmod_accumulate(self._ctx, 1 if on else 0))

    def qmod_sum(self, set_=0):
        """(per-node Qmod sum in reference node order, steps added)."""
        out = np.zeros(self.n_nodes)
        cnt = C.c_int64()
        self._chk(N.lib().mhmhip_get_qmod_sum(self._ctx, set_, N.dptr(out), C.byref(cnt)))
        return out, cnt.value

    def output_accumulate(self, flags):
        """Start summing the selected L1 outputs every step on the device
        (mHM_updateDataset; flags: 21 outputFlxState switches, index k-1 for
        output k); None stops."""
        if flags is None:
            self._chk(N.lib().mhmhip_output_accumulate(self._ctx, None))
            return
        f = np.ascontiguousarray([1 if v else 0 for v in flags], dtype=np.int32)
        if f.shape[0] != 21:
            raise ValueError("21 outputFlxState switches expected")
        self._chk(N.lib().mhmhip_output_accumulate(self._ctx, N.iptr(f)))

    def output_sum(self, k, horizon=1, set_=0):
        """(per-cell sum of output k in reference cell order, steps added)."""
        out = np.zeros(self.n_cells)
        cnt = C.c_int64()
        self._chk(N.lib().mhmhip_get_output_sum(self._ctx, int(k), int(horizon), set_,
                                                N.dptr(out), C.byref(cnt)))
        return out, cnt.value

    # -- stepping ----------------------------------------------------------------
    def mhm_step(self, step):
        self._chk(N.lib().mhmhip_mhm_step(self._ctx, C.byref(step)))

    def mrm_step(self, step):
        """Gauges (n_sets, n_gauges) of the step, or None when the step only
        accumulated runoff for a routing step longer than the model step."""
        g = np.zeros((self.n_sets, max(self.n_gauges, 1)))
        rc = N.lib().mhmhip_mrm_step(self._ctx, C.byref(step), N.dptr(g))
        if rc == N.PENDING:
            return None
        self._chk(rc)
        return g[:, :self.n_gauges]

    def run(self, steps, mode=N.RUN_FUSED, steps_per_launch=0, block=0, route_mode=0,
            want_gauges=True, use_graph=False, route_window=0, noalloc=False):
        """steps: ctypes array of mhmhip_step.  Returns gauges (n_steps, n_sets, n_gauges).
        noalloc: mhmhip_run_noalloc (no device allocation; per-call buffers in the
        library's fixed device scratch)."""
        n = len(steps)
        if not isinstance(steps, C.Array):
            steps = (N.Step * n)(*steps)
        opts = N.RunOpts()
        opts.mode = mode
        opts.steps_per_launch = steps_per_launch
        opts.block = block
        opts.route_mode = route_mode
        opts.use_graph = 1 if use_graph else 0
        opts.route_window = route_window
        g = np.zeros((n, self.n_sets, max(self.n_gauges, 1))) if want_gauges else None
        fn = N.lib().mhmhip_run_noalloc if noalloc else N.lib().mhmhip_run
        self._chk(fn(self._ctx, steps, n, C.byref(opts), N.dptr(g)))
        return None if g is None else g[:, :, :self.n_gauges]

    # -- state snapshots (restart-style round trip of the device states) ------------
    STATE_NAMES = ("INTER", "SNOWPACK", "SEALSTW", "UNSATSTW", "SATSTW", "SOILMOIST")

    def snapshot(self):
        """Host copy of every state (cells and routing) of every parameter set."""
        snap = {"cells": [{k: self.get(k, s) for k in self.STATE_NAMES} for s in range(self.n_sets)]}
        if self.route and getattr(self, "_net_arrays", None) is not None:
            snap["routing"] = [self.routing_state(s) for s in range(self.n_sets)]
        return snap

    def restore(self, snap, coefficients=False):
        """States back from a snapshot.  coefficients: also the routing
        coefficients C1/C2, as a restart file gives them (kept while steps
        run with read_states = 1, replaced by reg_rout at the first step
        with read_states = 0)."""
        for s, cells in enumerate(snap["cells"]):
            for k, v in cells.items():
                self.set_var(k, v, s)
        for s, rs in enumerate(snap.get("routing", [])):
            c1, c2 = (rs.get("C1"), rs.get("C2")) if coefficients else (None, None)
            self.set_routing_state(rs["qTIN"], rs["qTR"], s, c1, c2)

    # -- measurement ---------------------------------------------------------------
    def set_profiling(self, on=True, stride=1):
        """Kernel timing on/off; each kernel's launches timed one in stride."""
        self._chk(N.lib().mhmhip_set_profiling(self._ctx, max(1, int(stride)) if on else 0))

    def profile(self):
        """{kernel: (device ms, launches, work units)} accumulated since set_profiling(True)."""
        out = {}
        for k, name in enumerate(N.K_NAMES):
            ms, n, u = C.c_double(), C.c_int64(), C.c_double()
            self._chk(N.lib().mhmhip_get_profile(self._ctx, k, C.byref(ms), C.byref(n), C.byref(u)))
            out[name] = (ms.value, n.value, u.value)
        return out
