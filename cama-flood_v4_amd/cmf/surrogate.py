"""Learned spin-up warm start for the periodic climatological regime, trained on
equilibria that the FP64 device physics itself produces (:func:`cmf.spinup.spinup`).

Two layers. The physical prior comes from the network's linear periodic response
(:func:`cmf.spinup.periodic_state`): the steady state of the local-inertial
outflow under the annual-mean runoff (:func:`cmf.spinup.steady_state`), plus the
response of the mass balance, linearised in outflow and storage around it, to the
first six harmonics of the runoff climatology. Its outflows at the year start,
mapped back to storage through the steady state that carries them (the
quasi-steady map keeps the Manning nonlinearity), are the prior. It places the
lakes and other slow storage (what a cold start takes years to fill) within
~0.1% of their equilibrium; what it misses is the nonlinear seasonal response of
the cells with weeks to months of memory. The network learns that residual per
cell, log(equilibrium storage / prior storage), from catchment attributes, the
runoff climatology and the prior itself, on the cells that remember less than
60 days of the forcing (drawn in proportion to the storage they hold), and the
warm start is the prior times the predicted factor on those cells.

The network (24 -> 256 -> 256 -> 256 -> 2, ReLU) is trained and evaluated by the
project's HIP kernels (include/cmf_mlp.h: BF16 matrix-core products, FP32
accumulation, FP32 master weights, Adam). There is no CPU fallback: the library
must be built. The plain cold start (STORAGE_SEA_SURFACE) stays the fallback of
the spin-up.

Features (24 per cell, standardised with training-set statistics; q = upstream-
accumulated runoff, qm = its annual mean, S = steady storage under qm, tau = S/qm):
  0 log10 upstream area [km2]       1 log10 channel slope        2 log river width
  3 log river depth (bank height)   4 log river length           5 log tau [days]
  6 log(T + tau) [days], T = runoff-weighted travel time from the upstream sources
     (cmf.spinup.travel_time): how far back in the forcing the cell remembers
  7-10 log(mean q over the last w days of the year / qm), w = 1, 7, 30, 90
  11 log(prior storage / S)   12 asinh(linear-response elevation - steady elevation) [m]
  13 log(prior storage / (RIVLEN*RIVWTH))   14 log(linear-response year-start outflow / qm)
  15 log qm   16 CV of the monthly means of q   17 seasonal amplitude of q
  18 floodplain share of S
  19 log Manning depth at the cell's spill point of the accumulated runoff of the
     last 30 days of the year ((Q n / (W sqrt(slope)))^(3/5): the depth that carries
     the flow out of the flat reach)   20 log1p length [km] of the flat reach
     downstream of the cell (slope < 2e-5, up to the spill point)
  21 asinh of the cell's elevation above its spill point [m]
  22 log1p(top floodplain capacity / RIVSTOMAX)   23 log1p flat area [km2] upstream
  The spill point is the first cell downstream that is not flat (the cell itself
  when it is not flat). Features 19-21 and 23 describe filled depressions (lakes).
Targets: log(equilibrium RIVSTO + FLDSTO at the year start / prior storage) (used
for the warm start) and log(RIVOUT + FLDOUT at the year start / qm) (auxiliary).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _native
from . import maps, synth
from .maps import CMFMap
from .spinup import (Climatology, accumulate, days_to_equilibrium, make_climatology, periodic_state, spinup,
                     steady_state, travel_time)

N_IN, HIDDEN, N_OUT = 24, 256, 2
FLAT_SLOPE = 2e-5
WINDOWS = (1, 7, 30, 90)          # days before the year start, features 7-10
NHARM = 6                         # runoff harmonics of the linear periodic response
Q_FLOOR = 1e-3                    # flow ratios are clipped to >= this
STO_FLOOR = 1.0                   # m3, storages are floored at this in the log targets
FAST_DAYS = 60.0                  # the network corrects cells that remember less of the forcing


class SurrogateError(RuntimeError):
    pass


@dataclasses.dataclass
class Prior:
    """Annual-mean accumulated runoff qm, the steady state under it, the linear
    periodic response at the year start, and the physical prior: the steady
    state of the response's year-start outflows (the quasi-steady map from flows
    to storage keeps the Manning nonlinearity the linear levels miss)."""
    qm: np.ndarray
    steady: dict
    periodic: dict
    q0: np.ndarray
    state: dict

    @property
    def storage(self) -> np.ndarray:
        return np.maximum(self.state["rivsto"] + self.state["fldsto"], STO_FLOOR)


def prior(m: CMFMap, clim: Climatology) -> Prior:
    qm = accumulate(m, clim.monthly).mean(axis=0)
    st = steady_state(m, qm)
    ps = periodic_state(m, clim, st["sfcelv"], 0.0, NHARM)
    q0 = np.maximum(qm + ps["dq"], Q_FLOOR * qm)
    return Prior(qm, st, ps, q0, steady_state(m, q0))


def features(m: CMFMap, clim: Climatology, pr: Optional[Prior] = None) -> np.ndarray:
    n = m.n
    pr = pr or prior(m, clim)
    up = accumulate(m, m.grarea) * 1e-6
    nxt = m.next
    riv = nxt >= 0
    j = np.where(riv, nxt, 0)
    slope = np.where(riv, (m.elevtn - m.elevtn[j]) / m.nxtdst, m.elevtn / m.opts.pdstmth)
    slope = np.maximum(slope, 1e-5)
    acc = accumulate(m, clim.monthly)                       # [12, n]
    qm = np.maximum(pr.qm, 1e-12)
    sto = np.maximum(pr.steady["rivsto"] + pr.steady["fldsto"], STO_FLOOR)
    tau = sto / qm / 86400.0                                # days
    T = travel_time(m, qm, tau)
    daily = np.stack([clim.daily(d) for d in range(365 - max(WINDOWS), 365)])   # local runoff
    ps = pr.storage
    X = np.empty((n, N_IN), np.float64)
    X[:, 0] = np.log10(up)
    X[:, 1] = np.log10(slope)
    X[:, 2] = np.log(m.rivwth)
    X[:, 3] = np.log(m.rivhgt)
    X[:, 4] = np.log(m.rivlen)
    X[:, 5] = np.log(np.maximum(tau, 1e-6))
    X[:, 6] = np.log(np.maximum(T + tau, 1e-6))
    for k, w in enumerate(WINDOWS):
        qw = accumulate(m, daily[-w:].mean(axis=0))
        X[:, 7 + k] = np.log(np.maximum(qw, 1e-12) / qm)
    X[:, 11] = np.log(ps / sto)
    X[:, 12] = np.arcsinh(pr.periodic["sfcelv"] - pr.steady["sfcelv"])
    X[:, 13] = np.log(ps / (m.rivlen * m.rivwth))
    X[:, 14] = np.log(pr.q0 / qm)
    X[:, 15] = np.log(qm)
    X[:, 16] = acc.std(axis=0) / qm
    X[:, 17] = (acc.max(axis=0) - acc.min(axis=0)) / (2.0 * qm)
    X[:, 18] = pr.steady["fldsto"] / sto
    q30 = accumulate(m, daily[-30:].mean(axis=0))
    flat = slope < FLAT_SLOPE
    reach, spill = flat_reach_km(m, flat)
    hman = (q30[spill] * m.rivman[spill] / (m.rivwth[spill] * np.sqrt(slope[spill]))) ** 0.6
    X[:, 19] = np.log(hman + 1e-3)
    X[:, 20] = np.log1p(reach)
    X[:, 21] = np.arcsinh(m.elevtn - m.elevtn[spill])
    X[:, 22] = np.log1p(m.fldstomax[-1] / m.rivstomax)
    X[:, 23] = np.log1p(accumulate(m, np.where(flat, m.grarea, 0.0)) * 1e-6)
    return X.astype(np.float32)


def flat_reach_km(m: CMFMap, flat: np.ndarray):
    """(length [km] of the flat reach from each cell down to its spill point, the
    spill point): the spill point is the first cell downstream that is not flat,
    or the mouth; a cell that is not flat is its own spill point with length 0.
    NSEQ order puts every cell before its downstream cell, so one backward pass
    suffices."""
    nxt = m.next
    L = np.zeros(m.n)
    spill = np.arange(m.n)
    d = m.nxtdst * 1e-3
    for i in range(m.n - 1, -1, -1):
        if flat[i]:
            j = nxt[i]
            if j >= 0:
                L[i] = d[i] + L[j]
                spill[i] = spill[j]
            else:
                L[i] = d[i]
    return L, spill


def targets(pr: Prior, sto_eq: np.ndarray, q_start: np.ndarray) -> np.ndarray:
    """Network targets from the equilibrium's storage and outflow at the year start."""
    qm = np.maximum(pr.qm, 1e-12)
    return np.stack([np.log(np.maximum(sto_eq, STO_FLOOR) / pr.storage),
                     np.log(np.maximum(q_start / qm, Q_FLOOR))], axis=1).astype(np.float32)


def fast_cells(X: np.ndarray) -> np.ndarray:
    """Cells whose memory of the forcing (feature 6, travel plus residence time) is
    shorter than FAST_DAYS: where the network's correction applies. Slower cells
    (lakes, long flat reaches) keep the physical prior, which already places them
    within ~0.1% of their equilibrium."""
    return X[:, 6] < np.log(FAST_DAYS)


def storage(pr: Prior, y: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Total storage per cell [m3]: the prior times the network's correction
    (first output) on the fast cells."""
    corr = np.where(fast_cells(X), y[:, 0].astype(np.float64), 0.0)
    return pr.storage * np.exp(corr)


def training_set(X: np.ndarray, Y: np.ndarray, sto_eq: np.ndarray):
    """(features, targets, sampling weights) of the cells the network corrects.
    The weights (equilibrium storage over its mean, clipped to [0.05, 50]) draw
    cells in proportion to the storage they hold: what a spin-up's annual change
    measures is the total."""
    sel = fast_cells(X)
    s = np.maximum(sto_eq[sel], STO_FLOOR)
    return X[sel], Y[sel], np.clip(s / s.mean(), 0.05, 50.0)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases; packed as the library expects."""
    rng = np.random.default_rng(seed)
    parts = []
    for di, do in ((N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / di), di * do))
        parts.append(np.zeros(do))
    return np.concatenate(parts).astype(np.float32)


class MLP:
    """Handle of the device network (include/cmf_mlp.h)."""

    def __init__(self, max_batch: int = 65536, seed: int = 0, n_in=N_IN, hidden=HIDDEN, n_out=N_OUT):
        self.lib = _native.hip()
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        self.max_batch = max_batch
        h = C.c_void_p()
        self._check(self.lib.cmf_mlp_create(n_in, hidden, n_out, max_batch, C.byref(h)))
        self.h = h
        self.np = int(self.lib.cmf_mlp_param_count(h))
        self.step = 0
        if n_in == N_IN and hidden == HIDDEN and n_out == N_OUT:
            self.set_params(init_params(seed))

    def _check(self, rc):
        if rc < 0:
            msg = self.lib.cmf_mlp_last_error()
            raise SurrogateError(f"cmf_mlp error {rc}: {msg.decode() if msg else ''}")
        return rc

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, np.float32)
        if a.size != self.np:
            raise SurrogateError(f"{a.size} parameters, expected {self.np}")
        self._check(self.lib.cmf_mlp_set_params(self.h, a.ctypes.data))
        self.step = 0

    def params(self) -> np.ndarray:
        p = np.empty(self.np, np.float32)
        self._check(self.lib.cmf_mlp_get_params(self.h, p.ctypes.data, None))
        return p

    def grads(self) -> np.ndarray:
        g = np.empty(self.np, np.float32)
        self._check(self.lib.cmf_mlp_get_params(self.h, None, g.ctypes.data))
        return g

    def forward(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, np.float32)
        y = np.empty((a.shape[0], self.n_out), np.float32)
        self._check(self.lib.cmf_mlp_forward(self.h, a.ctypes.data, a.shape[0], y.ctypes.data))
        return y

    def set_loss(self, w_over: float) -> None:
        """Weight of over-predictions in the squared-error loss (1 = plain MSE)."""
        self._check(self.lib.cmf_mlp_set_loss(self.h, float(w_over)))

    def grad(self, x: np.ndarray, t: np.ndarray) -> float:
        a = np.ascontiguousarray(x, np.float32)
        b = np.ascontiguousarray(t, np.float32)
        loss = C.c_float(0.0)
        self._check(self.lib.cmf_mlp_grad(self.h, a.ctypes.data, b.ctypes.data, a.shape[0], C.byref(loss)))
        return float(loss.value)

    def fit(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 1024, lr: float = 1e-3,
            seed: int = 0, log=None, weights: Optional[np.ndarray] = None) -> list:
        """Adam over the resident data set; every epoch visits N samples, a
        permutation, or with `weights` a draw with replacement in proportion to
        them (a weighted loss in expectation)."""
        a = np.ascontiguousarray(x, np.float32)
        b = np.ascontiguousarray(t, np.float32)
        self._check(self.lib.cmf_mlp_set_data(self.h, a.ctypes.data, b.ctypes.data, a.shape[0]))
        rng = np.random.default_rng(seed)
        p = None if weights is None else np.asarray(weights, np.float64) / np.sum(weights)
        hist = []
        for ep in range(epochs):
            if p is None:
                order = rng.permutation(a.shape[0]).astype(np.int64)
            else:
                order = rng.choice(a.shape[0], a.shape[0], p=p).astype(np.int64)
            loss = C.c_float(0.0)
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * ep / max(epochs, 1)))     # cosine decay
            nb = self._check(self.lib.cmf_mlp_train(self.h, order.ctypes.data, order.size, batch, float(rate),
                                                    0.9, 0.999, 1e-8, self.step, C.byref(loss)))
            self.step += nb
            hist.append(float(loss.value))
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log(f"epoch {ep + 1}/{epochs}: loss {loss.value:.5f}")
        return hist

    def close(self):
        if getattr(self, "h", None):
            self.lib.cmf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclasses.dataclass
class Surrogate:
    """Trained network plus the feature/target standardisation."""
    params: np.ndarray
    x_mu: np.ndarray
    x_sd: np.ndarray
    y_mu: np.ndarray
    y_sd: np.ndarray

    def save(self, path: str) -> None:
        np.savez(path, params=self.params, x_mu=self.x_mu, x_sd=self.x_sd, y_mu=self.y_mu, y_sd=self.y_sd)

    @staticmethod
    def load(path: str) -> "Surrogate":
        z = np.load(path, allow_pickle=False)
        return Surrogate(z["params"], z["x_mu"], z["x_sd"], z["y_mu"], z["y_sd"])

    def predict(self, X: np.ndarray, net: Optional[MLP] = None) -> np.ndarray:
        own = net is None
        net = net or MLP(max_batch=1 << 18)
        net.set_params(self.params)
        y = net.forward((X - self.x_mu) / self.x_sd)
        if own:
            net.close()
        return y * self.y_sd + self.y_mu

    def warm_start(self, m: CMFMap, clim: Climatology, pr: Optional[Prior] = None,
                   net: Optional[MLP] = None) -> np.ndarray:
        """Predicted equilibrium storage (RIVSTO + FLDSTO) of every cell at the year start."""
        pr = pr or prior(m, clim)
        X = features(m, clim, pr)
        return storage(pr, self.predict(X, net), X)


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 40, batch: int = 1024, lr: float = 1e-3, seed: int = 0,
          log=None, w_over: float = 1.0, weights: Optional[np.ndarray] = None):
    """Standardise, train on the device, return (Surrogate, loss history).
    w_over > 1 weights over-predictions more than under-predictions; `weights`
    (per sample) sets how often each sample is drawn."""
    x_mu, x_sd = X.mean(axis=0), X.std(axis=0) + 1e-6
    y_mu, y_sd = Y.mean(axis=0), Y.std(axis=0) + 1e-6
    net = MLP(max_batch=max(batch, 1024), seed=seed)
    net.set_loss(w_over)
    hist = net.fit((X - x_mu) / x_sd, (Y - y_mu) / y_sd, epochs, batch, lr, seed, log, weights)
    s = Surrogate(net.params(), x_mu.astype(np.float32), x_sd.astype(np.float32), y_mu.astype(np.float32),
                  y_sd.astype(np.float32))
    net.close()
    return s, hist


def load_training_set(path: str):
    """(features, targets [n, 2], sampling weights, basin index) of a stored
    training set (tests/golden/make_surrogate_set.py): features and the storage
    target as stored, the auxiliary flow target replaced by the storage target (the
    set does not keep it), weights from the equilibrium storage as in
    training_set."""
    z = np.load(path, allow_pickle=False)
    X = z["X"].astype(np.float32)
    y0 = z["Y0"].astype(np.float32)
    s = np.maximum(z["sto"].astype(np.float64), STO_FLOOR)
    return X, np.stack([y0, y0], axis=1), np.clip(s / s.mean(), 0.05, 50.0), z["bid"].astype(np.int64)


def correction_error(pred: np.ndarray, X: np.ndarray, sto_eq: np.ndarray, prior_sto: np.ndarray) -> dict:
    """How far the prior and the warm start (prior x the predicted correction on
    the fast cells) lie from the equilibrium on the cells the network corrects:
    mean |log(start / equilibrium)| weighted by equilibrium storage, and the
    storage-weighted L1 share of the whole network."""
    fast = fast_cells(X)
    eq = np.maximum(sto_eq, STO_FLOOR)
    truth = np.log(eq / prior_sto)
    y = pred[:, 0].astype(np.float64)
    w = eq[fast] / max(eq[fast].sum(), STO_FLOOR)
    warm = prior_sto * np.exp(np.where(fast, y, 0.0))
    return {"prior_log_err": float((np.abs(truth[fast]) * w).sum()),
            "warm_log_err": float((np.abs(truth[fast] - y[fast]) * w).sum()),
            "prior_l1": float(np.abs(prior_sto - eq).sum() / eq.sum()),
            "warm_l1": float(np.abs(warm - eq).sum() / eq.sum()),
            "fast_storage_share": float(eq[fast].sum() / eq.sum())}


# ---------------------------------------------------------------------------
# experiment helpers: synthetic basins, equilibrium labels, cold vs warm spin-up

@dataclasses.dataclass
class Basin:
    net: object
    m: CMFMap
    clim: Climatology


def basins_from_networks(nets: int, per_net: int, seed: int, nx: int = 1440, ny: int = 720,
                         min_cells: int = 30, max_lat: Optional[float] = None):
    """River basins cut from glb_15min-shaped global networks: the `per_net`
    largest basins of each, so the set spans small coastal basins to continental
    ones with filled depressions. Every basin gets its own runoff climatology on
    the network's gamma(k=0.5) spatial field: mean 0.3-6 mm/day, seasonal
    amplitude 0-0.9, random phase and month-to-month factors."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nets):
        s_ = seed * 100 + k
        net = synth.make_network(nx, ny, seed=s_, max_lat=max_lat)
        m = maps.build_map(net, maps.Options())
        base = m.vec(synth.make_runoff(net, 1, s_, mean_mm_day=1.0)[0])        # m3/s at 1 mm/day mean
        b = maps.basin_of(m)
        ids, cnt = np.unique(b, return_counts=True)
        order = np.argsort(-cnt)
        for i in order[:per_net]:
            if cnt[i] < min_cells:
                break
            cells = np.flatnonzero(b == ids[i])
            sub = maps.subset_map(m, cells)
            mean = float(np.exp(rng.uniform(np.log(0.3), np.log(6.0))))
            amp, phase = float(rng.uniform(0.0, 0.9)), float(rng.uniform(0.0, 12.0))
            cyc = 1.0 + amp * np.sin(2.0 * np.pi * (np.arange(12) - phase) / 12.0)
            fac = np.exp(rng.normal(-0.02, 0.2, 12))
            monthly = np.maximum(cyc * fac, 0.0)[:, None] * (mean * base[cells])[None, :]
            out.append(Basin(net, sub, Climatology(np.ascontiguousarray(monthly), mean, amp, phase)))
    return out


def make_basins(count: int, seed: int, cells=(400, 3000), res: float = 0.25, opts=None):
    """glb_15min-shaped regional basins (0.25 degree cells) of random size with
    varied gamma-runoff climatologies (mean 0.3-6 mm/day, seasonal amplitude
    0-0.9, random phase)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        target = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        side = max(8, int(np.sqrt(target / 0.6 / 1.5)))
        nx, ny = int(side * 1.5), side
        north = float(rng.uniform(-30.0, 60.0)) + res * ny / 2
        west = float(rng.uniform(-180.0, 180.0 - res * nx))
        net = synth.make_network(nx, ny, seed=seed * 1000 + b, land_frac=0.6, west=west, east=west + res * nx,
                                 north=north, south=north - res * ny, regional=True)
        m = maps.build_map(net, opts or maps.Options())
        clim = make_climatology(net, m, seed * 1000 + b, mean_mm_day=float(np.exp(rng.uniform(np.log(0.3),
                                                                                                np.log(6.0)))),
                                amp=float(rng.uniform(0.0, 0.9)), phase=float(rng.uniform(0.0, 12.0)))
        out.append(Basin(net, m, clim))
    return out


def cold_state(m: CMFMap, precision: str = "fp64") -> dict:
    """STORAGE_SEA_SURFACE cold start followed by the initial CMF_PHYSICS_FLDSTG
    on the device."""
    from .model import CMFModel
    st = maps.cold_start(m)
    g = CMFModel(m, precision)
    g.set_state(st)
    g.physics_fldstg()
    st.update(g.get_state(["rivsto", "fldsto", "rivdph"]))
    g.close()
    return st


def warm_state(m: CMFMap, rivsto: np.ndarray, fldsto: np.ndarray, precision: str = "fp64") -> dict:
    """Predicted storages entered as the reference enters a storage-only restart
    (LRESTART + LSTOONLY, src/cmf_drv_control_mod.F90:273-278): flows zero,
    D2FLDSTO_PRE = P2FLDSTO (src/cmf_ctrl_restart_mod.F90:128-131), then on the
    device CMF_PHYSICS_FLDSTG and CMF_CALC_OUTPRE, which gives the previous-step
    flows (RIVOUT_PRE, FLDOUT_PRE) and RIVDPH_PRE of the predicted surface."""
    from .model import CMFModel
    st = maps.cold_start(m)
    st["rivsto"], st["fldsto"] = rivsto.copy(), fldsto.copy()
    g = CMFModel(m, precision)
    g.set_state(st)
    g.physics_fldstg()
    g.calc_outpre()
    st.update(g.get_state(["rivsto", "fldsto", "rivdph", "rivout", "fldout", "rivout_pre", "fldout_pre",
                           "rivdph_pre"]))
    st["fldsto_pre"] = st["fldsto"].copy()
    g.close()
    return st


def prior_state(m: CMFMap, clim: Climatology, pr: Optional[Prior] = None, which: str = "prior") -> dict:
    """Warm start from the physics alone: the physical prior the network refines
    ("prior") or the steady state under the annual-mean runoff ("steady")."""
    pr = pr or prior(m, clim)
    st = pr.state if which == "prior" else pr.steady
    return warm_state(m, st["rivsto"], st["fldsto"])


def equilibrium_labels(basins, tol: float = 1e-6, max_years: int = 10, log=None, start: str = "prior"):
    """Batched FP64 spin-up of every basin (one handle, one domain per basin) to
    an annual storage change below tol, from the steady-state prior (start =
    "prior") or from the cold start ("cold"); returns the feature matrix, the
    targets and the per-basin spin-up years. The equilibrium does not depend on
    the start; the prior only gets there sooner."""
    from .model import CMFModel
    cm, dom_cells = maps.concat_maps([b.m for b in basins])
    offs = np.cumsum([0] + dom_cells)
    priors = [prior(b.m, b.clim) for b in basins]
    X = np.concatenate([features(b.m, b.clim, pr) for b, pr in zip(basins, priors)])
    init = {}
    if start == "prior":
        sts = [prior_state(b.m, b.clim, pr) for b, pr in zip(basins, priors)]
    elif start == "cold":
        sts = [cold_state(b.m) for b in basins]
    else:
        raise ValueError(f"start must be 'prior' or 'cold', not {start!r}")
    for k in START_KEYS:
        init[k] = np.concatenate([s[k] for s in sts])
    model = CMFModel(cm, "fp64", members=1, dom_cells=dom_cells)
    model.set_state(init)
    res = spinup(model, [b.clim for b in basins], offs, tol=1e-3, stop_tol=tol, max_years=max_years, log=log)
    model.close()
    st = res.state
    s0 = st["rivsto"] + st["fldsto"]
    q0 = st["rivout"] + st["fldout"]
    Y = np.concatenate([targets(pr, s0[offs[d]:offs[d + 1]], q0[offs[d]:offs[d + 1]])
                        for d, pr in enumerate(priors)])
    return X, Y, res


RESUME_KEYS = ("rivsto", "fldsto", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre", "rivout")
# what a start hands to the spin-up: storages, the previous-step flows and stage,
# and the current flows (zero for every start, as after a restart)
START_KEYS = ("rivsto", "fldsto", "rivdph_pre", "rivout", "fldout", "rivout_pre", "fldout_pre", "fldsto_pre")


def spinup_from_starts(m: CMFMap, clim: Climatology, starts: Optional[dict], years: int = 20, tol: float = 1e-3,
                       log=None, names=None, resume: Optional[dict] = None,
                       time_budget: Optional[float] = None, progress_days: int = 0) -> dict:
    """FP64 spin-up of one network from several starts, run as the members of one
    handle (same map and forcing, own adaptive steps; the member count is padded
    to a power of two by repeating the last start), until every start's annual
    storage change is below tol, `years` pass, or (time_budget) the year that
    passes time_budget seconds ends. The total storage of every member is summed
    on the device at the end of every day (progress_days > 0: a log line every
    that many days). The result's "resume" dict (the
    restart state of every member plus the history) continues the run exactly
    when passed back as `resume`."""
    import time
    from .model import CMFModel
    from .spinup import DAYS
    names = list(names if names is not None else starts)
    M = 1 << max(len(names) - 1, 0).bit_length()
    single = M == 1                     # one start: the single-model kernels
    model = CMFModel(m, "fp64") if single else CMFModel(m, "fp64", members=M, dom_cells=[m.n])
    if resume is None:
        sts = [starts[nm] for nm in names] + [starts[names[-1]]] * (M - len(names))
        model.set_state({k: np.stack([st[k] for st in sts]) for k in START_KEYS})
        model.physics_fldstg()
        hist, daily = [model.instance_storage()[:len(names)]], []
        conv = [0] * len(names)
        substeps, spent = 0, 0.0
    else:
        model.set_state({k: resume[k] for k in RESUME_KEYS})
        hist, daily = list(resume["hist"]), list(resume["daily"])
        conv = [int(c) for c in resume["conv"]]
        substeps, spent = int(resume["member_substeps"]), float(resume["seconds"])
    t0 = time.time()
    one = np.ones(M)
    y0 = len(hist) - 1
    for y in range(y0 + 1, years + 1):
        for d in range(DAYS):
            if single:
                model.set_runoff(clim.daily(d))
                substeps += int(model.physics_advance(86400.0))
            else:
                model.set_runoff_scaled(clim.daily(d), one)
                nt, _ = model.ensemble_advance(86400.0)
                substeps += int(nt[:len(names)].sum())
            daily.append(model.instance_storage()[:len(names)])
            if log and progress_days and (d + 1) % progress_days == 0 and d + 1 < DAYS:
                log(f"year {y} day {d + 1} ({time.time() - t0:.1f}s)")
        now = daily[-1]
        ch = np.abs(now - hist[-1]) / now
        hist.append(now)
        for k in range(len(names)):
            if conv[k] == 0 and ch[k] < tol:
                conv[k] = y
        if log:
            log(f"year {y}: annual change " + " ".join(f"{nm} {ch[k]:.3e}" for k, nm in enumerate(names)) +
                f" ({time.time() - t0:.1f}s)")
        if all(conv) or (time_budget is not None and time.time() - t0 > time_budget):
            break
    seconds = spent + (time.time() - t0)
    st = {k: np.reshape(v, (M, -1)) for k, v in model.get_state(list(RESUME_KEYS)).items()}
    model.close()
    hist_a, daily_a = np.asarray(hist), np.asarray(daily)
    res = dict(names=names, hist=hist_a, daily=daily_a, conv=conv, seconds=round(seconds, 1),
               member_substeps=substeps, cells=m.n)
    res["resume"] = dict(hist=hist_a, daily=daily_a, conv=np.asarray(conv), seconds=seconds,
                         member_substeps=substeps, cells=m.n, **st)
    return res


def summarize_spinups(runs, tol: float = 1e-3) -> dict:
    """Years (first year with an annual change below tol) and days to the
    periodic equilibrium of every start of one or more spinup_from_starts runs
    on the same network. The equilibrium all starts are measured against is the
    last year of the start whose last annual change is smallest."""
    from .spinup import DAYS
    best = None
    for r in runs:
        last = np.abs(r["hist"][-1] - r["hist"][-2]) / r["hist"][-1]
        for k, nm in enumerate(r["names"]):
            if best is None or last[k] < best[0]:
                best = (float(last[k]), nm, r["daily"][-DAYS:, k], float(r["hist"][-1, k]))
    _, ref_name, eq, ref_total = best
    res = {"reference_member": ref_name}
    for r in runs:
        for k, nm in enumerate(r["names"]):
            h = r["hist"][:, k]
            res[nm] = dict(years=int(r["conv"][k]), days=days_to_equilibrium(r["daily"][:, k], eq, tol),
                           init_rel_err=float(abs(h[0] - ref_total) / ref_total),
                           annual_change=[float(abs(h[y] - h[y - 1]) / h[y]) for y in range(1, h.size)],
                           years_run=int(h.size - 1))
    return res


def compare_spinup(sur: Surrogate, basin: Basin, years: int = 20, tol: float = 1e-3, log=None,
                   groups=(("cold", "steady", "prior", "warm"),)) -> dict:
    """FP64 spin-ups of one network from the cold start, from the steady state
    under the annual-mean runoff, from the physical prior (linear periodic
    response) and from the surrogate's warm start. Each group of starts runs as
    the members of one handle (spinup_from_starts), until every annual storage
    change is below tol or `years` pass; summarize_spinups measures them against
    one equilibrium."""
    import time
    m, clim = basin.m, basin.clim
    pr = prior(m, clim)
    make = {"cold": lambda: cold_state(m), "steady": lambda: prior_state(m, clim, pr, "steady"),
            "prior": lambda: prior_state(m, clim, pr),
            "warm": lambda: warm_state(m, sur.warm_start(m, clim, pr), np.zeros(m.n))}
    t0 = time.time()
    runs = [spinup_from_starts(m, clim, {nm: make[nm]() for nm in g}, years=years, tol=tol, log=log)
            for g in groups]
    res = summarize_spinups(runs, tol)
    res["years_run"] = max(int(r["hist"].shape[0] - 1) for r in runs)
    res["member_years"] = int(sum((r["hist"].shape[0] - 1) * len(r["names"]) for r in runs))
    res["spinup_seconds"] = round(sum(r["seconds"] for r in runs), 1)
    res["seconds"] = round(time.time() - t0, 1)
    return res
