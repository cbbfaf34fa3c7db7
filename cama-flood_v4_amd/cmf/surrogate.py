"""Learned spin-up warm start: a per-cell network that predicts the equilibrium
river and floodplain storage of the periodic climatological regime from
catchment attributes and the runoff climatology, trained on equilibria that the
FP64 device physics itself produces (:func:`cmf.spinup.spinup`).

The network (24 -> 256 -> 256 -> 256 -> 2, ReLU) is trained and evaluated by the
project's HIP kernels (include/cmf_mlp.h: BF16 matrix-core products, FP32
accumulation, FP32 master weights, Adam). There is no CPU fallback: the library
must be built. The plain cold start (STORAGE_SEA_SURFACE) stays the fallback of
the spin-up and the judge of the warm start's drift.

Features (24 per cell, all standardised with training-set statistics):
  0 log10 upstream area [km2]       1 log10 channel slope        2 log river width
  3 log river depth (bank height)   4 log river length           5-14 log1p floodplain layer heights
  15 log mean upstream-accumulated runoff   16 CV of its monthly means   17 its seasonal amplitude
  18 log of its value on the first day of the year over its mean (the equilibrium is
     taken at the start of the climatological year)
  19 log Manning depth of the accumulated runoff of the last 30 days of the year
     ((Q n / (W sqrt(slope)))^(3/5): the steady uniform-flow depth the channel
     relaxes to)   20 log downstream distance   21 log RIVSTOMAX
  22 log1p(top floodplain capacity / RIVSTOMAX)   23 mouth flag
Targets: log(RIVSTO / (RIVLEN*RIVWTH) + 1e-3) and log1p(FLDSTO / (RIVLEN*RIVWTH)).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _native
from . import maps, synth
from .maps import CMFMap
from .spinup import Climatology, accumulate, days_to_equilibrium, make_climatology, spinup

N_IN, HIDDEN, N_OUT = 24, 256, 2
EPS_DEPTH = 1e-3


class SurrogateError(RuntimeError):
    pass


def features(m: CMFMap, clim: Climatology) -> np.ndarray:
    n = m.n
    up = accumulate(m, m.grarea) * 1e-6
    nxt = m.next
    riv = nxt >= 0
    j = np.where(riv, nxt, 0)
    slope = np.where(riv, (m.elevtn - m.elevtn[j]) / m.nxtdst, m.elevtn / m.opts.pdstmth)
    slope = np.maximum(slope, 1e-5)
    acc = accumulate(m, clim.monthly)                       # [12, n]
    mean = acc.mean(axis=0)
    cv = acc.std(axis=0) / np.maximum(mean, 1e-12)
    amp = (acc.max(axis=0) - acc.min(axis=0)) / np.maximum(2.0 * mean, 1e-12)
    fl = m.fldhgt                                           # [nlfp, n]
    X = np.empty((n, N_IN), np.float64)
    X[:, 0] = np.log10(up)
    X[:, 1] = np.log10(slope)
    X[:, 2] = np.log(m.rivwth)
    X[:, 3] = np.log(m.rivhgt)
    X[:, 4] = np.log(m.rivlen)
    for l in range(10):
        X[:, 5 + l] = np.log1p(fl[min(l, fl.shape[0] - 1)])
    X[:, 15] = np.log(np.maximum(mean, 1e-12))
    X[:, 16] = cv
    X[:, 17] = amp
    day0 = accumulate(m, clim.daily(0))
    X[:, 18] = np.log(np.maximum(day0, 1e-12) / np.maximum(mean, 1e-12))
    q30 = accumulate(m, np.mean([clim.daily(d) for d in range(335, 365)], axis=0))
    hman = (q30 * m.rivman / (m.rivwth * np.sqrt(slope))) ** 0.6
    X[:, 19] = np.log(hman + EPS_DEPTH)
    X[:, 20] = np.log(m.nxtdst)
    X[:, 21] = np.log(m.rivstomax)
    X[:, 22] = np.log1p(m.fldstomax[-1] / m.rivstomax)
    X[:, 23] = (~riv).astype(np.float64)
    return X.astype(np.float32)


def targets(m: CMFMap, rivsto: np.ndarray, fldsto: np.ndarray) -> np.ndarray:
    a = m.rivlen * m.rivwth
    return np.stack([np.log(rivsto / a + EPS_DEPTH), np.log1p(np.maximum(fldsto, 0.0) / a)], axis=1).astype(np.float32)


def storages(m: CMFMap, y: np.ndarray):
    a = m.rivlen * m.rivwth
    riv = np.maximum(np.exp(y[:, 0].astype(np.float64)) - EPS_DEPTH, 0.0) * a
    fld = np.maximum(np.expm1(y[:, 1].astype(np.float64)), 0.0) * a
    return riv, fld


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

    def grad(self, x: np.ndarray, t: np.ndarray) -> float:
        a = np.ascontiguousarray(x, np.float32)
        b = np.ascontiguousarray(t, np.float32)
        loss = C.c_float(0.0)
        self._check(self.lib.cmf_mlp_grad(self.h, a.ctypes.data, b.ctypes.data, a.shape[0], C.byref(loss)))
        return float(loss.value)

    def fit(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 1024, lr: float = 1e-3,
            seed: int = 0, log=None) -> list:
        a = np.ascontiguousarray(x, np.float32)
        b = np.ascontiguousarray(t, np.float32)
        self._check(self.lib.cmf_mlp_set_data(self.h, a.ctypes.data, b.ctypes.data, a.shape[0]))
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(a.shape[0]).astype(np.int64)
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

    def warm_start(self, m: CMFMap, clim: Climatology, net: Optional[MLP] = None):
        """Predicted equilibrium (rivsto, fldsto) of every cell."""
        return storages(m, self.predict(features(m, clim), net))


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 40, batch: int = 1024, lr: float = 1e-3, seed: int = 0,
          log=None):
    """Standardise, train on the device, return (Surrogate, loss history)."""
    x_mu, x_sd = X.mean(axis=0), X.std(axis=0) + 1e-6
    y_mu, y_sd = Y.mean(axis=0), Y.std(axis=0) + 1e-6
    net = MLP(max_batch=max(batch, 1024), seed=seed)
    hist = net.fit((X - x_mu) / x_sd, (Y - y_mu) / y_sd, epochs, batch, lr, seed, log)
    s = Surrogate(net.params(), x_mu.astype(np.float32), x_sd.astype(np.float32), y_mu.astype(np.float32),
                  y_sd.astype(np.float32))
    net.close()
    return s, hist


# ---------------------------------------------------------------------------
# experiment helpers: synthetic basins, equilibrium labels, cold vs warm spin-up

@dataclasses.dataclass
class Basin:
    net: object
    m: CMFMap
    clim: Climatology


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
    """Predicted storages, repartitioned by CMF_PHYSICS_FLDSTG on the device."""
    from .model import CMFModel
    st = maps.cold_start(m)
    st["rivsto"], st["fldsto"] = rivsto.copy(), fldsto.copy()
    g = CMFModel(m, precision)
    g.set_state(st)
    g.physics_fldstg()
    st.update(g.get_state(["rivsto", "fldsto", "rivdph"]))
    st["rivdph_pre"] = st["rivdph"].copy()
    g.close()
    return st


def equilibrium_labels(basins, tol: float = 1e-6, max_years: int = 10, log=None):
    """Batched FP64 spin-up of every basin from its cold start (one handle, one
    domain per basin) to an annual storage change below tol; returns the
    feature matrix, the targets and the per-basin spin-up years."""
    from .model import CMFModel
    cm, dom_cells = maps.concat_maps([b.m for b in basins])
    offs = np.cumsum([0] + dom_cells)
    init = {}
    sts = [cold_state(b.m) for b in basins]
    for k in ("rivsto", "fldsto", "rivdph_pre"):
        init[k] = np.concatenate([s[k] for s in sts])
    model = CMFModel(cm, "fp64", members=1, dom_cells=dom_cells)
    model.set_state(init)
    res = spinup(model, [b.clim for b in basins], offs, tol=1e-3, stop_tol=tol, max_years=max_years, log=log)
    model.close()
    X = np.concatenate([features(b.m, b.clim) for b in basins])
    Y = np.concatenate([targets(b.m, res.state["rivsto"][offs[d]:offs[d + 1]],
                                res.state["fldsto"][offs[d]:offs[d + 1]]) for d, b in enumerate(basins)])
    return X, Y, res


def compare_spinup(sur: Surrogate, basin: Basin, years: int = 3, tol: float = 1e-3, log=None) -> dict:
    """Cold-start and warm-start FP64 spin-ups of one basin. The cold start runs
    `years` years; its last year is the periodic equilibrium trajectory that both
    runs are judged against."""
    from .model import CMFModel
    m, clim = basin.m, basin.clim
    out = {}
    riv, fld = sur.warm_start(m, clim)
    for name, st in (("cold", cold_state(m)), ("warm", warm_state(m, riv, fld))):
        g = CMFModel(m, "fp64")
        g.set_state(st)
        r = spinup(g, [clim], [0, m.n], tol=tol, max_years=years, stop_tol=0.0, log=log)
        g.close()
        out[name] = r
    eq = out["cold"].daily[-365:]
    s_eq0 = out["cold"].storage[-1, 0]
    res = {}
    for name in ("cold", "warm"):
        r = out[name]
        res[name] = dict(years=int(r.years[0]), days=days_to_equilibrium(r.daily, eq, tol),
                         init_rel_err=float(abs(r.storage[0, 0] - s_eq0) / s_eq0),
                         annual_change=[float(abs(r.storage[y, 0] - r.storage[y - 1, 0]) / r.storage[y, 0])
                                        for y in range(1, r.storage.shape[0])])
    return res
