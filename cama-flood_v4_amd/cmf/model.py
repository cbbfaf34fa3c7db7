"""Device-resident CaMa-Flood physics: the host-side mirror of CMF_CTRL_PHYSICS_MOD.

``CMFModel`` owns one handle of the HIP library (include/cmf_hip.h) and exposes the
reference's two physics entry points with their names and meaning:

* ``physics_fldstg()``  -- CMF_PHYSICS_FLDSTG (src/cmf_ctrl_physics_mod.F90:288-305)
* ``physics_advance(dt)`` -- CMF_PHYSICS_ADVANCE (:21-155); returns NT

plus the state movement a device-resident model needs (``set_state``,
``get_state``, ``set_runoff``). Failures raise ``CMFError`` with the library's
message; there is no CPU fallback.

Ensembles: ``CMFModel(concatenated_map, members=M, dom_cells=[...])`` runs M
runoff realisations of every domain of a concatenated map
(:func:`cmf.maps.concat_maps`) in one handle; per-cell arrays then have shape
(M, NSEQALL) and path arrays (M, NPTHLEV, NPTHOUT) / (M, NPTHOUT).
``ensemble_advance`` advances every instance with its own adaptive step.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional

import numpy as np

from . import _native
from ._abi import (OUTPUT_FIELDS, STATE_FIELDS, TUNE_FIELDS, cmf_dims, cmf_ensemble, cmf_map, cmf_options,
                   cmf_output, cmf_state, cmf_tune, ptr)
from .maps import CMFMap

PREC_FP64 = 0
PREC_MIXED = 1
NGLB = 11
GLB_NAMES = ("stopre", "stonxt", "stonew", "rivinf", "rivout", "stopre2", "stonew2",
             "rivsto", "fldsto", "fldare", "levsto")


class CMFError(RuntimeError):
    pass


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class CMFModel:
    def __init__(self, m: CMFMap, precision: str = "fp64", opts=None, members: int = 1,
                 dom_cells: Optional[Iterable[int]] = None):
        self.lib = _native.hip()
        self.map = m
        o = opts if opts is not None else m.opts
        self.opts = o
        self.members = int(members)
        self.dom_cells = [m.n] if dom_cells is None else [int(c) for c in dom_cells]
        self.ninst = len(self.dom_cells) * self.members
        self.ensemble = self.ninst > 1 or dom_cells is not None
        self.prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[precision]
        d = cmf_dims(m.n, m.nseqriv, m.nlfp, m.npth if o.lpthout else 0,
                     m.npthlev if o.lpthout else 0, m.ndam if o.ldamout else 0)
        op = cmf_options(int(o.ladpstp), int(o.lfplain), int(o.lfldout), int(o.lpthout),
                         int(o.ldamout), int(o.llevee), self.prec,
                         o.pgrv, o.pmanfld, o.pdstmth, o.pcadp, o.pminslp, 8,
                         int(bool(getattr(o, "ldamh22", False))))
        h = C.c_void_p()
        if self.ensemble:
            dc = np.ascontiguousarray(self.dom_cells, np.int32)
            ens = cmf_ensemble(len(self.dom_cells), dc.ctypes.data, self.members)
            self._check(self.lib.cmf_hip_create_ensemble(C.byref(d), C.byref(op), C.byref(ens), C.byref(h)))
        else:
            self._check(self.lib.cmf_hip_create(C.byref(d), C.byref(op), C.byref(h)))
        self.h = h
        self._keep = []
        self._set_map(m)
        if getattr(m.opts, "lkine", False):
            self.set_kinematic(True)
        if getattr(m.opts, "lslpmix", False):
            self.set_slopemix(True)

    # ------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.cmf_hip_last_error()
            raise CMFError(f"cmf_hip error {rc}: {msg.decode() if msg else ''}")

    def _p(self, a):
        self._keep.append(a)
        return ptr(a)

    def _set_map(self, m: CMFMap):
        o = self.opts
        mp = cmf_map()
        i1next = np.where(m.next >= 0, m.next + 1, m.next).astype(np.int32)
        mp.i1next = self._p(i1next)
        for k in ("grarea", "elevtn", "nxtdst", "rivlen", "rivwth", "rivhgt", "rivman",
                  "rivstomax", "rivelv", "dwnelv", "fldhgt", "fldstomax", "fldgrd"):
            setattr(mp, k, self._p(_f64(getattr(m, k))))
        mp.dfrcinc = m.dfrcinc
        mp.i2mask = self._p(_i32(m.i2mask))
        if o.lpthout:
            mp.pth_upst = self._p(_i32(np.where(m.pth_upst >= 0, m.pth_upst + 1, 0)))
            mp.pth_down = self._p(_i32(np.where(m.pth_down >= 0, m.pth_down + 1, 0)))
            for k in ("pth_dst", "pth_elv", "pth_wth", "pth_man"):
                setattr(mp, k, self._p(_f64(getattr(m, k))))
        if o.llevee:
            for k in ("levhgt", "levfrc", "bashgt", "levdst", "levbassto", "levtopsto", "levfilsto"):
                setattr(mp, k, self._p(_f64(getattr(m, k))))
        if o.ldamout:
            mp.i1dam = self._p(_i32(m.i1dam))
            mp.damseq = self._p(_i32(np.where(m.damseq >= 0, m.damseq + 1, -9999)))
            mp.damstat = self._p(_i32(m.damstat))
            for k in ("dam_convol", "dam_emevol", "dam_adjvol", "dam_qa", "dam_qn", "dam_qf"):
                setattr(mp, k, self._p(_f64(getattr(m, k))))
            if getattr(o, "ldamh22", False):
                mp.dam_norvol = self._p(_f64(m.dam_norvol))
                mp.dam_r_volupa = self._p(_f64(m.dam_r_volupa))
        self._check(self.lib.cmf_hip_set_map(self.h, C.byref(mp)))
        self._keep.clear()

    # ------------------------------------------------------------------
    def set_state(self, st: dict, shared: bool = False) -> None:
        """Host -> device copy of the given arrays (names of cmf_state). shared:
        each array holds one plane that every member of an ensemble receives."""
        s = cmf_state()
        nplanes = 1 if shared else self.members
        keep = []
        m = self.map
        for k, v in st.items():
            if k in STATE_FIELDS and v is not None:
                if not self._available(k):
                    continue
                a = _f64(v)
                if k in ("pthflw", "pthflw_pre", "pthflw_avg"):
                    want = m.npthlev * m.npth
                elif k in ("pthflwsum", "pthflwsum_avg"):
                    want = m.npth
                else:
                    want = m.n
                if a.size != want * nplanes:
                    raise CMFError(f"set_state: {k} has {a.size} values, expected {want * nplanes}")
                keep.append(a)
                setattr(s, k, a.ctypes.data)
        if shared:
            self._check(self.lib.cmf_hip_set_state_shared(self.h, C.byref(s)))
        else:
            self._check(self.lib.cmf_hip_set_state(self.h, C.byref(s)))

    def get_state(self, names: Optional[Iterable[str]] = None) -> dict:
        m, o = self.map, self.opts
        names = list(names) if names is not None else [
            k for k in STATE_FIELDS if self._available(k)]
        out, s = {}, cmf_state()
        for k in names:
            if not self._available(k):
                continue
            lead = (self.members,) if self.members > 1 else ()
            if k in ("pthflw", "pthflw_pre", "pthflw_avg"):
                a = np.zeros(lead + (m.npthlev, m.npth))
            elif k in ("pthflwsum", "pthflwsum_avg"):
                a = np.zeros(lead + (m.npth,))
            else:
                a = np.zeros(lead + (m.n,))
            out[k] = a
            setattr(s, k, a.ctypes.data)
        self._check(self.lib.cmf_hip_get_state(self.h, C.byref(s)))
        return out

    def _available(self, k: str) -> bool:
        o = self.opts
        if k in ("levsto", "levdph"):
            return o.llevee
        if k in ("damsto", "daminf", "daminf_avg"):
            return o.ldamout
        if k.startswith("pthflw"):
            return o.lpthout
        return True

    def set_runoff(self, runoff: np.ndarray) -> None:
        """D2RUNOFF; shape (NSEQALL,) or (members, NSEQALL) for ensembles."""
        a = _f64(runoff)
        if a.size != self.members * self.map.n:
            raise CMFError(f"runoff has {a.size} values, expected {self.members * self.map.n}")
        self._check(self.lib.cmf_hip_set_runoff(self.h, a.ctypes.data))

    def set_runoff_scaled(self, base: np.ndarray, factor: np.ndarray) -> None:
        """Runoff of instance k (domain d, member m; k = d*members + m) =
        base[cell] * factor[k]."""
        b, f = _f64(base), _f64(factor)
        if b.size != self.map.n or f.size != self.ninst:
            raise CMFError("set_runoff_scaled: base needs NSEQALL values, factor one per instance")
        self._check(self.lib.cmf_hip_set_runoff_scaled(self.h, b.ctypes.data, f.ctypes.data))

    def set_inpmat(self, inpx: np.ndarray, inpy: np.ndarray, inpa: np.ndarray, nxin: int, nyin: int,
                   drofunit: float = 86400.0 * 1000.0, rmis: float = float(np.float32(1e20))) -> None:
        """Input matrix of the device runoff interpolation (ROFF_INTERP): INPX,
        INPY (1-based) and INPA [m2] as [inpn, NSEQALL] arrays."""
        ix = np.ascontiguousarray(inpx, np.int32)
        iy = np.ascontiguousarray(inpy, np.int32)
        a = _f64(inpa)
        if ix.ndim != 2 or ix.shape[1] != self.map.n or iy.shape != ix.shape or a.shape != ix.shape:
            raise CMFError("set_inpmat: inpx, inpy, inpa must be [inpn, NSEQALL]")
        self.nxin, self.nyin = int(nxin), int(nyin)
        self._check(self.lib.cmf_hip_set_inpmat(self.h, ix.shape[0], self.nxin, self.nyin, ix.ctypes.data,
                                                iy.ctypes.data, a.ctypes.data, float(drofunit), float(rmis)))

    def set_runoff_grid(self, grid: np.ndarray) -> None:
        """Interpolate one input-grid runoff field [nyin, nxin] (input units,
        e.g. mm/day) into D2RUNOFF on the device."""
        g = _f64(grid)
        if g.size != self.nxin * self.nyin:
            raise CMFError("set_runoff_grid: grid must hold NXIN*NYIN values")
        self._check(self.lib.cmf_hip_set_runoff_grid(self.h, g.ctypes.data))

    def set_runoff_records(self, grids: np.ndarray) -> None:
        """Keep a run's input-grid runoff fields [nrec, nyin, nxin] resident in
        HBM (one copy); use_runoff_record(k) then forms D2RUNOFF from field k."""
        g = _f64(grids)
        if g.size == 0 or g.size % (self.nxin * self.nyin):
            raise CMFError("set_runoff_records: grids must hold nrec*NXIN*NYIN values")
        self._check(self.lib.cmf_hip_set_runoff_records(self.h, g.size // (self.nxin * self.nyin), g.ctypes.data))

    def use_runoff_record(self, k: int) -> None:
        """Interpolate resident field k (0-based) into D2RUNOFF on the device."""
        self._check(self.lib.cmf_hip_use_runoff_record(self.h, int(k)))

    def interp_runoff(self, device_ptr: int) -> None:
        """Interpolate a runoff field that is already in device memory."""
        self._check(self.lib.cmf_hip_interp_runoff(self.h, C.c_void_p(device_ptr)))

    # binary restart from / into device state (WRTE_REST_BIN / READ_REST_BIN) --
    def _grid(self):
        m = self.map
        return (m.nx, m.ny, np.ascontiguousarray(m.seqx, np.int32), np.ascontiguousarray(m.seqy, np.int32))

    def restart_write(self, path: str, lrestdbl: bool = True, lstoonly: bool = False, dmis: float = 1e20,
                      rmis: float = float(np.float32(1e20))) -> None:
        nx, ny, sx, sy = self._grid()
        self._check(self.lib.cmf_hip_restart_write(self.h, path.encode(), nx, ny, sx.ctypes.data, sy.ctypes.data,
                                                   int(lrestdbl), int(lstoonly), float(dmis), float(rmis)))

    def output_write(self, var: str, path: str, loutvec: bool = False,
                     rmis: float = float(np.float32(1e20))) -> None:
        """Append one binary output record of `var` (CMF_OUTPUT_WRITE) to `path`."""
        nx, ny, sx, sy = self._grid()
        self._check(self.lib.cmf_hip_output_write(self.h, var.encode(), path.encode(), nx, ny, sx.ctypes.data,
                                                  sy.ctypes.data, int(loutvec), float(rmis)))

    def restart_read(self, path: str, lrestdbl: bool = True, lstoonly: bool = False) -> None:
        nx, ny, sx, sy = self._grid()
        self._check(self.lib.cmf_hip_restart_read(self.h, path.encode(), nx, ny, sx.ctypes.data, sy.ctypes.data,
                                                  int(lrestdbl), int(lstoonly)))

    # sea-level boundary (CMF_BOUNDARY_UPDATE) ----------------------------
    def set_dwnelv(self, dwnelv: np.ndarray) -> None:
        a = _f64(dwnelv)
        if a.size != self.map.n:
            raise CMFError("set_dwnelv: NSEQALL values expected")
        self._check(self.lib.cmf_hip_set_dwnelv(self.h, a.ctypes.data))

    def set_upsinf(self, upsinf: np.ndarray) -> None:
        """LUPSINF: upstream inflow D2UPSINF [m3/s] added to the river in STONXT."""
        a = _f64(upsinf)
        if a.size != self.members * self.map.n:
            raise CMFError(f"set_upsinf: {self.members * self.map.n} values expected")
        self._check(self.lib.cmf_hip_set_upsinf(self.h, a.ctypes.data))

    def set_kinematic(self, on: bool) -> None:
        """LKINE: kinematic-wave outflow (CMF_CALC_OUTFLW_KINE) instead of the
        local inertial one; applied from maps.Options.lkine at construction."""
        self._check(self.lib.cmf_hip_set_kinematic(self.h, int(bool(on))))

    def set_slopemix(self, on: bool) -> None:
        """LSLPMIX: CMF_CALC_OUTFLW_KINEMIX, the kinematic wave on the cells with
        I2MASK != 0 of the map (maps.apply_slope_mask), local inertial elsewhere;
        applied from maps.Options.lslpmix at construction."""
        self._check(self.lib.cmf_hip_set_slopemix(self.h, int(bool(on))))

    def update_boundary(self, meansl: Optional[np.ndarray] = None, sealev: Optional[np.ndarray] = None) -> None:
        """D2DWNELV = D2ELEVTN (+ meansl) (+ sealev) on the device."""
        arrs = [None if a is None else _f64(a) for a in (meansl, sealev)]
        for a in arrs:
            if a is not None and a.size != self.map.n:
                raise CMFError("update_boundary: NSEQALL values expected")
        self._check(self.lib.cmf_hip_update_boundary(self.h, *(None if a is None else a.ctypes.data for a in arrs)))

    # output-period diagnostics (CMF_DIAG_*_OUTPUT) ----------------------
    def diag_reset_output(self) -> None:
        self._check(self.lib.cmf_hip_diag_reset_output(self.h))

    def diag_avemax_output(self, dt: float) -> None:
        self._check(self.lib.cmf_hip_diag_avemax_output(self.h, float(dt)))

    def diag_getave_output(self) -> float:
        """Turns the accumulated sums into averages; returns NADD_out."""
        nadd = C.c_double(0.0)
        self._check(self.lib.cmf_hip_diag_getave_output(self.h, C.byref(nadd)))
        return nadd.value

    def get_output(self, names: Optional[Iterable[str]] = None) -> dict:
        m = self.map
        names = list(names) if names is not None else [
            k for k in OUTPUT_FIELDS if (k != "pthflw" or self.opts.lpthout) and (k != "daminf" or self.opts.ldamout)]
        lead = (self.members,) if self.members > 1 else ()
        out, s = {}, cmf_output()
        for k in names:
            a = np.zeros(lead + ((m.npthlev, m.npth) if k == "pthflw" else (m.n,)))
            out[k] = a
            setattr(s, k, a.ctypes.data)
        self._check(self.lib.cmf_hip_get_output(self.h, C.byref(s)))
        return out

    # ------------------------------------------------------------------
    def physics_fldstg(self) -> None:
        self._check(self.lib.cmf_hip_physics_fldstg(self.h))

    def calc_outpre(self) -> None:
        """CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421): previous-step flows
        of a storage-only restart from the current stage."""
        self._check(self.lib.cmf_hip_calc_outpre(self.h))

    def physics_advance(self, dt: float, nt_cap: int = 0, budget: bool = False):
        """One outer step of length dt [s]. Returns NT, or (NT, budget array
        [NT, 11]) when budget=True."""
        d = C.c_double(dt)
        nt = C.c_int32(0)
        if budget:
            cap = 4096
            g = np.zeros((cap, NGLB))
            self._check(self.lib.cmf_hip_physics_advance_capped(
                self.h, C.byref(d), nt_cap, C.byref(nt), g.ctypes.data, cap))
            nrun = min(nt.value if nt_cap <= 0 else min(nt.value, nt_cap), cap)
            return nt.value, g[:nrun].copy()
        self._check(self.lib.cmf_hip_physics_advance_capped(
            self.h, C.byref(d), nt_cap, C.byref(nt), None, 0))
        return nt.value

    def physics_advance_noalloc(self, dt: float):
        """CMF_PHYSICS_ADVANCE through the allocation-free entry point (no device
        allocation during the call; scratch in the library's fixed per-GPU array).
        Returns (NT, budget sums [NT, 11])."""
        d = C.c_double(dt)
        nt = C.c_int32(0)
        cap = 4096
        g = np.zeros((cap, NGLB))
        self._check(self.lib.cmf_hip_physics_advance_noalloc(self.h, C.byref(d), C.byref(nt), g.ctypes.data, cap))
        return nt.value, g[:min(nt.value, cap)].copy()

    def ensemble_advance(self, dt: float, nt_cap: int = 0):
        """One outer step of length dt of every instance. Returns (NT, DT) arrays
        of length ninst (instance k = domain*members + member)."""
        nt = np.zeros(self.ninst, np.int32)
        dto = np.zeros(self.ninst)
        self._check(self.lib.cmf_hip_ensemble_advance(self.h, float(dt), int(nt_cap), nt.ctypes.data,
                                                      dto.ctypes.data))
        return nt, dto

    def instance_storage(self) -> np.ndarray:
        """Total storage (P2RIVSTO + P2FLDSTO (+ P2LEVSTO)) of every instance,
        summed on the device."""
        out = np.zeros(self.ninst)
        self._check(self.lib.cmf_hip_instance_storage(self.h, out.ctypes.data))
        return out

    def dt_min(self) -> np.ndarray:
        """DT_MIN of CALC_ADPSTP of the last advance, per instance (0 without LADPSTP)."""
        out = np.zeros(self.ninst)
        self._check(self.lib.cmf_hip_get_dt_min(self.h, out.ctypes.data))
        return out

    def instance_times(self) -> np.ndarray:
        """Per-instance completion time [ms] of the last timed advance."""
        out = np.zeros(self.ninst)
        self._check(self.lib.cmf_hip_instance_times(self.h, out.ctypes.data))
        return out

    def set_tune(self, **kw) -> dict:
        t = self.get_tune()
        t.update(kw)
        ct = cmf_tune(*[int(t[k]) for k in TUNE_FIELDS])
        self._check(self.lib.cmf_hip_set_tune(self.h, C.byref(ct)))
        return self.get_tune()

    def get_tune(self) -> dict:
        ct = cmf_tune()
        self._check(self.lib.cmf_hip_get_tune(self.h, C.byref(ct)))
        return {k: getattr(ct, k) for k in TUNE_FIELDS}

    def autotune(self, dt: float, nsub: int = 8):
        """Time nsub substeps under each launch configuration, keep the fastest
        and restore the state. Returns (best configuration, trial table [k, 8])."""
        cap = 64
        rep = np.zeros((cap, 8))
        ct = cmf_tune()
        n = self.lib.cmf_hip_autotune(self.h, float(dt), int(nsub), C.byref(ct), rep.ctypes.data, cap)
        if n < 0:
            self._check(n)
        return {k: getattr(ct, k) for k in TUNE_FIELDS}, rep[:min(n, cap)].copy()

    def synchronize(self) -> None:
        self._check(self.lib.cmf_hip_synchronize(self.h))

    def kernel_times(self, enable: bool) -> np.ndarray:
        """Accumulated kernel milliseconds since the last call: [outflow, rate,
        update, stage, substeps]; then enable/disable event timing."""
        out = np.zeros(5)
        self._check(self.lib.cmf_hip_kernel_times(self.h, int(enable), out.ctypes.data))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.cmf_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
