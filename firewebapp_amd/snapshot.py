"""Periodic snapshots of the spread run in the reference's file format.

``run_3d_spread(..., snapshot_dir=D, snapshot_interval_s=s)`` writes
``D/snap_%04d.npz`` after every outer step that ends at least ``s`` seconds
of simulated time after the previous snapshot, the first after step 1
(spread_3d.py:1806-1810, :3537-3670).  Every key, dtype and shape is the
reference's: the state and the turbulence fields as float32, the passive
level set, the flame signed distance, the DOM forward flux in the ahead band
and its z sum, the per-cell bed-particle aggregates, the grid axes and the
run's wind speed and closure.  The derived fields are computed on the device
by libfirewave (fw_phi_flame, fw_ahead_band_mask, fw_dom_forward_flux,
fw_q_in_at_front_2d, fw_tau_mix, fw_bed_snapshot); only the finished arrays
are copied to the host, into pinned buffers, and a worker thread compresses
and writes the file while the next steps run.

Fields that are identically zero on this configuration are written as zeros,
as the reference writes them: q_frankman_3d (the Frankman flux is disabled,
spread_3d.py:2586), v_n_2d with a passive level set (:3234-3240) and the six
ft_* tendril inventories (finney_tendril_enable is false).  With an active
level set the band, the forward flux and v_n are the step's own.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr

DX_VN_BAND_M = 0.20          # flame_front_3d.py:64
LEVELSET_REINIT_INTERVAL = 10  # flame_front_3d.py:107
_G = 9.81
_FT_KEYS = ("ft_sink_M", "ft_sink_E", "ft_dep_M", "ft_dep_E", "ft_sink_t_rem",
            "ft_dep_t_rem")
# snapshot key -> solver field, for the state fields written as float32
_STATE_SRC = (("T_g", "T_g"), ("T_s", "T_s"), ("Y_fuel", "Y_fuel"), ("Y_O2", "Y_O2"),
              ("Y_H2O", "Y_H2O"), ("u", "u"), ("v", "v"), ("w", "w"), ("rho", "rho"),
              ("omega", "omega"), ("S_pyro", "S_pyro"), ("k_turb", "k"), ("eps_turb", "eps"),
              ("nu_t", "nu_t"))
# the keys in the order the reference writes them (spread_3d.py:3611-3668)
SNAPSHOT_ORDER = (
    "t", "step", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "u", "v", "w", "rho", "omega",
    "S_pyro", "k_turb", "eps_turb", "nu_t", "tau_mix", "phi_bed", "phi_flame", "front_x",
    "q_dom_fwd_3d", "q_frankman_3d", "q_in_at_front_2d", "v_n_2d") + _FT_KEYS + (
    "bp_count", "bp_m_solid", "bp_m_water", "bp_m_char", "bp_T_s_avg", "x_mid", "z_mid",
    "dx", "dy", "dz_arr", "wind_U", "closure")


def snapshot_path(snapshot_dir, index):
    return Path(snapshot_dir) / f"snap_{index:04d}.npz"


class SnapshotWriter:
    """Device buffers and the derived fields of one solver's snapshots."""

    def __init__(self, solver, wind_speed_m_s, closure="edc", workers=2):
        g = solver.grid
        dev = solver.dev
        self.sv = solver
        self.wind = float(wind_speed_m_s)
        self.closure = closure
        f64 = dict(dtype=torch.float64, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        self.phi_flame = torch.empty(solver.shape, **f64)
        self.edt_work = torch.empty(solver.shape, **f64)
        self.mask = torch.empty(solver.shape, dtype=torch.uint8, device=dev)
        self.q_fwd = torch.empty(solver.shape, dtype=solver.dtype, device=dev)
        self.q_in_2d = torch.empty((g.Ny, g.Nx), **f64)
        self.tau_mix = torch.empty(solver.shape, dtype=solver.dtype, device=dev)
        self.bp = {k: torch.empty(solver.shape, **f32)
                   for k in ("count", "m_solid", "m_water", "m_char", "T_s_avg")}
        # sampling of the flame distance: the mean dz (flame_front_3d.py:788)
        self.dz_eff = float(g.dz_arr.mean())
        # spread_3d.py:1559-1561 and :1601
        self.tau_buoy = math.sqrt(2.0 * max(g.dx, g.dy, g.dz) / _G)
        self.band = max(g.dx, DX_VN_BAND_M)
        self.active = not solver.cfg.level_set_passive
        self.v_n_2d = torch.zeros((g.Ny, g.Nx), **f64)
        self._pool = None
        self._workers = max(int(workers), 1)
        nslot = self._workers + 1
        self._pending = [None] * nslot
        self._stage = [{} for _ in range(nslot)]
        self._next = 0
        self._const_cache = None

    def _band_phi(self):
        """The passive level set the step's band mask was taken from: the one
        before that step's reinitialisation (spread_3d.py:2555 precedes
        :3260-3261); phi0 holds it on a reinitialising step, phi is unchanged
        otherwise."""
        sv = self.sv
        if sv.state.lset_step % LEVELSET_REINIT_INTERVAL == 0:
            return sv.phi0
        return sv.phi

    def _device_fields(self):
        """Launch the derived-field kernels; return the snapshot's arrays as
        device tensors (float32) and its host values."""
        sv = self.sv
        g = sv.grid
        f = sv.f
        s = stream_ptr()
        _lib.call("fw_tau_mix", sv.dg.ref, ptr(f["k"]), ptr(f["eps"]), self.tau_buoy,
                  ptr(self.tau_mix), s)
        _lib.call("fw_phi_flame", sv.dg.ref, ptr(f["omega"]), ptr(f["T_g"]),
                  ptr(f["Y_fuel"]), self.dz_eff, ptr(self.phi_flame), ptr(self.edt_work), s)
        if self.active:
            # the step's own band, forward flux and front speed
            mask, q_fwd = sv.lset_mask, sv.q_dom_fwd
            _lib.call("fw_v_n_2d", sv.dg.ref, ptr(sv.v_n), ptr(self.v_n_2d), s)
        else:
            mask, q_fwd = self.mask, self.q_fwd
            _lib.call("fw_ahead_band_mask", sv.dg.ref, ptr(self._band_phi()), self.band,
                      int(g.n_z_bed), ptr(mask), s)
            _lib.call("fw_dom_forward_flux", sv.dom_h, ptr(mask), ptr(q_fwd), s)
        _lib.call("fw_q_in_at_front_2d", sv.dg.ref, ptr(q_fwd), ptr(mask),
                  ptr(self.q_in_2d), s)
        bp = self.bp
        if sv.ix is not None:
            _lib.call("fw_bed_snapshot", sv.dg.ref, C.byref(sv.P.c), C.byref(sv.ix.c),
                      ptr(bp["count"]), ptr(bp["m_solid"]), ptr(bp["m_water"]),
                      ptr(bp["m_char"]), ptr(bp["T_s_avg"]), s)
        else:
            # no particles (Eulerian bed): the reference writes zeros
            for t in bp.values():
                t.zero_()
        dev = {k: f[src].float() for k, src in _STATE_SRC}
        dev.update(tau_mix=self.tau_mix.float(), phi_bed=sv.phi.float(),
                   phi_flame=self.phi_flame.float(), q_dom_fwd_3d=q_fwd.float(),
                   q_in_at_front_2d=self.q_in_2d.float(), bp_count=bp["count"],
                   bp_m_solid=bp["m_solid"], bp_m_water=bp["m_water"],
                   bp_m_char=bp["m_char"], bp_T_s_avg=bp["T_s_avg"])
        if self.active:
            dev["v_n_2d"] = self.v_n_2d.float()
        host = dict(t=np.float64(sv.t), step=np.int64(sv.step_no),
                    front_x=np.float64(sv.front_x[-1] if sv.front_x else 0.0))
        return dev, host

    def _assemble(self, arrays, host):
        """The snapshot dict in the reference's key order."""
        g = self.sv.grid
        out = {}
        for k in SNAPSHOT_ORDER:
            if k in arrays:
                out[k] = arrays[k]
            elif k in host:
                out[k] = host[k]
            elif k in self._const:
                out[k] = self._const[k]
            else:
                raise KeyError(k)
        return out

    @property
    def _const(self):
        if self._const_cache is None:
            g = self.sv.grid
            zeros = np.zeros(self.sv.shape, np.float32)
            c = dict(q_frankman_3d=zeros, v_n_2d=np.zeros((g.Ny, g.Nx), np.float32),
                     x_mid=g.x_mid.astype(np.float32), z_mid=g.z_mid.astype(np.float32),
                     dx=np.float64(g.dx), dy=np.float64(g.dy),
                     dz_arr=g.dz_arr.astype(np.float32), wind_U=np.float64(self.wind),
                     closure=np.array([self.closure], dtype=object))
            for k in _FT_KEYS:
                c[k] = zeros
            self._const_cache = c
        return self._const_cache

    def fields(self):
        """The snapshot of the solver's current state: {key: numpy array}
        (synchronous)."""
        dev, host = self._device_fields()
        return self._assemble({k: t.cpu().numpy() for k, t in dev.items()}, host)

    def write(self, path):
        """Write one snapshot now (synchronous)."""
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, **self.fields())
        return path

    # -- overlapped writes --------------------------------------------------------
    # The deflate of a 256 x 64 x 48 snapshot takes about half a second of host
    # time, longer than the device needs for a second of simulated time at
    # that size.  submit() therefore copies the arrays into pinned host
    # buffers on the compute stream and hands the compression to worker
    # threads (zlib runs without the interpreter lock), so the next steps run
    # meanwhile.  workers + 1 buffer sets: one more snapshot waits for the
    # oldest write to finish.

    def submit(self, path):
        import concurrent.futures as cf
        import torch
        if self._pool is None:
            self._pool = cf.ThreadPoolExecutor(self._workers, thread_name_prefix="fw-snapshot")
        slot = self._next
        self._next = (slot + 1) % len(self._stage)
        pend = self._pending[slot]
        if pend is not None:
            pend.result()          # the buffers of this slot are free again
        dev, host = self._device_fields()
        stage = self._stage[slot]
        for k, t in dev.items():
            if k not in stage:
                stage[k] = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
            stage[k].copy_(t, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        arrays = {k: stage[k].numpy() for k in dev}
        snap = self._assemble(arrays, host)
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)

        def job():
            ev.synchronize()
            np.savez_compressed(path, **snap)
            return path

        self._pending[slot] = self._pool.submit(job)
        return path

    def flush(self):
        """Wait for every submitted snapshot; re-raise a failed write."""
        for i, pend in enumerate(self._pending):
            if pend is not None:
                self._pending[i] = None
                pend.result()

    def close(self):
        try:
            self.flush()
        finally:
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
