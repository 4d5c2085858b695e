"""Continuous-batching MD service: jobs stream into one running ensemble.

Jobs are admitted at tick boundaries in submission order while their device
footprint fits the budget; every running job advances by one tick of steps
per tick in shared launches, its correlators (blocking-scheme MSD and ISF)
sample on the device, and it leaves the ensemble when it reaches its step
count.  Over include/halmd_hip.h (halmd_server_*).
"""
import ctypes

import numpy as np

from ._lib import Handle, call, dptr, f64, iptr, out_handle
from .ensemble import potential_matrices

__all__ = ["server"]

STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class server(Handle):
    _destroy = "halmd_server_destroy"

    def __init__(self, budget_bytes=64 << 30, max_jobs=256, tick_steps=50):
        super().__init__()
        self._h = out_handle()
        call("halmd_server_create", ctypes.c_size_t(int(budget_bytes)), int(max_jobs), int(tick_steps),
             ctypes.byref(self._h))
        self.tick_steps = tick_steps
        self._n = {}

    def now(self):
        t = ctypes.c_double()
        call("halmd_server_now", self.handle, ctypes.byref(t))
        return t.value

    def submit(self, position, velocity, length, steps, potential="lj", truncation="shifted", timestep=0.005,
               skin=0.3, species=None, arrival=None, block_size=10, shift=2, separation=0, wavenumber=(),
               tolerance=0.01, max_count=7, keep_state=False):
        """Queue one job; returns its id.  arrival: server-clock seconds (now by default)."""
        r = f64(position)
        v = f64(velocity)
        kind, ntype, eps, sigma, rc, _, _ = potential_matrices(potential)
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.uint32)
        k = np.ascontiguousarray(wavenumber, dtype=np.float64)
        jid = ctypes.c_uint64()
        call("halmd_server_submit", self.handle, len(r), float(length), int(kind), truncation.encode(), int(ntype),
             dptr(eps), dptr(sigma), dptr(rc), float(timestep), float(skin), dptr(r), dptr(v),
             None if sp is None else sp.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), int(steps),
             float(self.now() if arrival is None else arrival), int(block_size), int(shift), int(separation),
             dptr(k) if len(k) else None, len(k), float(tolerance), int(max_count), int(bool(keep_state)),
             ctypes.byref(jid))
        self._n[jid.value] = len(r)
        return jid.value

    def tick(self):
        """One tick; returns the number of running jobs."""
        a = ctypes.c_uint()
        call("halmd_server_tick", self.handle, ctypes.byref(a))
        return a.value

    def stats(self):
        a, q, u, p = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_size_t(), ctypes.c_uint64()
        call("halmd_server_stats", self.handle, ctypes.byref(a), ctypes.byref(q), ctypes.byref(u), ctypes.byref(p))
        return dict(active=a.value, queued=q.value, used_bytes=u.value, particle_steps=p.value)

    def phases(self):
        """Host seconds over all ticks: admission/placement, steps, samples, completion."""
        t = np.zeros(4)
        call("halmd_server_phases", self.handle, dptr(t))
        return dict(zip(("admit", "step", "sample", "complete"), t))

    def info(self, jid):
        st = ctypes.c_int()
        t = np.zeros(3)
        e = np.zeros(2)
        call("halmd_server_job_info", self.handle, int(jid), ctypes.byref(st), dptr(t), dptr(e))
        return dict(status=STATUS[st.value], arrival=t[0], admitted=t[1], done=t[2], en_pot=e[0], en_kin=e[1])

    def error(self, jid):
        """Why a job failed at admission ('' otherwise)."""
        buf = ctypes.create_string_buffer(1024)
        call("halmd_server_job_error", self.handle, int(jid), buf, ctypes.c_size_t(len(buf)))
        return buf.value.decode()

    def _levels(self, jid):
        lv, bs, ns = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        call("halmd_server_job_levels", self.handle, int(jid), ctypes.byref(lv), ctypes.byref(bs), ctypes.byref(ns))
        return lv.value, bs.value, ns.value

    def msd(self, jid):
        """(lag time, mean, count), each levels x block_size."""
        lv, bs, _ = self._levels(jid)
        t = np.zeros(lv * bs)
        m = np.zeros(lv * bs)
        c = np.zeros(lv * bs, dtype=np.uint64)
        call("halmd_server_job_msd", self.handle, int(jid), dptr(t), dptr(m),
             c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return t.reshape(lv, bs), m.reshape(lv, bs), c.reshape(lv, bs)

    def isf(self, jid):
        """(mean, count), each levels x block_size x shells."""
        lv, bs, ns = self._levels(jid)
        m = np.zeros(lv * bs * ns)
        c = np.zeros(lv * bs * ns, dtype=np.uint64)
        call("halmd_server_job_isf", self.handle, int(jid), dptr(m), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return m.reshape(lv, bs, ns), c.reshape(lv, bs, ns)

    def state(self, jid):
        """Final (positions, velocities, images) of a job submitted with keep_state."""
        n = self._n[jid]
        r = np.zeros((n, 3))
        v = np.zeros((n, 3))
        im = np.zeros((n, 3), dtype=np.int32)
        call("halmd_server_job_state", self.handle, int(jid), dptr(r), dptr(v), iptr(im))
        return r, v, im

    def release(self, jid):
        call("halmd_server_job_release", self.handle, int(jid))
        self._n.pop(jid, None)
