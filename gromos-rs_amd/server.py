"""Continuous-batching server over the session C ABI.

Clients submit jobs (topology, state, step budget) at any time; the server
admits them against the device's HBM budget (the session reserves each job's
footprint; a job that does not fit waits in the server's FIFO until running
jobs finish), steps the resident batch and returns every finished job's end
state and per-step energies.  `replay` drives it with an arrival schedule in
md steps of the session clock, the form a serving benchmark takes.
"""
import collections
import time

import numpy as np

from . import runtime as rt


class Server:
    def __init__(self, precision=rt.MIXED, nsnb=5, admit_interval=0, hbm_budget=0):
        self.sess = rt.Session(precision=precision, nsnb=nsnb, admit_interval=admit_interval,
                               hbm_budget=hbm_budget)
        self.nsnb = nsnb
        self.admit_interval = admit_interval or nsnb
        self.waiting = collections.deque()
        self.running = {}
        self.results = {}

    @property
    def clock(self):
        return self.sess.info()["clock"]

    def submit(self, key, system, nsteps):
        self.waiting.append(dict(key=key, system=system, nsteps=int(nsteps), t_submit=time.perf_counter(),
                                 clock_submit=self.clock))

    def _admit(self):
        while self.waiting:
            w = self.waiting[0]
            try:
                jid = self.sess.admit(w["system"], w["nsteps"])
            except rt.Busy:
                break
            self.waiting.popleft()
            w["job"] = jid
            w["clock_admit"] = self.clock
            w["t_admit"] = time.perf_counter()
            self.running[jid] = w

    def _collect(self):
        for jid in list(self.running):
            st, done = self.sess.query(jid)
            if st not in (rt.DONE, rt.FAILED):
                continue
            err = rt.lib().gromos_amd_last_error().decode() if st == rt.FAILED else None
            w = self.running.pop(jid)
            e = self.sess.energies(jid)
            x, v = self.sess.fetch(jid)
            self.sess.evict(jid)
            w.pop("system")
            w.update(t_done=time.perf_counter(), clock_done=self.clock, energies=e, pos=x, vel=v, failed=err,
                     steps_done=done)
            self.results[w["key"]] = w

    def step(self, nsteps):
        self._admit()
        self.sess.step(nsteps)
        self._collect()

    def idle(self):
        return not self.waiting and not self.running


def replay(server, schedule, chunk=None):
    """Run `schedule` (dicts with key, system, nsteps, arrival step) to completion,
    stepping the session an admission interval at a time (arrivals inside an
    interval are admitted at its end, as the session would).  Returns wall
    seconds, pair interactions, per-job latency and device step times."""
    chunk = chunk or server.admit_interval
    pending = sorted(schedule, key=lambda d: d["arrival"])
    server.sess.step_times(enable=1)
    info0 = server.sess.info()
    t0 = time.perf_counter()
    while pending or not server.idle():
        clock = server.clock
        while pending and pending[0]["arrival"] <= clock:
            d = pending.pop(0)
            server.submit(d["key"], d["system"], d["nsteps"])
        if server.idle() and pending:
            # nothing to run before the next arrival: let the clock jump
            gap = pending[0]["arrival"] - clock
            server.sess.step(max(gap, 1))
            continue
        server.step(chunk - clock % chunk)   # up to the next admission boundary
    wall = time.perf_counter() - t0
    info = server.sess.info()
    step_ms = server.sess.step_times(enable=0)
    lat = np.array([r["t_done"] - r["t_submit"] for r in server.results.values()])
    lat_steps = np.array([r["clock_done"] - r["clock_submit"] for r in server.results.values()])
    return dict(wall=wall, pairs=info["pair_count"] - info0["pair_count"], layouts=info["layouts"] - info0["layouts"],
                layout_ms=info["layout_ms"] - info0["layout_ms"],
                layout_concat_ms=info["layout_concat_ms"] - info0["layout_concat_ms"], step_ms=step_ms, job_latency_s=lat,
                job_latency_steps=lat_steps, clock=info["clock"])


class ThreadedServer:
    """The server loop on a worker thread that owns the session (one host thread
    per session); clients call submit() from any thread at any time.  While the
    worker is inside a session step (the GPU runs the resident batch; the
    library call releases the interpreter lock), new requests queue up and are
    admitted at the next admission boundary as memory allows, so arrivals
    overlap the work in flight.  Results carry wall-clock submit, admit and
    completion times."""

    def __init__(self, precision=rt.MIXED, nsnb=5, admit_interval=0, hbm_budget=0, kernel_timing=False):
        import queue
        import threading
        self._q = queue.Queue()
        self._stop = threading.Event()
        self._err = None
        self.results = {}
        self.layouts = 0
        self.pairs = 0
        self.kernel = (0.0, 0, 0)
        self._args = (precision, nsnb, admit_interval, hbm_budget, kernel_timing)
        self._ready = threading.Event()
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()
        self._ready.wait()

    def submit(self, key, system, nsteps):
        self._q.put(dict(key=key, system=system, nsteps=int(nsteps), t_submit=time.perf_counter()))

    def close(self, expected):
        """wait until `expected` results are in, stop the worker, re-raise its error"""
        while len(self.results) < expected and self._t.is_alive():
            time.sleep(0.002)
        self._stop.set()
        self._t.join()
        if self._err is not None:
            raise self._err

    def _run(self):
        precision, nsnb, admit_interval, hbm_budget, kernel_timing = self._args
        try:
            srv = Server(precision=precision, nsnb=nsnb, admit_interval=admit_interval, hbm_budget=hbm_budget)
        except Exception as e:   # noqa: BLE001 - handed to the caller
            self._err = e
            self._ready.set()
            return
        info0 = srv.sess.info()
        if kernel_timing:
            srv.sess.kernel_timing(enable=1)
        self._ready.set()
        try:
            while not self._stop.is_set():
                while True:
                    try:
                        w = self._q.get_nowait()
                    except Exception:   # queue.Empty
                        break
                    srv.waiting.append(dict(key=w["key"], system=w["system"], nsteps=w["nsteps"],
                                            t_submit=w["t_submit"], clock_submit=srv.clock))
                if srv.idle():
                    time.sleep(0.0005)
                    continue
                srv.step(srv.admit_interval - srv.clock % srv.admit_interval)
                for k, r in srv.results.items():
                    if k not in self.results:
                        self.results[k] = r
        except Exception as e:   # noqa: BLE001 - handed to the caller
            self._err = e
        finally:
            info = srv.sess.info()
            self.layouts = info["layouts"] - info0["layouts"]
            self.pairs = info["pair_count"] - info0["pair_count"]
            if kernel_timing:
                self.kernel = srv.sess.kernel_timing(enable=0)
            srv.sess.close()


def replay_realtime(server, schedule, seconds_per_step):
    """Submit `schedule` (dicts with key, system, nsteps, arrival in md steps)
    to a ThreadedServer in wall-clock time, arrival k at arrival *
    seconds_per_step seconds, from this (client) thread, while the server
    works; returns wall seconds until the last result and per-job latencies."""
    pending = sorted(schedule, key=lambda d: d["arrival"])
    t0 = time.perf_counter()
    for d in pending:
        dt = t0 + d["arrival"] * seconds_per_step - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        server.submit(d["key"], d["system"], d["nsteps"])
    server.close(len(pending))
    wall = max(r["t_done"] for r in server.results.values()) - t0
    lat = np.array([r["t_done"] - r["t_submit"] for r in server.results.values()])
    queued = np.array([r["t_admit"] - r["t_submit"] for r in server.results.values()])
    return dict(wall=wall, pairs=server.pairs, layouts=server.layouts, job_latency_s=lat, admit_wait_s=queued)


def equilibrate(systems, temperatures, steps=300, tau=0.01, precision=rt.MIXED, nsnb=5):
    """Relax synthetic starting structures (lattice waters with random orientations
    release a lot of potential energy in their first picoseconds): all systems
    run together in one session under strong Berendsen coupling to their
    temperatures, then their end states replace the starting ones.  Returns new
    System objects (thermostat setting of the originals kept)."""
    sess = rt.Session(precision=precision, nsnb=nsnb)
    jobs = []
    for s, t in zip(systems, temperatures):
        q = s.copy()
        q.thermostat = (float(t), float(tau), 3 * q.n_atoms - 3 * q.nsm)
        jobs.append(sess.admit(q, steps))
    sess.step(steps)
    out = []
    for s, j in zip(systems, jobs):
        st, _ = sess.query(j)
        if st != rt.DONE:
            raise rt.GromosAmdError(f"equilibration of {s.name} did not finish (state {st})")
        q = s.copy()
        q.pos, q.vel = sess.fetch(j)
        out.append(q)
    sess.close()
    return out
