"""Temperature replica exchange over a session: every replica is a session job,
all replicas step together in one batch, exchanges happen between segments.

Follows ReplicaExchange (crates/gromos-integrators/src/remd.rs):
  * run: segments of exchange_interval steps, exchanges after each full
    segment (remd.rs:395-417);
  * determine_exchange_pairs: sequential (0,1),(2,3),... or odd-even
    alternation starting with the even pairs (remd.rs:204-235);
  * calculate_exchange_probability, temperature exchange:
    delta = (beta_1 - beta_2) (E_2 - E_1), p = 1 if delta < 0 else exp(-delta)
    (remd.rs:261-305), beta = 1 / (kB T) (replica.rs:76-78);
  * perform_exchange: the two configurations swap, then each replica's
    velocities are rescaled to its own temperature by sqrt(T_target / T_current)
    (remd.rs:342-362, replica.rs:350-371), T = 2 E_kin / (kB 3N)
    (configuration.rs:339-341, replica.rs:343-347).

Random numbers come from numpy's generator (the reference uses rand's StdRng;
the streams differ, the acceptance rule is the same).  The potential energy of
a replica is the last step's record (force evaluation of that step), and
T_current uses that step's leap-frog averaged kinetic energy.
"""
import numpy as np

from . import runtime as rt

KB = 0.00831441          # kJ/(mol K), gromos-core units.rs:47


def exchange_probability(t1, t2, e1, e2):
    """(probability, delta) of exchanging replicas at temperatures t1, t2 with
    potential energies e1, e2 (remd.rs:261-305, temperature exchange)."""
    b1, b2 = 1.0 / (KB * t1), 1.0 / (KB * t2)
    delta = (b1 - b2) * (e2 - e1)
    p = 1.0 if delta < 0.0 else float(np.exp(-delta))
    return p, delta


def exchange_pairs(n, scheme, odd_even_state):
    """pairs to attempt and the next odd-even state (remd.rs:204-235)"""
    if scheme == "sequential":
        return [(i, i + 1) for i in range(0, n - 1, 2)], odd_even_state
    if scheme == "odd_even":
        start = 0 if odd_even_state else 1
        return [(i, i + 1) for i in range(start, n - 1, 2)], not odd_even_state
    raise ValueError(f"unknown exchange scheme {scheme!r}")


class ReplicaExchange:
    def __init__(self, session, systems, temperatures, exchange_interval=500, scheme="odd_even", seed=0,
                 rescale_velocities=True, total_steps=None):
        assert len(systems) == len(temperatures) and len(systems) >= 2
        self.sess = session
        self.T = [float(t) for t in temperatures]
        self.n_atoms = systems[0].n_atoms
        assert all(s.n_atoms == self.n_atoms for s in systems), "replicas share one topology"
        self.interval = int(exchange_interval)
        self.scheme = scheme
        self.odd_even_state = True
        self.rng = np.random.default_rng(seed)
        self.rescale = rescale_velocities
        budget = total_steps if total_steps is not None else 1 << 40
        self.jobs = [session.admit(s, budget) for s in systems]
        self.step_count = 0
        self.attempts = 0
        self.accepted = 0
        self.history = []          # (step, i, k, probability, accepted)

    def temperature_from_ekin(self, ekin):
        return 2.0 * ekin / (KB * 3 * self.n_atoms)

    def last_records(self):
        out = []
        for j in self.jobs:
            done = self.sess.query(j)[1]
            out.append(self.sess.energies(j, first=done - 1, count=1)[0] if done > 0 else np.zeros(5))
        return np.array(out)

    def attempt_all_exchanges(self):
        rec = self.last_records()
        epot, ekin = rec[:, 2], rec[:, 3]
        pairs, self.odd_even_state = exchange_pairs(len(self.jobs), self.scheme, self.odd_even_state)
        factors = np.ones(len(self.jobs))
        changed = []
        states = [self.sess.query(j)[0] for j in self.jobs]
        for i, k in pairs:
            if states[i] != rt.RUNNING or states[k] != rt.RUNNING:
                continue     # a replica that stopped (failed or done) takes no part
            p, _ = exchange_probability(self.T[i], self.T[k], epot[i], epot[k])
            r = self.rng.random()
            acc = bool(r < p)
            self.attempts += 1
            self.history.append((self.step_count, i, k, p, acc))
            if not acc:
                continue
            self.accepted += 1
            self.sess.swap_configurations(self.jobs[i], self.jobs[k])
            epot[i], epot[k] = epot[k], epot[i]
            ekin[i], ekin[k] = ekin[k], ekin[i]
            if self.rescale:
                for m in (i, k):
                    tc = self.temperature_from_ekin(ekin[m])
                    if tc > 0.0:
                        factors[m] = np.sqrt(self.T[m] / tc)
                        changed.append(m)
        if changed:
            self.sess.scale_velocities([self.jobs[m] for m in changed], factors[changed])

    def run(self, total_steps):
        done = 0
        while done < total_steps:
            n = min(self.interval, total_steps - done)
            self.sess.step(n)
            self.step_count += n
            done += n
            if self.step_count % self.interval == 0:
                self.attempt_all_exchanges()
        return self
