"""Continuous-batching energy+force serving, shaped like Fortnet's socket driver.

The reference's i-PI driver (``prg_fnet/fortnet.F90:126-169``) keeps the networks resident and,
per request, receives one structure, recomputes its ACSF values and derivatives, runs the
networks and returns energy and forces -- one request at a time.  Here requests that arrive while
others are on the GPU are queued and admitted into the next device batch as soon as a batch slot
frees up, up to an atom budget set by the HBM footprint: the device bytes per atom are measured
once (free device memory before and after a calibration batch is built and run), and each
admission takes as many queued atoms as the slot's resident buffers plus its share of the
currently free device memory hold.  Two slots (each a context with its own HIP stream and a
resident batch whose buffers are reused from batch to batch) let one batch be assembled and
launched while the other computes, so their work overlaps on the device.  ``serial_loop`` is the
one-at-a-time loop for comparison.

Latency of a request = time its results are on the host - its arrival time.
"""
import collections
import time

import numpy as np

from .core import Context
from ._lib import FNET_FP32


# first guess of the device bytes per atom before a calibration: neighbour entries in padded rows
# (index, key, displacement, mirror, orphan slot, FP64 gradient vector: ~72 bytes each, ~160 per
# atom), coordinates, ACSF values and their gradient (2 x nG FP64), energies and forces
def bytes_per_atom(n_features, n_neighbours=160):
    return n_neighbours * 72 + 64 + 16 * n_features + 64


class Request:
    __slots__ = ("rid", "geo", "arrival", "done", "energy", "forces")

    def __init__(self, rid, geo, arrival):
        self.rid = rid
        self.geo = geo
        self.arrival = arrival
        self.done = None
        self.energy = None
        self.forces = None


class _Slot:
    def __init__(self, netstat, precision, device):
        self.ctx = Context(device)
        self.ctx.upload_model(netstat)
        self.batch = None
        self.precision = precision
        self.reqs = []

    def launch(self, reqs):
        geos = [r.geo for r in reqs]
        if self.batch is None:
            self.batch = self.ctx.batch(geos, self.precision)
        else:
            self.batch.reload(geos)
            self.batch.set_mode(self.precision)
        self.batch.run()
        self.reqs = reqs

    def busy(self):
        return bool(self.reqs) and self.batch.query()

    def finish(self, now):
        """Downloads the results; ``now()`` is read once they are on the host."""
        E, _ = self.batch.energy()
        F = self.batch.forces()
        t = now()
        off = self.batch.offsets
        for s, r in enumerate(self.reqs):
            r.energy = E[s].copy()
            r.forces = F[off[s]:off[s + 1]].copy()
            r.done = t
        done = self.reqs
        self.reqs = []
        return done

    def close(self):
        if self.batch is not None:
            self.batch.close()
        self.ctx.close()


class BatchingServer:
    """Continuous batching over ``n_slots`` device slots, admission by free device memory."""

    def __init__(self, netstat, precision=FNET_FP32, hbm_fraction=0.5,
                 max_atoms_per_batch=1 << 18, n_slots=2, device=0):
        self.slots = [_Slot(netstat, precision, device) for _ in range(n_slots)]
        self.per_atom = float(bytes_per_atom(len(netstat.g)))
        self.fraction = hbm_fraction
        self.max_atoms = int(max_atoms_per_batch)
        self.held = [0] * n_slots   # atoms the slot's resident buffers were sized for
        self.batch_sizes = []
        self.budgets = []

    def close(self):
        for s in self.slots:
            s.close()

    def calibrate(self, geos):
        """Measures the device bytes per atom: free device memory before and after the first
        slot's batch is built and run on ``geos`` (a few representative structures)."""
        slot = self.slots[0]
        free0, _ = slot.ctx.device_memory()
        slot.launch([Request(-1, g, 0.0) for g in geos])
        slot.batch.sync()
        slot.finish(time.perf_counter)
        free1, _ = slot.ctx.device_memory()
        atoms = sum(g.n for g in geos)
        if atoms > 0 and free0 > free1:
            self.per_atom = max(self.per_atom * 0.25, (free0 - free1) / atoms)
        self.held[0] = atoms
        return self.per_atom

    def warm(self, geos):
        """Runs every slot once on ``geos`` (untimed), so that each slot's device buffers are
        sized for a batch that large before requests are measured: a slot's first batch of a new
        largest size allocates, which would otherwise land in the first latencies of a run."""
        for k, slot in enumerate(self.slots):
            slot.launch([Request(-1, g, 0.0) for g in geos])
            slot.batch.sync()
            slot.finish(time.perf_counter)
            self.held[k] = max(self.held[k], sum(g.n for g in geos))

    def _budget(self, k):
        """Atoms slot k may take now: its resident buffers plus its share of the free memory."""
        free, _ = self.slots[k].ctx.device_memory()
        share = free * self.fraction / len(self.slots)
        return int(min(self.max_atoms, self.held[k] + share / self.per_atom))

    def _admit(self, queue, k):
        """Requests for slot k's next batch: FIFO while the atom budget allows (at least one)."""
        budget = self._budget(k)
        self.budgets.append(budget)
        take, atoms = [], 0
        while queue and (not take or atoms + queue[0].geo.n <= budget):
            r = queue.popleft()
            take.append(r)
            atoms += r.geo.n
        self.held[k] = max(self.held[k], atoms)
        return take

    def run(self, requests, t0=None):
        """Serves ``requests`` (arrival times in seconds from t0, sorted); returns them with
        results and completion times filled in."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        queue = collections.deque()
        clock = time.perf_counter
        t0 = clock() if t0 is None else t0
        nxt = 0
        ndone = 0
        while ndone < len(reqs):
            now = clock() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                queue.append(reqs[nxt])
                nxt += 1
            progressed = False
            for s in self.slots:
                if s.reqs and not s.batch.query():
                    ndone += len(s.finish(lambda: clock() - t0))
                    progressed = True
            for k, s in enumerate(self.slots):
                if not s.reqs and queue:
                    take = self._admit(queue, k)
                    self.batch_sizes.append(len(take))
                    s.launch(take)
                    progressed = True
            if not progressed:
                if nxt < len(reqs) and not queue and not any(s.reqs for s in self.slots):
                    wait = reqs[nxt].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.002))
                else:
                    time.sleep(20e-6)
        return reqs


def serial_loop(netstat, requests, precision=FNET_FP32, device=0, t0=None):
    """One request at a time: wait for its arrival, evaluate it alone, return its results."""
    slot = _Slot(netstat, precision, device)
    clock = time.perf_counter
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = clock() if t0 is None else t0
    try:
        for r in reqs:
            wait = r.arrival - (clock() - t0)
            if wait > 0:
                time.sleep(wait)
            slot.launch([r])
            slot.batch.sync()
            slot.finish(lambda: clock() - t0)
    finally:
        slot.close()
    return reqs


def poisson_arrivals(rate, n, seed):
    """Arrival times (s) of a Poisson process of the given rate (requests/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def latency_summary(reqs):
    lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    atoms = sum(r.geo.n for r in reqs)
    return {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "throughput_req_s": len(reqs) / span, "atoms_per_s": atoms / span}
