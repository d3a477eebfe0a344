"""Serving loop: continuous batching of world requests on one device-resident batch.

A request is a world (bodies, optional joints) and a horizon in steps.  Requests arrive at any
step; the server keeps a FIFO of waiting requests and admits each one as soon as the batch has
room for it (its bodies, joints and manifold slots fit the ranges still free in a batch sized
once from an HBM byte budget, `avbd_batch_footprint`).  Every admitted world is stepped with all
other resident worlds by one batch step per tick (one captured HIP graph), stays in HBM between
ticks, and is retired (its ranges freed for the next admission) when its horizon is done.

A request that could not fit even the empty batch is rejected at submission (it would otherwise
block the FIFO for ever).  A world whose step fails on the device (its manifold slots overflow,
its broad-phase walk fails, or its state stops being finite) is evicted after that step and
reported in `evicted`; the other worlds of the batch stepped normally and go on.

Admission overlaps the device work: a tick launches the step of the resident worlds, admits
waiting requests while it runs (their arrays go to the device on a second stream from pinned
host memory, `avbd_batch_add_world_async`), then waits for the step; the worlds admitted this
way are stepped from the next tick on.  (overlap=False admits before the step, synchronously.)

Latency is recorded per world per step: from the moment the step was admitted (the request's
arrival for its first step, the start of the tick otherwise) to the moment its result is ready
(`avbd_batch_step_wait` returns after the step has completed on the device).
"""
import ctypes as C
import time
from collections import deque

import numpy as np

from . import abi
from .world import Batch


def footprint(params, max_worlds, max_bodies, max_joints=0, max_manifolds=0):
    """Device bytes of a batch with these capacities."""
    lib = abi.load_library()
    out = C.c_int64()
    abi.check(lib.avbd_batch_footprint(C.byref(params), int(max_worlds), int(max_bodies), int(max_joints),
                                       int(max_manifolds), C.byref(out)))
    return out.value


def capacity_for_budget(params, hbm_bytes, max_worlds, joints_per_body=1.0):
    """Largest body capacity (with joints_per_body joints per body) whose batch fits the budget."""
    lo, hi = 1, 1 << 27
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if footprint(params, max_worlds, mid, int(mid * joints_per_body)) <= hbm_bytes:
            lo = mid
        else:
            hi = mid - 1
    return lo, int(lo * joints_per_body)


class Request:
    __slots__ = ("rid", "bodies", "joints", "horizon", "arrival", "wid", "done", "movable", "first_step")

    def __init__(self, rid, bodies, joints, horizon, arrival):
        self.rid, self.bodies, self.joints, self.horizon, self.arrival = rid, bodies, joints, horizon, arrival
        self.wid, self.done, self.first_step = -1, 0, True
        self.movable = int((np.asarray(bodies["mass"]) > 0).sum())


class Server:
    """Continuous batching over one `Batch` sized from `hbm_bytes`."""

    def __init__(self, params, hbm_bytes, max_worlds=1024, joints_per_body=1.0, use_graph=True, overlap=True):
        self.overlap = overlap
        self.max_bodies, self.max_joints = capacity_for_budget(params, hbm_bytes, max_worlds, joints_per_body)
        self.batch = Batch(params, max_worlds, self.max_bodies, self.max_joints)
        self.batch.use_graph(use_graph)
        b = C.c_int64()
        abi.check(self.batch.lib.avbd_batch_device_bytes(self.batch.h, C.byref(b)))
        self.device_bytes = b.value
        self.waiting = deque()
        self.resident = {}
        self.latencies = []          # seconds, one per world-step
        self.admit_waits = []        # seconds from arrival to admission
        self.body_steps = 0
        self.finished = []           # (request, final state)
        self.keep_final = False
        self.rejected = []           # requests larger than the whole batch
        self.evicted = []            # (request, AVBD_WORLD_* reason bits, steps done)

    def submit(self, req):
        """Queue a request; False (and the request in `rejected`) when it can never fit."""
        req.arrival = time.perf_counter()
        nj = 0 if req.joints is None else len(req.joints["body_b"])
        if not self.batch.fits(len(req.bodies["mass"]), nj)[1]:
            self.rejected.append(req)
            return False
        self.waiting.append(req)
        return True

    def _admit(self):
        now = time.perf_counter()
        while self.waiting:
            req = self.waiting[0]
            try:
                req.wid = self.batch.add(req.bodies, req.joints, overlap=self.overlap)
            except abi.AvbdError as e:
                if e.code == abi.AVBD_ERR_CAPACITY:
                    break            # no room until a resident world retires (FIFO: no overtaking)
                raise
            self.waiting.popleft()
            self.admit_waits.append(now - req.arrival)
            self.resident[req.wid] = req

    def tick(self):
        """Step every resident world once, admitting what fits meanwhile (or before, without
        overlap), then retire finished worlds."""
        if not self.overlap:
            self._admit()
        if not self.resident:
            if self.overlap:
                self._admit()   # nothing to overlap with: these worlds step from the next tick
            return 0
        t0 = time.perf_counter()
        stepping = list(self.resident.items())
        try:
            if self.overlap:
                self.batch.step_launch()
                self._admit()
                self.batch.step_wait()
            else:
                self.batch.step()
        except abi.AvbdError as e:
            if e.code != abi.AVBD_ERR_WORLD:
                raise
            for wid, why in self.batch.failed_worlds():
                req = self.resident.pop(wid, None)
                if req is not None:
                    self.evicted.append((req, why, req.done))
                    self.batch.remove(wid)
        t1 = time.perf_counter()
        stepped = 0
        for wid, req in stepping:
            if wid not in self.resident:   # evicted
                continue
            self.latencies.append(t1 - (req.arrival if req.first_step else t0))
            req.first_step = False
            req.done += 1
            stepped += req.movable
            if req.done >= req.horizon:
                if self.keep_final:
                    self.finished.append((req, self.batch.download(wid)))
                self.batch.remove(wid)
                del self.resident[wid]
        self.body_steps += stepped
        return stepped

    def close(self):
        self.batch.close()


def run_alone(params, bodies, joints, horizon):
    """The reference result for `test: same alone and in the mix`: one world, one batch."""
    b = Batch(params, 1, len(bodies["mass"]), 0 if joints is None else len(joints["body_b"]))
    wid = b.add(bodies, joints)
    for _ in range(horizon):
        b.step()
    out = b.download(wid)
    b.close()
    return out
