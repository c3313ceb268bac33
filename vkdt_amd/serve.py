"""A request loop over the batched develop path: continuous batching of
stills and raw-video frames.

Requests arrive over time (each names a profile -- the planned graph of its
camera stream or still -- and a mosaic in pinned host memory).  The loop
keeps `slots` batches in flight on their own HIP streams.  Whenever a slot
is free it admits the waiting requests, oldest first, as far as the memory
budget allows, uploads their mosaics and launches them as one batch; their
kernels overlap on the device with the batches already running.  A batch
retires when its event completes: its requests are done and its memory
returns to the pool.  Profiles (descriptors, noise parameters, curve
tables) are planned once and stay resident; device memory is one arena
allocated up front and never returned to the driver while the server runs.

Each request's result does not depend on what it was batched with (every
image's output equals its own per-node graph run), which tests check.
"""
import collections
import ctypes as C
import time

import numpy as np

from . import batch, hip

ALIGN = 256


def _al(n):
    return (n + ALIGN - 1) // ALIGN * ALIGN


class FirstFit:
    """First-fit allocation of 256-byte aligned blocks in [0, size)."""

    def __init__(self, size):
        self.size = size
        self.free_blocks = [(0, size)]   # sorted (offset, size), never adjacent
        self.used = {}
        self.peak = 0

    def largest_free(self):
        return max((s for _, s in self.free_blocks), default=0)

    def alloc(self, n):
        n = _al(max(1, n))
        for i, (o, s) in enumerate(self.free_blocks):
            if s >= n:
                if s == n:
                    del self.free_blocks[i]
                else:
                    self.free_blocks[i] = (o + n, s - n)
                self.used[o] = n
                self.peak = max(self.peak, self.in_use())
                return o
        return None

    def free(self, o):
        n = self.used.pop(o)
        fb = self.free_blocks
        i = 0
        while i < len(fb) and fb[i][0] < o:
            i += 1
        fb.insert(i, (o, n))
        if i + 1 < len(fb) and fb[i][0] + fb[i][1] == fb[i + 1][0]:   # merge with the next block
            fb[i] = (fb[i][0], fb[i][1] + fb[i + 1][1])
            del fb[i + 1]
        if i > 0 and fb[i - 1][0] + fb[i - 1][1] == fb[i][0]:         # and the previous one
            fb[i - 1] = (fb[i - 1][0], fb[i - 1][1] + fb[i][1])
            del fb[i]

    def in_use(self):
        return sum(self.used.values())


class Arena(FirstFit):
    """FirstFit over one device allocation made up front."""

    def __init__(self, nbytes):
        super().__init__(nbytes)
        p = C.c_void_p()
        hip.check(hip.lib().vkdt_hip_alloc(C.byref(p), nbytes), "arena alloc")
        self.base = p.value

    def close(self):
        if self.base:
            hip.check(hip.lib().vkdt_hip_free(self.base), "arena free")
            self.base = None


class Profile:
    """A resident image template: the descriptor of one planned graph."""

    def __init__(self, desc, name=""):
        self.desc = batch.DevelopImage()
        C.memmove(C.byref(self.desc), C.byref(desc), C.sizeof(desc))
        self.name = name
        self.cfa_bytes = desc.cfa_wd * desc.cfa_ht * 2
        self.out_bytes = desc.wd * desc.ht * 16
        self.pixels = desc.wd * desc.ht
        # upper bound of one request's share of a batch allocation
        self.bytes = _al(self.cfa_bytes) + _al(self.out_bytes) + batch.scratch_bytes([self.desc])


class Request:
    __slots__ = ("id", "profile", "host", "arrival", "admitted", "done", "out", "tag")

    def __init__(self, rid, profile, host_ptr, arrival, tag=None):
        self.id, self.profile, self.host, self.arrival, self.tag = rid, profile, host_ptr, arrival, tag
        self.admitted = self.done = None
        self.out = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Slot:
    def __init__(self):
        self.stream = hip.Stream()
        self.event = hip.Event()
        self.batch = None
        self.reqs = None
        self.block = None


class Server:
    """Continuous batching over `slots` streams within `budget` bytes of HBM.

    tune None keeps the batches' built-in launch shape, one tile per work
    group, under which the small batches in flight on the slots' streams
    overlap on the device; persistent work groups (a large batch's autotuned
    shape) fill every CU per batch and serialise the streams.

    on_done(request) is called when a request's batch has completed, with
    request.out the device address of its rgba f32 result (valid during the
    call only)."""

    def __init__(self, budget, slots=3, max_batch=32, tune=None, on_done=None):
        self.arena = Arena(budget)
        self.slots = [_Slot() for _ in range(slots)]
        self.max_batch = max_batch
        self.tune = tune
        self.on_done = on_done
        self.batches = []   # (size, launch time, retire time) per batch

    def close(self):
        for s in self.slots:
            if s.batch is not None:
                s.stream.sync()
                self._retire(s, time.perf_counter())
        self.arena.close()

    def _launch(self, slot, reqs, now):
        descs = []
        for r in reqs:
            d = batch.DevelopImage()
            C.memmove(C.byref(d), C.byref(r.profile.desc), C.sizeof(d))
            descs.append(d)
        scratch = batch.scratch_bytes(descs)
        total = sum(_al(r.profile.cfa_bytes) + _al(r.profile.out_bytes) for r in reqs) + scratch
        off = self.arena.alloc(total)
        if off is None:
            return False
        p = self.arena.base + off
        for r, d in zip(reqs, descs):
            d.cfa, d.cfa_pitch = p, d.cfa_wd
            hip.check(hip.lib().vkdt_hip_upload(p, r.host, r.profile.cfa_bytes, slot.stream.ptr), "upload")
            p += _al(r.profile.cfa_bytes)
            d.out = r.out = p
            p += _al(r.profile.out_bytes)
            r.admitted = now
        slot.batch = batch.Batch(descs, p, scratch, self.tune)
        slot.batch.run(slot.stream.ptr)
        slot.event.record(slot.stream.ptr)
        slot.reqs, slot.block = reqs, off
        self.batches.append([len(reqs), now, None])
        slot.bidx = len(self.batches) - 1
        return True

    def _retire(self, slot, now):
        for r in slot.reqs:
            r.done = now
            if self.on_done:
                self.on_done(r)
            r.out = None
        slot.batch.close()
        self.arena.free(slot.block)
        self.batches[slot.bidx][2] = now
        slot.batch = slot.reqs = slot.block = None

    def run(self, requests, poll=1e-4):
        """Serve `requests` (sorted by arrival, in seconds from the start)
        in real time; returns the wall time from start to the last retire."""
        t0 = time.perf_counter()
        n, i = len(requests), 0
        waiting = collections.deque()
        largest = max((r.profile.bytes for r in requests), default=0)
        if largest > self.arena.size:
            raise ValueError(f"a request needs {largest} bytes, more than the {self.arena.size}-byte budget")
        while i < n or waiting or any(s.batch is not None for s in self.slots):
            now = time.perf_counter() - t0
            while i < n and requests[i].arrival <= now:
                waiting.append(requests[i])
                i += 1
            busy = False
            for s in self.slots:
                if s.batch is not None and s.event.done():
                    self._retire(s, time.perf_counter() - t0)
            for s in self.slots:
                if s.batch is not None:
                    busy = True
                    continue
                if not waiting:
                    continue
                room, take = self.arena.largest_free(), []
                while waiting and len(take) < self.max_batch and room >= waiting[0].profile.bytes:
                    room -= waiting[0].profile.bytes
                    take.append(waiting.popleft())
                if not take:
                    continue
                if not self._launch(s, take, time.perf_counter() - t0):
                    waiting.extendleft(reversed(take))
                    continue
                busy = True
            if not busy and i < n and not waiting:
                time.sleep(min(poll, max(0.0, requests[i].arrival - (time.perf_counter() - t0))))
            elif busy:
                time.sleep(poll * 0.25)
        return time.perf_counter() - t0


def summary(requests, wall):
    """p50/p99 latency (ms) and throughput of served requests"""
    lat = np.array([r.latency for r in requests if r.latency is not None]) * 1e3
    px = sum(r.profile.pixels for r in requests if r.done is not None)
    return {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3),
            "mpix_per_s": round(px / wall / 1e6, 2), "wall_s": round(wall, 3)}
