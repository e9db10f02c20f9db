"""Many-object stepping and the serving loop over the batched API.

``step_frame`` advances every object of one batch by one frame of its own
gripper script.  ``Server`` is the continuous-batching request loop: requests
(objects with a lifetime in frames) arrive at any tick, are admitted while the
HBM budget (and, optionally, a tick-time budget) allows, are built into a new
batch on a background thread while the live batches keep stepping, join at the
next tick boundary, and leave when their lifetime is over.  Batches whose
objects have mostly left are merged: the survivors' state moves into a new
batch with ``copy_object`` and continues bit for bit.  Every tick advances all
live objects by one frame; its latency is the wall time from the tick's start
to every live object's frame (positions and shapes) being done.

The reference runs one object through a global instance and has no serving
path (native/xpbd_capi.cu:6, 17-29; src/runtime_loop.py:874-982 steps it from
the simulator callbacks); a plain per-object loop over single-object batches
stands in for it where the two are compared.
"""
import collections
import threading
import time

import numpy as np

from batch import XPBDBatch
import scenes

SUBSTEPS = 10
ITERS = 4


def step_frame(batch, entries, substeps=SUBSTEPS, shapes=True):
    """Queue one frame of every object of `batch`.  entries[k] is (script,
    frame) for a live object, None for one that only keeps its slot.  As in
    the single-object drivers, attachments are set before the first substep,
    except on an object's first frame, where they follow it (the reference
    refuses attachments before the first step, native/xpbd_capi.cu:195-200).
    Returns without waiting for the device (batch.sync() does)."""
    before, after = [], []
    for k, e in enumerate(entries):
        if e is None:
            continue
        script, f = e
        t, p, d = script.colliders(f)
        batch.set_colliders(k, t, p, d)
        att = script.attached(f)
        if att is not None:
            (after if f == 0 else before).append((k, att[0], att[1]))
        elif f == script.ramp:
            before.append((k, (), ()))
    if before:
        batch.set_attached(before)
    if after:
        batch.step_async(1)
        batch.set_attached(after)
        batch.step_async(substeps - 1)
    else:
        batch.step_async(substeps)
    if shapes:
        batch.shapes_device()


def build_batch(requests, lib=None, iters=ITERS, substeps=SUBSTEPS, tuning=None):
    """A batch of `requests` (objects with .scene and .recipe) configured with
    the full recipe and prepared (device buffers and graph ready)."""
    b = XPBDBatch([r.scene for r in requests], lib)
    for k, r in enumerate(requests):
        scenes.configure_full_batch(b, k, r.recipe, iters=iters, substeps=substeps)
    if tuning is not None:
        b.set_tuning(tuning)
    b.prepare()
    return b


def freeze(batch, k):
    """Switch every constraint, contact and force of object k off: a departed
    object keeps its slot until its batch is merged away, at the cost of a
    pass over its nodes that does no work."""
    batch.set_constraints(k, distance=False, shape=False, angle=False, volume=False)
    batch.set_self_collision(k, False)
    batch.set_ground(k, False, (0.0, 0.0, 1.0), 0.0, gravity=0.0)
    batch.set_colliders(k, [], [], [])
    batch.set_attached([(k, (), ())])


class Request:
    """One object to simulate for `life` frames (scene, recipe, script)."""

    def __init__(self, rid, scene, recipe, script, life, arrival_tick=None):
        self.id = rid
        self.scene, self.recipe, self.script = scene, recipe, script
        self.n = int(scene["pos"].shape[0])
        self.table = int(recipe["table"])
        self.life = int(life)
        self.arrival = arrival_tick
        self.admitted = None      # tick
        self.started = None       # tick of its first frame
        self.frame = 0
        self.done = None


def hbm_bytes(n, table):
    """Device memory of one object: ~2.5 KB per Gaussian (state, sliced
    distance entries, volume and shape clusters, self-collision buffers and
    candidate lists, the lane scan's 160 unsorted (key, distance) entries of
    8 B, the predicted-position grid's two records, shapes) plus 8 B per cell
    of its two hash tables (the coarse one and the nextpow2(2n)
    predicted-position one)."""
    fine = 1024
    while fine < 2 * n:
        fine <<= 1
    return 2500 * n + 8 * (table + fine)


class _Segment:
    def __init__(self, batch, objs):
        self.batch = batch
        self.objs = list(objs)          # Request per slot
        self.alive = [True] * len(objs)
        self.merging = False

    def nodes(self):
        return sum(r.n for r in self.objs)

    def live_nodes(self):
        return sum(r.n for r, a in zip(self.objs, self.alive) if a)


class Server:
    """Continuous-batching serving loop (one GPU)."""

    def __init__(self, hbm_budget=200e9, tick_budget_ms=None, max_segments=6, waste_limit=0.35, shapes=True,
                 record=False, lib=None, tuning=None):
        self.hbm_budget = hbm_budget
        self.tick_budget_ms = tick_budget_ms
        self.max_segments = max_segments
        self.waste_limit = waste_limit
        self.shapes = shapes
        self.record = record
        self.lib = lib
        self.tuning = tuning
        self.queue = collections.deque()
        self.segments = []
        self.building = []              # requests inside a pending build
        self.hbm_used = 0.0
        self.tick_no = 0
        self.rate = None                # measured Gaussian-substeps per ms
        self.trajectories = {}
        self.finished = []
        self.log = []
        self.merges = 0
        self._jobs = []
        self._pending_jobs = {"admit": collections.deque(), "merge": collections.deque()}
        self._lock = threading.Lock()
        self._worker = {}
        self._closers = []

    # -- background builds -----------------------------------------------------
    # One worker thread per kind of build (admissions, merges), so a merge
    # never holds new requests back.
    def _run_jobs(self, kind):
        while True:
            with self._lock:
                q = self._pending_jobs[kind]
                if not q:
                    self._worker[kind] = None
                    return
                payload = q.popleft()
            try:
                payload["batch"] = build_batch(payload["objs"], self.lib, tuning=self.tuning)
            except Exception as e:  # surfaced on the next tick
                payload["error"] = e
            with self._lock:
                self._jobs.append((kind, payload))

    def _close_later(self, batches):
        t = threading.Thread(target=lambda: [b.close() for b in batches], daemon=True)
        t.start()
        self._closers.append(t)

    def _submit_job(self, kind, payload):
        with self._lock:
            self._pending_jobs[kind].append(payload)
            if self._worker.get(kind) is None:
                self._worker[kind] = threading.Thread(target=self._run_jobs, args=(kind,), daemon=True)
                self._worker[kind].start()

    def _collect_jobs(self):
        with self._lock:
            done, self._jobs = self._jobs, []
        for kind, payload in done:
            if "error" in payload:
                raise RuntimeError(f"background {kind} build failed") from payload["error"]
            b = payload["batch"]
            if kind == "admit":
                for r in payload["objs"]:
                    self.building.remove(r)
                self.segments.append(_Segment(b, payload["objs"]))
            else:
                olds = payload["segments"]
                seg = _Segment(b, payload["objs"])
                for k, r in enumerate(payload["objs"]):
                    src, slot = payload["from"][k]
                    if src.alive[slot]:
                        b.copy_object(k, src.batch, slot)
                    else:
                        seg.alive[k] = False
                        freeze(b, k)
                for o in olds:
                    self.segments.remove(o)
                # freeing device memory can wait for the device: off the tick path
                self._close_later([o.batch for o in olds])
                self.segments.append(seg)

    # -- admission ---------------------------------------------------------------
    def submit(self, req):
        """Queue a request; admission runs at the tick boundary, so the
        requests that arrive together are built into one batch."""
        req.arrival = self.tick_no if req.arrival is None else req.arrival
        self.queue.append(req)

    def _live_nodes(self):
        return sum(s.live_nodes() for s in self.segments) + sum(r.n for r in self.building)

    def _admit(self):
        take = []
        while self.queue:
            r = self.queue[0]
            need = hbm_bytes(r.n, r.table)
            if self.hbm_used + need > self.hbm_budget:
                break
            if self.tick_budget_ms is not None and self.rate is not None:
                live = self._live_nodes() + sum(x.n for x in take) + r.n
                if live * SUBSTEPS / self.rate > 0.9 * self.tick_budget_ms and (self._live_nodes() or take):
                    break
            self.queue.popleft()
            self.hbm_used += need
            r.admitted = self.tick_no
            take.append(r)
        if take:
            self.building.extend(take)
            self._submit_job("admit", {"objs": take})

    def _maybe_merge(self):
        cand = [s for s in self.segments if not s.merging]
        if not cand:
            return
        wasteful = [s for s in cand if s.nodes() and 1.0 - s.live_nodes() / s.nodes() > self.waste_limit]
        pick = list(wasteful)
        if len(self.segments) > self.max_segments:
            small = sorted((s for s in cand if s not in pick), key=lambda s: s.live_nodes())
            pick += small[:max(2, len(self.segments) - self.max_segments + 1)]
        pick = [s for s in pick if any(s.alive)] + [s for s in pick if not any(s.alive)]
        if not pick:
            return
        empty = [s for s in pick if not any(s.alive)]
        for s in empty:                       # nothing to carry over
            self.segments.remove(s)
        if empty:
            self._close_later([s.batch for s in empty])
        pick = [s for s in pick if any(s.alive)]
        if len(pick) < 2 and not any(s in wasteful for s in pick):
            return
        objs, frm = [], []
        for s in pick:
            s.merging = True
            for k, (r, a) in enumerate(zip(s.objs, s.alive)):
                if a:
                    objs.append(r)
                    frm.append((s, k))
        self.merges += 1
        self._submit_job("merge", {"objs": objs, "from": frm, "segments": pick})

    # -- one tick ------------------------------------------------------------------
    def tick(self):
        t0 = time.perf_counter()
        self._collect_jobs()
        live_gs = 0
        for s in self.segments:
            entries = []
            for k, (r, a) in enumerate(zip(s.objs, s.alive)):
                if not a:
                    entries.append(None)
                    continue
                if r.started is None:
                    r.started = self.tick_no
                entries.append((r.script, r.frame))
                live_gs += r.n
            step_frame(s.batch, entries, shapes=self.shapes)
        for s in self.segments:
            s.batch.sync()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3
        if live_gs:
            rate = live_gs * SUBSTEPS / ms
            self.rate = rate if self.rate is None else 0.8 * self.rate + 0.2 * rate
        # bookkeeping after the results are ready
        for s in self.segments:
            for k, (r, a) in enumerate(zip(s.objs, s.alive)):
                if not a:
                    continue
                if self.record:
                    self.trajectories.setdefault(r.id, []).append(s.batch.positions(k))
                r.frame += 1
                if r.frame >= r.life:
                    r.done = self.tick_no
                    s.alive[k] = False
                    freeze(s.batch, k)
                    self.hbm_used -= hbm_bytes(r.n, r.table)
                    self.finished.append(r)
        self._admit()
        self._maybe_merge()
        rec = {"tick": self.tick_no, "ms": ms, "live_objects": sum(sum(s.alive) for s in self.segments),
               "live_gaussians": live_gs, "segments": len(self.segments), "queued": len(self.queue),
               "building": len(self.building)}
        self.log.append(rec)
        self.tick_no += 1
        return rec

    def _any_alive(self):
        return any(any(s.alive) for s in self.segments)

    def run(self, requests, max_ticks=100000, on_tick=None):
        """Submit each request at its arrival tick and tick until all have
        finished (or max_ticks).  While nothing is live and a build is
        outstanding the loop waits for it instead of ticking empty."""
        pending = sorted(requests, key=lambda r: r.arrival)
        while max_ticks > 0 and (pending or self.queue or self.building or self._any_alive()):
            while pending and pending[0].arrival <= self.tick_no:
                self.submit(pending.pop(0))
            self._admit()
            if not self._any_alive() and self.building:
                self._collect_jobs()
                if not self._any_alive():
                    time.sleep(0.002)
                    continue
            rec = self.tick()
            if on_tick is not None:
                on_tick(rec)
            max_ticks -= 1

    def close(self):
        while True:
            with self._lock:
                busy = any(w is not None for w in self._worker.values())
            if not busy:
                break
            time.sleep(0.01)
        self._collect_jobs()
        for t in self._closers:
            t.join()
        for s in self.segments:
            s.batch.close()
        self.segments = []


def run_alone(req, lib=None, shapes=True, tuning=None):
    """The request simulated alone in a one-object batch, same script and
    frames: positions after every frame."""
    b = build_batch([req], lib, tuning=tuning)
    out = []
    for f in range(req.life):
        step_frame(b, [(req.script, f)], shapes=shapes)
        b.sync()
        out.append(b.positions(0))
    b.close()
    return out


class TemplatePool:
    """Request payloads drawn from a pool of prepared objects: `count` objects
    with sizes log-uniform in [nmin, nmax] (each with its full recipe and a
    shared k-d tree); a request takes a random template and its own gripper
    script.  Stands in for clients that upload their own graphs."""

    def __init__(self, count, nmin, nmax, seed=5):
        from scipy.spatial import cKDTree
        rng = np.random.Generator(np.random.PCG64(seed))
        sizes = scenes.log_uniform_sizes(count, nmin, nmax, rng)
        self.items = []
        for j, n in enumerate(sizes):
            scene, recipe, _ = scenes.object_request(int(n), seed * 7919 + j)
            self.items.append((scene, recipe, cKDTree(scene["pos"].astype(np.float64))))

    def request(self, rid, rng, life, arrival):
        scene, recipe, tree = self.items[int(rng.integers(len(self.items)))]
        n = scene["pos"].shape[0]
        script = scenes.GripperScript(scene, seed=int(rng.integers(1 << 30)), attach=min(500, max(4, n // 10)),
                                      tree=tree)
        return Request(rid, scene, recipe, script, life, arrival)


def poisson_stream(pool, rate_per_s, ticks, life=(60, 600), tick_hz=60.0, seed=3, steady_start=True):
    """Requests arriving as a Poisson process of `rate_per_s` over `ticks`
    ticks, lifetimes uniform in `life` frames.  steady_start adds the
    population a stream that had been running for long would hold at tick 0
    (Little's law: rate x mean life objects, each with a residual lifetime
    drawn from the equilibrium distribution, density proportional to
    P(life > t))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    lo, hi = life
    if steady_start:
        mean_life = 0.5 * (lo + hi)
        count = rng.poisson(rate_per_s * mean_life / tick_hz)
        for _ in range(count):
            while True:        # residual life t ~ P(L > t) / E[L] by rejection
                t = rng.uniform(1, hi)
                if rng.uniform() * (hi - lo) <= min(hi - lo, hi - t):
                    break
            out.append(pool.request(len(out), rng, int(t), 0))
    t = 0.0
    while True:
        t += rng.exponential(tick_hz / rate_per_s)
        if t >= ticks:
            break
        out.append(pool.request(len(out), rng, int(rng.integers(lo, hi + 1)), int(t)))
    return out
