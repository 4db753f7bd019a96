"""Request loop for independent scenes on one GPU: continuous batching over a HipBatch context.

A request is one scene (parameters, initial state) and a number of IISPH steps. Requests are
admitted in arrival order while their device footprint (tsp_batch_footprint) fits the HBM
budget; admitted scenes stay resident and advance together, one launch per kernel for all of
them; a scene leaves the batch after its last step, its final state is copied out, and the
freed memory admits the next waiting requests. Scenes join and leave only between groups of
steps (at most `chunk_steps`, cut short by the scene that finishes first), so every scene
takes exactly its own number of steps, and its result is bit-identical to the scene run alone.

The plain form (`serve_plain`) runs the same stream one scene at a time, one context per scene,
in arrival order: the reference's one-simulation-per-backend model.

A request's scene is either a host state or a scene spec; a spec is generated on the device when
the request is admitted (tsp_batch_update_specs / tsp_pull_spec), so no host state is built or
copied for it.

Scenes of very different sizes are paced by the largest when they share a context: every step
of the batch waits for the slowest scene's sweeps. serve_batched therefore sorts requests into
size buckets (by fluid count), each with its own context, stream and host thread, so that the
small scenes' steps overlap with the large ones' on the GPU instead of waiting for them; the
buckets share the HBM budget.

Besides HBM, a context's cell tables are an admission resource: every resident scene holds the
cells of its cell frame (24 bytes each, dense over the frame's volume), and a context takes at
most its cell limit of them (tsp_set_cell_limit; 2^32 - 1 by default). A context refuses scenes
whose frames do not fit (TSP_ERR_CAPACITY) and keeps its resident scenes; the loop then admits
fewer of the waiting requests, or holds them until a resident scene leaves.
"""
import collections
import threading
import time

from . import _capi
from .backend import HipBackend, HipBatch, TspError, spec_counts


class Request:
    """One scene and its step count; `state` is a HostState, or None with `spec` a tsp_scene_spec."""

    def __init__(self, rid, arrival_s, params, state, steps, spec=None):
        if (state is None) == (spec is None):
            raise ValueError("a request carries either a state or a scene spec")
        self.rid = rid
        self.arrival_s = float(arrival_s)
        self.params = params
        self.state = state
        self.spec = spec
        self.steps = int(steps)
        self.n = state.n if state is not None else spec_counts(spec)
        self.footprint = HipBatch.footprint(self.n)
        # filled in by the server
        self.admitted_s = None
        self.done_s = None
        self.result = None

    @property
    def particle_steps(self):
        return self.n[0] * self.steps

    @property
    def latency_s(self):
        return None if self.done_s is None else self.done_s - self.arrival_s


def _wait_until(t0, t):
    dt = t - (time.perf_counter() - t0)
    if dt > 0:
        time.sleep(dt)


def _admit(batch, admit):
    """One admission of host states and scene specs; either all of them join or none does.
    Returns the requests in the order of their new scene ids, and the ids."""
    hosted = [r for r in admit if r.state is not None]
    specced = [r for r in admit if r.state is None]
    ids = batch.update(add=[(r.params, r.state) for r in hosted]) if hosted else []
    try:
        if specced:
            ids += batch.update_specs(add=[(r.params, r.spec) for r in specced])
    except Exception:
        if ids:
            batch.update(remove=ids)
        raise
    return hosted + specced, ids


class _Budget:
    """The HBM budget the size buckets share."""

    def __init__(self, total):
        self.total = total
        self.used = 0
        self.lock = threading.Lock()

    def take(self, n):
        with self.lock:
            if self.used + n > self.total:
                return False
            self.used += n
            return True

    def give(self, n):
        with self.lock:
            self.used -= n


def _serve_bucket(requests, budget, chunk_steps, device, keep_results, t0, cell_limit=None):
    """Continuous batching of one size bucket in its own context (arrival-order admission)."""
    batch = HipBatch(device)
    if cell_limit is not None:
        batch.set_cell_limit(cell_limit)
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival_s))
    waiting = collections.deque()
    active = {}            # scene id -> [request, steps left]
    hold = False           # the context refused the head of the queue: wait for a completion
    try:
        while pending or waiting or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival_s <= now:
                waiting.append(pending.popleft())
            admit = []
            while not hold and waiting and budget.take(waiting[0].footprint):
                admit.append(waiting.popleft())
            if waiting and not active and not admit and waiting[0].footprint > budget.total:
                r = waiting[0]
                raise RuntimeError(f"request {r.rid} needs {r.footprint} B, more than the {budget.total} B budget")
            while admit:
                try:
                    admitted, ids = _admit(batch, admit)
                except TspError as e:
                    if e.code != _capi.TSP_ERR_CAPACITY or (len(admit) == 1 and not active):
                        raise
                    # put the requests back at the head of the queue, in order; try the first
                    # alone, and hold the queue when even that one does not fit
                    for r in reversed(admit):
                        waiting.appendleft(r)
                        budget.give(r.footprint)
                    if len(admit) == 1:
                        hold = True
                        admit = []
                    else:
                        r = waiting.popleft()
                        admit = [r] if budget.take(r.footprint) else []
                        if not admit:
                            waiting.appendleft(r)
                    continue
                now = time.perf_counter() - t0
                for sid, r in zip(ids, admitted):
                    r.admitted_s = now
                    active[sid] = [r, r.steps]
                admit = []
            if not active:
                # nothing resident: wait for the next arrival, or for another bucket to free memory
                if pending and not waiting:
                    dt = pending[0].arrival_s - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                else:
                    time.sleep(0.001)
                continue
            n = min(chunk_steps, min(left for _, left in active.values()))
            batch.step(n)
            done = []
            for sid, entry in active.items():
                entry[1] -= n
                if entry[1] == 0:
                    done.append(sid)
            if done:
                for sid in done:
                    r = active[sid][0]
                    if keep_results:
                        r.result = batch.push(sid)
                batch.update(remove=done)
                now = time.perf_counter() - t0
                for sid in done:
                    r = active.pop(sid)[0]
                    r.done_s = now
                    budget.give(r.footprint)
                hold = False
    finally:
        batch.close()


# fluid counts that separate the size buckets of serve_batched
DEFAULT_BUCKETS = (60_000, 400_000)


def serve_batched(requests, hbm_budget_bytes, chunk_steps=10, device=0, keep_results=True, buckets=DEFAULT_BUCKETS,
                  cell_limit=None):
    """Serves `requests` (arrival times relative to the call) with continuous batching, one
    context per size bucket (fluid counts split at `buckets`), the buckets' host threads running
    side by side; returns the wall time from the start to the last completion. Fills admitted_s,
    done_s and result. `cell_limit` caps each context's cell tables (tsp_set_cell_limit)."""
    edges = sorted(buckets)
    groups = [[] for _ in range(len(edges) + 1)]
    for r in requests:
        groups[sum(r.n[0] >= e for e in edges)].append(r)
    budget = _Budget(hbm_budget_bytes)
    for r in requests:
        if r.footprint > hbm_budget_bytes:
            raise RuntimeError(f"request {r.rid} needs {r.footprint} B, more than the {hbm_budget_bytes} B budget")
    errors = []
    t0 = time.perf_counter()

    def run(g):
        try:
            _serve_bucket(g, budget, chunk_steps, device, keep_results, t0, cell_limit)
        except BaseException as e:   # re-raised in the caller
            errors.append(e)

    threads = [threading.Thread(target=run, args=(g,)) for g in groups if g]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return time.perf_counter() - t0


def serve_plain(requests, device=0, keep_results=True):
    """The same stream one scene at a time (one context per scene, arrival order)."""
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival_s):
        _wait_until(t0, r.arrival_s)
        r.admitted_s = time.perf_counter() - t0
        g = HipBackend(device)
        try:
            g.set_params(r.params)
            if r.state is not None:
                g.pull(r.state)
            else:
                g.pull_spec(r.spec)
            g.step(r.steps)
            out = g.push()
            if keep_results:
                r.result = out
        finally:
            g.close()
        r.done_s = time.perf_counter() - t0
    return time.perf_counter() - t0


def latency_summary(requests):
    """p50 / p99 completion latency (seconds, nearest rank) and the mean."""
    lat = sorted(r.latency_s for r in requests)
    if not lat:
        return {"p50_s": None, "p99_s": None, "mean_s": None}

    def rank(p):
        k = max(0, min(len(lat) - 1, int(-(-p * len(lat) // 1)) - 1))
        return lat[k]
    return {"p50_s": round(rank(0.50), 3), "p99_s": round(rank(0.99), 3), "mean_s": round(sum(lat) / len(lat), 3)}
