"""Serving: continuous batching of rollout requests on one GPU.

Requests (a particle object, its material and parameters, a grid size, a
horizon, and whether the gradient of a final-position loss is wanted)
arrive at any time.  The engine keeps every admitted request's state on the
device and advances them in ticks:

- admission: queued requests are admitted first come, first served while
  their device-memory estimate fits the HBM budget and the batch context's
  capacity (particles, simulations, grid size); a request that could never
  fit is rejected at once;
- forward work: every request in its forward phase advances by the next
  segment of its horizon (at most `chunk` substeps).  Requests with equal
  segment lengths run in one batched call: their particles are concatenated
  into one arena and every stage of every substep is one launch over all of
  them (empm_batch_rollout_fwd);
- backward work: a request that wants a gradient keeps its state at every
  segment start (a checkpoint) and, once its forward horizon is done,
  sweeps its segments in reverse: each segment is re-run recording its tape
  and reverse-swept (empm_batch_rollout_fwd + empm_batch_rollout_adj), again
  batched with every other request whose current segment has the same
  length.

So requests that arrive while others are on the GPU join the next tick and
share its launches with them; a request's result does not depend on what
it was batched with (the batched kernels give each simulation its own
particle range, grid and blocks).

Ticks are asynchronous: a tick's launches are queued and the host goes on
to plan the next one while the device runs them.  Each tick records an
event; the requests that finished in it are timestamped (and their loss and
failure flag read back, through pinned host copies queued before the
event) when that event completes.  At most `max_inflight` ticks are queued
ahead of the device, so a request that arrives is planned within that many
ticks.  A request whose state turned non-finite (an inverted element, say)
is reported as failed in its result; the others are unaffected.  The reference has no serving path: its
rollouts run one at a time through WarpMPMWrapper.advance_to
(phys_sim/diff_simulator/warp_mpm/warp_mpm_wrapper.py:278-306).

The co-scheduled requests share dt and the collider table; the engine uses
the grid-relative bounding box (add_bounding_box,
mpm_solver_diff.py:871-938: the velocity component pointing out of the
domain is zeroed within 3 cells of every face), which applies to any grid
size and does not depend on time.

`Scheduler` holds the admission and planning logic (no device work);
`ServingEngine` executes its plans.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from . import runtime as rt
from .diff_simulator.warp_mpm.mpm_data_structure import friction_alpha

F32 = torch.float32

QUEUED, FORWARD, BACKWARD, DONE, REJECTED = "queued", "forward", "backward", "done", "rejected"


@dataclass
class RolloutRequest:
    """One rollout: particles x (N, 3) with velocities v (N, 3, default 0),
    volumes vol (N,) and density rho, on an n_grid^3 grid over the unit cube,
    for n_steps substeps.  With want_grad, the loss is the mean over
    particles of |x_T - target|^2 and the result carries its gradient with
    respect to (log E, nu, log yield_stress)."""
    x: np.ndarray
    vol: np.ndarray
    material: int
    E: float
    nu: float
    n_grid: int
    n_steps: int
    v: np.ndarray | None = None
    rho: float = 1000.0
    yield_stress: float = 1e3
    friction_angle: float = 30.0
    softening: float = 0.1
    want_grad: bool = False
    target: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])


@dataclass
class RolloutResult:
    x: torch.Tensor                  # (N, 3) final positions, request order
    v: torch.Tensor                  # (N, 3) final velocities
    loss: float | None = None
    grad: torch.Tensor | None = None  # d loss / d(log E, nu, log Y), float64
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    failed: bool = False              # the final state holds non-finite values

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def bin_ints_estimate(n_particles: int, n_grid: int) -> int:
    """Ints of one binning record (perm, per-cell offsets, block lists)."""
    nb = (n_grid + 1) // 2
    return int(n_particles) + 13 * nb ** 3 + 24


def request_bytes(n_particles: int, n_grid: int, n_steps: int, want_grad: bool, chunk: int) -> int:
    """Device bytes a request holds while admitted: its state and parameters,
    its copies in the batched arena (input, two ping-pong buffers, the
    parameter arrays), and with want_grad its segment checkpoints, cotangent
    and parameter gradients plus its share of one recorded segment (state
    with Fe 132 B/particle, grid 32 B/node and the binning, per substep)."""
    N, G = int(n_particles), int(n_grid) ** 3
    b = (96 + 40) * N + (3 * 96 + 40) * N
    if want_grad:
        segs = -(-int(n_steps) // int(chunk))
        b += segs * 96 * N + 2 * 96 * N + 2 * 20 * N + 12 * N
        b += (chunk + 1) * 132 * N + chunk * (32 * G + 4 * bin_ints_estimate(N, n_grid))
    return b


@dataclass
class Slot:
    rid: int
    req: RolloutRequest
    arrival: float
    nbytes: int
    phase: str = QUEUED
    k: int = 0                                        # forward substeps done
    starts: list = field(default_factory=list)        # segment starts (want_grad)
    seg_end: int = 0                                  # backward: end of the next segment to sweep
    admitted: float = 0.0
    dev: object = None                                # device state (ServingEngine)


@dataclass
class Work:
    kind: str          # FORWARD or BACKWARD
    length: int        # substeps
    rids: list


class Scheduler:
    """Admission and tick planning (host logic only)."""

    def __init__(self, budget_bytes: int, max_particles: int, max_cand: int, max_grid: int, chunk: int = 50):
        self.budget = int(budget_bytes)
        self.max_particles = int(max_particles)
        self.max_cand = int(max_cand)
        self.max_grid = int(max_grid)
        self.chunk = int(chunk)
        self.queue: collections.deque = collections.deque()
        self.slots: dict[int, Slot] = {}
        self.used = 0

    def submit(self, rid: int, req: RolloutRequest, arrival: float) -> Slot:
        s = Slot(rid, req, arrival, request_bytes(req.n, req.n_grid, req.n_steps, req.want_grad, self.chunk))
        self.slots[rid] = s
        if (s.nbytes > self.budget or req.n > self.max_particles or req.n_grid > self.max_grid or req.n_steps < 1
                or req.n < 1):
            s.phase = REJECTED
        else:
            self.queue.append(rid)
        return s

    def active(self) -> list[Slot]:
        return [s for s in self.slots.values() if s.phase in (FORWARD, BACKWARD)]

    def admit(self, now: float) -> list[Slot]:
        """Admit queued requests in arrival order while they fit."""
        out = []
        n_active = len(self.active())
        while self.queue:
            s = self.slots[self.queue[0]]
            if self.used + s.nbytes > self.budget or n_active >= self.max_cand:
                break
            self.queue.popleft()
            s.phase = FORWARD
            s.admitted = now
            s.starts = [0] if s.req.want_grad else []
            self.used += s.nbytes
            n_active += 1
            out.append(s)
        return out

    def _groups(self, kind: str, items: list[tuple[int, Slot]]) -> list[Work]:
        """Requests by equal segment length, each group split so that one call
        fits the context's particle and simulation capacity."""
        by_len: dict[int, list[Slot]] = collections.defaultdict(list)
        for length, s in items:
            by_len[length].append(s)
        work = []
        for length in sorted(by_len, reverse=True):
            cur, n = [], 0
            for s in by_len[length]:
                if cur and (n + s.req.n > self.max_particles or len(cur) >= self.max_cand):
                    work.append(Work(kind, length, [x.rid for x in cur]))
                    cur, n = [], 0
                cur.append(s)
                n += s.req.n
            if cur:
                work.append(Work(kind, length, [x.rid for x in cur]))
        return work

    def plan(self) -> list[Work]:
        fwd = [(min(self.chunk, s.req.n_steps - s.k), s) for s in self.active() if s.phase == FORWARD]
        bwd = [(s.seg_end - s.starts[-1], s) for s in self.active() if s.phase == BACKWARD]
        return self._groups(FORWARD, fwd) + self._groups(BACKWARD, bwd)

    def forward_done(self, rid: int, length: int) -> bool:
        """Record a forward segment; True when the forward horizon is done."""
        s = self.slots[rid]
        s.k += length
        if s.k < s.req.n_steps:
            if s.req.want_grad:
                s.starts.append(s.k)
            return False
        if s.req.want_grad:
            s.phase = BACKWARD
            s.seg_end = s.k
        else:
            self._finish(s)
        return True

    def backward_done(self, rid: int) -> bool:
        """Record a swept segment; True when the reverse sweep is done."""
        s = self.slots[rid]
        s.seg_end = s.starts.pop()
        if s.seg_end == 0:
            self._finish(s)
            return True
        return False

    def _finish(self, s: Slot):
        s.phase = DONE
        self.used -= s.nbytes

    def idle(self) -> bool:
        return not self.queue and not self.active()


def bounding_box_collider():
    """add_bounding_box with its default activity window (mpm_solver_diff.py:871-938)."""
    c = _lib.Collider()
    c.kind = 2
    c.start_time, c.end_time = 0.0, 999.0
    return c


@dataclass
class _Dev:
    perm: torch.Tensor
    model: rt.ModelArrays
    E: torch.Tensor
    nu: torch.Tensor
    state: rt.ParticleState
    ckpt: dict
    target: torch.Tensor | None = None
    ys0: torch.Tensor | None = None  # yield stress at admission (the gradient is w.r.t. its log)
    cot: rt.ParticleState | None = None
    grads: rt.ParamGrads | None = None
    loss: float | None = None


class ServingEngine:
    """Executes Scheduler plans on one device.  `submit` enqueues a request
    (arrival time defaults to now), `step` queues one tick (admission, then
    every planned call) and returns the ids of the requests it finishes;
    their results appear in `results` when the device has run the tick
    (`drain` waits for every queued tick)."""

    def __init__(self, device="cuda:0", max_particles=2_000_000, max_cand=32, max_grid=128, dt=1e-4, chunk=50,
                 hbm_budget_bytes: int | None = None, gravity=(0.0, 0.0, -9.8), max_inflight: int = 2):
        self.device = torch.device(device)
        self.max_inflight = max(1, int(max_inflight))
        # queued ticks: (event, [(rid, result, pinned loss, pinned finite flag)])
        self._inflight: collections.deque = collections.deque()
        self.dt = float(dt)
        self.gravity = tuple(gravity)
        self.ctx = rt.Context(max_grid, max_particles, self.device, max_cand=max_cand)
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget_bytes = int(0.6 * free)
        self.sched = Scheduler(hbm_budget_bytes, max_particles, max_cand, max_grid, chunk)
        self.cols = [bounding_box_collider()]
        self.results: dict[int, RolloutResult] = {}
        self._next = 0
        self.ticks = 0
        self.calls = 0

    # ------------------------------------------------------------ requests
    def submit(self, req: RolloutRequest, arrival: float | None = None) -> int:
        if req.want_grad and req.target is None:
            raise ValueError("a gradient request needs a target")
        rid = self._next
        self._next += 1
        s = self.sched.submit(rid, req, time.perf_counter() if arrival is None else arrival)
        if s.phase == REJECTED:
            raise ValueError(f"request {rid} can never fit: {s.nbytes / 1e9:.2f} GB of {self.sched.budget / 1e9:.2f} "
                             f"GB, N={req.n} (max {self.sched.max_particles}), grid {req.n_grid} "
                             f"(max {self.sched.max_grid})")
        return rid

    def _admit_device(self, s: Slot):
        r, dev = s.req, self.device

        def t(a):
            # through pinned memory: the upload is queued behind the ticks in
            # flight instead of waiting for them
            h = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).pin_memory()
            return h.to(dev, non_blocking=True)

        x = t(r.x)
        n = r.n
        # cell order of the initial positions: neighbouring particles share blocks
        key = torch.floor(x * float(r.n_grid) - 0.5).clamp_(0, r.n_grid - 1).long()
        key = (key[:, 0] * r.n_grid + key[:, 1]) * r.n_grid + key[:, 2]
        perm = torch.argsort(key, stable=True)
        g = lambda a: a.index_select(0, perm).contiguous()
        E = torch.full((n,), float(r.E), dtype=F32, device=dev)
        nu = torch.full((n,), float(r.nu), dtype=F32, device=dev)
        mu, lam = torch.empty_like(E), torch.empty_like(E)
        rt.mu_lam(E, nu, mu, lam)
        vol = g(t(r.vol))
        model = rt.ModelArrays(
            material=int(r.material), n_grid=int(r.n_grid), grid_lim=1.0, dx=float(np.float32(1.0 / r.n_grid)),
            inv_dx=float(np.float32(r.n_grid)), mu=mu, lam=lam,
            yield_stress=torch.full((n,), float(r.yield_stress), dtype=F32, device=dev), mass=vol * float(r.rho),
            vol=vol, gravity=self.gravity, alpha=friction_alpha(r.friction_angle), softening=float(r.softening))
        v = g(t(r.v)) if r.v is not None else torch.zeros((n, 3), device=dev)
        state = rt.ParticleState(g(x), v, torch.zeros((n, 3, 3), device=dev),
                                 torch.eye(3, device=dev).repeat(n, 1, 1).contiguous())
        s.dev = _Dev(perm=perm, model=model, E=E, nu=nu, state=state, ckpt={},
                     target=g(t(r.target)) if r.want_grad else None)
        if r.want_grad:
            s.dev.ys0 = model.yield_stress.clone()
            s.dev.ckpt[0] = self._checkpoint(state, model)

    @staticmethod
    def _checkpoint(state: rt.ParticleState, model: rt.ModelArrays) -> rt.ParticleState:
        """A segment start of a gradient request: its state and, for a
        material that rewrites (yield stress, mu, lam) in its return map,
        their values there (the segment's re-run starts from them)."""
        return rt.ParticleState(state.x, state.v, state.C, state.F,
                                mstate=model.material_state() if model.mutates else None)

    # ------------------------------------------------------------ ticks
    def _batch(self, slots: list[Slot]):
        return rt.BatchModel.concat([s.dev.model for s in slots])

    @staticmethod
    def _cat(states):
        return rt.ParticleState(*(torch.cat([getattr(s, f) for s in states]).contiguous()
                                  for f in ("x", "v", "C", "F")))

    @staticmethod
    def _split(st: rt.ParticleState, counts):
        parts = [torch.split(getattr(st, f), counts) for f in ("x", "v", "C", "F")]
        return [rt.ParticleState(*(p[i].clone() for p in parts)) for i in range(len(counts))]

    def _forward(self, w: Work, slots: list[Slot]):
        bm = self._batch(slots)
        counts = [s.req.n for s in slots]
        ro = rt.Rollout(self.ctx, bm, bm.n, sort=False)
        final, _, _, _ = ro.forward(self._cat([s.dev.state for s in slots]), list(self.cols), 0.0, self.dt, w.length)
        for s, st, mu, lam, ys in zip(slots, self._split(final, counts), torch.split(bm.mu, counts),
                                      torch.split(bm.lam, counts), torch.split(bm.yield_stress, counts)):
            s.dev.state = st
            if s.dev.model.mutates:  # plasticine rewrites its parameters: they are state
                s.dev.model.mu.copy_(mu)
                s.dev.model.lam.copy_(lam)
                s.dev.model.yield_stress.copy_(ys)
            if self.sched.forward_done(s.rid, w.length):
                if s.req.want_grad:
                    d = s.dev
                    diff = st.x - d.target
                    d.loss = (diff * diff).sum(1).double().mean()
                    d.cot = rt.ParticleState.empty(s.req.n, self.device)
                    d.cot.zero_()
                    d.cot.x.copy_(diff * (2.0 / s.req.n))
                    d.grads = rt.ParamGrads.zeros(s.req.n, self.device)
            elif s.req.want_grad:
                s.dev.ckpt[s.k] = self._checkpoint(st, s.dev.model)

    def _backward(self, w: Work, slots: list[Slot]):
        # the segment is re-run from its start: a material that rewrites its
        # parameters starts from their values there (the checkpoint's mstate)
        import dataclasses

        models = []
        for s in slots:
            ck = s.dev.ckpt[s.starts[-1]]
            m = s.dev.model
            if ck.mstate is not None:
                m = dataclasses.replace(m, yield_stress=ck.mstate[0], mu=ck.mstate[1], lam=ck.mstate[2])
            models.append(m)
        bm = rt.BatchModel.concat(models)
        counts = [s.req.n for s in slots]
        ro = rt.Rollout(self.ctx, bm, bm.n, sort=False)
        start = self._cat([s.dev.ckpt[s.starts[-1]] for s in slots])
        _, _, _, tape = ro.forward(start, list(self.cols), 0.0, self.dt, w.length, tape=True, ckpt_every=1)
        # the parameter cotangents run through the segments: for materials
        # that rewrite (yield stress, mu, lam) they are those of the state at
        # the segment's end on entry and at its start on return; for the
        # others the segment adds its share
        g_in = rt.ParamGrads(*(torch.cat([getattr(s.dev.grads, f) for s in slots]).contiguous()
                               for f in ("mu", "lam", "ys", "mass", "vol")))
        d0, g = ro.backward(tape, self._cat([s.dev.cot for s in slots]), flags=_lib.EMPM_ADJ_EXACT, grads=g_in)
        gparts = [torch.split(t, counts) for t in (g.mu, g.lam, g.ys, g.mass, g.vol)]
        for i, (s, cot) in enumerate(zip(slots, self._split(d0, counts))):
            d = s.dev
            d.cot = cot
            d.grads = rt.ParamGrads(*(part[i].clone() for part in gparts))
            d.ckpt.pop(s.starts[-1], None)
            self.sched.backward_done(s.rid)

    def _result(self, s: Slot):
        """The result of a finished request (device work queued; the loss and
        the failure flag come back through pinned host copies)."""
        d = s.dev
        inv = torch.empty_like(d.perm)
        inv[d.perm] = torch.arange(d.perm.numel(), device=self.device)
        res = RolloutResult(x=d.state.x.index_select(0, inv), v=d.state.v.index_select(0, inv),
                            arrival=s.arrival, admitted=s.admitted)
        finite = torch.empty((), dtype=torch.bool).pin_memory()
        finite.copy_(torch.isfinite(d.state.x).all() & torch.isfinite(d.state.v).all(), non_blocking=True)
        loss = None
        if s.req.want_grad:
            dE, dnu = torch.zeros_like(d.E), torch.zeros_like(d.E)
            rt.mu_lam_adj(d.E, d.nu, d.grads.mu, d.grads.lam, dE, dnu)
            loss = torch.empty((), dtype=torch.float64).pin_memory()
            loss.copy_(d.loss, non_blocking=True)
            res.grad = torch.stack([(dE * d.E).double().sum(), dnu.double().sum(),
                                    (d.grads.ys * d.ys0).double().sum()])
        return res, loss, finite

    def _poll(self, block: bool = False):
        """Complete the queued ticks whose events have passed, oldest first
        (with block, wait for the oldest): timestamp their finished requests
        and read back their losses and failure flags."""
        while self._inflight:
            ev, items = self._inflight[0]
            if not ev.query():
                if not block:
                    return
                ev.synchronize()
            block = False
            now = time.perf_counter()
            for rid, res, loss, finite in items:
                res.finished = now
                res.failed = not bool(finite)
                if loss is not None:
                    res.loss = float(loss)
                self.results[rid] = res
            self._inflight.popleft()

    def step(self) -> list[int]:
        """One tick: admission, then every planned call, queued on the
        device; returns the ids of the requests whose last work it queued
        (their results appear in `results` once the device has run it)."""
        self._poll()
        while len(self._inflight) >= self.max_inflight:
            self._poll(block=True)
        now = time.perf_counter()
        for s in self.sched.admit(now):
            self._admit_device(s)
        plan = self.sched.plan()
        for w in plan:
            slots = [self.sched.slots[r] for r in w.rids]
            (self._forward if w.kind == FORWARD else self._backward)(w, slots)
            self.calls += 1
        items = []
        for s in self.sched.slots.values():
            if s.phase == DONE and s.dev is not None:
                items.append((s.rid, *self._result(s)))
                s.dev = None
        ev = torch.cuda.Event()
        ev.record()
        self._inflight.append((ev, items))
        self.ticks += 1
        return [it[0] for it in items]

    def drain(self):
        """Wait for every queued tick and complete its results."""
        while self._inflight:
            self._poll(block=True)

    def run_until_idle(self) -> dict[int, RolloutResult]:
        while not self.sched.idle():
            self.step()
        self.drain()
        return self.results

    def close(self):
        self.drain()
        self.ctx.close()


def serve_poisson(engine: ServingEngine, requests: list[RolloutRequest], qps: float, seed: int = 6,
                  log=None) -> dict:
    """Submit `requests` with Poisson arrivals at `qps` (exponential gaps
    from the seed) into a running engine, tick until all are served, and
    return latency percentiles and throughput."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / qps, len(requests))
    offsets = np.cumsum(gaps) - gaps[0]
    t0 = time.perf_counter()
    pending = collections.deque(zip(offsets, requests))
    ids = []
    while pending or not engine.sched.idle():
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            off, req = pending.popleft()
            ids.append(engine.submit(req, arrival=t0 + off))
        if engine.sched.idle():
            # complete what is queued first, so that finish times are not
            # taken after the sleep
            engine.drain()
            if pending:
                time.sleep(max(0.0, pending[0][0] - (time.perf_counter() - t0)))
            continue
        engine.step()
        if log and engine.ticks % 20 == 0:
            log(f"[serve] t={time.perf_counter() - t0:.1f}s tick {engine.ticks}: {len(engine.sched.active())} active, "
                f"{len(engine.sched.queue)} queued, {len(engine.results)} done")
    engine.drain()
    wall = time.perf_counter() - t0
    lat = np.array([engine.results[i].latency for i in ids])
    work = sum(r.n * r.n_steps for r in requests)
    return {"requests": len(ids), "qps": qps, "wall_s": wall, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
            "particle_substeps_per_s": work / wall, "ticks": engine.ticks, "calls": engine.calls,
            "grad_requests": sum(1 for r in requests if r.want_grad),
            "failed_requests": sum(1 for i in ids if engine.results[i].failed)}
