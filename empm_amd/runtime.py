"""Runtime around libempm: device context, particle states, and the
checkpointed forward / reverse rollout drivers.

Everything numerical runs in the HIP kernels behind ``empm_amd._lib``; this
module only owns buffers (torch tensors used as device memory), builds the
C-ABI descriptors and sequences the calls.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import torch

from . import _lib
from ._lib import Batch, BatchCreateDesc, Collider, CreateDesc, GridRec, Model, Particles, check

F32 = torch.float32


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if t.dtype not in (torch.float32, torch.int32):
        raise RuntimeError(f"expected float32/int32 tensor, got {t.dtype}")
    if not t.is_contiguous():
        raise RuntimeError("tensor must be contiguous")
    if t.device.type != "cuda":
        raise RuntimeError(f"tensor must live on a GPU, got {t.device}")
    return t.data_ptr()


def stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class Context:
    """An empm_ctx: grid buffers + per-particle scratch on one device.  With
    max_cand given it is a batch context for up to max_cand simulations of at
    most max_particles particles in all, each grid at most n_grid^3."""

    def __init__(self, n_grid: int, max_particles: int, device="cuda:0", max_cand: int | None = None):
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"libempm runs on a GPU; got device {device}")
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.device = torch.device("cuda", idx)
        self.n_grid = int(n_grid)
        self.max_particles = int(max_particles)
        self.max_cand = max_cand
        h = C.c_void_p()
        if max_cand is None:
            desc = CreateDesc(device=idx, n_grid=self.n_grid, max_particles=self.max_particles)
            check(self.lib.empm_create(C.byref(h), C.byref(desc)))
        else:
            desc = BatchCreateDesc(device=idx, max_cand=int(max_cand), max_particles=self.max_particles,
                                   max_grid=self.n_grid)
            check(self.lib.empm_create_batch(C.byref(h), C.byref(desc)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def set_graphs(self, enable: bool = True):
        """Replay every rollout call of this context as a cached HIP graph."""
        check(self.lib.empm_set_graphs(self._h, int(bool(enable))))
        return self

    def set_tuning(self, tuning: dict):
        """Launch parameters of the binned kernels (empm_tuning fields; see
        empm_amd.autotune)."""
        check(self.lib.empm_set_tuning(self._h, C.byref(_lib.Tuning.of(tuning))))
        return self

    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        check(self.lib.empm_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_deterministic(self, enable: bool = True):
        """Bitwise-reproducible results: float sums in an order fixed by the
        particles' indices instead of by atomic arrival (slower; for tests
        and debugging)."""
        check(self.lib.empm_set_deterministic(self._h, int(bool(enable))))
        return self

    def graph_stats(self):
        """(captures, replays) of this context's graph cache."""
        a, b = C.c_longlong(), C.c_longlong()
        check(self.lib.empm_graph_stats(self._h, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def read_status(self, clear: bool = True):
        """(return code, per-simulation EMPM_STATUS_* bits) since the last read:
        one synchronisation of the device's current stream."""
        n = int(self.max_cand or 1)
        arr = (C.c_int * n)()
        rc = self.lib.empm_read_status(self._h, arr, n, int(bool(clear)), stream_handle(self.device))
        return rc, list(arr)

    def check_status(self):
        """Raise EmpmError (EMPM_E_NONFINITE) if a forward substep since the
        last check produced a non-finite state; warn on a CFL violation."""
        import warnings

        rc, st = self.read_status()
        fast = [c for c, b in enumerate(st) if b & _lib.EMPM_STATUS_CFL]
        if fast:
            warnings.warn(f"particles moved more than a cell in one substep (simulations {fast}): the time step is "
                          f"too large for the velocities", RuntimeWarning, stacklevel=3)
        check(rc)
        return st

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.empm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleState:
    """x (N,3), v (N,3), C (N,3,3), F (N,3,3) = F_trial, optional Fe (N,3,3)
    and optional mstate (3, N): the yield stress, mu and lam the next substep
    starts from, state for the materials that rewrite them (material 5,
    hardening == 1)."""

    __slots__ = ("x", "v", "C", "F", "Fe", "mstate")

    def __init__(self, x, v, C_, F, Fe=None, mstate=None):
        self.x, self.v, self.C, self.F, self.Fe, self.mstate = x, v, C_, F, Fe, mstate

    @staticmethod
    def empty(n, device, with_fe=False):
        z3 = lambda: torch.zeros((n, 3), dtype=F32, device=device)
        z9 = lambda: torch.zeros((n, 3, 3), dtype=F32, device=device)
        return ParticleState(z3(), z3(), z9(), z9(), z9() if with_fe else None)

    def zero_(self):
        for t in (self.x, self.v, self.C, self.F):
            t.zero_()
        return self

    def copy_(self, o: "ParticleState"):
        self.x.copy_(o.x)
        self.v.copy_(o.v)
        self.C.copy_(o.C)
        self.F.copy_(o.F)
        if self.mstate is not None and o.mstate is not None:
            self.mstate.copy_(o.mstate)
        return self

    def clone(self):
        return ParticleState(self.x.clone(), self.v.clone(), self.C.clone(), self.F.clone(),
                             mstate=self.mstate.clone() if self.mstate is not None else None)

    def desc(self) -> Particles:
        return Particles(x=_ptr(self.x), v=_ptr(self.v), C=_ptr(self.C), F=_ptr(self.F),
                         Fe=_ptr(self.Fe) if self.Fe is not None else None,
                         mstate=_ptr(self.mstate) if self.mstate is not None else None)


@dataclass
class ModelArrays:
    """Per-particle arrays + scalars that become an empm_model."""
    material: int
    n_grid: int
    grid_lim: float
    dx: float
    inv_dx: float
    mu: torch.Tensor
    lam: torch.Tensor
    yield_stress: torch.Tensor
    mass: torch.Tensor
    vol: torch.Tensor
    selection: torch.Tensor | None = None
    rpic_damping: float = 0.0
    grid_v_damping_scale: float = 1.1
    gravity: tuple = (0.0, 0.0, 0.0)
    alpha: float = 0.0
    hardening: float = 0.0
    xi: float = 0.0
    plastic_viscosity: float = 0.0
    softening: float = 0.1

    @property
    def mutates(self) -> bool:
        """Whether the return map rewrites yield stress, mu and lam (material
        5's softening, hardening == 1 of the von Mises maps)."""
        return _mutates(self)

    def material_state(self) -> torch.Tensor:
        """(3, N) copy of (yield stress, mu, lam): the current material state."""
        return torch.stack([self.yield_stress, self.mu, self.lam]).contiguous()

    def with_private_params(self) -> "ModelArrays":
        """The same model with its own copies of mu, lam and yield stress."""
        import dataclasses
        return dataclasses.replace(self, mu=self.mu.clone(), lam=self.lam.clone(),
                                   yield_stress=self.yield_stress.clone())

    def struct(self) -> Model:
        m = Model()
        m.n_particles = int(self.mu.shape[0])
        m.n_grid = int(self.n_grid)
        m.material = int(self.material)
        m.grid_lim = float(self.grid_lim)
        m.dx = float(self.dx)
        m.inv_dx = float(self.inv_dx)
        m.rpic_damping = float(self.rpic_damping)
        m.grid_v_damping_scale = float(self.grid_v_damping_scale)
        for i in range(3):
            m.gravity[i] = float(self.gravity[i])
        m.alpha = float(self.alpha)
        m.hardening = float(self.hardening)
        m.xi = float(self.xi)
        m.plastic_viscosity = float(self.plastic_viscosity)
        m.softening = float(self.softening)
        m.mu = _ptr(self.mu)
        m.lam = _ptr(self.lam)
        m.yield_stress = _ptr(self.yield_stress)
        m.mass = _ptr(self.mass)
        m.vol = _ptr(self.vol)
        m.selection = _ptr(self.selection) if self.selection is not None else None
        return m


def _mutates(m) -> bool:
    return int(m.material) == 5 or (int(m.material) in (1, 7) and float(m.hardening) == 1.0)


class BatchModel:
    """Several simulations advanced together: per-simulation scalars (the
    ModelArrays in `sims`, whose own per-particle arrays are ignored) and
    per-particle arrays over the concatenated particle arena, simulation c
    owning the next sims[c].n particles."""

    def __init__(self, sims: list[ModelArrays], counts: list[int], mu, lam, yield_stress, mass, vol,
                 selection=None):
        self.sims = list(sims)
        self.counts = [int(k) for k in counts]
        self.mu, self.lam, self.yield_stress, self.mass, self.vol = mu, lam, yield_stress, mass, vol
        self.selection = selection
        self.n = int(sum(self.counts))
        self.offsets = [0]
        for k in self.counts:
            self.offsets.append(self.offsets[-1] + k)
        self._models = None

    @staticmethod
    def concat(models: list[ModelArrays]) -> "BatchModel":
        """A batch from single-simulation models, copying their per-particle
        arrays into the arena in order."""
        cat = lambda ts: torch.cat(ts).contiguous()
        sel = None
        if any(m.selection is not None for m in models):
            sel = cat([m.selection if m.selection is not None else torch.zeros_like(m.mu, dtype=torch.int32)
                       for m in models])
        return BatchModel(models, [int(m.mu.shape[0]) for m in models], cat([m.mu for m in models]),
                          cat([m.lam for m in models]), cat([m.yield_stress for m in models]),
                          cat([m.mass for m in models]), cat([m.vol for m in models]), sel)

    def struct(self) -> Batch:
        arr = (Model * len(self.sims))()
        for i, (m, k) in enumerate(zip(self.sims, self.counts)):
            st = Model()
            st.n_particles = k
            st.n_grid = int(m.n_grid)
            st.material = int(m.material)
            st.grid_lim = float(m.grid_lim)
            st.dx = float(m.dx)
            st.inv_dx = float(m.inv_dx)
            st.rpic_damping = float(m.rpic_damping)
            st.grid_v_damping_scale = float(m.grid_v_damping_scale)
            for d in range(3):
                st.gravity[d] = float(m.gravity[d])
            st.alpha = float(m.alpha)
            st.hardening = float(m.hardening)
            st.xi = float(m.xi)
            st.plastic_viscosity = float(m.plastic_viscosity)
            st.softening = float(m.softening)
            arr[i] = st
        self._models = arr  # keep the host table alive with the struct
        b = Batch()
        b.n_cand = len(self.sims)
        b.n_particles = self.n
        b.models = C.cast(arr, C.c_void_p)
        b.mu, b.lam, b.yield_stress = _ptr(self.mu), _ptr(self.lam), _ptr(self.yield_stress)
        b.mass, b.vol = _ptr(self.mass), _ptr(self.vol)
        b.selection = _ptr(self.selection) if self.selection is not None else None
        return b

    @property
    def mutates(self) -> bool:
        return any(_mutates(m) for m in self.sims)

    def material_state(self) -> torch.Tensor:
        return torch.stack([self.yield_stress, self.mu, self.lam]).contiguous()

    def with_private_params(self) -> "BatchModel":
        return BatchModel(self.sims, self.counts, self.mu.clone(), self.lam.clone(), self.yield_stress.clone(),
                          self.mass, self.vol, self.selection)

    def sizes(self):
        """(total grid nodes, ints of one binning record)."""
        g, nb = C.c_longlong(), C.c_longlong()
        st = self.struct()
        check(_lib.load().empm_batch_sizes(C.byref(st), C.byref(g), C.byref(nb)))
        return int(g.value), int(nb.value)

    def permuted(self, perm) -> "BatchModel":
        g = lambda t: t.index_select(0, perm).contiguous() if t is not None else None
        return BatchModel(self.sims, self.counts, g(self.mu), g(self.lam), g(self.yield_stress), g(self.mass),
                          g(self.vol), g(self.selection))


def collider_array(cols: list[Collider]):
    n = len(cols)
    if n > _lib.EMPM_MAX_COLLIDERS:
        raise ValueError(f"at most {_lib.EMPM_MAX_COLLIDERS} grid post-processes are supported, got {n}")
    arr = (Collider * max(n, 1))()
    for i, c in enumerate(cols):
        arr[i] = c
    return arr, n


def mu_lam(E: torch.Tensor, nu: torch.Tensor, mu: torch.Tensor, lam: torch.Tensor):
    lib = _lib.load()
    check(lib.empm_mu_lam(int(E.shape[0]), _ptr(E), _ptr(nu), _ptr(mu), _ptr(lam), stream_handle(E.device)))


def mu_lam_adj(E, nu, dmu, dlam, dE, dnu):
    lib = _lib.load()
    check(lib.empm_mu_lam_adj(int(E.shape[0]), _ptr(E), _ptr(nu), _ptr(dmu), _ptr(dlam), _ptr(dE), _ptr(dnu),
                              stream_handle(E.device)))


def mass_from_density(rho, vol, mass):
    lib = _lib.load()
    check(lib.empm_mass_from_density(int(rho.shape[0]), _ptr(rho), _ptr(vol), _ptr(mass), stream_handle(rho.device)))


def substep_fwd(ctx: Context, model: ModelArrays, cols: list[Collider], time: float, dt: float,
                s_in: ParticleState, s_out: ParticleState):
    arr, n = collider_array(cols)
    m = model.struct()
    check(ctx.lib.empm_substep_fwd(ctx.handle, C.byref(m), arr, n, C.c_float(time), C.c_float(dt),
                                   C.byref(s_in.desc()), C.byref(s_out.desc()), stream_handle(ctx.device)))


def restart_F(ctx: Context, model: ModelArrays, x: torch.Tensor, target: torch.Tensor, F_trial: torch.Tensor):
    """restart_and_compute_F_C's deformation gradient (empm_restart_deformation): F_trial
    (N, 3, 3) of the particles with selection == 0 becomes I + the grid
    gradient of the mass-weighted displacement target - x.  Positions are
    not written."""
    n = int(model.mu.shape[0])
    for name, t, shape in (("x", x, (n, 3)), ("target", target, (n, 3)), ("F_trial", F_trial, (n, 3, 3))):
        if tuple(t.shape) != shape or t.dtype != F32 or not t.is_contiguous():
            raise RuntimeError(f"restart_F: {name} must be a contiguous float32 tensor of shape {shape}")
    m = model.struct()
    check(ctx.lib.empm_restart_deformation(ctx.handle, C.byref(m), _ptr(x), _ptr(target), _ptr(F_trial),
                                 stream_handle(ctx.device)))


def substep_adj(ctx: Context, model: ModelArrays, cols: list[Collider], time: float, dt: float,
                s_in: ParticleState, d_out: ParticleState, d_in: ParticleState, grads: "ParamGrads", flags=0):
    arr, n = collider_array(cols)
    m = model.struct()
    check(ctx.lib.empm_substep_adj(ctx.handle, C.byref(m), arr, n, C.c_float(time), C.c_float(dt),
                                   C.byref(s_in.desc()), C.byref(d_out.desc()), C.byref(d_in.desc()),
                                   *grads.ptrs(), int(flags), stream_handle(ctx.device)))


@dataclass
class ParamGrads:
    mu: torch.Tensor
    lam: torch.Tensor
    ys: torch.Tensor
    mass: torch.Tensor
    vol: torch.Tensor

    @staticmethod
    def zeros(n, device):
        z = lambda: torch.zeros(n, dtype=F32, device=device)
        return ParamGrads(z(), z(), z(), z(), z())

    def ptrs(self):
        return (_ptr(self.mu), _ptr(self.lam), _ptr(self.ys), _ptr(self.mass), _ptr(self.vol))


@dataclass
class RolloutTape:
    """What a reverse sweep needs.

    Recorded mode (ckpt_every == 1): every substep's input state (with its
    return-mapped F, Fe) and grid, so the reverse sweep recomputes nothing.
    Checkpointed mode (ckpt_every = K > 1): the state every K substeps; each
    segment is replayed (and recorded) before its reverse sweep.  Both keep
    the collider table and time of every substep."""
    n_steps: int
    dt: float
    ckpt_every: int
    ncol: int
    checkpoints: dict = field(default_factory=dict)  # step -> ParticleState
    cols_hist: object = None  # (Collider * (n_steps*ncol))
    # the collider table the reverse sweep masks with, when it is not the
    # one each substep ran with (None: cols_hist); replays always use cols_hist
    adj_cols_hist: object = None
    times_hist: object = None  # (c_double * n_steps)
    perm: object = None  # cell order used internally (None: caller order)
    model: object = None  # ModelArrays in the internal order
    rec_states: list | None = None  # recorded mode: input state of substep k (with Fe)
    rec_grids: list | None = None  # recorded mode: (grid (2, n^3, 4), bins) of substep k


class Records:
    """Device storage for `n` fully recorded substeps: states 0..n (x, v, C,
    F_trial and Fe; Fe of state n unused), the grid of substeps 0..n-1
    ((m v, m) and v_out, float4 per node) and their particle binnings, one
    tensor per field."""

    def __init__(self, n: int, n_particles: int, n_nodes: int, bin_ints: int, device, mstate: bool = False):
        N, G = int(n_particles), int(n_nodes)
        e = lambda *shape: torch.empty(shape, dtype=F32, device=device)
        self.n = int(n)
        self.x, self.v = e(n + 1, N, 3), e(n + 1, N, 3)
        self.C, self.F, self.Fe = e(n + 1, N, 3, 3), e(n + 1, N, 3, 3), e(n + 1, N, 3, 3)
        self.mstate = e(n + 1, 3, N) if mstate else None
        self.grid = e(max(n, 1), 2, G, 4)
        self.bins = torch.empty((max(n, 1), int(bin_ints)), dtype=torch.int32, device=device)

    def state(self, k: int) -> ParticleState:
        return ParticleState(self.x[k], self.v[k], self.C[k], self.F[k], self.Fe[k],
                             self.mstate[k] if self.mstate is not None else None)

    def record(self, k: int):
        """(grid, bins) of substep k, the entry of a grid-record list."""
        return (self.grid[k], self.bins[k])


def tape_bytes(n_particles: int, n_nodes: int, bin_ints: int, n_steps: int, ckpt_every: int,
               mutates: bool = False) -> int:
    """Device bytes a reverse-mode rollout holds for its tape: every
    substep's record (state with Fe 132 B/particle, grid 32 B/node, binning)
    when ckpt_every == 1, else a 96 B/particle checkpoint every ckpt_every
    substeps plus one segment of records; plus the reverse sweep's
    cotangents and parameter gradients.  Materials that rewrite (yield
    stress, mu, lam) add 12 B/particle to every record and checkpoint."""
    N, G, T, K = int(n_particles), int(n_nodes), int(n_steps), max(1, int(ckpt_every))
    ms = 12 if mutates else 0
    rec = lambda k: (k + 1) * N * (132 + ms) + k * (32 * G + 4 * int(bin_ints))
    sweep = 2 * 96 * N + 20 * N
    if K == 1:
        return rec(T) + sweep
    return (-(-T // K)) * (96 + ms) * N + rec(min(K, T)) + sweep


def _states_array(states: list[ParticleState]):
    arr = (Particles * len(states))()
    for i, st in enumerate(states):
        arr[i] = st.desc()
    return arr


def _grid_array(recs: list):
    """empm_grid_rec array from (grid (2, n^3, 4), bins) pairs."""
    arr = (GridRec * max(1, len(recs)))()
    for i, (g, b) in enumerate(recs):
        arr[i] = GridRec(mv_m=_ptr(g[0]), v_out=_ptr(g[1]), bins=_ptr(b))
    return arr


class Rollout:
    """Forward and reverse rollouts of one scene (ModelArrays) or of a batch
    of scenes (BatchModel, on a batch context) on one context.

    ``forward`` advances ``state`` (ParticleState, in place via two ping-pong
    buffers) by ``n_steps`` substeps, returning the copies of x requested by
    ``record`` (a sorted list of substep counts k >= 1: x after k substeps) and,
    when ``tape`` is requested, a RolloutTape for ``backward``.  For a batch
    the state spans the particle arena (simulation c owns rows
    offsets[c] .. offsets[c + 1]) and so do the parameter gradients."""

    def __init__(self, ctx: Context, model, n_particles: int, sort: bool = True):
        self.ctx = ctx
        self.model = model
        self.batch = isinstance(model, BatchModel)
        self.n = int(n_particles)
        self.sort = sort
        dev = ctx.device
        self._buf = [ParticleState.empty(self.n, dev), ParticleState.empty(self.n, dev)]
        if self.batch:
            if self.n != model.n:
                raise ValueError(f"n_particles={self.n} but the batch holds {model.n}")
            self._nodes, self._bin_ints = model.sizes()
        else:
            self._nodes = int(model.n_grid) ** 3
            self._bin_ints = int(_lib.load().empm_bin_record_ints(int(model.n_grid), self.n))

    # ------------------------------------------------------------ C-ABI dispatch
    def _fwd(self, m, *args):
        f = self.ctx.lib.empm_batch_rollout_fwd if self.batch else self.ctx.lib.empm_rollout_fwd
        check(f(self.ctx.handle, C.byref(m), *args))

    def _replay(self, m, *args):
        f = self.ctx.lib.empm_batch_rollout_fwd_replay if self.batch else self.ctx.lib.empm_rollout_fwd_replay
        check(f(self.ctx.handle, C.byref(m), *args))

    def _adj(self, m, *args):
        f = self.ctx.lib.empm_batch_rollout_adj if self.batch else self.ctx.lib.empm_rollout_adj
        check(f(self.ctx.handle, C.byref(m), *args))

    # ------------------------------------------------------------ ordering
    def cell_order(self, x: torch.Tensor) -> torch.Tensor:
        perm = torch.empty(self.n, dtype=torch.int32, device=x.device)
        if self.batch:
            b = self.model.struct()
            check(self.ctx.lib.empm_batch_sort_by_cell(self.ctx.handle, C.byref(b), _ptr(x.contiguous()), _ptr(perm),
                                                       stream_handle(x.device)))
        else:
            check(self.ctx.lib.empm_sort_by_cell(self.ctx.handle, self.n, _ptr(x.contiguous()),
                                                 C.c_float(self.model.inv_dx), _ptr(perm), stream_handle(x.device)))
        return perm.long()

    @staticmethod
    def _gather(perm, t):
        return t.index_select(0, perm).contiguous() if t is not None else None

    @staticmethod
    def _scatter(perm, t):
        out = torch.empty_like(t)
        out.index_copy_(0, perm, t)
        return out

    def _permuted_model(self, perm):
        m = self.model
        if self.batch:
            return m.permuted(perm)
        g = lambda t: self._gather(perm, t)
        return ModelArrays(material=m.material, n_grid=m.n_grid, grid_lim=m.grid_lim, dx=m.dx, inv_dx=m.inv_dx,
                           mu=g(m.mu), lam=g(m.lam), yield_stress=g(m.yield_stress), mass=g(m.mass), vol=g(m.vol),
                           selection=g(m.selection), rpic_damping=m.rpic_damping,
                           grid_v_damping_scale=m.grid_v_damping_scale, gravity=m.gravity, alpha=m.alpha,
                           hardening=m.hardening, xi=m.xi, plastic_viscosity=m.plastic_viscosity,
                           softening=m.softening)

    def _perm_state(self, perm, s: ParticleState, inverse=False):
        f = self._scatter if inverse else self._gather
        return ParticleState(f(perm, s.x), f(perm, s.v), f(perm, s.C), f(perm, s.F))

    def autotune(self, state: ParticleState, cols: list[Collider], dt: float, n_steps: int = 40,
                 backward: bool = True, cache: dict | None = None, force: bool = False, **kw) -> dict:
        """Tune this rollout's context (empm_amd.autotune) on n_steps substeps
        of this workload from `state` (forward, and with `backward` the
        reverse sweep of a recorded tape); the result is cached per workload
        signature and left set on the context."""
        from . import autotune as at

        dev = self.ctx.device
        m = self.model
        sims = m.sims if isinstance(m, BatchModel) else [m]
        key = at.workload_key(dev, [s.n_grid for s in sims], self.n, [s.material for s in sims], len(sims),
                              backward)
        cols = list(cols)

        def run():
            final, _, _, tape = self.forward(state, list(cols), 0.0, dt, n_steps, tape=backward, ckpt_every=1)
            if backward:
                d = ParticleState.empty(self.n, dev)
                d.zero_()
                d.x.fill_(1e-3)
                self.backward(tape, d, flags=0)

        return at.tune(self.ctx, run, key=key, cache=cache, force=force, **kw)

    def forward(self, state: ParticleState, cols: list[Collider], time: float, dt: float, n_steps: int,
                record=(), tape: bool = False, ckpt_every: int = 1, cols_out: list | None = None,
                check: bool = False):
        """Returns (final_state, final_time, recorded_x list, tape or None).
        `cols` (ctypes Collider structs) are advanced in place like the
        reference's modify_bc; `cols_out` receives the advanced table.  With
        sort=True the rollout runs on a cell-ordered copy of the particles
        and every output is returned in the caller's order.  With check, the
        device's failure flags are read once at the end (one
        synchronisation): a non-finite state raises EmpmError
        (EMPM_E_NONFINITE), a CFL violation warns.  Without it nothing
        synchronises; the flags stay set for the next read."""
        out = self._forward_sorted(state, cols, time, dt, n_steps, record, tape, ckpt_every, cols_out)
        if check:
            self.ctx.check_status()
        return out

    def _forward_sorted(self, state, cols, time, dt, n_steps, record, tape, ckpt_every, cols_out):
        if not self.sort or self.n == 0:
            return self._forward_core(self.model, state, cols, time, dt, n_steps, record, tape, ckpt_every,
                                      cols_out)
        perm = self.cell_order(state.x)
        pm = self._permuted_model(perm)
        # the cell-ordered copy lives in a buffer of the rollout's own, so that
        # repeated calls present the same pointers (graph replay)
        if getattr(self, "_sorted", None) is None:
            self._sorted = ParticleState.empty(self.n, self.ctx.device)
        s_sorted = self._sorted
        for f in ("x", "v", "C", "F"):
            torch.index_select(getattr(state, f), 0, perm, out=getattr(s_sorted, f))
        final, t1, xs, tp = self._forward_core(pm, s_sorted, cols, time, dt, n_steps, record, tape, ckpt_every,
                                               cols_out)
        sims = pm.sims if self.batch else [pm]
        if any(_mutates(s) for s in sims):  # in-place updates go back to the caller
            for a, b in ((self.model.mu, pm.mu), (self.model.lam, pm.lam), (self.model.yield_stress, pm.yield_stress)):
                a.index_copy_(0, perm, b)
        if tp is not None:
            tp.perm = perm
            tp.model = pm
        return self._perm_state(perm, final, inverse=True), t1, [self._scatter(perm, x) for x in xs], tp

    def _forward_core(self, model, state, cols, time, dt, n_steps, record, tape, ckpt_every, cols_out):
        arr, ncol = collider_array(cols)
        m = model.struct()
        t = C.c_double(time)
        rec = sorted(set(int(k) for k in record if 1 <= int(k) <= n_steps))
        tp = None
        if tape:
            tp = RolloutTape(n_steps=n_steps, dt=dt, ckpt_every=max(1, int(ckpt_every)), ncol=ncol)
            tp.cols_hist = (Collider * max(1, n_steps * ncol))()
            tp.times_hist = (C.c_double * max(1, n_steps))()
        stream = stream_handle(self.ctx.device)
        mut = tape and model.mutates
        if tape and tp.ckpt_every == 1:
            # recorded mode: every substep writes its own output slot, Fe and
            # grid (and the material state it starts from, for materials that
            # rewrite it)
            R = Records(n_steps, self.n, self._nodes, self._bin_ints, self.ctx.device, mstate=mut)
            R.state(0).copy_(state)
            states = [R.state(k) for k in range(n_steps + 1)]
            grids = [R.record(k) for k in range(n_steps)]
            self._fwd(m, arr, ncol, C.byref(t), C.c_double(dt), n_steps, _states_array(states), _grid_array(grids),
                      tp.cols_hist, tp.times_hist, stream)
            tp.rec_states = states[:n_steps]
            tp.rec_grids = grids
            tp.checkpoints[0] = states[0]
            if cols_out is not None:
                cols_out[:] = [arr[i] for i in range(ncol)]
            return states[n_steps], t.value, [states[k].x.clone() for k in rec], tp
        stops = set(rec)
        if tape:
            stops.update(range(tp.ckpt_every, n_steps, tp.ckpt_every))
        stops.add(n_steps)
        stops = sorted(s for s in stops if s >= 1)
        cur = state
        k = 0
        out_x = []
        if tape:
            tp.checkpoints[0] = state.clone()
            if mut:
                tp.checkpoints[0].mstate = model.material_state()
        for stop in stops:
            n = stop - k
            if n <= 0:
                continue
            # ping-pong: cur -> buf[0] -> buf[1] -> ...
            seq = [cur]
            for _ in range(n):
                seq.append(self._buf[0] if seq[-1] is not self._buf[0] else self._buf[1])
            sarr = _states_array(seq)
            ch = C.cast(C.byref(tp.cols_hist, k * ncol * C.sizeof(Collider)), C.POINTER(Collider)) if tape else None
            th = C.cast(C.byref(tp.times_hist, k * C.sizeof(C.c_double)), C.POINTER(C.c_double)) if tape else None
            self._fwd(m, arr, ncol, C.byref(t), C.c_double(dt), n, sarr, None, ch, th, stream)
            cur = seq[-1]
            k = stop
            if k in rec:
                out_x.append(cur.x.clone())
            if tape and k < n_steps and k % tp.ckpt_every == 0:
                tp.checkpoints[k] = cur.clone()
                if mut:
                    tp.checkpoints[k].mstate = model.material_state()
        if cols_out is not None:
            cols_out[:] = [arr[i] for i in range(ncol)]
        return cur, t.value, out_x, tp

    def backward(self, tp: RolloutTape, d_final: ParticleState, seeds: dict | None = None, flags: int = 0,
                 grads: ParamGrads | None = None):
        """Reverse sweep.  d_final: dL/d(state n_steps).  seeds: {k: ParticleState
        cotangent of state k} added when the sweep reaches state k (0 < k < n).
        Returns (dL/d(state 0) as ParticleState, ParamGrads); all in the
        caller's particle order.  `grads`, if given, is accumulated into."""
        seeds = seeds or {}
        if tp.perm is None:
            return self._backward_core(self.model, tp, d_final, seeds, flags, grads)
        perm = tp.perm
        d_sorted = self._perm_state(perm, d_final)
        seeds_sorted = {k: self._perm_state(perm, v) for k, v in seeds.items()}
        d0, g = self._backward_core(tp.model, tp, d_sorted, seeds_sorted, flags, None)
        gu = ParamGrads(*(self._scatter(perm, t) for t in (g.mu, g.lam, g.ys, g.mass, g.vol)))
        if grads is not None:
            for a, b in ((grads.mu, gu.mu), (grads.lam, gu.lam), (grads.ys, gu.ys), (grads.mass, gu.mass),
                         (grads.vol, gu.vol)):
                a.add_(b)
            gu = grads
        return self._perm_state(perm, d0, inverse=True), gu

    def _backward_core(self, model, tp, d_final, seeds, flags, grads):
        dev = self.ctx.device
        n = self.n
        if grads is None:
            grads = ParamGrads.zeros(n, dev)
        m = model.struct()
        stream = stream_handle(dev)
        ncol = tp.ncol
        adj = [ParticleState.empty(n, dev), ParticleState.empty(n, dev)]
        adj[0].copy_(d_final)

        def add_seed(k):
            sd = seeds[k]
            adj[0].x.add_(sd.x)
            adj[0].v.add_(sd.v)
            adj[0].C.add_(sd.C)
            adj[0].F.add_(sd.F)

        def reverse(s0, k_hi, states, grids):
            """Reverse over substeps [s0, k_hi) whose input states / grids are
            states[k - s0] / grids[k - s0], in chunks split at seed points."""
            nonlocal adj
            k = k_hi
            while k > s0:
                if k in seeds and k != tp.n_steps:
                    add_seed(k)
                j = max([s for s in seeds if s0 < s < k] + [s0])
                cnt = k - j
                sub = _states_array(states[j - s0:k - s0])
                garr = _grid_array(grids[j - s0:k - s0])
                adj_cols = tp.adj_cols_hist if tp.adj_cols_hist is not None else tp.cols_hist
                chj = C.cast(C.byref(adj_cols, j * ncol * C.sizeof(Collider)), C.POINTER(Collider))
                thj = C.cast(C.byref(tp.times_hist, j * C.sizeof(C.c_double)), C.POINTER(C.c_double))
                self._adj(m, chj, ncol, thj, C.c_double(tp.dt), cnt, sub, garr, _states_array(adj), *grads.ptrs(),
                          int(flags), stream)
                if cnt % 2 == 1:
                    adj = [adj[1], adj[0]]
                k = j

        if tp.rec_states is not None:
            reverse(0, tp.n_steps, tp.rec_states, tp.rec_grids)
        else:
            K = tp.ckpt_every
            mut = model.mutates
            if mut:
                # replays rewrite the model's parameter arrays: give them their own
                model = model.with_private_params()
                m = model.struct()
            seg = Records(min(K, tp.n_steps), n, self._nodes, self._bin_ints, dev, mstate=mut)
            seg_starts = sorted(tp.checkpoints.keys())
            k_hi = tp.n_steps
            for s0 in reversed(seg_starts):
                if s0 >= k_hi:
                    continue
                L = k_hi - s0
                seg.state(0).copy_(tp.checkpoints[s0])
                states = [seg.state(i) for i in range(L + 1)]
                grids = [seg.record(i) for i in range(L)]
                ch = C.cast(C.byref(tp.cols_hist, s0 * ncol * C.sizeof(Collider)), C.POINTER(Collider))
                th = C.cast(C.byref(tp.times_hist, s0 * C.sizeof(C.c_double)), C.POINTER(C.c_double))
                # replay the segment, recording every substep's Fe and grid
                self._replay(m, ch, ncol, th, C.c_double(tp.dt), L, _states_array(states), _grid_array(grids),
                             stream)
                reverse(s0, k_hi, states[:L], grids)
                k_hi = s0
            # the replays repeat the forward's failure flags: clear them
            self.ctx.read_status()
        if 0 in seeds:
            add_seed(0)
        return adj[0], grads
