"""Batched evaluation of a CMA-ES population (or a gradient batch) of
material-parameter candidates.

The reference evaluates a population one candidate at a time: CMA-ES calls
``error_func_cma`` per candidate, which builds a solver
(``initialize_mpm``) and rolls it out (``advance_to``)
(phys_sim/engine/trainer_warp_mpm.py:598-663).  Here the whole population is
one batch: each candidate's particles own a contiguous range of one particle
arena, each its own grid in one node arena, and every stage of every
substep is one launch over all of them (empm_batch_rollout_*).  The
candidates may differ in particle count, shape, grid and material; they
share dt, the substep count and the collider table.

``fitness`` is the per-candidate position MSE of the final state against a
target rollout (by default: the same objects with the nominal parameters),
``fitness_and_grad`` adds d(fitness)/d(log E, nu, log Y) per candidate
through the batched reverse sweep.  ``*_sequential`` runs the same work as
one plain rollout per candidate, one after the other: the baseline the
batch is measured against, and the reference for its parity tests.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import runtime as rt
from .diff_simulator.warp_mpm.mpm_data_structure import friction_alpha

F32 = torch.float32


@dataclass
class Candidate:
    """One candidate's material parameters."""
    E: float
    nu: float
    yield_stress: float
    friction_angle: float = 30.0

    @staticmethod
    def of_scene(sc) -> "Candidate":
        return Candidate(float(sc.params["E"]), float(sc.params["nu"]), float(sc.params["yield_stress"]),
                         float(sc.friction_angle))


def surface_colliders(dicts) -> list:
    """empm_collider structs of reference-style surface collider dicts
    ({"kind": "surface", "point", "normal"}: sticky planes, always active)."""
    out = []
    for c in dicts:
        if c.get("kind", "surface") != "surface":
            raise ValueError(f"only surface colliders are supported here, got {c.get('kind')}")
        g = _lib.Collider()
        g.kind = 0
        g.surface_type = {"sticky": 0, "slip": 1, "cut": 11}.get(c.get("surface", "sticky"), 2)
        g.friction = float(c.get("friction", 0.0))
        g.start_time, g.end_time = float(c.get("start_time", 0.0)), float(c.get("end_time", 999.0))
        nrm = np.asarray(c["normal"], np.float64)
        nrm = nrm / np.sqrt(float(np.sum(nrm * nrm)))
        for i in range(3):
            g.point[i] = float(c["point"][i])
            g.normal[i] = float(nrm[i])
        out.append(g)
    return out


class PopulationEvaluator:
    """The objects of a population (scenes.Scene list: geometry, material id,
    mass and volume) resident on one device, evaluated for any list of
    Candidate parameters with one batched rollout."""

    def __init__(self, scenes, device="cuda:0", targets: torch.Tensor | None = None):
        if not scenes:
            raise ValueError("empty population")
        s0 = scenes[0]
        for sc in scenes:
            if (sc.dt, sc.n_steps, sc.grid_lim) != (s0.dt, s0.n_steps, s0.grid_lim):
                raise ValueError("the candidates of one batch share dt, the substep count and grid_lim")
            if sc.colliders != s0.colliders:
                raise ValueError("the candidates of one batch share the collider table")
        self.scenes = list(scenes)
        self.device = torch.device(device)
        self.n_cand = len(scenes)
        self.counts = [sc.n for sc in scenes]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.dt, self.n_steps = float(s0.dt), int(s0.n_steps)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(np.concatenate(a), dtype=np.float32), device=self.device)
        self.x0 = t([sc.x for sc in scenes])
        self.v0 = t([sc.params["v0"] for sc in scenes])
        self.vol = t([sc.vol for sc in scenes])
        self.mass = t([sc.vol * sc.rho for sc in scenes])
        self.cand_id = torch.repeat_interleave(torch.arange(self.n_cand, device=self.device),
                                               torch.as_tensor(self.counts, device=self.device))
        self.counts_t = torch.as_tensor(self.counts, dtype=torch.float64, device=self.device)
        self._ends = torch.as_tensor(self.offsets[1:] - 1, device=self.device)
        self.cols = surface_colliders(s0.colliders)
        self.ctx = rt.Context(max(sc.n_grid for sc in scenes), self.n, self.device, max_cand=self.n_cand)
        self._solo_ctx = None
        self.targets = targets
        self.failed: list[int] = []

    # ------------------------------------------------------------ arena set-up
    def _per_particle(self, vals):
        v = torch.as_tensor(np.asarray(vals, np.float32), device=self.device)
        return torch.repeat_interleave(v, torch.as_tensor(self.counts, device=self.device)).contiguous()

    def model(self, cands: list[Candidate]) -> tuple[rt.BatchModel, torch.Tensor, torch.Tensor]:
        """(BatchModel, E, nu) for these candidates: mu, lam from (E, nu) by
        the library's kernel, yield stress per particle, the Drucker-Prager
        alpha of each candidate's friction angle."""
        if len(cands) != self.n_cand:
            raise ValueError(f"{len(cands)} candidates for a population of {self.n_cand}")
        E = self._per_particle([c.E for c in cands])
        nu = self._per_particle([c.nu for c in cands])
        ys = self._per_particle([c.yield_stress for c in cands])
        mu, lam = torch.empty_like(E), torch.empty_like(E)
        rt.mu_lam(E, nu, mu, lam)
        sims = []
        for c, (sc, cd) in enumerate(zip(self.scenes, cands)):
            a, b = int(self.offsets[c]), int(self.offsets[c + 1])
            sims.append(rt.ModelArrays(
                material=sc.material, n_grid=sc.n_grid, grid_lim=sc.grid_lim,
                dx=float(np.float32(sc.grid_lim / sc.n_grid)), inv_dx=float(np.float32(sc.n_grid / sc.grid_lim)),
                mu=mu[a:b], lam=lam[a:b], yield_stress=ys[a:b], mass=self.mass[a:b], vol=self.vol[a:b],
                rpic_damping=sc.rpic_damping, grid_v_damping_scale=sc.grid_v_damping_scale, gravity=sc.gravity,
                alpha=friction_alpha(cd.friction_angle), softening=sc.softening))
        return rt.BatchModel(sims, self.counts, mu, lam, ys, self.mass, self.vol), E, nu

    def initial_state(self) -> rt.ParticleState:
        N = self.n
        return rt.ParticleState(self.x0.clone(), self.v0.clone(), torch.zeros((N, 3, 3), device=self.device),
                                torch.eye(3, device=self.device).repeat(N, 1, 1).contiguous())

    def _seg_sum(self, v: torch.Tensor) -> torch.Tensor:
        """Per-candidate sums of a per-particle array (float64): differences of
        an inclusive scan at the candidates' range ends (candidates own
        contiguous ranges; a scatter-add into n_cand slots would serialise
        on a few addresses)."""
        cs = torch.cumsum(v.double(), 0)
        ends = cs[self._ends]
        return torch.diff(ends, prepend=torch.zeros(1, dtype=torch.float64, device=self.device))

    # ------------------------------------------------------------ batched
    def rollout(self, cands: list[Candidate]) -> rt.ParticleState:
        """Final state of every candidate's rollout (arena order); the
        candidates whose rollout produced a non-finite state are in
        self.failed afterwards."""
        bm, _, _ = self.model(cands)
        ro = rt.Rollout(self.ctx, bm, self.n)
        final, _, _, _ = ro.forward(self.initial_state(), list(self.cols), 0.0, self.dt, self.n_steps, check=False)
        self._read_failures()
        return final

    def _read_failures(self):
        """Per-candidate failure flags of the last batched forward (one read):
        a candidate whose parameters drive its rollout to a non-finite state
        gets an infinite fitness instead of failing the whole population."""
        _, st = self.ctx.read_status()
        self.failed = [c for c in range(self.n_cand) if st[c] & _lib.EMPM_STATUS_NONFINITE]

    def _mask_failed(self, fit: torch.Tensor) -> torch.Tensor:
        if self.failed:
            fit = fit.clone()
            fit[torch.as_tensor(self.failed, device=fit.device)] = float("inf")
        return fit

    def set_targets(self, cands: list[Candidate]):
        """Target final positions: the rollout of these candidates."""
        self.targets = self.rollout(cands).x.clone()

    def fitness(self, cands: list[Candidate]) -> torch.Tensor:
        """Per-candidate mean over its particles of |x_T - x*_T|^2 (float64)."""
        d = self.rollout(cands).x - self.targets
        return self._mask_failed(self._seg_sum((d * d).sum(1)) / self.counts_t)

    def tape_bytes(self, ckpt_every: int) -> int:
        bm, _, _ = self.model([Candidate(1e5, 0.3, 1e3)] * self.n_cand)
        G, bins = bm.sizes()
        return rt.tape_bytes(self.n, G, bins, self.n_steps, ckpt_every)

    def fitness_and_grad(self, cands: list[Candidate], ckpt_every: int = 1):
        """(fitness (n_cand,), d fitness / d(log E, nu, log Y) (n_cand, 3)),
        float64, through one batched forward recording its tape and one
        batched reverse sweep of the summed fitness (candidates do not
        interact, so each candidate's gradient is that of its own fitness)."""
        bm, E, nu = self.model(cands)
        ys0 = bm.yield_stress.clone()  # plasticine rewrites it during the rollout
        ro = rt.Rollout(self.ctx, bm, self.n)
        final, _, _, tape = ro.forward(self.initial_state(), list(self.cols), 0.0, self.dt, self.n_steps,
                                       tape=True, ckpt_every=ckpt_every, check=False)
        self._read_failures()
        diff = final.x - self.targets
        fit = self._mask_failed(self._seg_sum((diff * diff).sum(1)) / self.counts_t)
        d = rt.ParticleState.empty(self.n, self.device)
        d.zero_()
        scale = (2.0 / self.counts_t.float())[self.cand_id]
        d.x.copy_(diff * scale[:, None])
        _, g = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
        dE, dnu = torch.zeros_like(E), torch.zeros_like(E)
        rt.mu_lam_adj(E, nu, g.mu, g.lam, dE, dnu)
        grad = torch.stack([self._seg_sum(dE * E), self._seg_sum(dnu), self._seg_sum(g.ys * ys0)], 1)
        return fit, grad

    # ------------------------------------------------------------ tuning
    def autotune(self, cands: list[Candidate], backward: bool = False, n_steps: int = 20, sequential: bool = False,
                 cache: dict | None = None, force: bool = False, **kw) -> dict:
        """Tune the launch parameters (empm_amd.autotune) of the batched
        evaluation (fitness, or with backward fitness_and_grad) on n_steps
        substeps, or with `sequential` those of the one-at-a-time context on
        the first four candidates; the targets must be set.  The result is
        cached per workload signature and left set on the context."""
        from . import autotune as at

        if self.targets is None:
            raise ValueError("autotune: set_targets first")
        saved = self.n_steps
        self.n_steps = int(n_steps)
        try:
            grids = [sc.n_grid for sc in self.scenes]
            mats = [sc.material for sc in self.scenes]
            if sequential:
                sub = list(range(min(4, self.n_cand)))
                bm, _, _ = self.model(cands)
                self._solo(0, bm)  # creates the one-at-a-time context
                ctx = self._solo_ctx

                def run():
                    for c in sub:
                        ro, s0, a, b = self._solo(c, bm)
                        final, _, _, tape = ro.forward(s0, list(self.cols), 0.0, self.dt, self.n_steps,
                                                       tape=backward)
                        if backward:
                            d = rt.ParticleState.empty(b - a, self.device)
                            d.zero_()
                            d.x.copy_(final.x - self.targets[a:b])
                            ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)

                key = at.workload_key(self.device, grids, max(self.counts), mats, 1, backward) + "|one-at-a-time"
            else:
                ctx = self.ctx

                def run():
                    if backward:
                        self.fitness_and_grad(cands)
                    else:
                        self.fitness(cands)

                key = at.workload_key(self.device, grids, self.n, mats, self.n_cand, backward)
            return at.tune(ctx, run, key=key, cache=cache, force=force, **kw)
        finally:
            self.n_steps = saved

    # ------------------------------------------------------------ one at a time
    def _solo(self, c: int, bm: rt.BatchModel):
        if self._solo_ctx is None:
            self._solo_ctx = rt.Context(max(sc.n_grid for sc in self.scenes), max(self.counts), self.device)
        a, b = int(self.offsets[c]), int(self.offsets[c + 1])
        m = bm.sims[c]
        s0 = rt.ParticleState(self.x0[a:b].clone(), self.v0[a:b].clone(),
                              torch.zeros((b - a, 3, 3), device=self.device),
                              torch.eye(3, device=self.device).repeat(b - a, 1, 1).contiguous())
        return rt.Rollout(self._solo_ctx, m, b - a), s0, a, b

    def fitness_sequential(self, cands: list[Candidate]) -> torch.Tensor:
        """fitness() as one plain rollout per candidate, one after another."""
        bm, _, _ = self.model(cands)
        out = []
        for c in range(self.n_cand):
            ro, s0, a, b = self._solo(c, bm)
            final, _, _, _ = ro.forward(s0, list(self.cols), 0.0, self.dt, self.n_steps)
            d = final.x - self.targets[a:b]
            out.append((d * d).sum(1).double().mean())
        return torch.stack(out)

    def fitness_and_grad_sequential(self, cands: list[Candidate], ckpt_every: int = 1):
        """fitness_and_grad() as one plain rollout + reverse sweep per
        candidate, one after another."""
        bm, E, nu = self.model(cands)
        ys0 = bm.yield_stress.clone()
        fits, grads = [], []
        for c in range(self.n_cand):
            ro, s0, a, b = self._solo(c, bm)
            final, _, _, tape = ro.forward(s0, list(self.cols), 0.0, self.dt, self.n_steps, tape=True,
                                           ckpt_every=ckpt_every)
            diff = final.x - self.targets[a:b]
            fits.append((diff * diff).sum(1).double().mean())
            d = rt.ParticleState.empty(b - a, self.device)
            d.zero_()
            d.x.copy_(diff * (2.0 / (b - a)))
            _, g = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
            dE, dnu = torch.zeros(b - a, device=self.device), torch.zeros(b - a, device=self.device)
            rt.mu_lam_adj(E[a:b], nu[a:b], g.mu, g.lam, dE, dnu)
            grads.append(torch.stack([(dE * E[a:b]).double().sum(), dnu.double().sum(),
                                      (g.ys * ys0[a:b]).double().sum()]))
        return torch.stack(fits), torch.stack(grads)
