"""The four torch.autograd.Functions of the reference
(phys_sim/diff_simulator/warp_mpm/interface.py:20-986) with their exact
positional signatures and backward return tuples, backed by libempm's
hand-written adjoint instead of a Warp Tape.

Memory: the reference keeps every substep's full state alive for its Tape
(interface.py:137-145).  Here, with ``mpm_solver.checkpoint_every`` = 1
(the default) the forward records every substep's input state, Fe and grid
(132 B per particle + 32 B per node per substep) and the backward recomputes
nothing; with checkpoint_every = K > 1 it keeps a state every K substeps and
the backward replays each segment before reversing it.

Adjoint mode (``mpm_solver.adjoint_mode``):
  "reference_compat" (default) reproduces the reference Tape's gradient
      conventions: the grid-mass adjoint is dropped
      (mpm_data_structure.py:100-105 + warp_utils.py:116-120), the Trail
      sums its per-frame output gradients onto the last state
      (interface.py:176-197), and the collider masks of the reverse sweep
      are those of the collider parameters as they stand when backward runs:
      the reference's Tape records each grid_postprocess launch with the
      solver's collider_params object itself (mpm_solver_diff.py:265-278),
      whose point modify_bc moves on after the launch (:860-866), so its
      replay reads the parameters of the object's last state (times are
      recorded by value per launch);
  "exact" is the derivative of the discrete rollout (what finite
      differences measure): grid-mass adjoint kept, each frame's gradient
      injected at the substep it was taken from.
"""
from __future__ import annotations

from typing import Optional

import torch
import torch.autograd as autograd
from torch import Tensor

from ... import runtime as rt
from ..._lib import EMPM_ADJ_DROP_GRID_MASS, EMPM_ADJ_EXACT, Collider


def _reset_or_restart(mpm_solver, mpm_state, mpm_model, init_pos, init_velocity, static_pos, device,
                      requires_grad):
    """interface.py:47-67,281-301: reset the state at init_pos, or, with
    static_pos, reset it at static_pos and restart it towards init_pos
    (restart_and_compute_F_C: F_trial from the displacement, x = init_pos)."""
    if static_pos is None:
        mpm_state.reset_state(init_pos.clone(), None, init_velocity, device=device, requires_grad=requires_grad)
        return
    mpm_state.reset_state(static_pos.clone(), None, init_velocity, device=device, requires_grad=requires_grad)
    mpm_solver.restart_and_compute_F_C(mpm_model, mpm_state, init_pos.clone().detach().contiguous(), device)


def _flags(solver):
    return EMPM_ADJ_EXACT if getattr(solver, "adjoint_mode", "reference_compat") == "exact" else EMPM_ADJ_DROP_GRID_MASS


def _compat_colliders(ctx):
    """reference_compat: the reverse sweep masks every substep with the
    collider table as it stands now (see the module docstring); the replays
    of a checkpointed tape keep the table each substep ran with."""
    solver, tp = ctx.mpm_solver, ctx.tape
    if _flags(solver) != EMPM_ADJ_DROP_GRID_MASS or tp is None or tp.ncol == 0:
        return
    cur = solver.collider_table()
    if len(cur) != tp.ncol:
        return
    arr = (Collider * (tp.n_steps * tp.ncol))()
    for k in range(tp.n_steps):
        for c in range(tp.ncol):
            arr[k * tp.ncol + c] = cur[c]
    tp.adj_cols_hist = arr


def _set_E_nu(solver, model, E, nu, scale_scalar_E):
    """interface.py:68-82: a 0-d E/nu enters as a float (the Trail multiplies
    a scalar E by 1e6) and its gradient is later the mean over particles."""
    agg_E = E.ndim == 0
    agg_nu = nu.ndim == 0
    E_inp = E.item() * (1e6 if scale_scalar_E else 1.0) if agg_E else E.detach()
    nu_inp = nu.item() if agg_nu else nu.detach()
    solver.set_E_nu(model, E_inp, nu_inp)
    return agg_E, agg_nu


def _run_forward(solver, state, model, dt, n_steps, record, want_tape, step_callback=True):
    """Advance `state` by n_steps substeps through libempm; returns
    (final ParticleState (a library buffer), recorded x list, tape, rollout)."""
    solver._check_modifiers()
    n = state.n_particles
    ctx = solver.context(model.n_grid, n, state.particle_x.device)
    arrays = solver.model_arrays(model, state)
    ro = rt.Rollout(ctx, arrays, n)
    s0 = rt.ParticleState(state.particle_x.contiguous(), state.particle_v.contiguous(),
                          state.particle_C.contiguous(), state.particle_F_trial.contiguous())
    cb = solver.step_callback if step_callback else None
    K = int(getattr(solver, "checkpoint_every", 1))
    if cb is None:
        cols = solver.collider_table()
        cols_out = []
        # the user-facing API reads the device's failure flags once per call
        final, t1, xs, tape = ro.forward(s0, cols, solver.time, dt, n_steps, record=record, tape=want_tape,
                                         ckpt_every=K, cols_out=cols_out, check=True)
        solver.sync_colliders_from(cols_out)
        solver.time = t1
        return final, xs, tape, ro
    # a per-substep Python callback may mutate colliders: one call per substep,
    # in the caller's particle order so that the per-substep tapes merge
    ro = rt.Rollout(ctx, arrays, n, sort=False)
    xs = []
    tapes = []
    cur = s0
    rec = set(record)
    for k in range(n_steps):
        cb(k)
        cols = solver.collider_table()
        cols_out = []
        nxt, t1, _, tp = ro.forward(cur, cols, solver.time, dt, 1, tape=want_tape, ckpt_every=1, cols_out=cols_out,
                                    check=(k == n_steps - 1))
        solver.sync_colliders_from(cols_out)
        solver.time = t1
        if want_tape:
            tapes.append(tp)
        cur = nxt.clone() if want_tape else nxt
        if (k + 1) in rec:
            xs.append(cur.x.clone())
    tape = _merge_tapes(tapes, dt) if want_tape else None
    return cur, xs, tape, ro


def _merge_tapes(tapes, dt):
    import ctypes as C

    from ..._lib import Collider

    n = len(tapes)
    ncol = tapes[0].ncol if tapes else 0
    tp = rt.RolloutTape(n_steps=n, dt=dt, ckpt_every=1, ncol=ncol)
    tp.cols_hist = (Collider * max(1, n * ncol))()
    tp.times_hist = (C.c_double * max(1, n))()
    tp.rec_states, tp.rec_grids = [], []
    for k, t in enumerate(tapes):
        tp.checkpoints[k] = t.checkpoints[0]
        tp.rec_states.append(t.rec_states[0])
        tp.rec_grids.append(t.rec_grids[0])
        for j in range(ncol):
            tp.cols_hist[k * ncol + j] = t.cols_hist[j]
        tp.times_hist[k] = t.times_hist[0]
    return tp


def _zeros_like_state(n, dev):
    return rt.ParticleState.empty(n, dev)


def _param_grads(model, grads, agg_E, agg_nu, n):
    dE = torch.zeros_like(model.E)
    dnu = torch.zeros_like(model.nu)
    rt.mu_lam_adj(model.E, model.nu, grads.mu, grads.lam, dE, dnu)
    E_grad = dE.sum() / n if agg_E else dE.clone()
    nu_grad = dnu.sum() / n if agg_nu else dnu.clone()
    return E_grad, nu_grad


class MPMDifferentiableSimulationTrail(autograd.Function):
    """interface.py:20-251.  Returns the positions after substeps
    1, 1+substep, 1+2*substep, ... (next_state_list[::substep])."""

    @staticmethod
    def forward(ctx, mpm_solver, mpm_state, mpm_model, substep: int, substep_size: float, num_substeps: int,
                init_pos: Tensor, init_velocity: Tensor, E: Tensor, nu: Tensor,
                particle_density: Optional[Tensor] = None, density_change_mask: Optional[Tensor] = None,
                static_pos: Optional[Tensor] = None, device: str = "cuda:0", requires_grad: bool = True,
                extra_no_grad_steps: int = 0):
        _reset_or_restart(mpm_solver, mpm_state, mpm_model, init_pos, init_velocity, static_pos, device,
                          requires_grad)
        agg_E, agg_nu = _set_E_nu(mpm_solver, mpm_model, E, nu, scale_scalar_E=True)
        mpm_state.reset_density(tensor_density=particle_density, selection_mask=density_change_mask, device=device,
                                requires_grad=requires_grad)
        state = mpm_state
        if extra_no_grad_steps > 0:
            state.reset_density(None, update_mass=True)
            mpm_solver.prepare_mu_lam(mpm_model, state)
            final, _, _, _ = _run_forward(mpm_solver, state, mpm_model, substep_size, extra_no_grad_steps, (), False,
                                          step_callback=False)
            state = _state_from(state, final)
        state.reset_density(None, update_mass=True)
        mpm_solver.prepare_mu_lam(mpm_model, state)
        record = list(range(1, num_substeps + 1, substep))
        final, xs, tape, ro = _run_forward(mpm_solver, state, mpm_model, substep_size, num_substeps, record,
                                           requires_grad)
        ctx.mpm_solver, ctx.mpm_model, ctx.state, ctx.tape, ctx.ro = mpm_solver, mpm_model, state, tape, ro
        ctx.agg_E, ctx.agg_nu, ctx.record = agg_E, agg_nu, record
        ctx.n = state.n_particles
        ctx.warmup = int(extra_no_grad_steps)
        ctx.save_for_backward(density_change_mask)
        return torch.stack(xs, dim=0)

    @staticmethod
    def backward(ctx, out_pos_grad: Tensor):
        if ctx.tape is None:
            raise RuntimeError("forward ran with requires_grad=False")
        solver, model, n = ctx.mpm_solver, ctx.mpm_model, ctx.n
        density_change_mask = ctx.saved_tensors[0]
        dev = out_pos_grad.device
        d_final = _zeros_like_state(n, dev)
        seeds = {}
        if _flags(solver) == EMPM_ADJ_DROP_GRID_MASS:
            d_final.x.copy_(out_pos_grad.sum(dim=0))  # interface.py:176-197
        else:
            for f, k in enumerate(ctx.record):
                g = out_pos_grad[f].contiguous()
                if k == ctx.tape.n_steps:
                    d_final.x.add_(g)
                else:
                    sd = seeds.setdefault(k, _zeros_like_state(n, dev))
                    sd.x.add_(g)
        _compat_colliders(ctx)
        d0, grads = ctx.ro.backward(ctx.tape, d_final, seeds=seeds, flags=_flags(solver))
        solver.last_grads = grads
        E_grad, nu_grad = _param_grads(model, grads, ctx.agg_E, ctx.agg_nu, n)
        velo_grad, density_grad = _initial_state_grads(ctx, d0, grads, density_change_mask)
        return (None, None, None, None, None, None, None, velo_grad, E_grad, nu_grad, density_grad, None, None,
                None, None, None)


def _initial_state_grads(ctx, d0, grads, density_change_mask):
    """init_velocity's and the density's gradients of the Trail and
    Simulation variants.  The reference reads them from the caller's state
    (mpm_state.particle_v.grad, .particle_density.grad: interface.py:201-204,
    236-241, 421-424, 456-461).  With extra_no_grad_steps > 0 the tape starts
    at the clone left by the warm-up steps (:112-117, 331-336), so nothing
    reaches the caller's state and both gradients stay zero."""
    if ctx.warmup > 0:
        velo_grad = torch.zeros_like(d0.v)
        density_grad = torch.zeros_like(grads.mass)
    else:
        velo_grad = d0.v.clone()
        density_grad = grads.mass * ctx.state.particle_vol
    if density_change_mask is not None:
        density_grad = density_grad[density_change_mask.to(torch.bool)]
    return velo_grad, density_grad


def _state_from(template, ps: rt.ParticleState):
    st = template.partial_clone(device=template.particle_x.device)
    st.particle_x = ps.x.clone()
    st.particle_v = ps.v.clone()
    st.particle_C = ps.C.clone()
    st.particle_F_trial = ps.F.clone()
    return st


class MPMDifferentiableSimulation(autograd.Function):
    """interface.py:254-472.  Returns the final positions (N, 3)."""

    @staticmethod
    def forward(ctx, mpm_solver, mpm_state, mpm_model, substep: int, substep_size: float, num_substeps: int,
                init_pos: Tensor, init_velocity: Tensor, E: Tensor, nu: Tensor,
                particle_density: Optional[Tensor] = None, density_change_mask: Optional[Tensor] = None,
                static_pos: Optional[Tensor] = None, device: str = "cuda:0", requires_grad: bool = True,
                extra_no_grad_steps: int = 0):
        _reset_or_restart(mpm_solver, mpm_state, mpm_model, init_pos, init_velocity, static_pos, device,
                          requires_grad)
        agg_E, agg_nu = _set_E_nu(mpm_solver, mpm_model, E, nu, scale_scalar_E=False)
        mpm_state.reset_density(tensor_density=particle_density, selection_mask=density_change_mask, device=device,
                                requires_grad=requires_grad)
        state = mpm_state
        if extra_no_grad_steps > 0:
            state.reset_density(None, update_mass=True)
            mpm_solver.prepare_mu_lam(mpm_model, state)
            final, _, _, _ = _run_forward(mpm_solver, state, mpm_model, substep_size, extra_no_grad_steps, (), False,
                                          step_callback=False)
            state = _state_from(state, final)
        state.reset_density(None, update_mass=True)
        mpm_solver.prepare_mu_lam(mpm_model, state)
        final, _, tape, ro = _run_forward(mpm_solver, state, mpm_model, substep_size, num_substeps, (),
                                          requires_grad, step_callback=False)
        ctx.mpm_solver, ctx.mpm_model, ctx.state, ctx.tape, ctx.ro = mpm_solver, mpm_model, state, tape, ro
        ctx.agg_E, ctx.agg_nu, ctx.n = agg_E, agg_nu, state.n_particles
        ctx.warmup = int(extra_no_grad_steps)
        ctx.save_for_backward(density_change_mask)
        return final.x.clone()

    @staticmethod
    def backward(ctx, out_pos_grad: Tensor):
        solver, model, n = ctx.mpm_solver, ctx.mpm_model, ctx.n
        density_change_mask = ctx.saved_tensors[0]
        d_final = _zeros_like_state(n, out_pos_grad.device)
        d_final.x.copy_(out_pos_grad)
        _compat_colliders(ctx)
        d0, grads = ctx.ro.backward(ctx.tape, d_final, flags=_flags(solver))
        solver.last_grads = grads
        E_grad, nu_grad = _param_grads(model, grads, ctx.agg_E, ctx.agg_nu, n)
        velo_grad, density_grad = _initial_state_grads(ctx, d0, grads, density_change_mask)
        return (None, None, None, None, None, None, None, velo_grad, E_grad, nu_grad, density_grad, None, None,
                None, None, None)


def _continue_forward(ctx, mpm_solver, mpm_state, mpm_model, substep_size, num_substeps, particle_x, particle_v,
                      particle_F, particle_C, E, nu, particle_density, query_mask, device, extra_no_grad_steps):
    """Shared forward of the WCheckpoint and Clean variants
    (interface.py:484-575,785-895)."""
    mpm_state.continue_from_torch(particle_x, particle_v, particle_F, particle_C, device=device, requires_grad=True)
    agg_E, agg_nu = _set_E_nu(mpm_solver, mpm_model, E, nu, scale_scalar_E=False)
    mpm_solver.prepare_mu_lam(mpm_model, mpm_state)
    mpm_state.reset_density(tensor_density=particle_density, selection_mask=query_mask, device=device,
                            requires_grad=True, update_mass=True)
    state = mpm_state
    if extra_no_grad_steps > 0:
        final, _, _, _ = _run_forward(mpm_solver, state, mpm_model, substep_size, extra_no_grad_steps, (), False,
                                      step_callback=False)
        state = _state_from(state, final)
    final, _, tape, ro = _run_forward(mpm_solver, state, mpm_model, substep_size, num_substeps, (), True,
                                      step_callback=False)
    ctx.mpm_solver, ctx.mpm_model, ctx.state, ctx.tape, ctx.ro = mpm_solver, mpm_model, state, tape, ro
    ctx.agg_E, ctx.agg_nu, ctx.n = agg_E, agg_nu, state.n_particles
    return final


def _continue_backward(ctx, out_pos_grad, out_velo_grad, out_F_grad, out_C_grad, seed_position_only=False):
    """Shared backward of WCheckpoint and Clean.  WCheckpoint seeds the
    cotangents of all four outputs (interface.py:621-720); the reference's
    Clean seeds only the position (:907-927) and ignores out_velo_grad,
    out_F_grad and out_C_grad, which seed_position_only reproduces."""
    solver, model, n = ctx.mpm_solver, ctx.mpm_model, ctx.n
    dev = out_pos_grad.device
    d = _zeros_like_state(n, dev)
    d.x.copy_(out_pos_grad)
    if seed_position_only:
        out_velo_grad = out_F_grad = out_C_grad = None
    if out_velo_grad is not None:
        d.v.copy_(out_velo_grad)
    if out_F_grad is not None:
        d.F.copy_(out_F_grad.reshape(n, 3, 3))
    if out_C_grad is not None:
        d.C.copy_(out_C_grad.reshape(n, 3, 3))
    if seed_position_only:
        # Clean's Tape is recorded in its forward (interface.py:853-872), so
        # its reverse sweep reads the colliders' last state as well
        _compat_colliders(ctx)
    d0, grads = ctx.ro.backward(ctx.tape, d, flags=_flags(solver))
    solver.last_grads = grads
    E_grad, nu_grad = _param_grads(model, grads, ctx.agg_E, ctx.agg_nu, n)
    density_grad = grads.mass * ctx.state.particle_vol
    return (None, None, None, None, None, d0.x.clone(), d0.v.clone(), d0.F.clone(), d0.C.clone(), E_grad, nu_grad,
            density_grad, None, None, None, None)


class MPMDifferentiableSimulationWCheckpoint(autograd.Function):
    """interface.py:476-774.  Returns (x, v, F_trial, C) of the final state."""

    @staticmethod
    def forward(ctx, mpm_solver, mpm_state, mpm_model, substep_size: float, num_substeps: int, particle_x: Tensor,
                particle_v: Tensor, particle_F: Tensor, particle_C: Tensor, E: Tensor, nu: Tensor,
                particle_density: Optional[Tensor] = None, query_mask: Optional[Tensor] = None,
                device: str = "cuda:0", requires_grad: bool = True, extra_no_grad_steps: int = 0):
        with torch.no_grad():
            final = _continue_forward(ctx, mpm_solver, mpm_state, mpm_model, substep_size, num_substeps, particle_x,
                                      particle_v, particle_F, particle_C, E, nu, particle_density, query_mask, device,
                                      extra_no_grad_steps)
        return final.x.clone(), final.v.clone(), final.F.clone(), final.C.clone()

    @staticmethod
    def backward(ctx, out_pos_grad, out_velo_grad, out_F_grad, out_C_grad):
        return _continue_backward(ctx, out_pos_grad, out_velo_grad, out_F_grad, out_C_grad)


class MPMDifferentiableSimulationClean(autograd.Function):
    """interface.py:777-986.  Returns (x, v, F_trial, C, cov (N, 6))."""

    @staticmethod
    def forward(ctx, mpm_solver, mpm_state, mpm_model, substep_size: float, num_substeps: int, particle_x: Tensor,
                particle_v: Tensor, particle_F: Tensor, particle_C: Tensor, E: Tensor, nu: Tensor,
                particle_density: Optional[Tensor] = None, query_mask: Optional[Tensor] = None,
                device: str = "cuda:0", requires_grad: bool = True, extra_no_grad_steps: int = 0):
        with torch.no_grad():
            final = _continue_forward(ctx, mpm_solver, mpm_state, mpm_model, substep_size, num_substeps, particle_x,
                                      particle_v, particle_F, particle_C, E, nu, particle_density, query_mask, device,
                                      extra_no_grad_steps)
        # the covariance update is disabled in the reference (mpm_utils.py:578)
        cov = mpm_state.particle_cov.view(-1, 6).clone()
        return final.x.clone(), final.v.clone(), final.F.clone(), final.C.clone(), cov

    @staticmethod
    def backward(ctx, out_pos_grad, out_velo_grad, out_F_grad, out_C_grad, out_cov_grad):
        # reference_compat: only the position seed, as the reference (interface.py:907-927);
        # exact: the derivative of everything the caller differentiated
        compat = _flags(ctx.mpm_solver) == EMPM_ADJ_DROP_GRID_MASS
        return _continue_backward(ctx, out_pos_grad, out_velo_grad, out_F_grad, out_C_grad,
                                  seed_position_only=compat)
