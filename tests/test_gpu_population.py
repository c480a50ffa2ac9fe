"""Batched population evaluation (empm_amd.population) against the same
candidates evaluated one plain rollout at a time.

A scaled-down population of the CMA-ES workload (6 candidates, 2k-6k
particles each, cube / rod / ellipsoid, 32^3 grids, 60 substeps): final
positions per candidate within 1e-5 relative, fitness and the per-candidate
gradients d fitness / d(log E, nu, log Y) within 1e-3 relative.  Fitness
is a mean squared difference of two rollouts, so round-off of the atomic
sums (~1e-7 of a position) enters it relative to the (small) position
differences: 1e-3 is its bar, as for the gradients."""
import numpy as np
import pytest
import torch

from empm_amd import population as pop
from empm_amd import scenes
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def small_population(second_material):
    kw = dict(n_cand=6, n_grid=32, n_steps=60, n_range=(2000, 6000), second_material=second_material)
    return scenes.population_c3(**kw), scenes.population_c3(nominal=True, **kw)


@pytest.mark.parametrize("second_material", [5, 7])
def test_population_forward_matches_sequential(dev, second_material):
    sc, nom = small_population(second_material)
    print(f"[population fwd] materials {[s.material for s in sc]} N {[s.n for s in sc]}", flush=True)
    ev = pop.PopulationEvaluator(sc, dev)
    ev.set_targets([pop.Candidate.of_scene(s) for s in nom])
    cands = [pop.Candidate.of_scene(s) for s in sc]
    final = ev.rollout(cands)
    bm, _, _ = ev.model(cands)
    for c in range(ev.n_cand):
        ro, s0, a, b = ev._solo(c, bm)
        f1, _, _, _ = ro.forward(s0, list(ev.cols), 0.0, ev.dt, ev.n_steps)
        e = rel_err(final.x[a:b].cpu().numpy(), f1.x.cpu().numpy())
        eF = rel_err(final.F[a:b].cpu().numpy(), f1.F.cpu().numpy())
        print(f"  candidate {c} ({sc[c].name}): rel err x {e:.1e} F {eF:.1e}", flush=True)
        assert e < 1e-5 and eF < 1e-4
    fb = ev.fitness(cands).cpu().numpy()
    fs = ev.fitness_sequential(cands).cpu().numpy()
    print(f"  fitness batched {fb}\n  fitness sequential {fs}", flush=True)
    assert np.all(fb > 0)
    np.testing.assert_allclose(fb, fs, rtol=1e-3)
    # the nominal candidates reproduce their own targets
    f0 = ev.fitness([pop.Candidate.of_scene(s) for s in nom]).cpu().numpy()
    assert np.all(f0 < 1e-6 * fb.max()), f0


@pytest.mark.parametrize("ckpt,second_material", [(1, 5), (7, 5), (1, 7)])
def test_population_gradients_match_sequential(dev, ckpt, second_material):
    """Sand and plasticine (the workload's mix: plasticine's rewritten yield
    stress, mu and lam are tape state) or sand and Neo-Hookean + von Mises."""
    sc, nom = small_population(second_material)
    ev = pop.PopulationEvaluator(sc, dev)
    ev.set_targets([pop.Candidate.of_scene(s) for s in nom])
    cands = [pop.Candidate.of_scene(s) for s in sc]
    fb, gb = ev.fitness_and_grad(cands, ckpt_every=ckpt)
    fs, gs = ev.fitness_and_grad_sequential(cands, ckpt_every=ckpt)
    fb, gb, fs, gs = (t.cpu().numpy() for t in (fb, gb, fs, gs))
    print(f"[population grad ckpt {ckpt}]\n  batched {gb}\n  sequential {gs}", flush=True)
    np.testing.assert_allclose(fb, fs, rtol=1e-3)
    # per component, relative to the candidate's largest gradient component
    err = np.abs(gb - gs).max(1) / np.abs(gs).max(1)
    print(f"  rel err per candidate {err}", flush=True)
    assert np.all(err < 1e-3), err
    assert np.all(np.isfinite(gb)) and np.all(np.abs(gb).max(1) > 0)
