"""Lockstep ensembles (ecsprt_hip_ensemble_*): one host thread; the tile passes and the
solve phases one launch for all runs, each run's small launches on its own stream between
a fork and a join.  From the same state, every run's ions, moments and collision counts
equal those of stepping its context alone (ecsprt_hip_step) bit for bit, and its
potential meets the solve tolerance against the oracle's operator and forcing."""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


# (grid edge, ions per cell, particle precision, B0, collisions): 4^3 tiles on the small
# grids, 8x8x4 (FP64) and 8x8x8 (mixed) tiles at 64^3, so the runs form four launches.
RUNS = [(12, 6, _lib.PARTICLES_MIXED, 0.05, True), (16, 4, _lib.PARTICLES_FP64, 0.08, True),
        (8, 10, _lib.PARTICLES_MIXED, 0.02, True), (64, 1, _lib.PARTICLES_FP64, 0.04, True),
        (64, 1, _lib.PARTICLES_MIXED, 0.06, True), (12, 5, _lib.PARTICLES_FP64, 0.05, False)]


def make(N, ppc, prec, B0, coll, seed, solve=_lib.SOLVE_FP64):
    """A run after its first solve, which fixes its operator (the density then), and that
    density's moments."""
    cli = configs.base_cli(N, B0=B0, nui=0.02, E0=0.05)
    o = pic.default_options(particle_precision=prec, solve_precision=solve, ksp_rtol=1e-10,
                            collide_mode=_lib.COLLIDE_PHILOX if coll else _lib.COLLIDE_OFF, seed=seed)
    c = pic.Context(pic.setup_context(**cli), o)
    c.init_particles(ppc * N ** 3, seed, 0.05, 1)
    c.collect()
    c.compute_potential()
    return c, (cli, c.get_moments())


def check_potential(run, m, phi, rtol):
    """||b(m) - A(m0) phi|| / ||b(m)|| with the operator of the density m0 it was built from."""
    cli, m0 = run
    po = oracle.setup_context(**cli)
    b = oracle.full_rhs(po, m)
    r = np.linalg.norm(b - oracle.lhs_apply(po, m0, phi)) / np.linalg.norm(b)
    assert r <= 1.01 * rtol, r
    assert abs(phi.mean()) <= 1e-12 * np.abs(phi).max()
    return r


@pytest.mark.parametrize("solve", [_lib.SOLVE_FP64, _lib.SOLVE_MIXED])
def test_lockstep_ensemble_matches_single_steps(solve):
    say(f"lockstep ensemble vs single steps, solve precision {solve}")
    ens = [make(*r, seed=q + 1, solve=solve) for q, r in enumerate(RUNS)]
    seq = [make(*r, seed=q + 1, solve=solve)[0] for q, r in enumerate(RUNS)]
    e = pic.Ensemble([c for c, _ in ens])
    for step in range(3):
        e.advance(1.0)
        for (a, run), b in zip(ens, seq):
            b.step(1.0)
            pa, va = a.get_particles()
            pb, vb = b.get_particles()
            assert np.array_equal(pa, pb) and np.array_equal(va, vb)
            ma = a.get_moments()
            assert np.array_equal(ma, b.get_moments())
            assert a.collision_stats() == b.collision_stats()
            phi_a, phi_b = a.get_potential(), b.get_potential()
            err = np.abs(phi_a - phi_b).max() / np.abs(phi_b).max()
            say(f"  step {step} grid {run[0]['Nx']}: phi vs single {err:.1e}")
            r = check_potential(run, ma, phi_a, 1e-10)
            say(f"    residual {r:.1e}")
            assert err <= 1e-7
            # the next step starts from the same state on both sides
            b.set_potential(phi_a)
    assert e.traffic() > 0
    e.close()
    # the contexts step alone afterwards, on their own streams
    for (a, _), b in zip(ens, seq):
        a.step(1.0)
        b.step(1.0)
        assert np.array_equal(a.get_particles()[1], b.get_particles()[1])
    for c in [c for c, _ in ens] + seq:
        c.close()


def test_lockstep_ensemble_failed_run_leaves_others():
    """A run that loses an ion (the reference's Abort on a count change) returns its code
    and message; the other run's step completes as it would alone."""
    say("lockstep ensemble with a failing run")
    (a, _), (b, run_b) = make(8, 2, _lib.PARTICLES_FP64, 0.05, True, 1), make(12, 3, _lib.PARTICLES_FP64, 0.05, True, 2)
    ref = make(12, 3, _lib.PARTICLES_FP64, 0.05, True, 2)[0]
    pos, vel = a.get_particles()
    vel[0] = [3 * 8.0, 0, 0]  # leaves the domain after one wrap
    a.set_particles(pos, vel)
    e = pic.Ensemble([a, b])
    with pytest.raises(pic.HipError) as err:
        e.advance(1.0)
    assert err.value.code == _lib.ERR_COUNT_CHANGED
    ref.step(1.0)
    assert np.array_equal(b.get_particles()[1], ref.get_particles()[1])
    check_potential(run_b, b.get_moments(), b.get_potential(), 1e-10)
    e.close()
    for c in (a, b, ref):
        c.close()


def test_lockstep_ensemble_rejects_unsupported_runs():
    c1, _ = make(8, 2, _lib.PARTICLES_FP64, 0.05, True, 1)
    c2, _ = make(8, 2, _lib.PARTICLES_FP64, 0.05, True, 2)
    o = pic.default_options(ksp_rtol=1e-8)
    c3 = pic.Context(pic.setup_context(**configs.base_cli(8)), o)
    c3.init_particles(1024, 3, 0.0, 1)
    with pytest.raises(pic.HipError, match="solver options differ"):
        pic.Ensemble([c1, c3])
    with pytest.raises(pic.HipError, match="appears twice"):
        pic.Ensemble([c1, c1])
    o = pic.default_options(ksp_rtol=1e-10, collide_mode=_lib.COLLIDE_PREDRAWN)
    c4 = pic.Context(pic.setup_context(**configs.base_cli(8)), o)
    with pytest.raises(pic.HipError, match="Philox collisions or none"):
        pic.Ensemble([c2, c4])
    for c in (c1, c2, c3, c4):
        c.close()
