"""The benchmark configurations at their full sizes, checked against the oracle.

c0  32^3 grid, 64 ions per cell (2,097,152 ions), cold neutrals, rtol 1e-8, 10 steps:
    bit-exact particles and cell counts against the oracle with collisions off; one step of
    pre-drawn collisions fed the reference's own NR stream (logged by the oracle); Philox
    collisions against that stream in distribution over 5 steps.
c1  128^3 grid, 100 ions per cell (209,715,200 ions), FP64, rtol 1e-10: conservation,
    the true residual by the oracle's operator, the fused pass equal to the staged slots,
    and a 1-in-1000 id sample of the push and drift against the oracle.
c2  the c1 workload in mixed precision: per-step errors against FP64 from a common state.
c4  the 16-run ensemble: concurrent steps equal sequential ones; per-run mixed errors.
(c3, the batched solves, is covered in test_gpu_batch.py.)
"""
import gc

import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def reference_ics(po, n):
    """The reference's initial ions (NormalDistribution + NormalVelocities, seed -12345)
    drawn by the oracle's generators."""
    s3 = oracle.nr_stream("ran3", -12345, 3 * n)
    sg = oracle.nr_stream("gasdev", -12345, 3 * n)
    pos = s3.astype(np.float64).reshape(n, 3) * np.array([po.Lx, po.Ly, po.Lz])
    vel = np.array([po.viTx, po.viTy, po.viTz]) * sg.astype(np.float64).reshape(n, 3) + \
        np.array([po.vi0x, po.vi0y, po.vi0z])
    return pos, vel


def residual_ratio(po, m_op, m_rhs, phi):
    """||b - A phi|| / ||b|| with A from the density m_op and b from the moments m_rhs."""
    b = oracle.full_rhs(po, m_rhs)
    r = b - oracle.lhs_apply(po, m_op, phi)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


# ----------------------------------------------------------------------------------- c0
def test_c0_fp64_ten_steps_bit_exact():
    """c0 with collisions off, from the reference's own initial ions: ten fused steps
    (push, drift/wrap, re-binning, deposit, solve) against the oracle's push, drift, wrap
    and deposit on the same potential.  Positions, velocities and cell counts bit-equal
    every step; moments within 1e-12; every solve meets rtol 1e-8 on the oracle's operator."""
    wl = configs.CPU32
    say(f"c0 fp64 10 steps, {wl.np} ions")
    p = pic.setup_context(**wl.cli)
    po = oracle.setup_context(**wl.cli)
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=wl.rtol,
                                             operator_mode=_lib.OPERATOR_REBUILD))
    ctx.init_reference(wl.np)
    pos, vel = reference_ics(po, wl.np)
    gpos, gvel = ctx.get_particles()
    np.testing.assert_array_equal(gpos, pos)
    np.testing.assert_array_equal(gvel, vel)
    ctx.collect()
    ctx.compute_potential()
    m = ctx.get_moments()
    assert rel(m, oracle.collect_moments(po, pos, vel)) <= 1e-12
    for step in range(1, wl.steps + 1):
        phi = ctx.get_potential()
        ctx.step(wl.dt)
        oracle.boris(po, wl.dt, pos, vel, phi)
        oracle.update_positions(wl.dt, pos, vel)
        oracle.apply_bc(po, pos)
        gpos, gvel = ctx.get_particles()
        np.testing.assert_array_equal(gvel, vel)
        np.testing.assert_array_equal(gpos, pos)
        m = ctx.get_moments()
        e = rel(m, oracle.collect_moments(po, pos, vel))
        rr = residual_ratio(po, m, m, ctx.get_potential())
        say(f"  step {step}: moments {e:.1e}, ||b - A phi||/||b|| {rr:.2e}, {ctx.solver_stats()[0]} iterations")
        assert e <= 1e-12
        assert rr <= 1.01 * wl.rtol
    np.testing.assert_array_equal(ctx.cell_counts(), oracle.cell_counts(po, pos))
    ctx.close()


def test_c0_predrawn_reference_stream_one_step():
    """c0's collision regime (cold neutrals: Nc = 0.4 np, 838,860 collisions) fed the
    reference's own NR stream: the oracle runs ScatterElastic with ran3/gasdev and logs
    every attempt; the GPU replays the log after its push.  Velocities agree per ion to
    1e-12 and Nc, successes, refusals and attempts consumed are equal."""
    wl = configs.CPU32
    say(f"c0 pre-drawn NR stream, {wl.np} ions")
    p = pic.setup_context(**wl.cli)
    po = oracle.setup_context(**wl.cli)
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_PREDRAWN, ksp_rtol=wl.rtol))
    ctx.init_reference(wl.np)
    pos, vel = reference_ics(po, wl.np)
    ctx.collect()
    ctx.compute_potential()
    phi = ctx.get_potential()
    ctx.push(wl.dt)
    oracle.boris(po, wl.dt, pos, vel, phi)
    rc, nl, ns, nf, log = oracle.scatter_elastic_nr(po, wl.dt, vel, oracle.NR(), log_cap=4 * wl.np)
    Nc = oracle.collision_count(po, wl.np, wl.dt)
    say(f"  Nc {Nc}, attempts {nl}, refusals {nf}")
    assert rc == 0 and ns == Nc and nl == len(log) and Nc > 0.3 * wl.np
    ctx.set_attempts(log)
    ctx.collide(wl.dt)
    assert ctx.collision_stats() == (Nc, ns, nf, nl)
    _, gvel = ctx.get_particles()
    # FP64 scattering from the same draws; sin/cos of the azimuth differ from the host's
    # libm in the last ulp, which a near-zero sin(theta) (collisions.c:111-114) magnifies
    assert rel(gvel, vel) <= 1e-12
    ctx.close()


def test_c0_philox_statistics_against_reference_stream():
    """c0 with Philox collisions against the reference's algorithm run on its NR stream,
    over 5 steps from the same initial ions and (step by step) the same potential.

    Step 1 runs the slots one by one so that the fraction of ions that collided at least
    once (the zero bin of the successes-per-ion histogram) is visible on both sides; it is
    compared with the reference's own stream (seed -12345).  Later steps are compared with
    the reference's loop on a stream that is independent from step to step (seed -12345 -
    7919 s): the reference re-seeds ran3 with the same seed on every call, and ran3 leaves
    the seed at 1 for ran1's LCG (src/collisions.c:41, src/random.c:103-134), so its
    collisions repeat from step to step; that correlation is a property of its generator
    use, which an independent counter-based stream does not reproduce (exact per-ion
    reproduction of the reference's stream is the pre-drawn mode's job).  After 5 steps of
    cooling on the cold background the velocity variances agree within 2%, and every step
    collides Nc ions within 1%."""
    wl = configs.CPU32
    say(f"c0 Philox vs NR stream, {wl.np} ions")
    p = pic.setup_context(**wl.cli)
    po = oracle.setup_context(**wl.cli)
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=wl.rtol))
    ctx.init_reference(wl.np)
    pos, vel = reference_ics(po, wl.np)
    ctx.collect()
    ctx.compute_potential()
    nr = oracle.NR()
    Nc = oracle.collision_count(po, wl.np, wl.dt)
    for step in range(1, 6):
        phi = ctx.get_potential()
        oracle.boris(po, wl.dt, pos, vel, phi)
        seed = -12345 - 7919 * (step - 1)
        if step == 1:
            ctx.push(wl.dt)
            _, va = ctx.get_particles()
            ctx.collide(wl.dt)
            _, vb = ctx.get_particles()
            ctx.update_positions(wl.dt)
            ctx.collect()
            ctx.compute_potential()
            before = vel.copy()
            rc, nl, ns, nf, _ = oracle.scatter_elastic_nr(po, wl.dt, vel, nr, seed=seed)
            assert rc == 0 and ns == Nc
            f_gpu = np.any(vb != va, axis=1).mean()
            f_ref = np.any(vel != before, axis=1).mean()
            sigma = np.sqrt(f_ref * (1 - f_ref) / wl.np)
            say(f"  step 1: ions collided at least once: GPU {f_gpu:.5f}, reference stream {f_ref:.5f}")
            assert abs(f_gpu - f_ref) <= 0.01 * f_ref + 6 * sigma
            del va, vb, before
        else:
            ctx.step(wl.dt)
            rc, nl, ns, nf, _ = oracle.scatter_elastic_nr(po, wl.dt, vel, nr, seed=seed)
            assert rc == 0
        oracle.update_positions(wl.dt, pos, vel)
        oracle.apply_bc(po, pos)
        gNc, gNs, gNf, gna = ctx.collision_stats()
        say(f"  step {step}: Nc {gNc}, GPU successes {gNs} (attempts {gna}); reference attempts {nl}")
        assert gNc == Nc and gNf == 0
        assert abs(gNs - Nc) <= 0.01 * Nc
        assert abs(gna - nl) <= 0.02 * nl
    _, gvel = ctx.get_particles()
    var_g, var_r = gvel.var(axis=0), vel.var(axis=0)
    say(f"  after 5 steps: variance GPU {var_g}, reference loop {var_r}")
    np.testing.assert_allclose(var_g, var_r, rtol=0.02)
    assert (np.abs(gvel.mean(axis=0) - vel.mean(axis=0)) <= 6 * np.sqrt(var_r / wl.np) + 1e-3).all()
    ctx.close()


# ----------------------------------------------------------------------------------- c1
def _c1_context(collide, operator=_lib.OPERATOR_FROZEN, prec=_lib.PARTICLES_FP64):
    wl = configs.PROD128
    p = pic.setup_context(**wl.cli)
    solve = _lib.SOLVE_MIXED if prec == _lib.PARTICLES_MIXED else _lib.SOLVE_FP64
    ctx = pic.Context(p, pic.default_options(collide_mode=collide, ksp_rtol=wl.rtol, operator_mode=operator,
                                             particle_precision=prec, solve_precision=solve))
    ctx.init_particles(wl.np, 2024, wl.amp, wl.k)
    ctx.collect()
    return wl, ctx


def test_c1_fp64_full_size_conservation_and_residual():
    """c1 at full size, Philox collisions, the operator frozen at t = 0 as in the bench: 3
    steps keep every ion (count, sum of the density = np to 1e-12), each solve meets rtol
    1e-10 on the oracle's operator built from the t = 0 density, the flux sums equal the
    velocity sums and the GPU's cell counts equal the oracle's histogram of the
    downloaded positions."""
    wl, ctx = _c1_context(_lib.COLLIDE_PHILOX)
    say(f"c1 fp64 full size, {wl.np} ions")
    po = oracle.setup_context(**wl.cli)
    m0 = ctx.get_moments()
    assert m0[..., 0].sum() == pytest.approx(wl.np, rel=1e-12)
    ctx.compute_potential()
    for step in range(1, 4):
        ctx.step(wl.dt)
        assert ctx.num_particles == wl.np
        m = ctx.get_moments()
        it, rn, bn = ctx.solver_stats()
        rr = residual_ratio(po, m0, m, ctx.get_potential())
        Nc, Ns, Nf, na = ctx.collision_stats()
        say(f"  step {step}: sum n - np = {m[..., 0].sum() - wl.np:.2e}, {it} iterations, "
            f"||b - A phi||/||b|| {rr:.2e}; collisions {Ns} of Nc {Nc}")
        assert m[..., 0].sum() == pytest.approx(wl.np, rel=1e-12)
        assert rn <= wl.rtol * bn and rr <= 1.01 * wl.rtol
        assert Nf == 0 and abs(Ns - Nc) <= 0.05 * Nc
    pos, vel = ctx.get_particles()
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    assert (pos >= 0).all() and (pos <= 128).all()
    G = m[..., 1:].reshape(-1, 3).sum(axis=0)
    np.testing.assert_allclose(G, vel.sum(axis=0), rtol=0, atol=1e-10 * np.abs(vel).sum())
    counts = ctx.cell_counts()
    assert counts.sum() == wl.np
    np.testing.assert_array_equal(counts, oracle.cell_counts(po, pos))
    ctx.close()
    del pos, vel
    gc.collect()


def test_c1_fp64_fused_equals_staged_and_sampled_oracle():
    """c1 at full size with collisions off: one fused step equals the staged slots (push,
    UpdatePositions, collect) from the same state bit for bit, and every 1000th ion of it
    equals the oracle's push, drift and wrap on the same potential bit for bit."""
    wl, ctx = _c1_context(_lib.COLLIDE_OFF)
    say(f"c1 fused vs staged, {wl.np} ions")
    po = oracle.setup_context(**wl.cli)
    ctx.compute_potential()
    ctx.step(wl.dt)
    pos0, vel0 = ctx.get_particles()
    phi0 = ctx.get_potential()
    ctx.step(wl.dt)
    pos1, vel1 = ctx.get_particles()
    m1 = ctx.get_moments()
    sample = np.arange(0, wl.np, 1000)
    sp, sv = pos0[sample].copy(), vel0[sample].copy()
    oracle.boris(po, wl.dt, sp, sv, phi0)
    oracle.update_positions(wl.dt, sp, sv)
    oracle.apply_bc(po, sp)
    np.testing.assert_array_equal(vel1[sample], sv)
    np.testing.assert_array_equal(pos1[sample], sp)
    ctx.set_particles(pos0, vel0)
    del pos0, vel0
    gc.collect()
    ctx.set_potential(phi0)
    ctx.push(wl.dt)
    ctx.update_positions(wl.dt)
    ctx.collect()
    pos2, vel2 = ctx.get_particles()
    np.testing.assert_array_equal(vel2, vel1)
    np.testing.assert_array_equal(pos2, pos1)
    assert rel(ctx.get_moments(), m1) <= 1e-12
    ctx.close()
    del pos1, vel1, pos2, vel2
    gc.collect()


def test_c2_mixed_per_step_errors_against_fp64():
    """c2: the mixed path (FP32 offset + cell, FP32 push/deposit, mixed solve) started each
    step from the FP64 run's state: density, flux and potential after one step agree with
    FP64 to 1e-6 of each field's maximum (the north star's potential bound).  Both solves
    build the operator from the step's own density, so the potentials answer the same
    question."""
    wl, c64 = _c1_context(_lib.COLLIDE_OFF, operator=_lib.OPERATOR_REBUILD)
    say(f"c2 mixed vs fp64 per step, {wl.np} ions")
    c64.compute_potential()
    p = pic.setup_context(**wl.cli)
    cmx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=wl.rtol,
                                             particle_precision=_lib.PARTICLES_MIXED,
                                             solve_precision=_lib.SOLVE_MIXED, operator_mode=_lib.OPERATOR_REBUILD))
    for step in range(1, 3):
        pos, vel = c64.get_particles()
        cmx.set_particles(pos, vel)
        del pos, vel
        gc.collect()
        cmx.set_potential(c64.get_potential())
        c64.step(wl.dt)
        cmx.step(wl.dt)
        m64, mmx = c64.get_moments(), cmx.get_moments()
        en = rel(mmx[..., 0], m64[..., 0])
        eg = max(rel(mmx[..., d], m64[..., d]) for d in (1, 2, 3))
        ef = rel(cmx.get_potential(), c64.get_potential())
        say(f"  step {step}: n {en:.2e}, G {eg:.2e}, phi {ef:.2e}")
        assert en <= 1e-6 and eg <= 1e-6 and ef <= 1e-6
    c64.close()
    cmx.close()


# ----------------------------------------------------------------------------------- c4
def _ensemble_member(run, extra, collide=_lib.COLLIDE_PHILOX):
    c = pic.Context(pic.setup_context(**configs.ensemble_cli(run)),
                    pic.default_options(ksp_rtol=1e-10, collide_mode=collide, **extra))
    c.init_particles(run["ppc"] * run["N"] ** 3, run["seed"], 1e-3, 2)
    c.collect()
    c.compute_potential()
    return c


MIXED = dict(particle_precision=_lib.PARTICLES_MIXED, solve_precision=_lib.SOLVE_MIXED)


def test_c4_ensemble_equals_sequential_steps():
    """c4's 16 uneven runs (mixed precision, Philox collisions): two concurrent ensemble
    steps equal the same runs stepped one by one, ion by ion and node by node."""
    runs = configs.ensemble_runs()
    say(f"c4 ensemble vs sequential, {sum(r['ppc'] * r['N'] ** 3 for r in runs)} ions in 16 runs")
    ens = [_ensemble_member(r, MIXED) for r in runs]
    seq = [_ensemble_member(r, MIXED) for r in runs]
    for _ in range(2):
        pic.ensemble_step(ens, 1.0)
        for c in seq:
            c.step(1.0)
    for r, a, b in zip(runs, ens, seq):
        pa, va = a.get_particles()
        pb, vb = b.get_particles()
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(va, vb)
        np.testing.assert_array_equal(a.get_moments(), b.get_moments())
        np.testing.assert_array_equal(a.get_potential(), b.get_potential())
        assert a.collision_stats() == b.collision_stats() and a.collision_stats()[1] > 0
        a.close()
        b.close()


def test_c4_lockstep_ensemble_equals_single_steps():
    """c4 as the bench steps it, one lockstep ensemble: per step from a common state, each
    run's ions, moments and collision counts equal its own single step bit for bit, and its
    potential (the lockstep solver's) is within 1e-8 of the single solve's."""
    runs = configs.ensemble_runs()
    say("c4 lockstep ensemble vs single steps")
    ens = [_ensemble_member(r, MIXED) for r in runs]
    seq = [_ensemble_member(r, MIXED) for r in runs]
    e = pic.Ensemble(ens)
    for step in range(2):
        e.advance(1.0)
        worst = 0.0
        for a, b in zip(ens, seq):
            b.step(1.0)
            pa, va = a.get_particles()
            pb, vb = b.get_particles()
            np.testing.assert_array_equal(pa, pb)
            np.testing.assert_array_equal(va, vb)
            np.testing.assert_array_equal(a.get_moments(), b.get_moments())
            assert a.collision_stats() == b.collision_stats() and a.collision_stats()[1] > 0
            fa = a.get_potential()
            worst = max(worst, rel(fa, b.get_potential()))
            b.set_potential(fa)
        say(f"  step {step}: potential vs single solves {worst:.1e}")
        assert worst <= 1e-8
    e.close()
    for c in ens + seq:
        c.close()


def test_c4_per_run_mixed_errors_against_fp64():
    """Each of c4's runs: one mixed step against one FP64 step from the same initial ions
    (collisions off): density, flux and potential within 1e-6 of each field's maximum."""
    for r in configs.ensemble_runs():
        a = _ensemble_member(r, {}, _lib.COLLIDE_OFF)
        b = _ensemble_member(r, MIXED, _lib.COLLIDE_OFF)
        a.step(1.0)
        b.step(1.0)
        ma, mb = a.get_moments(), b.get_moments()
        en = rel(mb[..., 0], ma[..., 0])
        eg = max(rel(mb[..., d], ma[..., d]) for d in (1, 2, 3))
        ef = rel(b.get_potential(), a.get_potential())
        say(f"  run N={r['N']} ppc={r['ppc']}: n {en:.1e}, G {eg:.1e}, phi {ef:.1e}")
        assert en <= 1e-6 and eg <= 1e-6 and ef <= 1e-6
        a.close()
        b.close()


# ----------------------------------------------------------------------------------- c3
def test_c3_batched_solves_full_size():
    """c3 at full size: 256 fields with edges drawn from {32, ..., 192} (2.0e8 nodes), the
    smoothed log-normal densities of the bench, solved in lockstep in mixed precision and in
    FP64 to rtol 1e-8.  Every problem's true residual, by the oracle's assembled operator,
    meets rtol, and the mixed potential agrees with the FP64 one to 1e-6 per problem."""
    from concurrent.futures import ThreadPoolExecutor
    dims = configs.batch_dims()
    say(f"c3 batched solves, {len(dims)} problems, {sum(a * b * c for a, b, c in dims)} nodes")
    p = pic.setup_context(**configs.base_cli(32, E0=0.05))
    mom = [pic.lognormal_moments(d, seed=1000 + q, sigma=0.3, smooth=3.0) for q, d in enumerate(dims)]
    res = {}
    for name, prec in (("mixed", _lib.SOLVE_MIXED), ("fp64", _lib.SOLVE_FP64)):
        phi = [pic.DeviceVector(d[0] * d[1] * d[2]) for d in dims]
        b = pic.Batch(p, pic.default_options(ksp_rtol=1e-8, solve_precision=prec), dims, slots=0)
        b.solve(mom, phi)
        it, _ = b.stats()
        say(f"  {name}: iterations {int(it.min())}-{int(it.max())}, traffic {b.traffic() / 1e9:.1f} GB")
        b.close()
        res[name] = [f.download() for f in phi]
        for f in phi:
            f.free()
    cli0 = configs.base_cli(32, E0=0.05)

    def check(q):
        d = dims[q]
        cli = dict(cli0, Nx=d[0], Ny=d[1], Nz=d[2])
        po = oracle.setup_context(**cli)
        m = mom[q].download().reshape(d[2], d[1], d[0], 4)
        bq = oracle.full_rhs(po, m)
        out = []
        for name in ("mixed", "fp64"):
            x = res[name][q].reshape(d[2], d[1], d[0])
            out.append(float(np.linalg.norm(bq - oracle.lhs_apply(po, m, x)) / np.linalg.norm(bq)))
        ref = res["fp64"][q]
        out.append(float(np.abs(res["mixed"][q] - ref).max() / np.abs(ref).max()))
        return out

    with ThreadPoolExecutor(16) as ex:
        rows = list(ex.map(check, range(len(dims))))
    rows = np.array(rows)
    say(f"  worst residual: mixed {rows[:, 0].max():.2e}, fp64 {rows[:, 1].max():.2e}; "
        f"mixed vs fp64 worst {rows[:, 2].max():.2e}, median {np.median(rows[:, 2]):.2e}")
    assert (rows[:, 0] <= 1.01e-8).all() and (rows[:, 1] <= 1.01e-8).all()
    assert (rows[:, 2] <= 1e-6).all()
    for v in mom:
        v.free()
