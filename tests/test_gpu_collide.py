"""GPU parity of ScatterElastic (src/collisions.c:9-167).

Pre-drawn mode must match the oracle per particle given the same records, both
for synthetic records and for the records the reference's own NR stream produces
(logged by the oracle).  Philox mode is checked statistically on velocity moments.
"""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def records(n, rng):
    att = np.zeros(n, dtype=oracle.ATTEMPT_DTYPE)
    att["u_ip"] = rng.random(n, dtype=np.float32)
    att["g"] = rng.normal(size=(n, 3)).astype(np.float32)
    att["u_acc1"] = rng.random(n, dtype=np.float32)
    att["u_acc2"] = rng.random(n, dtype=np.float32)
    att["u_z"] = rng.random(n, dtype=np.float32)
    att["u_phi"] = rng.random(n, dtype=np.float32)
    return att


def setup(N, ppc, nui, Tn_over_Ti, seed=1, mn=1.0):
    cli = configs.base_cli(N, nui=nui, Tn_over_Ti=Tn_over_Ti)
    cli["mn"] = mn
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(tile=4, collide_mode=_lib.COLLIDE_PREDRAWN))
    rng = np.random.default_rng(seed)
    n = ppc * N ** 3
    pos = rng.uniform(0, N, (n, 3))
    vel = rng.normal(size=(n, 3))
    return ctx, p, po, pos, vel


@pytest.mark.parametrize("N,ppc,nui,Tn,mn", [(8, 8, 0.05, 1.0, 1.0), (12, 4, 0.3, 1.0, 2.0), (8, 16, 0.01, 0.01, 1.0)])
def test_predrawn_matches_oracle(N, ppc, nui, Tn, mn):
    say(f"predrawn N={N} ppc={ppc} nui={nui} Tn={Tn} mn={mn}")
    ctx, p, po, pos, vel = setup(N, ppc, nui, Tn, mn=mn)
    n = pos.shape[0]
    Nc = oracle.collision_count(po, n, 1.0)
    assert Nc > 0
    att = records(12 * Nc + 1000, np.random.default_rng(5))
    ctx.set_particles(pos, vel)
    ctx.set_attempts(att)
    ctx.collide(1.0)
    ovel = vel.copy()
    rc, used, ns, nf = oracle.scatter_elastic_predrawn(po, 1.0, ovel, att, Nc)
    assert rc == 0
    gNc, gNs, gNf, gatt = ctx.collision_stats()
    assert (gNc, gNs, gNf, gatt) == (Nc, ns, nf, used)
    _, gvel = ctx.get_particles()
    np.testing.assert_allclose(gvel, ovel, rtol=1e-12, atol=1e-13)
    changed = np.any(ovel != vel, axis=1).sum()
    say(f"  Nc={Nc} attempts={used} particles changed={changed}")
    ctx.close()


def test_nr_stream_log_reproduced_per_particle():
    """The oracle's reference-order NR stream, replayed from its log on the GPU."""
    say("NR log replay")
    ctx, p, po, pos, vel = setup(8, 8, 0.05, 1.0)
    n = pos.shape[0]
    nr = oracle.NR()
    ovel = vel.copy()
    rc, nl, ns, nf, log = oracle.scatter_elastic_nr(po, 1.0, ovel, nr, log_cap=10 ** 6)
    assert rc == 0 and nl == len(log)
    ctx.set_particles(pos, vel)
    ctx.set_attempts(log)
    ctx.collide(1.0)
    _, gvel = ctx.get_particles()
    np.testing.assert_allclose(gvel, ovel, rtol=1e-12, atol=1e-13)
    assert ctx.collision_stats()[1:] == (ns, nf, nl)
    ctx.close()


def test_predrawn_refusals_abort():
    say("predrawn abort")
    ctx, p, po, pos, vel = setup(8, 2, 0.05, 1.0)
    att = records(2000, np.random.default_rng(6))
    att["u_acc1"] = 0.0
    att["g"] = 1e3
    ctx.set_particles(pos, vel)
    ctx.set_attempts(att)
    with pytest.raises(pic.HipError) as e:
        ctx.collide(1.0)
    assert e.value.code == _lib.ERR_COLLISION
    ctx.close()


def test_predrawn_exhausted_records():
    say("predrawn exhausted")
    ctx, p, po, pos, vel = setup(8, 2, 0.05, 1.0)
    att = records(5, np.random.default_rng(7))
    ctx.set_particles(pos, vel)
    ctx.set_attempts(att)
    with pytest.raises(pic.HipError) as e:
        ctx.collide(1.0)
    assert e.value.code == _lib.ERR_ATTEMPTS
    ctx.close()


def test_philox_statistics_match_nr_stream():
    """Velocity moments after one collision step: Philox mode (fused step) against the
    reference's NR stream on the oracle, same initial state, no field."""
    say("philox statistics")
    N, ppc = 16, 64
    cli = configs.base_cli(N, nui=0.1, Tn_over_Ti=1.0, B0=0.0)
    cli["vn0x"] = 0.7  # drifting neutrals give the collisions a visible signature
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_PHILOX, ksp_rtol=1e-6,
                                             operator_mode=_lib.OPERATOR_REBUILD))
    rng = np.random.default_rng(8)
    n = ppc * N ** 3
    pos = rng.uniform(0, N, (n, 3))
    vel = rng.normal(size=(n, 3))
    ctx.set_particles(pos, vel)
    ctx.set_potential(np.zeros((N, N, N)))
    ctx.step(1.0)
    _, gvel = ctx.get_particles()
    Nc, Ns, Nf, na = ctx.collision_stats()
    ovel = vel.copy()
    oracle.boris(po, 1.0, pos, ovel, np.zeros((N, N, N)))
    rc, nl, ns, nf, _ = oracle.scatter_elastic_nr(po, 1.0, ovel, oracle.NR())
    assert rc == 0
    say(f"  Nc={Nc} gpu successes={Ns} attempts={na}; oracle attempts={nl}")
    assert abs(Ns - ns) < 5 * np.sqrt(ns)
    # mean drift picked up from the neutrals and the temperature, per component
    dm = np.abs(gvel.mean(axis=0) - ovel.mean(axis=0))
    assert (dm < 5 * ovel.std(axis=0) / np.sqrt(n) * 2).all(), dm
    dv = np.abs(gvel.var(axis=0) - ovel.var(axis=0))
    assert (dv < 0.01).all(), dv
    ctx.close()
