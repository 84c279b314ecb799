"""The plugin-slot boundary used as the reference's loop uses it (src/pic.c:246-258,
src/velocities.c:375-390, src/boundaries.c:57-92): the staged slot sequence in the Philox
collision mode, and the BCFunc slot on the host arrays the reference's caller passes."""
import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, pic, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _pair(prec, N=16, ppc=16, seed=3):
    cli = configs.base_cli(N, nui=0.05, Tn_over_Ti=1.0, E0=0.05)
    p = pic.setup_context(**cli)
    solve = _lib.SOLVE_MIXED if prec == _lib.PARTICLES_MIXED else _lib.SOLVE_FP64
    opts = dict(collide_mode=_lib.COLLIDE_PHILOX, particle_precision=prec, solve_precision=solve, ksp_rtol=1e-10,
                operator_mode=_lib.OPERATOR_REBUILD)
    rng = np.random.default_rng(seed)
    n = ppc * N ** 3
    pos = rng.uniform(0, N, (n, 3))
    vel = rng.normal(size=(n, 3))
    ctxs = []
    for _ in range(2):
        c = pic.Context(p, pic.default_options(**opts))
        c.set_particles(pos, vel)
        c.collect()
        c.compute_potential()
        ctxs.append(c)
    return ctxs


@pytest.mark.parametrize("prec", [_lib.PARTICLES_FP64, _lib.PARTICLES_MIXED])
def test_staged_philox_slots_equal_fused_step(prec):
    """push -> collide -> UpdatePositions -> collect -> ComputePotential through the slots
    equals ecsprt_hip_step: the same Philox draws, the same collision counts, and on FP64
    ions the same velocities and positions bit for bit (mixed ions: the two kernels' FP32
    scattering may contract differently, so within FP32 rounding).  The first step is
    fused on both contexts, so the attempt rate starts from the same pilot."""
    say(f"staged Philox slots vs fused step, precision {prec}")
    fused, staged = _pair(prec)
    fused.step(1.0)
    staged.step(1.0)
    for step in range(3):
        fused.step(1.0)
        if step == 0:
            staged.update_velocities(1.0)  # UpdateVelocities: push, then collide (mn > 0)
        else:
            staged.push(1.0)
            staged.collide(1.0)
        staged.update_positions(1.0)
        staged.collect()
        staged.compute_potential()
        sf, ss = fused.collision_stats(), staged.collision_stats()
        say(f"  step {step}: fused (Nc, Ns, Nf, attempts) {sf}, staged {ss}")
        assert ss[0] > 0 and ss[1] > 0
        pf, vf = fused.get_particles()
        ps, vs = staged.get_particles()
        if prec == _lib.PARTICLES_FP64:
            assert sf == ss
            np.testing.assert_array_equal(vs, vf)
            np.testing.assert_array_equal(ps, pf)
            mf = fused.get_moments()
            assert np.abs(staged.get_moments() - mf).max() <= 1e-12 * np.abs(mf).max()
        else:
            assert abs(ss[1] - sf[1]) <= 0.01 * sf[1] + 5
            differ = np.any(np.abs(vs - vf) > 1e-5 * np.abs(vf).max(), axis=1).mean()
            say(f"  ions whose velocity differs beyond FP32 rounding: {differ:.2e}")
            assert differ <= 2e-3
    fused.close()
    staged.close()


def test_collide_slot_off_mode_and_counts():
    """With the collide mode off the slot does nothing; with mn = 0 UpdateVelocities does
    not collide (src/velocities.c:384-386)."""
    cli = configs.base_cli(12, nui=0.05)
    p = pic.setup_context(**cli)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 12, (12 ** 3 * 4, 3))
    vel = rng.normal(size=pos.shape)
    c = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_OFF))
    c.set_particles(pos, vel)
    c.collide(1.0)
    _, v = c.get_particles()
    np.testing.assert_array_equal(v, vel)
    c.close()
    cli0 = dict(cli, mn=0.0)
    c = pic.Context(pic.setup_context(**cli0), pic.default_options(collide_mode=_lib.COLLIDE_PHILOX))
    c.set_particles(pos, vel)
    c.set_potential(np.zeros((12, 12, 12)))
    c.update_velocities(1.0)
    assert c.collision_stats()[1] == 0
    c.close()


def test_pointer_classification():
    """Host memory (pageable numpy) is not device memory; a device allocation is."""
    host = np.zeros(1024)
    assert not pic.pointer_is_device(host.ctypes.data)
    dv = pic.DeviceVector(1024)
    assert pic.pointer_is_device(dv.ptr)
    dv.free()


def test_apply_bc_slot_on_host_arrays():
    """BCFunc with the swarm's host array, as src/boundaries.c:67-80 passes it: staged
    through the device, equal to ApplyBC on the oracle, including the edge cases."""
    say("apply_bc on host arrays")
    cli = configs.base_cli(10)
    cli.update(dx=0.5, dz=0.25)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    c = pic.Context(p, pic.default_options(tile=4))
    rng = np.random.default_rng(5)
    L = np.array([po.Lx, po.Ly, po.Lz])
    pos = rng.uniform(-0.9, 1.9, (5000, 3)) * L
    edges = np.array([[L[0], 0.0, -1e-17], [L[0] - 1e-15, L[1], L[2]], [-L[0] / 2, 1.5 * L[1], 0.0]])
    pos = np.ascontiguousarray(np.vstack([pos, edges]))
    # never hand a host array to the slot unless the library classifies it as host memory
    assert not pic.pointer_is_device(pos.ctypes.data)
    ref = pos.copy()
    oracle.apply_bc(po, ref)
    c.apply_bc_host(pos)
    np.testing.assert_array_equal(pos, ref)
    c.close()
