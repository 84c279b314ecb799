"""Autotuner candidate selection (host logic) and a tuning pass on a small grid."""
import pytest

from ecsprt_amd import autotune, configs, pic, _lib


class P:
    def __init__(self, n):
        self.Nx = self.Ny = self.Nz = n


def test_candidates_fit_the_grid():
    assert autotune.valid_candidates(P(128)) == autotune.CANDIDATES
    small = autotune.valid_candidates(P(6))
    assert small and all(t == 4 and tz == 4 for t, tz, _ in small)
    assert all(b % 64 == 0 and 128 <= b <= 1024 for _, _, b in autotune.CANDIDATES)


@pytest.mark.gpu
def test_tune_particles_small_grid():
    wl = configs.small(16, 20)
    p = pic.setup_context(**wl.cli)
    for prec in (_lib.PARTICLES_FP64, _lib.PARTICLES_MIXED):
        o, table = autotune.tune_particles(p, pic.default_options(particle_precision=prec, ksp_rtol=1e-8), wl.np,
                                           steps=2)
        assert set(table) == set(autotune.valid_candidates(p))
        assert (o.tile, o.tile_z, o.block) == min(table, key=table.get)
        assert o.particle_precision == prec
        assert all(ms > 0 for ms in table.values())
