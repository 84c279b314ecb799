"""Workload definitions in the reference's command-line parameters (src/parameters.c).

Ions are in grid units (q = m = 1, dx = 1, dt = 1), so the ion gyrofrequency
times dt is qi*B0/mi*dt and the thermal speed is viT cells per step.  The
electron fluid is fixed by kappa_e = Omega_e/nu_e and by Te: Te = Q/KB makes the
pressure coefficient cth = KB*Te/Q of ComputeFullRHS3D equal to 1.

Neutral temperature sets vnT = sqrt(2*KB*Tn/mn) (src/setup.c:268) and with it the
reference's collision count Nc = np*nu*dt*vrm/((vnT+vn0)*sqrt(mn/mi)).
"""
from dataclasses import dataclass, field

Q = 1.6022e-19
ME = 9.1094e-31
KB = 1.3807e-23


def electron_cli(B0, kappa_e):
    """nu_e giving kappa_e = Omega_e/nu_e; an unmagnetised case (B0 = 0) keeps nu_e = 1."""
    omega_e = Q * B0 / ME
    nue = omega_e / kappa_e if omega_e > 0 and kappa_e > 0 else 1.0
    return {"nue": nue, "Te": Q / KB}


def ion_temperature(viT, mi=1.0):
    """Ti that SetUpContext derives from per-component thermal speed viT (src/setup.c:250-252)."""
    return (0.5 * mi / KB) * (3.0 * viT * viT)


@dataclass
class Workload:
    name: str
    cli: dict
    ppc: int            # ions per cell
    dt: float
    steps: int
    rtol: float
    amp: float = 0.0    # x-density perturbation 1 + amp*sin(2 pi k x / Lx)
    k: int = 0
    notes: str = ""
    options: dict = field(default_factory=dict)

    @property
    def ncell(self):
        return self.cli["Nx"] * self.cli["Ny"] * self.cli["Nz"]

    @property
    def np(self):
        return self.ppc * self.ncell


def base_cli(N, B0=0.05, kappa_e=4.0, nui=0.01, Tn_over_Ti=1.0, E0=0.0, viT=1.0):
    Ti = ion_temperature(viT)
    cli = dict(Nx=N, Ny=N, Nz=N, dx=1.0, dy=1.0, dz=1.0, qi=1.0, mi=1.0, mn=1.0, B0=B0, E0=E0,
               nui=nui, viTx=viT, viTy=viT, viTz=viT, Tn=Tn_over_Ti * Ti, n0=1.0)
    cli.update(electron_cli(B0, kappa_e))
    return cli


# Plain FP64 hybrid PIC, CPU-runnable: 32^3, 64 ions/cell, Omega_i*dt = 0.05,
# nu*dt = 0.01 with a cold neutral background (Tn = Ti/100, vnT = 0.1*|vT|).
CPU32 = Workload("fp64_32cube_64ppc", base_cli(32, nui=0.01, Tn_over_Ti=0.01), ppc=64, dt=1.0, steps=10,
              rtol=1e-8)

# Plain FP64 at production size: 128^3, 100 ions/cell, E0 along x, k=4 density
# perturbation of 1e-3, nu*dt = 0.02 with Tn = Ti (the reference's default).
PROD128 = Workload("fp64_128cube_100ppc", base_cli(128, nui=0.02, Tn_over_Ti=1.0, E0=0.05), ppc=100, dt=1.0,
              steps=200, rtol=1e-10, amp=1e-3, k=4)


def small(N=12, ppc=8, **kw):
    """Small parity case (seconds on the oracle)."""
    args = dict(nui=0.01, Tn_over_Ti=1.0, E0=0.05)
    args.update(kw)
    return Workload(f"parity_{N}", base_cli(N, **args), ppc=ppc, dt=1.0, steps=3, rtol=1e-10)


def ensemble_runs(n=16, seed=4):
    """The 16-run ensemble: per run a cubic grid edge from {48, 64, 80, 96}, 16-64 ions per
    cell, Omega_i dt in [0.02, 0.1], nu dt in [0.005, 0.05], Philox seed r + 1."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [dict(N=int(rng.choice([48, 64, 80, 96])), ppc=int(rng.integers(16, 65)), B0=float(rng.uniform(0.02, 0.1)),
                 nui=float(rng.uniform(0.005, 0.05)), seed=r + 1) for r in range(n)]


def ensemble_cli(run):
    """Command-line parameters of one ensemble run (E0 = 0.05 along x, as the 128^3 case)."""
    return base_cli(run["N"], B0=run["B0"], nui=run["nui"], E0=0.05)


def batch_dims(nprob=256, seed=3):
    """Grid shapes of the batched standalone solves: each edge drawn uniformly from
    {32, 48, 64, 96, 128, 192}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice([32, 48, 64, 96, 128, 192], size=3)) for _ in range(nprob)]
