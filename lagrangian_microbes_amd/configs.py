"""Named workloads: particle lattice, synthetic field, radius and game settings.

Every input is generated from a seed, so a workload is reproducible anywhere.
Choices the workload descriptions leave open are fixed here:
- "1 km" is 1000 m / (1852*60 m per degree) = 0.0089993 degrees, applied in
  degree space as the reference's cKDTree query does (the radius is a plain
  lon/lat distance, interaction_simulator.py:98);
- patches use 0-360 degrees east longitude (150 W = 210 E) as the reference's
  example does (rock_paper_scissors_example.py:22);
- the 10-degree patches are the reference's canonical 25-35 N, 205-215 E
  (rock_paper_scissors_example.py:22, sandbox/rock_paper_scissors_advect.py);
- fields extend a margin beyond the patch (2 degrees for the 1-degree patch,
  10 degrees for the 60-day run, 5 degrees otherwise) so that no particle
  reaches the field edge;
- the rock-paper-scissors key is (species seed, run index).
"""
import math
from dataclasses import dataclass, field as dc_field

import numpy as np

from .interactions import PAPER, ROCK, SCISSORS
from .particle_advecter import uniform_particle_locations
from .velocity_fields import synthetic_fourier_field

KM_DEG = 1000.0 / (1852.0 * 60.0)  # 1 km as degrees of arc
HOUR = 3600.0
DAY = 86400.0


@dataclass
class Workload:
    name: str
    n: int
    lat_range: tuple
    lon_range: tuple
    field_spacing: float
    field_margin: float
    field_levels: int
    field_level_interval: float  # seconds (ignored for a steady field)
    field_modes: int
    field_rms: float
    field_seed: int
    dt: float                     # seconds
    steps: int
    radius: float = KM_DEG
    norm: float = 2.0
    p: float = 0.5
    species_seed: int = 0
    run: int = 0
    precision: str = "f64"
    extra: dict = dc_field(default_factory=dict)

    @property
    def key(self):
        return (self.species_seed, self.run)

    @property
    def params(self):
        return {"pRS": self.p, "pPR": self.p, "pSP": self.p}

    def particles(self):
        return uniform_particle_locations(self.n, self.lat_range[0], self.lat_range[1], self.lon_range[0],
                                          self.lon_range[1])

    def species(self):
        return np.random.RandomState(self.species_seed).choice([ROCK, PAPER, SCISSORS], self.n).astype(np.int8)

    def velocity_field(self):
        m = self.field_margin
        return synthetic_fourier_field(self.lon_range[0] - m, self.lon_range[1] + m, self.lat_range[0] - m,
                                       self.lat_range[1] + m, self.field_spacing, self.field_modes, self.field_rms,
                                       self.field_seed, self.field_levels, self.field_level_interval)

    def target_velocity_field(self):
        """The second seeded field whose trajectories are an adjoint run's targets."""
        m = self.field_margin
        return synthetic_fourier_field(self.lon_range[0] - m, self.lon_range[1] + m, self.lat_range[0] - m,
                                       self.lat_range[1] + m, self.field_spacing, self.field_modes, self.field_rms,
                                       self.extra["target_seed"], self.field_levels, self.field_level_interval)

    def describe(self):
        return {"workload": self.name, "particles": self.n, "steps": self.steps, "dt_s": self.dt,
                "radius_deg": self.radius, "p": self.p, "precision": self.precision,
                "field": f"{self.field_levels} level(s) on {self.field_spacing:g} deg, "
                         f"{self.field_modes} modes, {self.field_rms:g} m/s rms"}


def rps_10k_steady():
    """10,000 microbes on a 100x100 lattice, 1x1 degree patch at 35 N 150 W, steady field, 24 h."""
    return Workload(name="rps_10k_steady_24h", n=10_000, lat_range=(34.5, 35.5), lon_range=(209.5, 210.5),
                    field_spacing=0.25, field_margin=2.0, field_levels=1, field_level_interval=0.0,
                    field_modes=8, field_rms=0.2, field_seed=0, dt=HOUR, steps=24, p=0.5, species_seed=0)


def rps_490k_30day():
    """490,000 microbes on 700x700 over 10x10 degrees, 7 field levels 5 days apart, 720 h."""
    return Workload(name="rps_490k_30day", n=490_000, lat_range=(25.0, 35.0), lon_range=(205.0, 215.0),
                    field_spacing=0.25, field_margin=5.0, field_levels=7, field_level_interval=5 * DAY,
                    field_modes=16, field_rms=0.25, field_seed=1, dt=HOUR, steps=720, p=0.55, species_seed=1)


def rps_4m_60day():
    """4,000,000 microbes on 2000x2000 over 20x20 degrees, 1/12 degree field, 21 levels 3 days apart.
    The field extends 10 degrees beyond the patch: over 60 days the eddies carry some particles more
    than 5 degrees."""
    return Workload(name="rps_4m_60day", n=4_000_000, lat_range=(20.0, 40.0), lon_range=(200.0, 220.0),
                    field_spacing=1.0 / 12.0, field_margin=10.0, field_levels=21, field_level_interval=3 * DAY,
                    field_modes=32, field_rms=0.3, field_seed=2, dt=0.5 * HOUR, steps=2880, p=0.6,
                    species_seed=2, precision="f32")


def adjoint_1m_10day():
    """1,000,000 particles on 1000x1000 over 10x10 degrees, 20 float32 field levels 13 h apart, 240 h;
    targets from a second field seeded 105."""
    return Workload(name="adjoint_1m_10day", n=1_000_000, lat_range=(25.0, 35.0), lon_range=(205.0, 215.0),
                    field_spacing=0.25, field_margin=5.0, field_levels=20, field_level_interval=13 * HOUR,
                    field_modes=16, field_rms=0.25, field_seed=5, dt=HOUR, steps=240, p=0.5, species_seed=5,
                    precision="f32 field, f64 positions", extra={"target_seed": 105})


def adjoint_10k_24h():
    """The same field and patch with 10,000 particles (100x100) and 24 steps: the size of the
    autograd cross-check."""
    w = adjoint_1m_10day()
    w.name, w.n, w.steps = "adjoint_10k_24h", 10_000, 24
    return w


WORKLOADS = {w().name: w for w in (rps_10k_steady, rps_490k_30day, rps_4m_60day, adjoint_1m_10day,
                                   adjoint_10k_24h)}


@dataclass
class EnsembleWorkload:
    """Independent runs sharing one velocity field: each run is a Workload
    (its lattice, species seed, p and Philox key) and all of them read
    ``field_of``'s field."""
    name: str
    runs: list
    field_of: Workload
    steps: int

    @property
    def n(self):
        return sum(r.n for r in self.runs)

    def velocity_field(self):
        return self.field_of.velocity_field()

    def particles(self):
        parts = [r.particles() for r in self.runs]
        return np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])

    def species(self):
        return np.concatenate([r.species() for r in self.runs])

    def describe(self):
        sizes = [r.n for r in self.runs]
        return {"workload": self.name, "runs": len(self.runs), "particles": self.n, "min_run": min(sizes),
                "max_run": max(sizes), "steps": self.steps, "p_range": [self.runs[0].p, self.runs[-1].p],
                "field": self.field_of.describe()["field"]}


ENSEMBLE_CENTRES = ((210.0, 30.0), (207.5, 27.5), (212.5, 27.5), (207.5, 32.5), (212.5, 32.5))


def ensemble_runs(sizes, steps, name, spacing=10.0 / 699.0):
    """Runs on square lattices of side m = sqrt(n) with the given spacing
    (default the 490k run's, 10/699 degrees), centred in turn on
    ENSEMBLE_CENTRES; p from 0.50 to 0.70 evenly over the runs; species seed
    100 + r, Philox key (100 + r, r); the 490k run's 30-day field, hourly
    steps, 1 km, FP32."""
    base = rps_490k_30day()
    runs = []
    R = len(sizes)
    for r, n in enumerate(sizes):
        m = int(round(math.sqrt(n)))
        half = 0.5 * (m - 1) * spacing
        lon_c, lat_c = ENSEMBLE_CENTRES[r % len(ENSEMBLE_CENTRES)]
        p = 0.50 + 0.20 * r / (R - 1) if R > 1 else 0.50
        runs.append(Workload(name=f"{name}_run{r}", n=m * m, lat_range=(lat_c - half, lat_c + half),
                             lon_range=(lon_c - half, lon_c + half), field_spacing=base.field_spacing,
                             field_margin=base.field_margin, field_levels=base.field_levels,
                             field_level_interval=base.field_level_interval, field_modes=base.field_modes,
                             field_rms=base.field_rms, field_seed=base.field_seed, dt=HOUR, steps=steps, p=p,
                             species_seed=100 + r, run=r, precision="f32"))
    return EnsembleWorkload(name=name, runs=runs, field_of=base, steps=steps)


def ensemble_64runs():
    """64 runs, particle counts log-uniform in [50k, 500k] (seed 4) rounded to
    square lattices, 240 hourly steps."""
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(math.log(50_000), math.log(500_000), 64))
    return ensemble_runs(sizes, 240, "ensemble_64runs_10day")


def ensemble_small():
    """Five small runs of uneven size for the per-run parity tests, on lattices
    finer than the radius (0.007 degrees) so that they interact from the
    first step."""
    return ensemble_runs([400, 1225, 2500, 1681, 784], 24, "ensemble_5runs_1day", spacing=0.007)
