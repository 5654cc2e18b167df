"""lagrangian-microbes hot path on AMD Instinct MI355X (gfx950).

RK4 advection, fixed-radius pair search and the rock-paper-scissors update as
HIP kernels (liblm.so, C ABI in include/lm.h) behind the reference's Python
API: ParticleAdvecter, InteractionSimulator, rock_paper_scissors.
"""
from .interactions import (PAPER, PAPER_COLOR, ROCK, ROCK_COLOR, SCISSORS, SCISSORS_COLOR,
                           rock_paper_scissors, rock_paper_scissors_interaction)
from .particle_advecter import ParticleAdvecter, distribute_particles_across_tiles, uniform_particle_locations
from .interaction_simulator import InteractionSimulator
from .simulator import FusedRunner, MicrobeSimulator
from .velocity_fields import VelocityField, oscar_dataset, synthetic_fourier_field
from .utils import factor, most_symmetric_integer_factorization, pretty_filesize, pretty_time

__all__ = [
    "ROCK", "PAPER", "SCISSORS", "ROCK_COLOR", "PAPER_COLOR", "SCISSORS_COLOR",
    "rock_paper_scissors", "rock_paper_scissors_interaction",
    "ParticleAdvecter", "uniform_particle_locations", "distribute_particles_across_tiles",
    "InteractionSimulator", "MicrobeSimulator", "FusedRunner",
    "VelocityField", "synthetic_fourier_field", "oscar_dataset",
    "factor", "most_symmetric_integer_factorization", "pretty_time", "pretty_filesize",
]
