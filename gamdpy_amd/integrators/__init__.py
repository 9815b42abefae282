"""Integrators (integrators/ of the reference): NVE, NVT (Nose-Hoover), NVT_Langevin."""
from .integrator import Integrator
from .NVE import NVE
from .NVT import NVT
from .NVT_Langevin import NVT_Langevin
