import numpy as np

from .integrator import Integrator, schedule_fields


class NVT_Langevin(Integrator):
    """Gronbech-Jensen/Farago Langevin leap-frog (integrators/NVT_Langevin.py:10-138).

    The noise comes from a counter-based generator (Philox-4x32-10 keyed by
    ``seed``, counter = particle index and step), so a new Simulation replays
    the same noise from the seed, as the reference does by re-creating its
    generator states in get_params.
    """
    kind = "NVT_Langevin"

    def __init__(self, temperature, alpha: float, dt: float, seed: int) -> None:
        self.temperature = temperature
        self.alpha = alpha
        self.dt = dt
        self.seed = seed

    def get_params(self, configuration, interactions_params, verbose=False):
        return {"integrator": 3, "dt": np.float32(self.dt), "alpha": np.float32(self.alpha),
                "seed": int(self.seed) & 0xFFFFFFFFFFFFFFFF, **schedule_fields(self.temperature)}
