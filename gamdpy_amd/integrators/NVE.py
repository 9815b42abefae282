import numpy as np

from .integrator import Integrator


class NVE(Integrator):
    """Leap-frog NVE (integrators/NVE.py:8-112).

    v(t+dt/2) = v(t-dt/2) + f(t) dt/m;  r(t+dt) = r(t) + v(t+dt/2) dt;
    K from the mean of the two half-step velocities.
    """
    kind = "NVE"

    def __init__(self, dt):
        self.dt = dt

    def get_params(self, configuration, interactions_params, verbose=False):
        return {"integrator": 1, "dt": np.float32(self.dt)}
