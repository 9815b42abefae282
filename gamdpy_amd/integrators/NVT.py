import numpy as np

from .integrator import Integrator, schedule_fields


class NVT(Integrator):
    """Leap-frog with a Nose-Hoover thermostat (integrators/NVT.py:8-141).

    Parameters
    ----------
    temperature : float or a function of time made by make_function_constant/ramp/sin
    tau : thermostat relaxation time
    dt : timestep
    """
    kind = "NVT"

    def __init__(self, temperature, tau: float, dt: float) -> None:
        self.temperature = temperature
        self.tau = tau
        self.dt = dt
        self.thermostat_state = np.zeros(2, dtype=np.float32)   # [xi, sum K] (NVT.py:143)

    def get_params(self, configuration, interactions_params, verbose=False):
        dt = np.float32(self.dt)
        omega2 = np.float32(4.0 * np.pi * np.pi / self.tau / self.tau)
        degrees = configuration.N * configuration.D - configuration.D
        return {"integrator": 2, "dt": dt, "omega2": omega2, "degrees": int(degrees),
                **schedule_fields(self.temperature)}

    def upload_state(self, ctx):
        ctx.set_thermostat(self.thermostat_state)

    def download_state(self, ctx):
        self.thermostat_state = ctx.get_thermostat()
