"""Integrator protocol (integrators/integrator.py:4-23)."""
import numpy as np

from ..misc import temperature_schedule


class Integrator:
    """An integrator describes one timestep of the fused step kernel.

    ``get_params`` returns the descriptor fields of the C ABI (gmd_desc) that
    select and parameterise the HIP integrator template.
    """

    kind = None

    def get_params(self, configuration, interactions_params, verbose=False):
        raise NotImplementedError

    def get_kernel(self, configuration, compute_plan, compute_flags, interactions_kernel, verbose=False):
        return {"integrator": self.kind}

    # hooks used by Simulation around each timeblock
    def upload_state(self, ctx):
        pass

    def download_state(self, ctx):
        pass


def schedule_fields(temperature):
    kind, params = temperature_schedule(temperature)
    return {"tfunc": kind, "tparams": params}
