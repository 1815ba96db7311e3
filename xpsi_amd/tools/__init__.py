from .energy_integrator import energy_integrator  # noqa: F401
