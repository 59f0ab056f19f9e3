"""Equation of state (src/eos.jl:1-22): ideal-gas polytrope, p = (gamma - 1) u."""
import math


class Polytrope:
    """Verona.EosTypes.Polytrope{T}(gamma) (src/eos.jl:7-9)."""

    def __init__(self, gamma):
        self.gamma = float(gamma)

    def __repr__(self):
        return f"Polytrope({self.gamma!r})"


def Pressure(u, eos):
    """src/eos.jl:11-13"""
    return (eos.gamma - 1) * u


def SoundSpeed(rho, u, eos):
    """src/eos.jl:15-17  c_s = sqrt(gamma (gamma - 1) u / (rho + gamma u))."""
    g = eos.gamma if isinstance(eos, Polytrope) else float(eos)
    return math.sqrt((g * (g - 1) * u) / (rho + g * u))
