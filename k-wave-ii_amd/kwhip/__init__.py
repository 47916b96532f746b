"""kwhip: MI355X-native k-space pseudospectral time stepping with the call
signature of k-Wave's kspaceFirstOrder1D / kspaceFirstOrder2D /
kspaceFirstOrder3D."""

from ._abi import KWError, load
from .engine import Engine, Ticket
from .kgrid import kWaveGrid
from .sim import Simulation, launch_tiles
from .solver import (cart2grid, kspaceFirstOrder1D, kspaceFirstOrder2D, kspaceFirstOrder3D,
                     kspaceFirstOrderAS, prepare, smooth)

__all__ = ["kWaveGrid", "kspaceFirstOrder1D", "kspaceFirstOrder2D", "kspaceFirstOrder3D",
           "kspaceFirstOrderAS",
           "Simulation", "launch_tiles", "Engine", "Ticket", "prepare", "smooth", "cart2grid", "KWError", "load"]
