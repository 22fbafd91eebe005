"""Physics specifications (jax_cfd/ml/physics_specifications.py)."""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional


@dataclasses.dataclass
class NavierStokesPhysicsSpecs:
  """Density, viscosity and the forcing module (grid -> forcing function)."""
  density: float
  viscosity: float
  forcing_module: Optional[Callable] = None
