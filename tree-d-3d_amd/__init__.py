"""tree-d-3d on MI355X: the field-evaluation path of tree-d-3d (direct sums, flat
quadtree, Barnes-Hut walk and FMM) as hand-written gfx950 HIP kernels behind a C ABI.

The Python surface mirrors the reference's ``tree_d_3d`` (flat_tree.py, quad.py,
constants.py, smooth.py): ``direct_field``, ``FlatQuadtree``, ``FlatNode``, ``Quad``,
``K_E``, and the module ``smooth`` (``BoxTree``, ``SmoothFMM``, ``direct_energy``,
``direct_fields``, ``comoving_fields``).
Import it as ``tree_d_3d_amd`` (see ``__graft_entry__.load_package``).
"""

from .constants import K_E
from .flat_tree import MAX_ORDER, FlatNode, FlatQuadtree, direct_field
from .quad import Quad
from .simulation import ChargedBody, ElectrostaticSimulation
from .adjoint import log_energy, log_fields, log_potentials
from .ensemble import Ensemble
from .repeated import RepeatedSums
from .strip import GroundedStripField
from . import smooth

__all__ = ["K_E", "MAX_ORDER", "ChargedBody", "Ensemble", "ElectrostaticSimulation", "FlatNode",
           "FlatQuadtree", "GroundedStripField", "Quad", "RepeatedSums", "direct_field", "log_energy", "log_fields", "log_potentials",
           "smooth"]
