"""MI355X-native explicit-dynamics internal-force path of Akantu's
SolidMechanicsModel (central difference, lumped mass), on hand-written gfx950
HIP kernels behind a C ABI (include/akantu_mi355x.h)."""
from . import bc as BC
from .mesh import (HEX8, QUAD4, TET4, TRI3, BoxMesh, box_mesh, clamped_face_x0, face_group,
                   initial_velocity)
from .model import (Material, MaterialFactory, SolidMechanicsModel, _explicit_lumped_mass,
                    _ghost, _not_ghost, _pat_modifiable, _pat_parsable, _pat_readable,
                    _pat_writable)

__all__ = ["BoxMesh", "box_mesh", "clamped_face_x0", "face_group", "initial_velocity",
           "SolidMechanicsModel", "Material", "MaterialFactory", "_explicit_lumped_mass",
           "_not_ghost", "_ghost", "_pat_parsable", "_pat_readable", "_pat_writable",
           "_pat_modifiable", "TET4", "HEX8", "TRI3", "QUAD4", "BC"]
