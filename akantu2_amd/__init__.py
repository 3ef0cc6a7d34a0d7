"""MI355X-native explicit-dynamics internal-force path of Akantu's
SolidMechanicsModel (central difference, lumped mass), on hand-written gfx950
HIP kernels behind a C ABI (include/akantu_mi355x.h)."""
from .mesh import HEX8, TET4, BoxMesh, box_mesh, clamped_face_x0, initial_velocity
from .model import Material, SolidMechanicsModel, _explicit_lumped_mass

__all__ = ["BoxMesh", "box_mesh", "clamped_face_x0", "initial_velocity",
           "SolidMechanicsModel", "Material", "_explicit_lumped_mass", "TET4", "HEX8"]
