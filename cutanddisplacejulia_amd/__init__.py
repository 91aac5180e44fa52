"""MI355X (gfx950) engine for CutAndDisplaceJulia's triangular-dislocation path.

`api` mirrors the reference's TD / CalculateInfluenceMatrices3D / SlipCalculator3D interface on top of the C ABI
in include/cdj.h (libcdjhip.so); `meshes` builds the seeded benchmark geometries.
"""
from . import api, meshes  # noqa: F401
from .api import (TD, CalculateInfluenceMatrices3D, SlipCalculator3D, SetupTractionVector,  # noqa: F401
                  CreateFaceNormalAndMidPoint, ElasticConstantsCheck, Stresses, Strains, Disps, Tractions,
                  MixedBoundaryConditions, TractionInf, StrainInf, DispInf, set_precision, solve)
