"""MI355X (gfx950) crystal-plasticity Newton step: the PRISMS-Plasticity crystalPlasticity<3> /
ellipticBVP<3> hot path as hand-written HIP kernels behind a C ABI (include/plasticity_amd.h)."""
from .native import PlasticityError, lib
from .model import (Batch, Context, crystalOrientationsIO, crystalPlasticity, main, make_material, make_problem,
                    userInputParameters, write_stressstrain)

__all__ = ["PlasticityError", "lib", "Batch", "Context", "crystalOrientationsIO", "crystalPlasticity", "main", "make_material",
           "make_problem", "userInputParameters", "write_stressstrain"]
