"""h2_amd: MI355X-native (gfx950) linear-solve path of Heat3ds (kenoogl/H2).

The compute lives in libh3d.so (HIP kernels + C ABI, include/h3d.h); this
package is the host-side mirror of the reference's q3d/PBiCGSTAB!/CG!
interface on top of it.
"""
from . import _lib
from .heat3ds import (BoundaryCondition, BoundaryConditionSet, CG, PBiCGSTAB, WorkBuffers, adiabatic_bc,
                      build_model, calRHS, convection_bc, create_boundary_conditions, genZ, heat_flux_bc,
                      isothermal_bc, load_model, pitch, q3d, set_BC_coef, set_mode3_bc_parameters)

__all__ = ["BoundaryCondition", "BoundaryConditionSet", "CG", "PBiCGSTAB", "WorkBuffers", "adiabatic_bc",
           "build_model", "calRHS", "convection_bc", "create_boundary_conditions", "genZ", "heat_flux_bc",
           "isothermal_bc", "load_model", "pitch", "q3d", "set_BC_coef", "set_mode3_bc_parameters"]
