"""elfel_hip -- Elfel.jl's element-integration and global-assembly path on MI355X.

Host-side mirror of the reference's API (RefShapes, FElements, FESpaces,
FEIterators/QPIterators data, LocalAssemblers, Assemblers) over the C ABI of
libelfel_hip.so; the element loop, the sparse build and the solve run on the
GPU.  There is no CPU fallback for any of them.
"""
from . import _lib
from .assemblers import (SysmatAssemblerHIP, SysvecAssemblerHIP, assemble, assemble_init, default_context, finish,
                         makematrix, makevector, start)
from .felements import FEH1_L2, FEH1_Q4, FEH1_T3, FEH1_T3_BUBBLE, FEH1_T4, FEH1_T6, FEL2_Q4
from .fespaces import FESpace, gathersysvec, ndofs, numberdofs, nunknowns, scattersysvec
from .integrate import (integrate_elasticity, integrate_heat, integrate_mass, integrate_stokes, integrate_stokes_reddy,
                        plane_strain_D, plane_stress_D)
from .localassemblers import LocalMatrixAssembler, LocalVectorAssembler, init
from .meshes import L2block, Mesh, Q4block, T3block, T4block, T6block, T6toT3, load
from .refshapes import quadrature

# Elfel-style aliases: the HIP assemblers are drop-ins for these.
SysmatAssemblerSparse = SysmatAssemblerHIP
SysvecAssembler = SysvecAssemblerHIP

__all__ = [
    "SysmatAssemblerHIP", "SysvecAssemblerHIP", "SysmatAssemblerSparse", "SysvecAssembler", "start", "assemble",
    "finish", "assemble_init", "makematrix", "makevector", "default_context", "FEH1_T3", "FEH1_T6", "FEH1_Q4",
    "FEH1_L2", "FEH1_T3_BUBBLE", "FEH1_T4", "FEL2_Q4", "integrate_mass", "integrate_stokes_reddy", "L2block", "T4block",
    "FESpace", "numberdofs", "gathersysvec", "scattersysvec", "ndofs", "nunknowns", "integrate_heat",
    "integrate_elasticity", "integrate_stokes", "plane_strain_D", "plane_stress_D", "LocalMatrixAssembler",
    "LocalVectorAssembler", "init", "Mesh", "T3block", "T6block", "Q4block", "T6toT3", "load", "quadrature", "_lib",
]
