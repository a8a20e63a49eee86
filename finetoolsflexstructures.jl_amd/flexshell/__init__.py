"""flexshell -- MI355X (gfx950) shell finite-element hot path of
FinEtoolsFlexStructures.jl: T3FF / Q4RS (and layered) stiffness scattered into
CSR, lumped mass, and the explicit central-difference loop, behind the
reference's FEMM / assembler interface.  All numeric work runs in
libflexshell.so on the GPU."""
from . import _lib, fe, material, femm, explicit  # noqa: F401
from ._lib import Context, FlexShellError, load  # noqa: F401
from .femm import (FEMMShellT3FF, FEMMShellQ4RS, FEMMShellT3FFComp, FEMMShellQ4RSComp,  # noqa: F401
                   SysmatAssemblerSparseSymm, SysmatAssemblerSparseCSRSymm, SysmatAssemblerSparseDiag,
                   SysmatAssemblerFFBlock, SysmatAssemblerHIP, make)
from .material import (MatDeforElastIso, MatDeforElastOrtho, Ply, CompositeLayup, lamina_material,  # noqa: F401
                       cartesian_csys, shell_material_stiffness)
