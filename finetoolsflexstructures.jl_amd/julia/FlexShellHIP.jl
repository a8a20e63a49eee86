"""
    FlexShellHIP

Drop-in GPU path for the shell FEMMs of FinEtoolsFlexStructures
(`FEMMShellT3FF`, `FEMMShellQ4RS`, `FEMMShellT3FFComp`, `FEMMShellQ4RSComp`)
on one AMD Instinct MI355X, through the C ABI of `libflexshell.so`
(`include/flexshell.h`).

The FEMM call sites stay as they are; only the assembler changes:

    associategeometry!(femm, geom0)                      # reference code, unchanged
    K  = stiffness(femm, SysmatAssemblerHIPCSR(), geom0, u1, Rfield1, dchi)
    Kf = stiffness(femm, SysmatAssemblerHIPCSR(; block = :ff), geom0, u1, Rfield1, dchi)
    M  = mass(femm, SysmatAssemblerHIPDiag(; block = :ff), geom0, dchi)

`stiffness` returns the same `SparseMatrixCSR{1,Float64,Int64}` that
`SysmatAssemblerSparseCSRSymm` produces (`AssemblyModule.jl:20-53`): identical
pattern (structural union of the element dof pairs, zeros retained), values
within 1e-12 relative.  The nodal normals computed by the reference
`associategeometry!` are uploaded, so custom `mcsys` closures keep working.

The explicit loop of `examples/shells/dynamics/homogeneous/explicit/
plate_expl_examples.jl:61-94` runs on the device with [`cd_loop!`](@ref).
"""
module FlexShellHIP

using SparseArrays
using SparseMatricesCSR
using FinEtools
using FinEtools.AssemblyModule: AbstractSysmatAssembler
import FinEtools.AssemblyModule: startassembly!, assemble!, makematrix!
using FinEtoolsFlexStructures.FEMMShellT3FFModule: FEMMShellT3FF
using FinEtoolsFlexStructures.FEMMShellQ4RSModule: FEMMShellQ4RS
using FinEtoolsFlexStructures.FEMMShellT3FFCompModule: FEMMShellT3FFComp
using FinEtoolsFlexStructures.FEMMShellQ4RSCompModule: FEMMShellQ4RSComp
using FinEtoolsFlexStructures.CompositeLayupModule: thickness, laminate_stiffnesses!,
    laminate_transverse_stiffness!, laminate_inertia!
import FinEtoolsFlexStructures.FEMMShellT3FFModule
import FinEtoolsFlexStructures.FEMMShellQ4RSModule
import FinEtoolsFlexStructures.FEMMShellT3FFCompModule
import FinEtoolsFlexStructures.FEMMShellQ4RSCompModule

const libflexshell = get(ENV, "FLEXSHELL_LIB", joinpath(@__DIR__, "..", "flexshell", "libflexshell.so"))

const FS_OK = Cint(0)
const FS_BLOCK_ALL = Cint(0)
const FS_BLOCK_FF = Cint(1)
const FS_PREC_FP64 = Int32(0)
const FS_PREC_MIXED = Int32(1)
const FS_PREC_MIXED_F32 = Int32(2)

# ---------------------------------------------------------------- ABI structs
# Field order and types follow include/flexshell.h exactly.
struct ShellParams               # fs_shell_params
    composite::Int32
    formulation::Int32
    thickness::Float64
    stab_c::Float64
    drilling_stiffness_scale::Float64
    Dps::NTuple{9,Float64}       # column-major 3 x 3
    Dt::NTuple{4,Float64}        # column-major 2 x 2 (without the 5/6 factor)
    ngroups::Int32
    reserved::Int32
    group_ABDH::Ptr{Float64}     # ngroups x 31: A 9 | B 9 | D 9 | H 4, column-major
    group_csys::Ptr{Float64}     # ngroups x 9, column-major
    group_thickness::Ptr{Float64}
    elem_group::Ptr{Int32}
    group_homog::Ptr{Float64}
end

struct MassParams                # fs_mass_params
    composite::Int32
    ngroups::Int32
    rho::Float64
    thickness::Float64
    group_inertia::Ptr{Float64}  # ngroups x 2
    elem_group::Ptr{Int32}
    group_rho_t::Ptr{Float64}
end

struct CDParams                  # fs_cd_params
    dt::Float64
    cdamp::Float64
    nsteps::Int64
    precision::Int32
    probe_stride::Int32
    ngroups::Int32
    use_graph::Int32
    row_group::Ptr{Int32}
    group_dt::Ptr{Float64}
    group_cdamp::Ptr{Float64}
    dt_mass::Float64               # dt of invMC (before the tend/dt adjustment)
end

# ---------------------------------------------------------------- handles
mutable struct Context
    h::Ptr{Cvoid}
    function Context(device::Integer = 0)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:fs_ctx_create, libflexshell), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, out)
        st == FS_OK || error("fs_ctx_create(device = $device) failed with status $st (a HIP device is required)")
        c = new(out[])
        finalizer(c -> ccall((:fs_ctx_destroy, libflexshell), Cvoid, (Ptr{Cvoid},), c.h), c)
        return c
    end
end

function check(ctx::Context, st::Cint, what::AbstractString)
    st == FS_OK && return nothing
    msg = unsafe_string(ccall((:fs_last_error, libflexshell), Cstring, (Ptr{Cvoid},), ctx.h))
    error("$what: status $st: $msg")   # the reference's error() convention
end

# Launch settings of a context (no counterpart in the reference): element
# ranges of the assembly pipeline, and the form of the explicit step's row
# products (0 = lanes per row; 8 or 16 lanes per node, positions per lane 0 =
# chosen from the longest row).  Every setting gives the same results up to
# the order of addition (bitwise the same for the pipeline).
set_pipeline!(ctx::Context, ranges::Integer) =
    check(ctx, ccall((:fs_ctx_set_pipeline, libflexshell), Cint, (Ptr{Cvoid}, Cint), ctx.h, ranges),
          "fs_ctx_set_pipeline")
set_cd_lanes!(ctx::Context, lanes_per_node::Integer, positions_per_lane::Integer = 0) =
    check(ctx, ccall((:fs_ctx_set_cd_lanes, libflexshell), Cint, (Ptr{Cvoid}, Cint, Cint), ctx.h,
                     lanes_per_node, positions_per_lane), "fs_ctx_set_cd_lanes")

# Assembly form of fs_model_stiffness: 1 = fused element tiles writing CSR
# values (default), 0 = element blocks + pair scatter (no tile plan to build:
# suits models assembled once).  Both give the same pattern; the values
# differ by rounding only (within 1e-13).
set_assembly!(ctx::Context, mode::Integer) =
    check(ctx, ccall((:fs_ctx_set_assembly, libflexshell), Cint, (Ptr{Cvoid}, Cint), ctx.h, mode),
          "fs_ctx_set_assembly")

const _default_ctx = Ref{Union{Nothing,Context}}(nothing)
default_context() = something(_default_ctx[], (_default_ctx[] = Context(parse(Int, get(ENV, "FLEXSHELL_DEVICE", "0")))))

mutable struct DeviceModel
    ctx::Context
    h::Ptr{Cvoid}
    nen::Int
    geom_key::Tuple{UInt,UInt,Int,Int}
    dofs_key::UInt
    normals_key::UInt
end

function _destroy!(m::DeviceModel)
    m.h == C_NULL || ccall((:fs_model_destroy, libflexshell), Cvoid, (Ptr{Cvoid},), m.h)
    m.h = C_NULL
end

# ---------------------------------------------------------------- assemblers
"""
    SysmatAssemblerHIPCSR(ctx = default_context(); block = :all)

Assembler whose `stiffness` runs on the GPU.  `block = :all` gives the
all-dof matrix of `SysmatAssemblerSparseCSRSymm`; `block = :ff` the free-free
block of `SysmatAssemblerFFBlock`.  The device model of each FEMM (mesh,
dofs, normals, CSR pattern, matrices) stays resident between calls.
"""
mutable struct SysmatAssemblerHIPCSR <: AbstractSysmatAssembler
    ctx::Context
    block::Symbol
    models::IdDict{Any,DeviceModel}
end
SysmatAssemblerHIPCSR(ctx::Context = default_context(); block::Symbol = :all) =
    SysmatAssemblerHIPCSR(ctx, block, IdDict{Any,DeviceModel}())

"""
    SysmatAssemblerHIPDiag(ctx = default_context(); block = :all)

Lumped-mass counterpart of `SysmatAssemblerSparseDiag` /
`SysmatAssemblerFFBlock(SysmatAssemblerSparseDiag(), nf, nf)`.
"""
mutable struct SysmatAssemblerHIPDiag <: AbstractSysmatAssembler
    ctx::Context
    block::Symbol
    models::IdDict{Any,DeviceModel}
end
SysmatAssemblerHIPDiag(ctx::Context = default_context(); block::Symbol = :all) =
    SysmatAssemblerHIPDiag(ctx, block, IdDict{Any,DeviceModel}())

_block(a) = a.block === :ff ? FS_BLOCK_FF : FS_BLOCK_ALL

"""
    SysmatAssemblerHIPGeneric(ctx = default_context())

The assembler protocol of `SysmatAssemblerSparseCSRSymm`
(`AssemblyModule.jl:28-52`) for any FEMM, whose element matrices stay on the
host: `assemble!` buffers each element matrix with its dof numbers,
`makematrix!` reduces the buffer on the device (`fs_coo_values`) into the
`SparseMatrixCSR{1}` that `SysmatAssemblerSparseCSRSymm` returns: the
pattern of `sparse(I, J, V)` with zeros retained, duplicates summed in
`assemble!` order.  Dof numbers beyond the `startassembly!` sizes are
dropped, so the sizes of the free-free block give the free-free matrix.
The reduction plan is kept while the dof numbers of a new assembly repeat
the previous ones.
"""
mutable struct SysmatAssemblerHIPGeneric <: AbstractSysmatAssembler
    ctx::Context
    sizes::NTuple{4,Int}              # kr, kc, nrows, ncols
    n::Int                            # matrices buffered
    mats::Vector{Float64}             # kr kc per matrix, column-major
    dr::Vector{Int64}                 # kr per matrix
    dc::Vector{Int64}                 # kc per matrix
    h::Ptr{Cvoid}                     # fs_coo handle of the current plan
    plan_key::Tuple{NTuple{4,Int},Int,Vector{Int64},Vector{Int64}}
    rowptr::Vector{Int64}
    colval::Vector{Int64}
    function SysmatAssemblerHIPGeneric(ctx::Context = default_context())
        a = new(ctx, (0, 0, 0, 0), 0, Float64[], Int64[], Int64[], C_NULL, ((0, 0, 0, 0), 0, Int64[], Int64[]),
                Int64[], Int64[])
        finalizer(_coo_destroy!, a)      # the handle holds its own reference to the context
        return a
    end
end

function _coo_destroy!(a::SysmatAssemblerHIPGeneric)
    a.h == C_NULL || ccall((:fs_coo_destroy, libflexshell), Cvoid, (Ptr{Cvoid},), a.h)
    a.h = C_NULL
end

function startassembly!(self::SysmatAssemblerHIPGeneric, elem_mat_nrows::Integer, elem_mat_ncols::Integer,
                        n_elem_mats::Integer, row_nalldofs::Integer, col_nalldofs::Integer; force_init = false)
    kr, kc = Int(elem_mat_nrows), Int(elem_mat_ncols)
    self.sizes = (kr, kc, Int(row_nalldofs), Int(col_nalldofs))
    self.n = 0
    self.mats = Vector{Float64}(undef, kr * kc * n_elem_mats)
    self.dr = Vector{Int64}(undef, kr * n_elem_mats)
    self.dc = Vector{Int64}(undef, kc * n_elem_mats)
    return self
end

function assemble!(self::SysmatAssemblerHIPGeneric, mat::AbstractMatrix, dofnums_row::AbstractVector,
                   dofnums_col::AbstractVector)
    kr, kc = self.sizes[1], self.sizes[2]
    size(mat) == (kr, kc) || error("assemble!: element matrix $(size(mat)), expected ($kr, $kc)")
    length(dofnums_row) == kr && length(dofnums_col) == kc || error("assemble!: dof numbers do not match the matrix")
    k = self.n
    (k + 1) * kr <= length(self.dr) || error("assemble!: more element matrices than startassembly! announced")
    copyto!(self.mats, k * kr * kc + 1, vec(mat), 1, kr * kc)
    copyto!(self.dr, k * kr + 1, dofnums_row, 1, kr)
    copyto!(self.dc, k * kc + 1, dofnums_col, 1, kc)
    self.n = k + 1
    return self
end

function makematrix!(self::SysmatAssemblerHIPGeneric)
    kr, kc, nrows, ncols = self.sizes
    n = self.n
    dr, dc = self.dr[1:(kr * n)], self.dc[1:(kc * n)]
    key = (self.sizes, n, dr, dc)
    ctx = self.ctx
    if self.h == C_NULL || key != self.plan_key
        _coo_destroy!(self)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ctx, ccall((:fs_coo_create, libflexshell), Cint,
                         (Ptr{Cvoid}, Int64, Int64, Int32, Int32, Int64, Ptr{Int64}, Ptr{Int64}, Cint, Ptr{Ptr{Cvoid}}),
                         ctx.h, nrows, ncols, kr, kc, n, dr, dc, 1, out), "fs_coo_create")
        self.h, self.plan_key = out[], key
        nz = Ref{Int64}(0)
        check(ctx, ccall((:fs_coo_pattern, libflexshell), Cint, (Ptr{Cvoid}, Ptr{Int64}), self.h, nz),
              "fs_coo_pattern")
        self.rowptr = Vector{Int64}(undef, nrows + 1)
        self.colval = Vector{Int64}(undef, nz[])
        check(ctx, ccall((:fs_coo_pattern_get, libflexshell), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Cint),
                         self.h, self.rowptr, self.colval, 1), "fs_coo_pattern_get")
    end
    nzval = Vector{Float64}(undef, length(self.colval))
    GC.@preserve self begin
        check(ctx, ccall((:fs_coo_values, libflexshell), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                         self.h, pointer(self.mats), nzval), "fs_coo_values")
    end
    return SparseMatrixCSR{1}(nrows, ncols, copy(self.rowptr), copy(self.colval), nzval)
end

# ---------------------------------------------------------------- model upload
_nen(::Union{FEMMShellT3FF,FEMMShellT3FFComp}) = 3
_nen(::Union{FEMMShellQ4RS,FEMMShellQ4RSComp}) = 4

# The device model of a FEMM is keyed by the identity and size of the
# coordinate and connectivity arrays (no pass over their contents per call).
# Coordinates changed in place need `invalidate!(assembler, femm)`; the
# dofs and the normals are compared by content, since numberdofs! and
# associategeometry! rewrite their arrays in place.
function _device_model!(a, self, geom0::NodalField{Float64}, dchi::NodalField)
    ctx = a.ctx
    nen = _nen(self)
    X = geom0.values                                   # N x 3, column-major (SoA)
    fconn = self.integdomain.fes.conn
    gkey = (UInt(objectid(X)), UInt(objectid(fconn)), size(X, 1), length(fconn))
    m = get(a.models, self, nothing)
    if m === nothing || m.geom_key != gkey
        m === nothing || _destroy!(m)
        conn = connasarray(self.integdomain.fes)       # E x nen, 1-based
        connrm = permutedims(conn)                     # row-major E x nen
        out = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve X connrm begin
            check(ctx, ccall((:fs_model_create, libflexshell), Cint,
                             (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Cint, Ptr{Ptr{Cvoid}}),
                             ctx.h, nen, size(X, 1), X, size(conn, 1), connrm, 1, out), "fs_model_create")
        end
        m = DeviceModel(ctx, out[], nen, gkey, UInt(0), UInt(0))
        finalizer(_destroy!, m)
        a.models[self] = m
    end
    dkey = hash(dchi.dofnums, hash(nfreedofs(dchi)))
    if m.dofs_key != dkey
        dn = dchi.dofnums                              # N x 6, column-major, 1-based
        GC.@preserve dn begin
            check(ctx, ccall((:fs_model_set_dofs, libflexshell), Cint,
                             (Ptr{Cvoid}, Ptr{Int64}, Int64, Int64, Cint),
                             m.h, dn, nalldofs(dchi), nfreedofs(dchi), 1), "fs_model_set_dofs")
        end
        m.dofs_key = dkey
    end
    # normals from the reference associategeometry! (any mcsys closure),
    # uploaded when they differ from the resident ones
    nkey = hash(self._normals, hash(self._normal_valid))
    if m.normals_key != nkey
        nrm = self._normals
        valid = UInt8.(self._normal_valid)
        GC.@preserve nrm valid begin
            check(ctx, ccall((:fs_model_normals_upload, libflexshell), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{UInt8}),
                             m.h, nrm, valid), "fs_model_normals_upload")
        end
        m.normals_key = nkey
    end
    return m
end

"""
    invalidate!(assembler, femm)

Drop the device model of `femm` (after its coordinates or connectivity were
changed in place); the next `stiffness`/`mass` call uploads the mesh again.
"""
function invalidate!(a, self)
    m = pop!(a.models, self, nothing)
    m === nothing || _destroy!(m)
    return a
end

# stab_fun(t, h) = t^2 / (t^2 + c h^2) crosses the ABI as c.  c is read at
# (t, h) = (1, 1) and the form is checked at three more points: a custom
# stab_fun of another form is refused rather than approximated.
function _stab_coefficient(f)
    c = 1 / f(1.0, 1.0) - 1.0
    for (t, h) in ((0.01, 0.1), (0.002, 0.5), (0.05, 0.013))
        v = f(t, h)
        ok = isapprox(v, t^2 / (t^2 + c * h^2); rtol = 1e-12)
        ok || error("FlexShellHIP: stab_fun(t, h) is not of the form t^2/(t^2 + c h^2) " *
                    "(c = $c from stab_fun(1, 1); stab_fun($t, $h) = $v)")
    end
    return c
end

# otherdimension evaluated at every element (FEMMShellT3FFModule.jl:677 and
# FEMMShellQ4RSModule.jl:924 per Gauss point); the device path takes
# one thickness per FEMM, so a thickness that varies is refused.
function _thickness(self)
    fes = self.integdomain.fes
    od = self.integdomain.otherdimension
    loc = [1/3 1/3 1/3]
    N = _nen(self) == 3 ? [1/3 1/3 1/3] : [1/4 1/4 1/4 1/4]
    t1 = od(loc, fes.conn[1], N)
    for e in 2:length(fes.conn)
        te = od(loc, fes.conn[e], N)
        te == t1 || error("FlexShellHIP: the thickness (otherdimension) differs between elements 1 and $e " *
                          "($t1, $te); the device path takes a constant thickness")
    end
    return t1
end

function _homog_params(self)
    Dps, Dt = FEMMShellT3FFModule._shell_material_stiffness(self.material)
    return ShellParams(0, self.transv_shear_formulation, _thickness(self), _stab_coefficient(self.stab_fun),
                       self.drilling_stiffness_scale, Tuple(vec(Dps)), Tuple(vec(Dt)), 0, 0,
                       C_NULL, C_NULL, C_NULL, C_NULL, C_NULL)
end

function _layup_arrays(self)
    ng = length(self.layup_groups)
    abdh = zeros(31, ng); csys = zeros(9, ng); tg = zeros(ng); inert = zeros(2, ng)
    for (g, (layup, _)) in enumerate(self.layup_groups)
        A = zeros(3, 3); B = zeros(3, 3); D = zeros(3, 3); H = zeros(2, 2)
        laminate_stiffnesses!(layup, A, B, D)
        laminate_transverse_stiffness!(layup, H)
        abdh[:, g] = vcat(vec(A), vec(B), vec(D), vec(H))
        # the device path takes one layup frame per group: a position-dependent
        # layup CSys (evaluated per quadrature point by the reference,
        # FEMMShellT3FFCompModule.jl:612, FEMMShellQ4RSCompModule.jl:929) is refused
        layup.csys.isconstant || error("FlexShellHIP: layup group $g has a position-dependent CSys; " *
                                       "the device path takes a constant layup CSys per group")
        csys[:, g] = vec(csmat(layup.csys))
        tg[g] = thickness(layup)
        inert[:, g] .= laminate_inertia!(layup)
    end
    grp = Int32.(self._layup_group_lookup .- 1)
    return abdh, csys, tg, inert, grp
end

# ---------------------------------------------------------------- stiffness
function _stiffness(self, assembler::SysmatAssemblerHIPCSR, geom0, dchi)
    @assert self._associatedgeometry == true
    m = _device_model!(assembler, self, geom0, dchi)
    ctx, block = assembler.ctx, _block(assembler)
    nr, nz = Ref{Int64}(0), Ref{Int64}(0)
    check(ctx, ccall((:fs_model_pattern, libflexshell), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}),
                     m.h, block, nr, nz), "fs_model_pattern")
    rowptr = Vector{Int64}(undef, nr[] + 1)
    colval = Vector{Int64}(undef, nz[])
    nzval = Vector{Float64}(undef, nz[])
    check(ctx, ccall((:fs_model_pattern_get, libflexshell), Cint,
                     (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Cint), m.h, block, rowptr, colval, 1),
          "fs_model_pattern_get")
    if self isa Union{FEMMShellT3FFComp,FEMMShellQ4RSComp}
        abdh, csys, tg, _, grp = _layup_arrays(self)
        GC.@preserve abdh csys tg grp begin
            p = Ref(ShellParams(1, self.transv_shear_formulation, 0.0, _stab_coefficient(self.stab_fun),
                                self.drilling_stiffness_scale, ntuple(_ -> 0.0, 9), ntuple(_ -> 0.0, 4),
                                length(tg), 0, pointer(abdh), pointer(csys), pointer(tg), pointer(grp), C_NULL))
            check(ctx, ccall((:fs_model_stiffness, libflexshell), Cint, (Ptr{Cvoid}, Ptr{ShellParams}, Cint, Ptr{Float64}),
                             m.h, p, block, nzval), "fs_model_stiffness")
        end
    else
        p = Ref(_homog_params(self))
        check(ctx, ccall((:fs_model_stiffness, libflexshell), Cint, (Ptr{Cvoid}, Ptr{ShellParams}, Cint, Ptr{Float64}),
                         m.h, p, block, nzval), "fs_model_stiffness")
    end
    n = Int(nr[])
    return SparseMatrixCSR{1}(n, n, rowptr, colval, nzval)
end

for (M, F) in ((FEMMShellT3FFModule, FEMMShellT3FF), (FEMMShellQ4RSModule, FEMMShellQ4RS),
               (FEMMShellT3FFCompModule, FEMMShellT3FFComp), (FEMMShellQ4RSCompModule, FEMMShellQ4RSComp))
    @eval function $M.stiffness(self::$F, assembler::SysmatAssemblerHIPCSR, geom0::NodalField{Float64},
                                u1::NodalField{TI}, Rfield1::NodalField{TI}, dchi::NodalField{TI}) where {TI<:Number}
        return _stiffness(self, assembler, geom0, dchi)
    end
    @eval function $M.mass(self::$F, assembler::SysmatAssemblerHIPDiag, geom0::NodalField{Float64},
                           dchi::NodalField{TI}) where {TI<:Number}
        return _mass(self, assembler, geom0, dchi)
    end
end

# ---------------------------------------------------------------- mass
function _mass(self, assembler::SysmatAssemblerHIPDiag, geom0, dchi)
    @assert self._associatedgeometry == true
    m = _device_model!(assembler, self, geom0, dchi)
    ctx, block = assembler.ctx, _block(assembler)
    n = block == FS_BLOCK_FF ? nfreedofs(dchi) : nalldofs(dchi)
    d = Vector{Float64}(undef, n)
    if self isa Union{FEMMShellT3FFComp,FEMMShellQ4RSComp}
        _, _, _, inert, grp = _layup_arrays(self)
        GC.@preserve inert grp begin
            p = Ref(MassParams(1, size(inert, 2), 0.0, 0.0, pointer(inert), pointer(grp), C_NULL))
            check(ctx, ccall((:fs_model_mass, libflexshell), Cint, (Ptr{Cvoid}, Ptr{MassParams}, Cint, Ptr{Float64}),
                             m.h, p, block, d), "fs_model_mass")
        end
    else
        p = Ref(MassParams(0, 0, massdensity(self.material), _thickness(self), C_NULL, C_NULL, C_NULL))
        check(ctx, ccall((:fs_model_mass, libflexshell), Cint, (Ptr{Cvoid}, Ptr{MassParams}, Cint, Ptr{Float64}),
                         m.h, p, block, d), "fs_model_mass")
    end
    return spdiagm(0 => d)
end

# ---------------------------------------------------------------- explicit loop
"""
    cd_loop!(femm, assembler, ksi_omega, U0, V0, tend, dt, Fmag, g; probes = Int[], stride = 1, mixed = false)

`_cd_loop!` (plate_expl_examples.jl:61-94) on the FF-block stiffness and mass
left resident by `stiffness(femm, SysmatAssemblerHIPCSR(; block = :ff), ...)`
and `mass(femm, SysmatAssemblerHIPDiag(; block = :ff), ...)` (same assembler
context).  The force closure is `F(t) = g(t) * Fmag`; `peek` becomes the probe
dofs (1-based) recorded every `stride` steps.  `mixed = true` runs the step in
FS_PREC_MIXED (FP32 membrane + bending, FP64 shear), `mixed = :f32` in
FS_PREC_MIXED_F32 (shear in FP32 too).  Returns (U, V, probe history).
"""
function cd_loop!(femm, assembler, cdamp, U0, V0, tend, dt, Fmag, g; probes = Int[], stride = 1, mixed = false)
    m = assembler.models[femm]
    dt_mass = dt                               # invMC uses the dt before the adjustment (:70-71)
    nsteps = Int(round(tend / dt))
    if nsteps * dt < tend                      # plate_expl_examples.jl:72-75
        dt = tend / (nsteps + 1)
    end
    t = 0.0
    gv = zeros(nsteps + 1)
    gv[1] = g(0.0)
    for n in 1:nsteps
        t = t + dt
        gv[n + 1] = g(t)
    end
    n = length(U0)
    U = similar(U0); V = similar(V0)
    pr = Int64.(probes .- 1)
    pout = zeros(length(pr), nsteps ÷ stride + 1)
    prec = mixed === :f32 ? FS_PREC_MIXED_F32 : (mixed === true ? FS_PREC_MIXED : FS_PREC_FP64)
    p = Ref(CDParams(dt, cdamp, nsteps, prec, stride, 0, 1, C_NULL, C_NULL, C_NULL, dt_mass))
    check(m.ctx, ccall((:fs_model_explicit, libflexshell), Cint,
                       (Ptr{Cvoid}, Ptr{CDParams}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                        Ptr{Int64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                       m.h, p, U0, V0, Fmag, gv, isempty(pr) ? C_NULL : pr, length(pr),
                       isempty(pr) ? C_NULL : pout, U, V), "fs_model_explicit")
    return U, V, permutedims(pout)
end

# ---------------------------------------------------------------- dofs and loads
const FS_LOAD_UNIFORM = Int32(0)
const FS_LOAD_PER_POINT = Int32(1)

"""
    numberdofs_hip!(femm, geom0, dchi[, perm])

`numberdofs!` of FinEtools on the device (fs_model_number_dofs): free dofs
first, then the fixed ones, in node order or in the order of `perm`; the
numbering is written into `dchi` and installed in the device model.
"""
function numberdofs_hip!(femm, geom0::NodalField{Float64}, dchi::NodalField, perm = nothing)
    _inspect[] === nothing && (_inspect[] = _InspectModels(default_context(), IdDict{Any,DeviceModel}()))
    m = _device_model!(_inspect[], femm, geom0, dchi)
    N = size(dchi.values, 1)
    fixed = permutedims(UInt8.(dchi.is_fixed))          # N x 6 row-major
    pm = perm === nothing ? C_NULL : Int64.(collect(perm) .- 1)
    dn = zeros(Int64, N, 6)                             # column-major, 0-based
    nf = Ref{Int64}(0)
    GC.@preserve fixed pm dn begin
        check(m.ctx, ccall((:fs_model_number_dofs, libflexshell), Cint,
                           (Ptr{Cvoid}, Ptr{UInt8}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                           m.h, fixed, pm, dn, nf), "fs_model_number_dofs")
    end
    dchi.dofnums .= dn .+ 1
    dchi.nfreedofs = nf[]
    m.dofs_key = hash(dchi.dofnums, hash(nfreedofs(dchi)))
    return dchi
end

"""
    distribloads_hip(femm, geom0, dchi, fi::ForceIntensity, npts; block = :ff)

`distribloads(FEMMBase(IntegDomain(fes, rule)), geom0, dchi, fi, 2)` on the
device (fs_model_surface_load) with rule = TriRule(npts) (npts 1, 3, 6) for
T3 and GaussRule(2, 2) (npts 4) for Q4.  A constant `fi` crosses as one
6-vector; otherwise `fi` is evaluated with `updateforce!` at the rule's points
(fs_model_load_points).  `block = :ff` returns the free rows.
"""
function distribloads_hip(femm, geom0::NodalField{Float64}, dchi::NodalField, fi, npts::Integer; block::Symbol = :ff)
    _inspect[] === nothing && (_inspect[] = _InspectModels(default_context(), IdDict{Any,DeviceModel}()))
    m = _device_model!(_inspect[], femm, geom0, dchi)
    E = count(femm.integdomain.fes)
    if fi.isconstant
        kind, f = FS_LOAD_UNIFORM, copy(fi._cache)
    else
        X = zeros(3, npts, E)
        check(m.ctx, ccall((:fs_model_load_points, libflexshell), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                           m.h, npts, X), "fs_model_load_points")
        f = zeros(6, npts, E)
        J = zeros(3, 2)
        for e in 1:E, q in 1:npts
            f[:, q, e] .= updateforce!(fi, reshape(X[:, q, e], 1, 3), J, e, q)
        end
        kind = FS_LOAD_PER_POINT
    end
    blk = block === :ff ? FS_BLOCK_FF : FS_BLOCK_ALL
    F = zeros(block === :ff ? nfreedofs(dchi) : nalldofs(dchi))
    GC.@preserve f F begin
        check(m.ctx, ccall((:fs_model_surface_load, libflexshell), Cint,
                           (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Cint, Ptr{Float64}),
                           m.h, npts, kind, f, blk, F), "fs_model_surface_load")
    end
    return F
end

# ---------------------------------------------------------------- stress resultants
const FS_OCSYS_ISOPARAMETRIC = Int32(0)
const FS_OCSYS_MATRIX = Int32(1)
const FS_OCSYS_PER_POINT = Int32(2)
const FS_RES_LAYERED_Q4_SINGLE_T = Int32(1)

struct _InspectModels                  # device models of inspectintegpoints, per FEMM
    ctx::Context
    models::IdDict{Any,DeviceModel}
end
const _inspect = Ref{Union{Nothing,_InspectModels}}(nothing)

_quantity_offset(q) = q in (:membrane_force, :membrane) ? 0 :
                      q in (:transverse_shear, :transverse, :shear) ? 6 : 3

"""
    inspectintegpoints_hip(femm, geom0, u, dT, felist, inspector, idat, quantity = :Cauchy;
                           outputcsys = nothing, transform_once = false)

`inspectintegpoints` of the four shell FEMMs (FEMMShellT3FFModule.jl:850-962,
FEMMShellQ4RSModule.jl:1061-1170, FEMMShellT3FFCompModule.jl:809-943,
FEMMShellQ4RSCompModule.jl:1077-1209) with all resultants of the listed
elements computed on the GPU in one launch (fs_model_resultants); the
inspector is then called per element and point as in the reference.
`outputcsys = nothing` is the FEMM's default (the isoparametric mcsys of the
homogeneous FEMMs, the first layup csys of the layered ones); a constant
`CSys` crosses the ABI as one matrix; any other `CSys` is evaluated per point
with `updatecsmat!` at the point's location and tangents, as the reference
does.  `transform_once = true` sets FS_RES_LAYERED_Q4_SINGLE_T.
"""
function inspectintegpoints_hip(self, geom0::NodalField{Float64}, u::NodalField{T}, dT, felist::AbstractVector,
                            inspector::F, idat, quantity = :Cauchy; outputcsys = nothing,
                            transform_once::Bool = false) where {T<:Number,F<:Function}
    @assert self._associatedgeometry == true
    _inspect[] === nothing && (_inspect[] = _InspectModels(default_context(), IdDict{Any,DeviceModel}()))
    m = _device_model!(_inspect[], self, geom0, u)
    ctx = m.ctx
    nen = _nen(self)
    npts = nen == 3 ? 1 : 4
    fes = self.integdomain.fes
    fl = Int32.(collect(felist) .- 1)
    vals = permutedims(u.values)                       # N x 6 row-major
    layered = self isa Union{FEMMShellT3FFComp,FEMMShellQ4RSComp}
    ocs = outputcsys === nothing ? (layered ? self.layup_groups[1][1].csys : nothing) : outputcsys
    kind, ocm = FS_OCSYS_ISOPARAMETRIC, Float64[]
    if ocs !== nothing && ocs.isconstant
        kind, ocm = FS_OCSYS_MATRIX, vec(copy(csmat(ocs)))
    elseif ocs !== nothing
        kind = FS_OCSYS_PER_POINT
        ocm = zeros(9, npts, length(fl))
        npr, Ns, gradNparams, _, _ = integrationdata(self.integdomain, self.integdomain.integration_rule)
        loc, J = zeros(1, 3), zeros(3, 2)
        for (l, e) in enumerate(felist), j in 1:npts
            X = geom0.values[collect(fes.conn[e]), :]
            if nen == 3                                # centroid and J0 (FEMMShellT3FFModule.jl:908-911)
                loc .= sum(X, dims = 1) ./ 3
                J .= hcat(X[2, :] - X[1, :], X[3, :] - X[1, :])
            else
                locjac!(loc, J, X, Ns[j], gradNparams[j])
            end
            updatecsmat!(ocs, loc, J, e, nen == 3 ? 0 : j)
            ocm[:, j, l] = vec(csmat(ocs))
        end
    end
    out = zeros(8, npts, length(fl))
    mis = Ref{Int32}(0)
    flags = transform_once ? FS_RES_LAYERED_Q4_SINGLE_T : Int32(0)
    if layered
        abdh, csys, tg, _, grp = _layup_arrays(self)
        GC.@preserve abdh csys tg grp vals fl ocm out begin
            p = Ref(ShellParams(1, self.transv_shear_formulation, 0.0, _stab_coefficient(self.stab_fun),
                                self.drilling_stiffness_scale, ntuple(_ -> 0.0, 9), ntuple(_ -> 0.0, 4),
                                length(tg), 0, pointer(abdh), pointer(csys), pointer(tg), pointer(grp), C_NULL))
            check(ctx, ccall((:fs_model_resultants, libflexshell), Cint,
                             (Ptr{Cvoid}, Ptr{ShellParams}, Ptr{Float64}, Int64, Ptr{Int32}, Int32, Ptr{Float64},
                              Int32, Ptr{Float64}, Ptr{Int32}),
                             m.h, p, vals, length(fl), fl, kind, isempty(ocm) ? C_NULL : pointer(ocm), flags, out,
                             mis), "fs_model_resultants")
        end
    else
        p = Ref(_homog_params(self))
        GC.@preserve vals fl ocm out begin
            check(ctx, ccall((:fs_model_resultants, libflexshell), Cint,
                             (Ptr{Cvoid}, Ptr{ShellParams}, Ptr{Float64}, Int64, Ptr{Int32}, Int32, Ptr{Float64},
                              Int32, Ptr{Float64}, Ptr{Int32}),
                             m.h, p, vals, length(fl), fl, kind, isempty(ocm) ? C_NULL : pointer(ocm), flags, out,
                             mis), "fs_model_resultants")
        end
    end
    mis[] != 0 && @warn "Coordinate systems mismatched?"
    k0 = _quantity_offset(quantity)
    nk = k0 == 6 ? 2 : 3
    o = zeros(3)
    npr, Ns, _, _, _ = integrationdata(self.integdomain, self.integdomain.integration_rule)
    for (l, e) in enumerate(felist)
        X = geom0.values[collect(fes.conn[e]), :]
        for j in 1:npts
            fill!(o, 0.0)
            o[1:nk] .= out[(k0 + 1):(k0 + nk), j, l]
            loc = nen == 3 ? sum(X, dims = 1) ./ 3 : Ns[j]' * X
            idat = inspector(idat, e, fes.conn[e], X, o, loc)
        end
    end
    return idat
end

export SysmatAssemblerHIPCSR, SysmatAssemblerHIPDiag, SysmatAssemblerHIPGeneric, cd_loop!, Context,
    default_context, inspectintegpoints_hip, numberdofs_hip!, distribloads_hip, set_pipeline!, set_cd_lanes!,
    set_assembly!

end # module
