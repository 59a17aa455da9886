# BLSAmd.jl -- ccall bindings of libblsamd.so for BlockLinearSystems.jl.
#
# Adds device methods behind the reference's own entry points, with their signatures and argument meaning:
#   reset_schur_sparse! / reset_schur_dense!   (src/schur_reduction.jl:389-436)
#   reduce_schur_sparse! / reduce_schur_dense! (src/schur_reduction.jl:465-507), conditioning scales s_r, s_e included
#   backsubstitute!(x_e, x_r, structure, C_data, ws, s_r, s_e)  (src/schur_reduction.jl:519-576)
#   reduce_qr!(ws, J, scaling_e, lambda, s)     (src/qr_reduction.jl:162-187)
#   backsubstitute!(x_e, x_r, ws::QRReductionWorkspace)  (src/qr_reduced.jl:97-110)
#   Base.copyto!(dest::SparseMatrixCSC, src, scatter_map)  (src/SparseBlockLinearSystem.jl:356-362)
# The device is selected by the workspace argument (a DeviceSchurWorkspace / DeviceQRWorkspace wrapping the handle of
# the C ABI, include/bls_amd.h) or, for copyto!, by a DeviceBlockData source. Host arrays are staged through device
# buffers owned by the workspace; results are copied back into the caller's arrays.
# Further down: dot(x, H, y), getdiag! and prepare_scale! over a device copy of a filled Hessian, the dense reduced
# solve (Cholesky with its Bunch-Kaufman fallback), and the tables' fill from row-block Jacobian blocks.
# Written against the C ABI; not executed in this repository (no Julia toolchain where it was built).
module BLSAmd

using SparseArrays
import BlockLinearSystems
const BLS = BlockLinearSystems

const libblsamd = joinpath(@__DIR__, "..", "libblsamd.so")
const BLS_OK = Cint(0)
const BLS_FAILED = Cint(1)

last_error() = unsafe_string(ccall((:bls_last_error, libblsamd), Cstring, ()))
check(rc::Cint) = rc < 0 ? error("libblsamd: " * last_error()) : rc

# device buffer of n elements of T
mutable struct DevBuf{T}
    ptr::Ptr{Cvoid}
    n::Int
    function DevBuf{T}(n::Integer) where T
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bls_alloc, libblsamd), Cint, (Ptr{Ptr{Cvoid}}, Int64), p, n * sizeof(T)))
        b = new{T}(p[], n)
        finalizer(x -> ccall((:bls_free, libblsamd), Cint, (Ptr{Cvoid},), x.ptr), b)
        return b
    end
end
# a conditioning scale as the C ABI takes it: NULL for UnitScale (the identity), else a device copy
scaleptr(d, key::Symbol, s::AbstractVector) = s isa BLS.UnitScale ? C_NULL : staged(d, key, collect(Float64, s)).ptr

upload!(d::DevBuf{T}, v::AbstractVector{T}) where T =
    check(ccall((:bls_upload, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}), d.ptr, v, length(v) * sizeof(T), C_NULL))
download!(v::AbstractVector{T}, d::DevBuf{T}) where T =
    check(ccall((:bls_download, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}), v, d.ptr, length(v) * sizeof(T), C_NULL))

"""
    DeviceSchurWorkspace(structure, ws::SchurReductionWorkspace{Float64}; pattern=nothing, r_to_s_map=Int[], dense=false)

Registers `structure` and the workspace index with the device (integer maps uploaded once) and owns device copies of
E_data, C_data, b, S, g_S, E_inv and z. Pass it where the reference takes `ws`.
"""
mutable struct DeviceSchurWorkspace
    handle::Ptr{Cvoid}
    host::BLS.SchurReductionWorkspace{Float64}
    dense::Bool
    bufs::Dict{Symbol, DevBuf{Float64}}
end

function DeviceSchurWorkspace(st::BLS.SchurStructure, ws::BLS.SchurReductionWorkspace{Float64};
                              pattern=nothing, r_to_s_map::Vector{Int}=Int[], dense::Bool=false)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bls_schur_create, libblsamd), Cint,
                (Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
                st.nr, st.nrblocks, st.neblocks, length(st.entry_retained), collect(Int64, st.r_dims),
                collect(Int64, st.e_dims), st.boffsets_r, st.boffsets_e, st.eoffsets, st.eliminated_ptr,
                st.entry_retained, st.coffsets, ws.retained_ptr, ws.retained_entry, h))
    if dense
        check(ccall((:bls_schur_set_dense, libblsamd), Cint, (Ptr{Cvoid},), h[]))
    else
        p = pattern
        check(ccall((:bls_schur_set_sparse, libblsamd), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64,
                     Ptr{Int64}, Ptr{Int64}),
                    h[], nnz(p.S_sparse), p.S_sparse.colptr, length(p.block_rowval), p.block_colptr, p.block_rowval,
                    p.block_row_offsets, p.col_heights, length(r_to_s_map), r_to_s_map, p.diag_inds))
    end
    d = DeviceSchurWorkspace(h[], ws, dense, Dict{Symbol, DevBuf{Float64}}())
    finalizer(x -> ccall((:bls_schur_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), x.handle), d)
    return d
end

# which phase-2 kernel the workspace runs (no reference counterpart): 0 the tiles of S on the matrix cores where the
# sparse layout fits them, else the work list; 1 the work list; 2 the tiles
set_phase2_mode!(d::DeviceSchurWorkspace, mode::Integer) =
    check(ccall((:bls_schur_set_phase2_mode, libblsamd), Cint, (Ptr{Cvoid}, Cint), d.handle, mode))

# (kernel in use: 1 work list, 2 tiles; tiles; landmark entries; matrix-core products)
function phase2_info(d::DeviceSchurWorkspace)
    info = zeros(Int64, 4)
    check(ccall((:bls_schur_phase2_info, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Int64}), d.handle, info))
    return (info[1], info[2], info[3], info[4])
end

function staged(d, key::Symbol, v::AbstractVector{Float64})
    b = get!(() -> DevBuf{Float64}(length(v)), d.bufs, key)
    b.n == length(v) || (b = d.bufs[key] = DevBuf{Float64}(length(v)))
    upload!(b, v)
    return b
end
devbuf(d, key::Symbol, n::Integer) = (b = get!(() -> DevBuf{Float64}(n), d.bufs, key); b.n == n ? b : (d.bufs[key] = DevBuf{Float64}(n)))

# reset_schur_sparse!(nzval, R_nzval, r_to_s_map, diag_inds, scaling_r, lambda, ws, s_r, S)
# (src/schur_reduction.jl:416-436); the map and diag_inds were registered with the workspace (bls_schur_set_sparse)
function BLS.reset_schur_sparse!(nzval::Vector{Float64}, R_nzval::Vector{Float64}, r_to_s_map::Vector{Int},
                                 diag_inds::Vector{Int}, scaling_r::AbstractVector, lambda::Real, ws::DeviceSchurWorkspace,
                                 s_r::AbstractVector=BLS.UnitScale(length(diag_inds)), S=nothing)
    BLS.check_damping(lambda, scaling_r, :scaling_r)
    Sd = devbuf(ws, :S, length(nzval))
    check(ccall((:bls_reset_schur_sparse_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}),
                ws.handle, Sd.ptr, staged(ws, :R, R_nzval).ptr, staged(ws, :scaling_r, collect(Float64, scaling_r)).ptr,
                Float64(lambda), scaleptr(ws, :s_r, s_r), C_NULL))
    download!(nzval, Sd)
    ws.host.failed_block[] = 0
    return nothing
end

# reset_schur_dense!(S, R, scaling_r, lambda, ws, s_r) (src/schur_reduction.jl:389-407): S = R (scaled) + damping
function BLS.reset_schur_dense!(S::Matrix{Float64}, R::Matrix{Float64}, scaling_r::AbstractVector, lambda::Real,
                                ws::DeviceSchurWorkspace, s_r::AbstractVector=BLS.UnitScale(size(S, 1)))
    BLS.check_damping(lambda, scaling_r, :scaling_r)
    Sd = devbuf(ws, :S, length(S))
    check(ccall((:bls_reset_schur_dense_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}),
                ws.handle, Sd.ptr, staged(ws, :R, vec(R)).ptr, staged(ws, :scaling_r, collect(Float64, scaling_r)).ptr,
                Float64(lambda), scaleptr(ws, :s_r, s_r), C_NULL))
    download!(vec(S), Sd)
    ws.host.failed_block[] = 0
    return nothing
end

# reduce_schur_sparse!(S, g_S, b_r, b_e, scaling_e, lambda, structure, E_data, C_data, block_colptr, block_rowval,
#                      block_row_offsets, col_heights, ws) (src/schur_reduction.jl:489-507)
function BLS.reduce_schur_sparse!(S::SparseMatrixCSC{Float64}, g_S::AbstractVector, b_r::AbstractVector,
                                  b_e::AbstractVector, scaling_e::AbstractVector, lambda::Real,
                                  structure::BLS.SchurStructure, E_data::Vector{Float64}, C_data::Vector{Float64},
                                  block_colptr, block_rowval, block_row_offsets, col_heights, ws::DeviceSchurWorkspace,
                                  s_r::AbstractVector=BLS.UnitScale(length(b_r)),
                                  s_e::AbstractVector=BLS.UnitScale(length(b_e)))
    ws.dense && throw(ArgumentError("the workspace was registered for a dense reduced system"))
    return device_reduce!(:bls_reduce_schur_sparse_f64, S.nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data,
                          ws, s_r, s_e)
end

# reduce_schur_dense!(S, g_S, b_r, b_e, scaling_e, lambda, structure, E_data, C_data, ws, s_r, s_e)
# (src/schur_reduction.jl:465-479): S (upper triangle) holds R + damping from reset_schur_dense!
function BLS.reduce_schur_dense!(S::Matrix{Float64}, g_S::AbstractVector, b_r::AbstractVector, b_e::AbstractVector,
                                 scaling_e::AbstractVector, lambda::Real, structure::BLS.SchurStructure,
                                 E_data::Vector{Float64}, C_data::Vector{Float64}, ws::DeviceSchurWorkspace,
                                 s_r::AbstractVector=BLS.UnitScale(length(b_r)),
                                 s_e::AbstractVector=BLS.UnitScale(length(b_e)))
    ws.dense || throw(ArgumentError("the workspace was registered for a sparse reduced system"))
    return device_reduce!(:bls_reduce_schur_dense_f64, vec(S), g_S, b_r, b_e, scaling_e, lambda, E_data, C_data,
                          ws, s_r, s_e)
end

# both reductions: stage the inputs, run, and on success copy S, g_S, E_inv and z back (on failure S and g_S are
# left as the caller's, as the reference leaves them; E_inv and z stay on the device for backsubstitute!)
function device_reduce!(sym::Symbol, Sv::AbstractVector{Float64}, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data,
                        ws::DeviceSchurWorkspace, s_r, s_e)
    BLS.check_damping(lambda, scaling_e, :scaling_e)
    Sd = staged(ws, :S, Sv)                # S holds R + damping (the reset)
    g = devbuf(ws, :g, length(g_S))
    Ei = devbuf(ws, :E_inv, length(ws.host.E_inv_data))
    z = devbuf(ws, :z, length(ws.host.z_data))
    fb = Ref{Int64}(0)
    rc = check(ccall((sym, libblsamd), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid},
                      Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cvoid}),
                     ws.handle, Sd.ptr, g.ptr, staged(ws, :b_r, collect(Float64, b_r)).ptr,
                     staged(ws, :b_e, collect(Float64, b_e)).ptr, staged(ws, :scaling_e, collect(Float64, scaling_e)).ptr,
                     Float64(lambda), staged(ws, :E, E_data).ptr, staged(ws, :C, C_data).ptr, Ei.ptr, z.ptr,
                     scaleptr(ws, :s_r, s_r), scaleptr(ws, :s_e, s_e), fb, C_NULL))
    ws.host.failed_block[] = fb[]
    rc == BLS_FAILED && return false
    download!(Sv, Sd)
    tmp = Vector{Float64}(undef, length(g_S)); download!(tmp, g); g_S .= tmp
    download!(ws.host.E_inv_data, Ei)
    download!(ws.host.z_data, z)
    return true
end

# backsubstitute!(x_e, x_r, structure, C_data, ws, s_r, s_e) (src/schur_reduction.jl:519-576): x_e = z - E_inv sum
# C x_r from the E_inv and z of the workspace's last reduction (kept on the device), C_data as staged by it
function BLS.backsubstitute!(x_e::AbstractVector, x_r::AbstractVector, structure::BLS.SchurStructure,
                             C_data::Vector{Float64}, ws::DeviceSchurWorkspace,
                             s_r::AbstractVector=BLS.UnitScale(length(x_r)),
                             s_e::AbstractVector=BLS.UnitScale(length(x_e)))
    haskey(ws.bufs, :E_inv) || throw(ArgumentError("backsubstitute! needs a reduction on this workspace first"))
    xe = devbuf(ws, :x_e, length(x_e))
    check(ccall((:bls_backsubstitute_schur_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ws.handle, xe.ptr, staged(ws, :x_r, collect(Float64, x_r)).ptr, staged(ws, :C, C_data).ptr,
                ws.bufs[:E_inv].ptr, ws.bufs[:z].ptr, scaleptr(ws, :s_r, s_r), scaleptr(ws, :s_e, s_e), C_NULL))
    tmp = Vector{Float64}(undef, length(x_e)); download!(tmp, xe); x_e .= tmp
    return nothing
end

# ---- copyto!(dest::SparseMatrixCSC, src, scatter_map) (src/SparseBlockLinearSystem.jl:356-362) ------------------

"""
    DeviceBlockData(data::Vector{Float64})
    DeviceBlockData(ws::DeviceSchurWorkspace)

The `data` of a SparseBlockLinearSystem on the device (a copy of `data`, or the reduced system a
DeviceSchurWorkspace's last reset/reduction left on the device), as the source of `copyto!`.
"""
struct DeviceBlockData
    buf::DevBuf{Float64}
end
DeviceBlockData(data::Vector{Float64}) = (b = DevBuf{Float64}(length(data)); upload!(b, data); DeviceBlockData(b))
DeviceBlockData(ws::DeviceSchurWorkspace) = DeviceBlockData(ws.bufs[:S])

"""
    DeviceScatterMap(scatter_map)

A scatter map (the second result of sparse_with_map) uploaded once, for repeated `copyto!`.
"""
mutable struct DeviceScatterMap
    ptr::Ptr{Int64}
    n::Int
    function DeviceScatterMap(scatter_map::AbstractVector{<:Integer})
        dm = Ref{Ptr{Int64}}(C_NULL)
        m = collect(Int64, scatter_map)
        check(ccall((:bls_scatter_map_create, libblsamd), Cint, (Int64, Ptr{Int64}, Ptr{Ptr{Int64}}), length(m), m, dm))
        s = new(dm[], length(m))
        finalizer(x -> ccall((:bls_free, libblsamd), Cint, (Ptr{Cvoid},), x.ptr), s)
        return s
    end
end

function Base.copyto!(dest::SparseMatrixCSC{Float64}, src::DeviceBlockData, scatter_map::DeviceScatterMap)
    length(dest.nzval) == scatter_map.n || throw(DimensionMismatch("Destination nzval length ($(length(dest.nzval))) does not match scatter map length ($(scatter_map.n))"))
    d = DevBuf{Float64}(length(dest.nzval))
    check(ccall((:bls_copyto_csc_f64, libblsamd), Cint, (Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cvoid}),
                length(dest.nzval), d.ptr, src.buf.ptr, scatter_map.ptr, C_NULL))
    download!(dest.nzval, d)
    return dest
end

Base.copyto!(dest::SparseMatrixCSC{Float64}, src::DeviceBlockData, scatter_map::AbstractVector{<:Integer}) =
    copyto!(dest, src, DeviceScatterMap(scatter_map))

# ---- reduce_qr! and backsubstitute! of the QR path (src/qr_reduction.jl:162-187, src/qr_reduced.jl:97-110) -------

"""
    DeviceQRWorkspace(ws::QRReductionWorkspace{Float64})

Registers the workspace's SchurJacobianStructure with the device (bls_qr_create + bls_qr_set_reduced) and owns
device copies of the raw and factored tables. Pass it where the reference takes the QR workspace.
"""
mutable struct DeviceQRWorkspace
    handle::Ptr{Cvoid}
    host::BLS.QRReductionWorkspace{Float64}
    bufs::Dict{Symbol, DevBuf{Float64}}
end

function DeviceQRWorkspace(ws::BLS.QRReductionWorkspace{Float64})
    jst = ws.structure
    st = jst.schur
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bls_qr_create, libblsamd), Cint,
                (Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Ptr{Cvoid}}),
                st.nr, st.nrblocks, st.neblocks, collect(Int64, st.e_dims), st.boffsets_e, jst.ntables,
                jst.table_nrows, jst.table_ncols, jst.raw_offsets, jst.residual_offsets, h))
    d = DeviceQRWorkspace(h[], ws, Dict{Symbol, DevBuf{Float64}}())
    finalizer(x -> ccall((:bls_qr_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), x.handle), d)
    check(ccall((:bls_qr_set_reduced, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                d.handle, jst.table_top, jst.reduced_offsets, jst.table_ptr, length(jst.entry_retained),
                jst.entry_retained, jst.entry_col, jst.entry_table, jst.retained_ptr, jst.retained_entry,
                collect(Int64, st.r_dims), st.boffsets_r, jst.doffsets))
    return d
end

# reduce_qr!(ws, J, scaling_e, lambda, s) (src/qr_reduction.jl:162-187): the structure and length checks first, then
# every table on the device; the factored tables and residual are copied into ws.host (factored, fres)
function BLS.reduce_qr!(ws::DeviceQRWorkspace, J::BLS.SchurBlockJacobian{Float64}, scaling_e::AbstractVector,
                        lambda::Real, s::AbstractVector=BLS.UnitScale(0))
    jst = ws.host.structure
    J.structure === jst || throw(ArgumentError("The Jacobian and the workspace were built for different structures"))
    ne = sum(Int, jst.schur.e_dims; init=0)
    length(scaling_e) == ne || throw(DimensionMismatch("scaling_e has length $(length(scaling_e)), expected $ne"))
    BLS.check_damping(lambda, scaling_e, :scaling_e)
    F = devbuf(ws, :factored, length(ws.host.factored))
    Fr = devbuf(ws, :fres, length(ws.host.fres))
    fb = Ref{Int64}(0)
    rc = check(ccall((:bls_reduce_qr_f64, libblsamd), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid},
                      Ptr{Int64}, Ptr{Cvoid}),
                     ws.handle, F.ptr, Fr.ptr, staged(ws, :raw, J.data).ptr, staged(ws, :r, J.r).ptr,
                     staged(ws, :scaling_e, collect(Float64, scaling_e)).ptr, Float64(lambda),
                     scaleptr(ws, :s, s), fb, C_NULL))
    ws.host.failed_block[] = fb[]
    rc == BLS_FAILED && return false
    download!(ws.host.factored, F)
    download!(ws.host.fres, Fr)
    return true
end

# backsubstitute!(x_e, x_r, ws::QRReductionWorkspace) (src/qr_reduced.jl:97-110): x_e = U^-1 (q - B x_r) per table,
# from the device's factored tables of the last reduce_qr!
function BLS.backsubstitute!(x_e::AbstractVector, x_r::AbstractVector, ws::DeviceQRWorkspace)
    haskey(ws.bufs, :factored) || throw(ArgumentError("backsubstitute! needs reduce_qr! on this workspace first"))
    xe = devbuf(ws, :x_e, length(x_e))
    check(ccall((:bls_qr_backsubstitute_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ws.handle, xe.ptr, staged(ws, :x_r, collect(Float64, x_r)).ptr, ws.bufs[:factored].ptr,
                ws.bufs[:fres].ptr, C_NULL))
    tmp = Vector{Float64}(undef, length(x_e)); download!(tmp, xe); x_e .= tmp
    return nothing
end

# ---- The storage of a filled Hessian: dot(x, H, y), getdiag! and the conditioning scale -----------------------

# The stored lower blocks of a SparseBlockLinearSystem R as the block CSC the C ABI takes: for block column q, the
# rows p >= q with a stored block (p, q), ascending, and the blocks' 1-based data offsets. R.indicestransposed lists,
# in its column i (a block row), the stored blocks (i, j) and their offsets (src/SparseBlockLinearSystem.jl:52-61).
function lower_block_csc(R::BLS.SparseBlockLinearSystem)
    It = R.indicestransposed
    nb = size(It, 2)
    cols = [Tuple{Int64, Int64}[] for _ in 1:nb]
    for i in 1:nb, k in nzrange(It, i)
        push!(cols[rowvals(It)[k]], (Int64(i), Int64(nonzeros(It)[k])))
    end
    colptr, rowval, offs = Int64[1], Int64[], Int64[]
    for q in 1:nb
        for (p, o) in sort(cols[q])
            push!(rowval, p); push!(offs, o)
        end
        push!(colptr, length(rowval) + 1)
    end
    return colptr, rowval, offs
end

"""
    DeviceHessianLayout(H::SchurBlockHessian{Float64})

The storage layout of `H` on the device (a bls_asm handle registered with no row blocks) and device copies of its
E_data, C_data and R, for the queries below. A sparse R passes its stored lower blocks as a block CSC.
"""
mutable struct DeviceHessianLayout
    handle::Ptr{Cvoid}
    E::DevBuf{Float64}
    C::DevBuf{Float64}
    R::DevBuf{Float64}
end
function DeviceHessianLayout(H::BLS.SchurBlockHessian{Float64})
    st = H.structure
    dense = H.R isa Matrix
    rc_, rr_, ro_ = dense ? (C_NULL, C_NULL, C_NULL) : lower_block_csc(H.R)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    none = Int64[1]
    check(ccall((:bls_asm_create, libblsamd), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Ptr{Cvoid}}),
                st.nrblocks, st.neblocks, collect(Int64, st.r_dims), collect(Int64, st.e_dims), st.boffsets_r,
                st.boffsets_e, st.eoffsets, st.eliminated_ptr, st.entry_retained, st.coffsets, rc_, rr_, ro_,
                0, none, C_NULL, C_NULL, h))
    Rv = dense ? vec(H.R) : H.R.data
    l = DeviceHessianLayout(h[], DevBuf{Float64}(length(H.E_data)), DevBuf{Float64}(length(H.C_data)),
                            DevBuf{Float64}(length(Rv)))
    upload!(l.E, H.E_data); upload!(l.C, H.C_data); upload!(l.R, Rv)
    finalizer(x -> ccall((:bls_asm_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), x.handle), l)
    return l
end

# dot(x, H, y) (src/linalg.jl:133-260) on the device copy of H
function device_dot(x::Vector{Float64}, l::DeviceHessianLayout, y::Vector{Float64})
    dx = DevBuf{Float64}(length(x)); upload!(dx, x)
    dy = DevBuf{Float64}(length(y)); upload!(dy, y)
    out = DevBuf{Float64}(1)
    check(ccall((:bls_asm_dot_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                l.handle, dx.ptr, dy.ptr, l.E.ptr, l.C.ptr, l.R.ptr, out.ptr, C_NULL))
    r = zeros(1); download!(r, out)
    return r[1]
end

"""
    DeviceConditioning(n)

EliminationConditioning (src/conditioning.jl:47-55) on the device: the stale flag and nrescale in a bls_cond handle,
s and d in device buffers. `device_prepare_scale!` is prepare_scale! (:60-87) with getdiag! of the layout as diag!.
"""
mutable struct DeviceConditioning
    handle::Ptr{Cvoid}
    s::DevBuf{Float64}
    d::DevBuf{Float64}
end
function DeviceConditioning(n::Integer)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bls_cond_create, libblsamd), Cint, (Int64, Ptr{Ptr{Cvoid}}), n, h))
    c = DeviceConditioning(h[], DevBuf{Float64}(n), DevBuf{Float64}(n))
    finalizer(x -> ccall((:bls_cond_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), x.handle), c)
    return c
end
function device_prepare_scale!(c::DeviceConditioning, l::DeviceHessianLayout, lambda::Float64, sc::DevBuf{Float64},
                               ::Type{T}) where T
    if check(ccall((:bls_cond_is_stale, libblsamd), Cint, (Ptr{Cvoid},), c.handle)) == 1
        check(ccall((:bls_asm_diag_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    l.handle, l.E.ptr, l.R.ptr, c.d.ptr, C_NULL))
    end
    check(ccall((:bls_prepare_scale_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Cint, Ptr{Cvoid}),
                c.handle, c.s.ptr, c.d.ptr, sc.ptr, lambda, T === Float32 ? Cint(1) : Cint(0), C_NULL))
    return c.s                       # device pointer: pass as s_r / s_e (offset by nr) to the reduction
end

# ---- The dense reduced solve: solve_reduced_dense! (src/schur_linear_systems.jl:312-319) --------------------------

# x_r = S \ g_S by the Cholesky branch alone: Success, or NotPositiveDefinite with S left as built
function device_cholesky_solve!(x_r::Vector{Float64}, S::Matrix{Float64}, g_S::Vector{Float64})
    n = size(S, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bls_chol_create, libblsamd), Cint, (Int64, Ptr{Ptr{Cvoid}}), n, h))
    try
        dS = DevBuf{Float64}(n * n); upload!(dS, vec(S))
        fc = Ref{Int64}(0)
        rc = check(ccall((:bls_chol_factor_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cvoid}),
                         h[], dS.ptr, fc, C_NULL))
        rc == BLS_FAILED && return :NotPositiveDefinite
        dg = DevBuf{Float64}(n); upload!(dg, g_S)
        dx = DevBuf{Float64}(n)
        check(ccall((:bls_chol_solve_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    h[], dg.ptr, dx.ptr, C_NULL))
        download!(x_r, dx)
        return :Success
    finally
        ccall((:bls_chol_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), h[])
    end
end

# x_r = S \ g_S for the dense reduced S (upper triangle) as solve_reduced_dense! does it: :auto and :cholesky try
# Cholesky and fall back to Bunch-Kaufman, :ldl goes to Bunch-Kaufman; :Success, :Indefinite (S has a negative
# eigenvalue; x_r is solved) or :NotPositiveDefinite (a zero pivot: S singular)
function device_solve_reduced_dense!(x_r::Vector{Float64}, S::Matrix{Float64}, g_S::Vector{Float64},
                                     solver::Symbol=:auto)
    solver in (:auto, :cholesky, :ldl) || throw(ArgumentError("$solver is not a reduced solver"))
    solver === :ldl || device_cholesky_solve!(x_r, S, g_S) === :Success && return :Success
    n = size(S, 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bls_ldl_create, libblsamd), Cint, (Int64, Ptr{Ptr{Cvoid}}), n, h))
    try
        dS = DevBuf{Float64}(n * n); upload!(dS, vec(S))
        zc, neg = Ref{Int64}(0), Ref{Int64}(0)
        rc = check(ccall((:bls_ldl_factor_f64, libblsamd), Cint,
                         (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}), h[], dS.ptr, zc, neg, C_NULL))
        rc == BLS_FAILED && return :NotPositiveDefinite
        dg = DevBuf{Float64}(n); upload!(dg, g_S)
        dx = DevBuf{Float64}(n)
        check(ccall((:bls_ldl_solve_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    h[], dg.ptr, dx.ptr, C_NULL))
        download!(x_r, dx)
        return neg[] > 0 ? :Indefinite : :Success
    finally
        ccall((:bls_ldl_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), h[])
    end
end

# ---- The tables' fill: block(J, k, blockentry(jst, k, j)) .= J_kj, residual_block(J, k) .= r_k (src/schur_jacobian.jl:264-387)

"""
    DeviceJacobianFill(jst::SchurJacobianStructure, rowblocks)

The row blocks of a structure built from a description (rowblocks: row_table, row_first, row_nrows, row_ptr,
row_blocks, as SchurJacobianStructure(blocksizes, rowblocks, rowblocksizes; eliminated) lays them out) registered
with the device; `nvals` / `nres` are the lengths `device_fill!` takes.
"""
mutable struct DeviceJacobianFill
    handle::Ptr{Cvoid}
    nvals::Int
    nres::Int
end

function DeviceJacobianFill(jst::BLS.SchurJacobianStructure, row_table::Vector{Int}, row_first::Vector{Int},
                            row_nrows::Vector{Int}, row_ptr::Vector{Int}, row_blocks::Vector{Int})
    st = jst.schur
    h = Ref{Ptr{Cvoid}}(C_NULL)
    nv, nr = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:bls_jfill_create, libblsamd), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
                st.nrblocks, st.neblocks, collect(Int64, st.r_dims), collect(Int64, st.e_dims), jst.ntables,
                jst.table_nrows, jst.table_ncols, jst.table_top, jst.raw_offsets, jst.residual_offsets, jst.table_ptr,
                jst.entry_retained, jst.entry_col, length(row_table), row_table, row_first, row_nrows, row_ptr,
                row_blocks, nv, nr, h))
    f = DeviceJacobianFill(h[], nv[], nr[])
    finalizer(x -> ccall((:bls_jfill_destroy, libblsamd), Cvoid, (Ptr{Cvoid},), x.handle), f)
    return f
end

# every row block's blocks (vals: row block after row block, its blocks in ascending column block, each m x d
# column-major) and residual rows (res) written into J.data and J.r through the device
function device_fill!(J::BLS.SchurBlockJacobian{Float64}, f::DeviceJacobianFill, vals::Vector{Float64},
                      res::Vector{Float64})
    length(vals) == f.nvals || throw(DimensionMismatch("vals has length $(length(vals)), expected $(f.nvals)"))
    length(res) == f.nres || throw(DimensionMismatch("res has length $(length(res)), expected $(f.nres)"))
    draw = DevBuf{Float64}(length(J.data)); upload!(draw, J.data)
    dr = DevBuf{Float64}(length(J.r)); upload!(dr, J.r)
    dv = DevBuf{Float64}(length(vals)); upload!(dv, vals)
    dres = DevBuf{Float64}(length(res)); upload!(dres, res)
    check(ccall((:bls_jfill_f64, libblsamd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                f.handle, draw.ptr, dr.ptr, dv.ptr, dres.ptr, C_NULL))
    download!(J.data, draw)
    download!(J.r, dr)
    return J
end

end # module BLSAmd
