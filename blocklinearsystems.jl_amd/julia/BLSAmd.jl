# BLSAmd.jl -- ccall bindings of libblsamd.so for BlockLinearSystems.jl.
#
# Adds device methods behind the reference's own entry points: reduce_schur_sparse!, reduce_schur_dense!,
# reset_schur_sparse!, reset_schur_dense!, backsubstitute!, reduce_qr! and copyto!(csc, src, scatter_map) keep their
# signatures (src/schur_reduction.jl:389-507, src/qr_reduction.jl:162-187, src/SparseBlockLinearSystem.jl:356-362);
# the device is selected by the workspace argument, a DeviceSchurWorkspace / DeviceQRWorkspace wrapping the
# handle of the C ABI (include/bls_amd.h). Host arrays are staged through device buffers owned by the workspace.
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

function staged(d, key::Symbol, v::AbstractVector{Float64})
    b = get!(() -> DevBuf{Float64}(length(v)), d.bufs, key)
    b.n == length(v) || (b = d.bufs[key] = DevBuf{Float64}(length(v)))
    upload!(b, v)
    return b
end
devbuf(d, key::Symbol, n::Integer) = (b = get!(() -> DevBuf{Float64}(n), d.bufs, key); b.n == n ? b : (d.bufs[key] = DevBuf{Float64}(n)))

# reset_schur_sparse!(nzval, R_nzval, r_to_s_map, diag_inds, scaling_r, lambda, ws) (src/schur_reduction.jl:416-436)
function BLS.reset_schur_sparse!(nzval::Vector{Float64}, R_nzval::Vector{Float64}, r_to_s_map::Vector{Int},
                                 diag_inds::Vector{Int}, scaling_r::AbstractVector, lambda::Real, ws::DeviceSchurWorkspace)
    BLS.check_damping(lambda, scaling_r, :scaling_r)
    S = devbuf(ws, :S, length(nzval))
    check(ccall((:bls_reset_schur_sparse_f64, libblsamd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}),
                ws.handle, S.ptr, staged(ws, :R, R_nzval).ptr, staged(ws, :scaling_r, collect(Float64, scaling_r)).ptr,
                Float64(lambda), C_NULL, C_NULL))
    download!(nzval, S)
    return nothing
end

# reduce_schur_sparse!(S, g_S, b_r, b_e, scaling_e, lambda, structure, E_data, C_data, block_colptr, block_rowval,
#                      block_row_offsets, col_heights, ws) (src/schur_reduction.jl:489-507)
function BLS.reduce_schur_sparse!(S::SparseMatrixCSC{Float64}, g_S::AbstractVector, b_r::AbstractVector,
                                  b_e::AbstractVector, scaling_e::AbstractVector, lambda::Real,
                                  structure::BLS.SchurStructure, E_data::Vector{Float64}, C_data::Vector{Float64},
                                  block_colptr, block_rowval, block_row_offsets, col_heights, ws::DeviceSchurWorkspace)
    BLS.check_damping(lambda, scaling_e, :scaling_e)
    Sd = staged(ws, :S, S.nzval)           # S holds R + damping (reset_schur_sparse!)
    g = devbuf(ws, :g, length(g_S))
    Ei = devbuf(ws, :E_inv, length(ws.host.E_inv_data))
    z = devbuf(ws, :z, length(ws.host.z_data))
    fb = Ref{Int64}(0)
    rc = check(ccall((:bls_reduce_schur_sparse_f64, libblsamd), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid},
                      Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cvoid}),
                     ws.handle, Sd.ptr, g.ptr, staged(ws, :b_r, collect(Float64, b_r)).ptr,
                     staged(ws, :b_e, collect(Float64, b_e)).ptr, staged(ws, :scaling_e, collect(Float64, scaling_e)).ptr,
                     Float64(lambda), staged(ws, :E, E_data).ptr, staged(ws, :C, C_data).ptr, Ei.ptr, z.ptr,
                     C_NULL, C_NULL, fb, C_NULL))
    ws.host.failed_block[] = fb[]
    rc == BLS_FAILED && return false
    download!(S.nzval, Sd)
    tmp = Vector{Float64}(undef, length(g_S)); download!(tmp, g); g_S .= tmp
    download!(ws.host.E_inv_data, Ei)
    download!(ws.host.z_data, z)
    return true
end

# copyto!(dest, src, scatter_map) (src/SparseBlockLinearSystem.jl:356-362) on device copies; for repeated use keep
# the map on the device (bls_scatter_map_create) and call bls_copyto_csc_f64 directly
function device_copyto!(dest::Vector{Float64}, src::Vector{Float64}, scatter_map::Vector{Int64})
    length(dest) == length(scatter_map) || throw(DimensionMismatch("Destination nzval length ($(length(dest))) does not match scatter map length ($(length(scatter_map)))"))
    dm = Ref{Ptr{Int64}}(C_NULL)
    check(ccall((:bls_scatter_map_create, libblsamd), Cint, (Int64, Ptr{Int64}, Ptr{Ptr{Int64}}), length(scatter_map), scatter_map, dm))
    s = DevBuf{Float64}(length(src)); upload!(s, src)
    d = DevBuf{Float64}(length(dest))
    check(ccall((:bls_copyto_csc_f64, libblsamd), Cint, (Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Cvoid}),
                length(dest), d.ptr, s.ptr, dm[], C_NULL))
    download!(dest, d)
    ccall((:bls_free, libblsamd), Cint, (Ptr{Cvoid},), dm[])
    return dest
end

end # module BLSAmd
