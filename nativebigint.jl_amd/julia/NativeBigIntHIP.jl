# NativeBigIntHIP.jl -- ccall binding of libnbh.so for NativeBigInt.jl.
#
# Replaces the bodies of the engine pair mul_fpntt2!/sqr_fpntt2!
# (src/fpntt.jl:982-1037) with the MI355X engine; the four call sites in
# src/mul.jl (134, 146, 161, 389) are unchanged.  Same signatures, same
# 0-based offset convention (element i of a buffer is buf[off+i], so the C
# pointer is pointer(buf, off+1)), same ArgumentError for lengths above
# 315*2^34 (src/fpntt.jl:948-949).  r must not alias a or b.
module NativeBigIntHIP

const libnbh = normpath(joinpath(@__DIR__, "..", "libnbh.so"))

const NBH_ERR_LENGTH = Int32(1)

errstr() = unsafe_string(ccall((:nbh_last_error_string, libnbh), Cstring, ()))

@inline function check(rc::Int32)
    rc == 0 && return nothing
    rc == NBH_ERR_LENGTH && throw(ArgumentError(errstr()))
    error("nbh error $rc: $(errstr())")
end

function __init__()
    check(ccall((:nbh_init, libnbh), Int32, (Int32,), 0))
end

function mul_fpntt2!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, m::Int,
                     b::Memory{UInt64}, bo::Int, n::Int)
    GC.@preserve r a b begin
        check(ccall((:nbh_mul, libnbh), Int32,
                    (Ptr{UInt64}, Ptr{UInt64}, Int64, Ptr{UInt64}, Int64),
                    pointer(r, ro + 1), pointer(a, ao + 1), m, pointer(b, bo + 1), n))
    end
    return nothing
end

function sqr_fpntt2!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, n::Int)
    GC.@preserve r a begin
        check(ccall((:nbh_sqr, libnbh), Int32, (Ptr{UInt64}, Ptr{UInt64}, Int64),
                    pointer(r, ro + 1), pointer(a, ao + 1), n))
    end
    return nothing
end

# Low short products, mullo!/sqrlo! (src/mul.jl:305-378) with the same
# signatures (the scratch arguments are accepted and unused): r[1..k] of
# a*b mod beta^k.  Writes exactly k limbs; the reference's two clobber limbs
# are left alone.  For k at or above MULLO_FULL_THRESHOLD the reference runs
# a full mul! and copies k limbs; here only the low k limbs are unpacked and
# copied back.
function mullo!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, la::Int,
                b::Memory{UInt64}, bo::Int, lb::Int, k::Int,
                scratch::Union{Memory{UInt64},Nothing}=nothing, so::Int=0)
    GC.@preserve r a b begin
        check(ccall((:nbh_mullo, libnbh), Int32,
                    (Ptr{UInt64}, Ptr{UInt64}, Int64, Ptr{UInt64}, Int64, Int64),
                    pointer(r, ro + 1), pointer(a, ao + 1), la, pointer(b, bo + 1), lb, k))
    end
    return nothing
end

function sqrlo!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, la::Int, k::Int,
                scratch::Union{Memory{UInt64},Nothing}=nothing, so::Int=0)
    GC.@preserve r a begin
        check(ccall((:nbh_sqrlo, libnbh), Int32, (Ptr{UInt64}, Ptr{UInt64}, Int64, Int64),
                    pointer(r, ro + 1), pointer(a, ao + 1), la, k))
    end
    return nothing
end

# Many independent products in one call (nbh_mul_batch): each element of
# `jobs` is (r, ro, a, ao, m, b, bo, n, k) as in mul_fpntt2!/mullo!, with
# b === nothing for a square and k == 0 for the full product.  Results must
# not overlap each other or any operand.
struct NbhJob
    r::Ptr{UInt64}
    a::Ptr{UInt64}
    b::Ptr{UInt64}
    m::Int64
    n::Int64
    bits_a::Int64
    bits_b::Int64
    k::Int64
end

function mul_batch!(jobs::AbstractVector)
    bufs = Any[]
    recs = Vector{NbhJob}(undef, length(jobs))
    for (i, (r, ro, a, ao, m, b, bo, n, k)) in enumerate(jobs)
        push!(bufs, r, a)
        b === nothing || push!(bufs, b)
        pb = b === nothing ? Ptr{UInt64}(0) : pointer(b, bo + 1)
        recs[i] = NbhJob(pointer(r, ro + 1), pointer(a, ao + 1), pb,
                         m, b === nothing ? m : n, 0, 0, k)
    end
    GC.@preserve bufs begin
        check(ccall((:nbh_mul_batch, libnbh), Int32, (Ptr{NbhJob}, Int64), recs, length(recs)))
    end
    return nothing
end

function ntt_len(T::Int)
    N = Ref{Int64}(0)
    check(ccall((:nbh_ntt_len, libnbh), Int32, (Int64, Ptr{Int64}), T, N))
    return Int(N[])
end

function fp_ntt_params(bits_a::Int, bits_b::Int)
    b = Ref{Int32}(0)
    N = Ref{Int64}(0)
    check(ccall((:nbh_ntt_params, libnbh), Int32, (Int64, Int64, Ptr{Int32}, Ptr{Int64}),
                bits_a, bits_b, b, N))
    return Int(b[]), Int(N[])
end

# Install into NativeBigInt: redefine the engine pair so mul!/sqr!/mul_bal!
# (src/mul.jl:134,146,161,389) dispatch to the GPU above the thresholds.
function install!(NB::Module)
    @eval NB begin
        mul_fpntt2!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, m::Int,
                    b::Memory{UInt64}, bo::Int, n::Int) =
            $(NativeBigIntHIP).mul_fpntt2!(r, ro, a, ao, m, b, bo, n)
        sqr_fpntt2!(r::Memory{UInt64}, ro::Int, a::Memory{UInt64}, ao::Int, n::Int) =
            $(NativeBigIntHIP).sqr_fpntt2!(r, ro, a, ao, n)
    end
    return nothing
end

end # module
