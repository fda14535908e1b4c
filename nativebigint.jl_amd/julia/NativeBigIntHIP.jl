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
    # nbh_shutdown refuses (and changes nothing) while handles are live; at
    # exit that only means the process frees them
    atexit(() -> ccall((:nbh_shutdown, libnbh), Int32, ()))
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

# Device buffers for callers with no GPU array type (nbh_alloc & co.).
function dev_alloc(nlimbs::Int)
    p = Ref{Ptr{UInt64}}(C_NULL)
    check(ccall((:nbh_alloc, libnbh), Int32, (Int64, Ref{Ptr{UInt64}}), 8 * nlimbs, p))
    return p[]
end
dev_free(p::Ptr{UInt64}) = ccall((:nbh_free, libnbh), Cvoid, (Ptr{UInt64},), p)
function upload!(d::Ptr{UInt64}, x::Memory{UInt64}, n::Int=length(x))
    GC.@preserve x check(ccall((:nbh_upload, libnbh), Int32, (Ptr{UInt64}, Ptr{UInt64}, Int64),
                               d, pointer(x), 8 * n))
end
function download!(x::Memory{UInt64}, d::Ptr{UInt64}, n::Int=length(x))
    GC.@preserve x check(ccall((:nbh_download, libnbh), Int32, (Ptr{UInt64}, Ptr{UInt64}, Int64),
                               pointer(x), d, 8 * n))
end

# The serving queue (nbh_server_*): submit device-resident products as they
# arrive, poll for finished tickets.  b == C_NULL is a square; bits <= 0
# reads the top limb.
mutable struct Server
    h::Ptr{Cvoid}
end
function Server(budget_bytes::Int; nstreams::Int=4, max_batch::Int=256)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:nbh_server_create, libnbh), Int32, (Int64, Int32, Int32, Ref{Ptr{Cvoid}}),
                budget_bytes, nstreams, max_batch, h))
    s = Server(h[])
    finalizer(x -> ccall((:nbh_server_destroy, libnbh), Cvoid, (Ptr{Cvoid},), x.h), s)
    return s
end
function submit!(s::Server, r::Ptr{UInt64}, a::Ptr{UInt64}, m::Int, b::Ptr{UInt64}, n::Int;
                 bits_a::Int=0, bits_b::Int=0)
    t = Ref{Int64}(0)
    job = Ref(NbhJob(r, a, b, m, b == C_NULL ? m : n, bits_a, bits_b, 0))
    check(ccall((:nbh_server_submit, libnbh), Int32, (Ptr{Cvoid}, Ref{NbhJob}, Ref{Int64}), s.h, job, t))
    return t[]
end
function poll!(s::Server, done::Vector{Int64})
    n = Ref{Int64}(0)
    check(ccall((:nbh_server_poll, libnbh), Int32, (Ptr{Cvoid}, Ptr{Int64}, Int64, Ref{Int64}),
                s.h, done, length(done), n))
    return Int(n[])
end
drain!(s::Server) = check(ccall((:nbh_server_drain, libnbh), Int32, (Ptr{Cvoid},), s.h))

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
