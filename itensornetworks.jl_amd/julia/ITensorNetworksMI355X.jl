# ITensorNetworksMI355X: ccall binding of libitnbp.so for ITensorNetworks.jl.
#
# Plugs the MI355X library into the reference's BP seams without changing
# user code:
#   * `MI355XCache <: AbstractBeliefPropagationCache` implements the required
#     cache interface (src/caches/abstractbeliefpropagationcache.jl:44-69) and
#     overrides `update(::Algorithm"bp", ::MI355XCache)` (:313) to run whole
#     sweeps on the device;
#   * `environment`, `region_scalar`, `logscalar` and the `expect` driver
#     (src/expect.jl:21-41) read the device messages;
#   * `apply(o, psi; envs::MI355XCache, ...)` runs simple_update_bp
#     (src/apply.jl:33-95) on the device;
#   * `ChainRulesCore.rrule(bp_energy, ...)` wraps the library's adjoint.
#
# Marshalling: every site tensor is permuted to [siteind; link inds in
# incident_edges order] and passed column-major; messages are chi x chi
# column-major M[b_ket, b_bra]; the library copies everything it is given.
#
# This file is source shipped with the build; the build image has no Julia
# runtime, so it is exercised only through the identical C ABI calls of the
# Python host package and its tests.
module ITensorNetworksMI355X

using ITensors: ITensors, ITensor, Index, inds, dim, array, itensor, permute, commonind, siteinds
using ITensorNetworks: ITensorNetworks, AbstractITensorNetwork, AbstractBeliefPropagationCache,
    ITensorNetwork, linkinds, vertices, edges, neighbors, incident_edges
using NDTensors: @Algorithm_str, Algorithm
using ChainRulesCore: ChainRulesCore, NoTangent

const libitnbp = get(ENV, "ITNBP_LIB", joinpath(@__DIR__, "..", "itnbp", "libitnbp.so"))

const ITNBP_OK = Cint(0)
const ITNBP_F64 = Cint(0)
const ITNBP_SCHED_SEQUENTIAL = Cint(0)
const ITNBP_SCHED_PARALLEL = Cint(1)
const ITNBP_PREC_F64 = Cint(0)

struct UpdateOpts
    schedule::Cint
    maxiter::Cint
    tol::Cdouble
    res_tol::Cdouble
    precision::Cint
    switch_tol::Cdouble
    normalize::Cint
    nseq::Cint
    seq::Ptr{Cint}
    verbose::Cint
end

struct UpdateResult
    iterations::Cint
    iterations_low::Cint
    mean_diff::Cdouble
    max_residual::Cdouble
    converged::Cint
    device_ms::Cdouble
end

mutable struct Context
    h::Ptr{Cvoid}
end

function check(ctx::Context, rc::Cint)
    rc == ITNBP_OK && return nothing
    msg = unsafe_string(ccall((:itnbp_last_error, libitnbp), Cstring, (Ptr{Cvoid},), ctx.h))
    return error("libitnbp error $rc: $msg")
end

function Context(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:itnbp_create, libitnbp), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
    rc == ITNBP_OK || error("itnbp_create failed ($rc)")
    ctx = Context(h[])
    finalizer(c -> ccall((:itnbp_destroy, libitnbp), Cint, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

const DEFAULT_CONTEXT = Ref{Union{Nothing, Context}}(nothing)
default_context() = something(DEFAULT_CONTEXT[], (DEFAULT_CONTEXT[] = Context(); DEFAULT_CONTEXT[]))

# -- network description --------------------------------------------------------

struct Layout
    verts::Vector                # vertex order
    vindex::Dict                 # vertex -> 0-based index
    edges::Vector                # edge list (src, dst)
    eindex::Dict                 # (src, dst) and (dst, src) -> 0-based edge id
    axes::Dict                   # vertex -> [siteind; link inds in incident order]
end

function Layout(psi::AbstractITensorNetwork)
    verts = collect(vertices(psi))
    vindex = Dict(v => i - 1 for (i, v) in enumerate(verts))
    es = [(ITensorNetworks.src(e), ITensorNetworks.dst(e)) for e in edges(psi)]
    eindex = Dict{Any, Int}()
    for (i, (a, b)) in enumerate(es)
        eindex[(a, b)] = i - 1
        eindex[(b, a)] = i - 1
    end
    axes = Dict(v => [only(siteinds(psi, v)); [only(linkinds(psi, e)) for e in incident_edges(psi, v)]] for v in verts)
    return Layout(verts, vindex, es, eindex, axes)
end

mutable struct MI355XCache{V} <: AbstractBeliefPropagationCache{V, V}
    ctx::Context
    net::Ptr{Cvoid}
    psi::ITensorNetwork{V}
    layout::Layout
end

function flat_sites(psi, L::Layout)
    return reduce(vcat, [vec(array(permute(psi[v], L.axes[v]...))) for v in L.verts]; init = Float64[])
end

function MI355XCache(psi::ITensorNetwork; ctx::Context = default_context(), identity_init::Bool = true)
    L = Layout(psi)
    nv, ne = length(L.verts), length(L.edges)
    e_flat = Cint[]
    for (a, b) in L.edges
        push!(e_flat, L.vindex[a], L.vindex[b])
    end
    inc_ptr, inc_edge = Cint[0], Cint[]
    for v in L.verts
        for w in neighbors(psi, v)
            push!(inc_edge, L.eindex[(v, w)])
        end
        push!(inc_ptr, length(inc_edge))
    end
    chis = Cint[dim(only(linkinds(psi, e))) for e in edges(psi)]
    phys = Cint[dim(only(siteinds(psi, v))) for v in L.verts]
    net = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:itnbp_net_create, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint, Ref{Ptr{Cvoid}}),
        ctx.h, nv, ne, e_flat, inc_ptr, inc_edge, chis, phys, ITNBP_F64, net))
    bpc = MI355XCache(ctx, net[], psi, L)
    finalizer(c -> ccall((:itnbp_net_destroy, libitnbp), Cint, (Ptr{Cvoid},), c.net), bpc)
    check(ctx, ccall((:itnbp_set_sites, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}), bpc.net, flat_sites(psi, L)))
    if identity_init && !ITensorNetworks.is_tree(psi)
        check(ctx, ccall((:itnbp_init_identity_messages, libitnbp), Cint, (Ptr{Cvoid},), bpc.net))
    end
    return bpc
end

function Base.copy(bpc::MI355XCache)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(bpc.ctx, ccall((:itnbp_net_copy, libitnbp), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), bpc.net, out))
    c = MI355XCache(bpc.ctx, out[], copy(bpc.psi), bpc.layout)
    finalizer(x -> ccall((:itnbp_net_destroy, libitnbp), Cint, (Ptr{Cvoid},), x.net), c)
    return c
end

message_id(bpc::MI355XCache, a, b) = begin
    e = bpc.layout.eindex[(a, b)]
    bpc.layout.edges[e + 1] == (a, b) ? 2e : 2e + 1
end

# message(bpc, e) (abstractbeliefpropagationcache.jl:173-175) as an ITensor on the
# link index and its bra-side partner
function message_array(bpc::MI355XCache, a, b)
    chi = dim(only(linkinds(bpc.psi, a => b)))
    m = Matrix{Float64}(undef, chi, chi)
    check(bpc.ctx, ccall((:itnbp_get_message, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
        bpc.net, message_id(bpc, a, b), m))
    return m
end

# update(alg::Algorithm"bp", bpc) (abstractbeliefpropagationcache.jl:313-329) on the device
function ITensorNetworks.update(alg::Algorithm"bp", bpc::MI355XCache)
    isnothing(alg.kwargs.maxiter) && error("You need to specify a number of iterations for BP!")
    out = copy(bpc)
    parallel = alg.kwargs.edge_sequence isa AbstractVector{<:AbstractVector}
    opts = UpdateOpts(parallel ? ITNBP_SCHED_PARALLEL : ITNBP_SCHED_SEQUENTIAL, alg.kwargs.maxiter,
        isnothing(alg.kwargs.tol) ? -1.0 : alg.kwargs.tol, -1.0, ITNBP_PREC_F64, 0.0,
        alg.kwargs.message_update_alg.kwargs.normalize, 0, C_NULL, alg.kwargs.verbose)
    res = Ref{UpdateResult}()
    check(out.ctx, ccall((:itnbp_update, libitnbp), Cint, (Ptr{Cvoid}, Ref{UpdateOpts}, Ref{UpdateResult}),
        out.net, Ref(opts), res))
    return out
end

# expect(psi, op; alg = "bp") driver (expect.jl:21-41): <O_v> for every vertex
function expect_mi355x(bpc::MI355XCache, op::AbstractMatrix{<:Real}, verts = bpc.layout.verts)
    ids = Cint[bpc.layout.vindex[v] for v in verts]
    ops = reduce(vcat, [vec(Float64.(op)) for _ in verts]; init = Float64[])
    out = Vector{Float64}(undef, length(verts))
    check(bpc.ctx, ccall((:itnbp_expect1, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}, Ptr{Float64}), bpc.net, length(ids), ids, ops, out))
    return Dict(zip(verts, out))
end

function ITensorNetworks.logscalar(bpc::MI355XCache)
    re, im, ninf = Ref(0.0), Ref(0.0), Ref{Cint}(0)
    check(bpc.ctx, ccall((:itnbp_logscalar, libitnbp), Cint,
        (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Cint}), bpc.net, re, im, ninf))
    ninf[] != 0 && return -Inf
    return im[] == 0 ? re[] : complex(re[], im[])
end

# apply(o, psi; envs = ::MI355XCache, maxdim, cutoff, normalize, callback) (apply.jl:97-146)
function apply_mi355x(gate::AbstractMatrix{<:Real}, (a, b), bpc::MI355XCache; maxdim = 0, cutoff = -1.0,
        normalize = false, callback = (; singular_values, truncation_error) -> nothing)
    e = bpc.layout.eindex[(a, b)]
    g = Float64.(gate)
    if bpc.layout.edges[e + 1] != (a, b)   # gate rows/cols ordered (s_a, s_b): swap to the edge order
        da, db = isqrt(size(g, 1)), isqrt(size(g, 1))
        g = reshape(permutedims(reshape(g, da, db, da, db), (2, 1, 4, 3)), da * db, da * db)
    end
    te, sv, kept = Ref(0.0), zeros(32), Ref{Cint}(0)
    check(bpc.ctx, ccall((:itnbp_apply_gates, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}, Cint, Cdouble, Cint, Ref{Float64}, Ptr{Float64}, Cint, Ref{Cint}),
        bpc.net, 1, Cint[e], vec(g), something(maxdim, 0), something(cutoff, -1.0), normalize, te, sv, 32, kept))
    callback(; singular_values = sv[1:kept[]], truncation_error = te[])
    return bpc
end

# BP energy and its pullback: ChainRulesCore.rrule over the library's adjoint
function bp_energy(psi::ITensorNetwork, nsweeps::Integer, ops1::Vector, ops2::Vector; checkpoint = 1)
    return first(bp_energy_and_gradient(psi, nsweeps, ops1, ops2; checkpoint))
end

function bp_energy_and_gradient(psi::ITensorNetwork, nsweeps::Integer, ops1::Vector, ops2::Vector; checkpoint = 1)
    bpc = MI355XCache(psi; identity_init = false)
    n = ccall((:itnbp_net_psi_numel, libitnbp), Int64, (Ptr{Cvoid},), bpc.net)
    grad = Vector{Float64}(undef, n)
    E = Ref(0.0)
    o1 = reduce(vcat, vec.(ops1); init = Float64[])
    o2 = reduce(vcat, vec.(ops2); init = Float64[])
    check(bpc.ctx, ccall((:itnbp_energy_grad, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ref{Float64}, Ptr{Float64}),
        bpc.net, nsweeps, checkpoint, isempty(o1) ? C_NULL : o1, isempty(o2) ? C_NULL : o2, E, grad))
    L = bpc.layout
    gts = Dict{Any, ITensor}()
    off = 0
    for v in L.verts
        ax = L.axes[v]
        len = prod(dim.(ax))
        gts[v] = permute(itensor(grad[(off + 1):(off + len)], ax...), inds(psi[v])...)
        off += len
    end
    return E[], gts
end

function ChainRulesCore.rrule(::typeof(bp_energy), psi::ITensorNetwork, nsweeps::Integer, ops1, ops2; kwargs...)
    E, g = bp_energy_and_gradient(psi, nsweeps, ops1, ops2; kwargs...)
    function bp_energy_pullback(Ebar)
        dpsi = ChainRulesCore.Tangent{typeof(psi)}(; Dict(v => Ebar * t for (v, t) in g)...)
        return (NoTangent(), dpsi, NoTangent(), NoTangent(), NoTangent())
    end
    return E, bp_energy_pullback
end

end # module
