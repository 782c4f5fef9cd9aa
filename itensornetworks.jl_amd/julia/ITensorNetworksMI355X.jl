# ITensorNetworksMI355X: ccall binding of libitnbp.so for ITensorNetworks.jl.
#
# Plugs the MI355X library into the reference's BP seams:
#   * `MI355XCache <: AbstractBeliefPropagationCache` implements the cache
#     interface the reference requires of a cache type
#     (src/caches/abstractbeliefpropagationcache.jl:44-69):
#     `partitioned_tensornetwork`, `messages`, `copy`, `default_update_alg`,
#     `default_message_update_alg`, `default_bp_maxiter`, `default_bp_edge_sequence`,
#     `environment`, `region_scalar` (vertex and edge), `partitions`,
#     `quotientedges`, `partitioned_vertices`; messages live on the device and
#     `message` / `set_message!` / `set_messages!` read and write them there.
#   * `update(::Algorithm"bp", ::MI355XCache)` (:313) runs whole sweeps on the
#     device, in the order of the algorithm's `edge_sequence` (a sequential
#     order, or edge groups as in :294-308).
#   * `environment(bpc, verts)`: the operator vertices of one site or of two
#     neighbouring sites give the environment already contracted on the device
#     (the one- or two-site BP density matrix), so the reference's
#     `expect(ψ, op; alg = "bp", cache!)` (src/expect.jl:5-41) and
#     `environment(tn, verts; cache!)` (src/environment.jl:21-39) reach the
#     device through `cache![]` unchanged; any other vertex set gives the
#     incoming messages (from the device) and the partition tensors.
#   * `region_scalar`, `logscalar` and `rescale` run on the device.
#   * `apply(o, ψ, bpc::MI355XCache; maxdim, cutoff, normalize, callback)`
#     runs `simple_update_bp` (src/apply.jl:33-95) on the device with the
#     cache's messages as the product environment; the vector form applies
#     gates on disjoint bonds in one launch.
#   * `ChainRulesCore.rrule(bp_energy, ...)` wraps the library's adjoint.
#   * `inner(ϕ, [A,] ψ; alg = "bp")` and `loginner` (src/inner.jl:100-160) run
#     BP on the LinearFormNetwork on the device's two-layer kernels (bra and
#     ket of their own bond dimensions, real or complex).
#
# Marshalling: every site tensor is permuted to [siteind; link inds in
# incident_edges order] and passed column-major; messages are chi x chi
# column-major M[b_ket, b_bra]; the library copies everything it is given.
#
# The build image has no Julia runtime: this module is exercised through the
# identical C ABI calls of the Python host package, and a CPU test checks
# every ccall below against include/itnbp.h (symbol, argument count, C types).
module ITensorNetworksMI355X

using ITensors: ITensors, ITensor, Index, inds, dim, array, itensor, permute, dag, commoninds, Op
using ITensorNetworks: ITensorNetworks, AbstractITensorNetwork, AbstractBeliefPropagationCache,
    ITensorNetwork, QuadraticFormNetwork, linkinds, siteinds, vertices, edges, neighbors, incident_edges,
    tensornetwork
using NamedGraphs.PartitionedGraphs: PartitionedGraphs, PartitionedGraph, QuotientEdge, QuotientVertex,
    quotientedges, quotientvertices, partitioned_vertices, quotient_graph
using NamedGraphs.GraphsExtensions: is_tree
using Graphs: src, dst
using DataGraphs: DataGraphs
using Dictionaries: Dictionary, set!
using NDTensors: @Algorithm_str, Algorithm
using ChainRulesCore: ChainRulesCore, NoTangent

const libitnbp = get(ENV, "ITNBP_LIB", joinpath(@__DIR__, "..", "itnbp", "libitnbp.so"))

const ITNBP_OK = Cint(0)
const ITNBP_F64 = Cint(0)
const ITNBP_SCHED_SEQUENTIAL = Cint(0)
const ITNBP_SCHED_PARALLEL = Cint(1)
# precision modes (include/itnbp.h)
const PRECISIONS = Dict("f64" => Cint(0), "f32" => Cint(1), "bf16_f64" => Cint(2), "bf16x3" => Cint(3),
    "bf16x3_f64" => Cint(4), "bf16_bf16x3" => Cint(5), "bf16_bf16x3_f64" => Cint(6), "f16x2" => Cint(7),
    "bf16_f16x2" => Cint(8))

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
function default_context()
    isnothing(DEFAULT_CONTEXT[]) && (DEFAULT_CONTEXT[] = Context())
    return DEFAULT_CONTEXT[]
end

# -- network description --------------------------------------------------------

struct Layout
    verts::Vector                # site vertex order (device vertex i = verts[i + 1])
    vindex::Dict                 # site vertex -> 0-based index
    edges::Vector                # edge list (a, b) of the site graph, device edge order
    eindex::Dict                 # (a, b) and (b, a) -> 0-based edge id
    axes::Dict                   # site vertex -> [siteind; link inds in incident order]
    links::Dict                  # (a, b) -> the ket link index between a and b
end

function Layout(psi::AbstractITensorNetwork)
    verts = collect(vertices(psi))
    vindex = Dict(v => i - 1 for (i, v) in enumerate(verts))
    es = [(src(e), dst(e)) for e in edges(psi)]
    eindex = Dict{Any, Int}()
    links = Dict{Any, Index}()
    for (i, (a, b)) in enumerate(es)
        eindex[(a, b)] = i - 1
        eindex[(b, a)] = i - 1
        l = only(linkinds(psi, a => b))
        links[(a, b)] = l
        links[(b, a)] = l
    end
    axes = Dict(v => [only(siteinds(psi, v)); [links[(v, w)] for w in neighbors(psi, v)]] for v in verts)
    return Layout(verts, vindex, es, eindex, axes, links)
end

# directed message id of the C ABI: 2e for edges[e] = (a, b), 2e + 1 reversed
function message_id(L::Layout, a, b)
    e = L.eindex[(a, b)]
    return L.edges[e + 1] == (a, b) ? 2e : 2e + 1
end

flat_sites(psi, L::Layout) =
    reduce(vcat, [vec(array(permute(psi[v], L.axes[v]...))) for v in L.verts]; init = Float64[])

# -- the cache --------------------------------------------------------------------

"""
    MI355XCache(ψ; ctx, precision = "f64", switch_tol = 1e-4)

BP cache of the norm network ⟨ψ|ψ⟩ (one {operator, bra, ket} partition per site)
whose messages live on an MI355X.  `precision` selects the arithmetic of
`update` (include/itnbp.h `ITNBP_PREC_*`, e.g. "bf16_f16x2").
"""
mutable struct MI355XCache{V, PV, PTN} <: AbstractBeliefPropagationCache{V, PV}
    ptn::PTN                 # PartitionedGraph of the QuadraticFormNetwork (host tensors)
    qf::QuadraticFormNetwork
    ctx::Context
    net::Ptr{Cvoid}
    layout::Layout
    site_of::Dict            # quotient vertex -> site vertex
    precision::String
    switch_tol::Float64
end

function attach_finalizer!(bpc::MI355XCache)
    finalizer(c -> ccall((:itnbp_net_destroy, libitnbp), Cint, (Ptr{Cvoid},), c.net), bpc)
    return bpc
end

function MI355XCache(psi::ITensorNetwork; ctx::Context = default_context(), precision::String = "f64",
        switch_tol::Real = 1.0e-4, identity_init::Bool = true)
    haskey(PRECISIONS, precision) || error("unknown precision $precision")
    qf = QuadraticFormNetwork(psi)
    pv = ITensorNetworks.default_partitioned_vertices(qf)
    ptn = PartitionedGraph(tensornetwork(qf), pv)
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
    chis = Cint[dim(L.links[e]) for e in L.edges]
    phys = Cint[dim(first(L.axes[v])) for v in L.verts]
    net = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:itnbp_net_create, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint, Ref{Ptr{Cvoid}}),
        ctx.h, nv, ne, e_flat, inc_ptr, inc_edge, chis, phys, ITNBP_F64, net))
    site_of = Dict(pvx => first(first(partitioned_vertices(ptn)[pvx])) for pvx in quotientvertices(ptn))
    bpc = attach_finalizer!(MI355XCache(ptn, qf, ctx, net[], L, site_of, precision, Float64(switch_tol)))
    check(ctx, ccall((:itnbp_set_sites, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}), bpc.net, flat_sites(psi, L)))
    # initialize_cache (src/initialize_cache.jl:14-29): identity messages on loopy graphs
    if identity_init && !is_tree(psi)
        check(ctx, ccall((:itnbp_init_identity_messages, libitnbp), Cint, (Ptr{Cvoid},), bpc.net))
    end
    return bpc
end

ITensorNetworks.initialize_cache(::Algorithm"mi355x", psi::ITensorNetwork; kwargs...) = MI355XCache(psi; kwargs...)

function Base.copy(bpc::MI355XCache)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(bpc.ctx, ccall((:itnbp_net_copy, libitnbp), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), bpc.net, out))
    c = MI355XCache(copy(bpc.ptn), bpc.qf, bpc.ctx, out[], bpc.layout, bpc.site_of, bpc.precision, bpc.switch_tol)
    return attach_finalizer!(c)
end

# the ket state held by the cache (site vertices, tensors as on the device)
ket_state(bpc::MI355XCache) = ITensorNetwork(Dict(v => bpc.ptn[(v, "ket")] for v in bpc.layout.verts))

# refresh the host tensors (ket and bra layers) from the device after a device-side change
function pull_sites!(bpc::MI355XCache)
    L = bpc.layout
    n = ccall((:itnbp_net_psi_numel, libitnbp), Int64, (Ptr{Cvoid},), bpc.net)
    flat = Vector{Float64}(undef, n)
    check(bpc.ctx, ccall((:itnbp_get_sites, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}), bpc.net, flat))
    off = 0
    tn = tensornetwork(bpc)
    for v in L.verts
        ax = L.axes[v]
        len = prod(dim.(ax))
        t = itensor(flat[(off + 1):(off + len)], ax...)
        old_ket = tn[(v, "ket")]
        old_bra = tn[(v, "bra")]
        DataGraphs.set_vertex_data!(tn, permute(t, inds(old_ket)...), (v, "ket"))
        bra = ITensors.replaceinds(dag(t), ax, ITensorNetworks.dual_index_map(bpc.qf).(ax))
        DataGraphs.set_vertex_data!(tn, permute(bra, inds(old_bra)...), (v, "bra"))
        off += len
    end
    return bpc
end

# -- the interface required of a cache type (abstractbeliefpropagationcache.jl:44-69)

ITensorNetworks.partitioned_tensornetwork(bpc::MI355XCache) = bpc.ptn
ITensorNetworks.default_update_alg(::MI355XCache) = "bp"
ITensorNetworks.default_message_update_alg(::MI355XCache) = "contract"
ITensorNetworks.default_bp_maxiter(bpc::MI355XCache) = ITensorNetworks.default_bp_maxiter(quotient_graph(bpc.ptn))
ITensorNetworks.default_bp_edge_sequence(bpc::MI355XCache) = ITensorNetworks.default_edge_sequence(bpc.ptn)
ITensorNetworks.partitions(bpc::MI355XCache) = quotientvertices(bpc.ptn)
PartitionedGraphs.quotientedges(bpc::MI355XCache) = quotientedges(bpc.ptn)
PartitionedGraphs.partitioned_vertices(bpc::MI355XCache) = partitioned_vertices(bpc.ptn)

# the quotient edge pe = (partition of site a) -> (partition of site b)
sites_of(bpc::MI355XCache, pe::QuotientEdge) = (bpc.site_of[src(pe)], bpc.site_of[dst(pe)])

function message_is_set(bpc::MI355XCache, a, b)
    return ccall((:itnbp_message_is_set, libitnbp), Cint, (Ptr{Cvoid}, Cint), bpc.net,
        message_id(bpc.layout, a, b)) == 1
end

# M[b_ket, b_bra] of a -> b as the reference's message: [M] on (ket link k, dual of k)
function message_tensor(bpc::MI355XCache, a, b)
    L = bpc.layout
    k = L.links[(a, b)]
    chi = dim(k)
    m = Matrix{Float64}(undef, chi, chi)
    check(bpc.ctx, ccall((:itnbp_get_message, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
        bpc.net, message_id(L, a, b), m))
    return itensor(m, k, dag(ITensorNetworks.dual_index_map(bpc.qf)(k)))
end

function ITensorNetworks.message(bpc::MI355XCache, pe::QuotientEdge)
    a, b = sites_of(bpc, pe)
    message_is_set(bpc, a, b) || error("message $pe is not set (messages have no default)")
    return ITensor[message_tensor(bpc, a, b)]
end

function ITensorNetworks.messages(bpc::MI355XCache)
    ms = Dictionary{QuotientEdge, Vector{ITensor}}()
    for pe in quotientedges(bpc.ptn), e in (pe, reverse(pe))
        a, b = sites_of(bpc, e)
        message_is_set(bpc, a, b) && set!(ms, e, ITensor[message_tensor(bpc, a, b)])
    end
    return ms
end

function ITensorNetworks.set_message!(bpc::MI355XCache, pe::QuotientEdge, message)
    a, b = sites_of(bpc, pe)
    L = bpc.layout
    k = L.links[(a, b)]
    t = ITensors.contract(message...)  # one tensor on (k, dual of k)
    m = array(permute(t, k, dag(ITensorNetworks.dual_index_map(bpc.qf)(k))))
    check(bpc.ctx, ccall((:itnbp_set_message, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
        bpc.net, message_id(L, a, b), Matrix{Float64}(m)))
    return bpc
end

function ITensorNetworks.set_messages!(bpc::MI355XCache, quotientedges_messages)
    for pe in eachindex(quotientedges_messages)
        ITensorNetworks.set_message!(bpc, pe, quotientedges_messages[pe])
    end
    return bpc
end

# bpc[v] = t on a ket vertex replaces the site tensor (the bra follows as its dual)
function DataGraphs.set_vertex_data!(bpc::MI355XCache, value, vertex)
    last(vertex) == "ket" || error("MI355XCache stores the ket layer; set the ket vertex (the bra follows)")
    v = first(vertex)
    L = bpc.layout
    data = vec(array(permute(value, L.axes[v]...)))
    check(bpc.ctx, ccall((:itnbp_set_site, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
        bpc.net, L.vindex[v], data))
    return pull_sites!(bpc)
end

# device-side region scalars: Z_v for every partition, Z_e for every quotient edge
function region_scalars(bpc::MI355XCache)
    L = bpc.layout
    vs = Vector{Float64}(undef, length(L.verts))
    es = Vector{Float64}(undef, max(length(L.edges), 1))
    check(bpc.ctx, ccall((:itnbp_region_scalars, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
        bpc.net, vs, es))
    return vs, es
end

# region_scalar(bpc, ::QuotientVertex) (src/caches/beliefpropagationcache.jl:107-113)
function ITensorNetworks.region_scalar(bpc::MI355XCache, pv::QuotientVertex; kwargs...)
    vs, _ = region_scalars(bpc)
    return vs[bpc.layout.vindex[bpc.site_of[pv]] + 1]
end

# region_scalar(bpc, ::QuotientEdge) (src/caches/beliefpropagationcache.jl:115-119)
function ITensorNetworks.region_scalar(bpc::MI355XCache, pe::QuotientEdge; kwargs...)
    _, es = region_scalars(bpc)
    a, b = sites_of(bpc, pe)
    return es[bpc.layout.eindex[(a, b)] + 1]
end

# logscalar(bpc) (abstractbeliefpropagationcache.jl:397-408) in one device call
function ITensorNetworks.logscalar(bpc::MI355XCache)
    re, im, ninf = Ref(0.0), Ref(0.0), Ref{Cint}(0)
    check(bpc.ctx, ccall((:itnbp_logscalar, libitnbp), Cint,
        (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ref{Cint}), bpc.net, re, im, ninf))
    ninf[] != 0 && return -Inf
    return im[] == 0 ? re[] : complex(re[], im[])
end

# rescale(bpc) = rescale_messages then rescale_partitions (:349-395) on the device
function ITensorNetworks.rescale(bpc::MI355XCache; kwargs...)
    out = copy(bpc)
    check(out.ctx, ccall((:itnbp_normalize, libitnbp), Cint, (Ptr{Cvoid},), out.net))
    return pull_sites!(out)
end

# environment(bpc, verts) (src/caches/beliefpropagationcache.jl:100-105).  The
# operator vertices of one site, or of two neighbouring sites, give the BP
# density matrix of that site or pair contracted on the device: the single
# tensor rho[s, s'] (rho[s_u, s_v, s_u', s_v']) on the operator's indices, which
# is the contraction of the reference's list of incoming messages and bra/ket
# tensors.  Every other vertex set gives that list itself.
function ITensorNetworks.environment(bpc::MI355XCache, verts::Vector; kwargs...)
    L = bpc.layout
    if all(v -> last(v) == "operator", verts) && length(verts) in (1, 2)
        sites = first.(verts)
        if length(sites) == 1
            return ITensor[site_rdm(bpc, only(sites))]
        elseif haskey(L.eindex, (sites[1], sites[2]))
            return ITensor[pair_rdm(bpc, sites[1], sites[2])]
        end
    end
    pverts = unique([PartitionedGraphs.quotientvertex(bpc.ptn, v) for v in verts])
    ms = ITensorNetworks.incoming_messages(bpc, pverts)
    inner_verts = setdiff(reduce(vcat, [partitioned_vertices(bpc.ptn)[pvx] for pvx in pverts]), verts)
    return vcat(ms, ITensor[tensornetwork(bpc)[v] for v in inner_verts])
end

function operator_inds(bpc::MI355XCache, v)
    s = first(bpc.layout.axes[v])
    return s, ITensorNetworks.dual_index_map(bpc.qf)(s)
end

function site_rdm(bpc::MI355XCache, v)
    s, sp = operator_inds(bpc, v)
    d = dim(s)
    out = Vector{Float64}(undef, d * d)
    check(bpc.ctx, ccall((:itnbp_rdm1, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}),
        bpc.net, 1, Cint[bpc.layout.vindex[v]], out))
    # rho[s, s'] with s the ket site index and s' its dual (the operator's other index)
    return itensor(out, dag(s), dag(sp))
end

function pair_rdm(bpc::MI355XCache, a, b)
    L = bpc.layout
    e = L.eindex[(a, b)]
    u, w = L.edges[e + 1]
    su, sup = operator_inds(bpc, u)
    sw, swp = operator_inds(bpc, w)
    out = Vector{Float64}(undef, (dim(su) * dim(sw))^2)
    check(bpc.ctx, ccall((:itnbp_rdm2, libitnbp), Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}),
        bpc.net, 1, Cint[e], out))
    # rho[(s_u, s_w), (s_u', s_w')] with the edge's first site fastest
    return itensor(out, dag(su), dag(sw), dag(sup), dag(swp))
end

# -- update ---------------------------------------------------------------------

# set_default_kwargs(::Algorithm"bp", cache) (src/caches/beliefpropagationcache.jl:64-77)
function ITensorNetworks.set_default_kwargs(alg::Algorithm"bp", bpc::MI355XCache)
    verbose = get(alg.kwargs, :verbose, false)
    maxiter = get(alg.kwargs, :maxiter, ITensorNetworks.default_bp_maxiter(bpc))
    edge_sequence = get(alg.kwargs, :edge_sequence, ITensorNetworks.default_bp_edge_sequence(bpc))
    tol = get(alg.kwargs, :tol, nothing)
    message_update_alg = ITensorNetworks.set_default_kwargs(
        get(alg.kwargs, :message_update_alg, Algorithm(ITensorNetworks.default_message_update_alg(bpc)))
    )
    return Algorithm("bp"; verbose, maxiter, edge_sequence, tol, message_update_alg)
end

ids_of(bpc::MI355XCache, pes) = Cint[message_id(bpc.layout, sites_of(bpc, pe)...) for pe in pes]

function run_update!(bpc::MI355XCache, schedule::Cint, seq::Vector{Cint}, maxiter, tol, normalize, verbose)
    opts = UpdateOpts(schedule, maxiter, isnothing(tol) ? -1.0 : tol, -1.0, PRECISIONS[bpc.precision],
        bpc.switch_tol, normalize ? 1 : 0, length(seq), isempty(seq) ? C_NULL : pointer(seq), verbose ? 1 : 0)
    res = Ref{UpdateResult}()
    GC.@preserve seq begin
        check(bpc.ctx, ccall((:itnbp_update, libitnbp), Cint,
            (Ptr{Cvoid}, Ref{UpdateOpts}, Ref{UpdateResult}), bpc.net, Ref(opts), res))
    end
    return res[]
end

# update(alg::Algorithm"bp", bpc) (abstractbeliefpropagationcache.jl:313-329) on the device
function ITensorNetworks.update(alg::Algorithm"bp", bpc::MI355XCache)
    isnothing(alg.kwargs.maxiter) && error("You need to specify a number of iterations for BP!")
    out = copy(bpc)
    seq = alg.kwargs.edge_sequence
    normalize = alg.kwargs.message_update_alg.kwargs.normalize
    if seq isa AbstractVector{<:QuotientEdge}
        run_update!(out, ITNBP_SCHED_SEQUENTIAL, ids_of(out, seq), alg.kwargs.maxiter, alg.kwargs.tol, normalize,
            alg.kwargs.verbose)
    elseif all(g -> length(g) == 1, seq)
        # singleton groups: one Jacobi sweep per iteration (:294-308)
        run_update!(out, ITNBP_SCHED_PARALLEL, ids_of(out, only.(seq)), alg.kwargs.maxiter, alg.kwargs.tol,
            normalize, alg.kwargs.verbose)
    else
        # groups of several edges: each group sequential from the sweep's start, merged
        for i in 1:alg.kwargs.maxiter
            merged = Dictionary{QuotientEdge, Vector{ITensor}}()
            diff = 0.0
            for group in seq
                work = copy(out)
                r = run_update!(work, ITNBP_SCHED_SEQUENTIAL, ids_of(work, group), 1, 0.0, normalize, false)
                diff += r.mean_diff * length(group)
                for pe in group
                    set!(merged, pe, ITensorNetworks.message(work, pe))
                end
            end
            ITensorNetworks.set_messages!(out, merged)
            if !isnothing(alg.kwargs.tol) && diff / length(seq) <= alg.kwargs.tol
                alg.kwargs.verbose && println("BP converged to desired precision after $i iterations.")
                break
            end
        end
    end
    return out
end

# -- apply ----------------------------------------------------------------------

# the 4-index gate o on sites (a, b) as O[(s_a', s_b'), (s_a, s_b)], first site fastest
function gate_matrix(o::ITensor, bpc::MI355XCache, a, b)
    sa = first(bpc.layout.axes[a])
    sb = first(bpc.layout.axes[b])
    t = permute(o, ITensors.prime(sa), ITensors.prime(sb), sa, sb)
    da, db = dim(sa), dim(sb)
    return reshape(array(t), da * db, da * db)
end

# the gate's two sites, in the order of the bond's edge (a, b) in the layout
function gate_sites(o::ITensor, bpc::MI355XCache)
    L = bpc.layout
    vs = [v for v in L.verts if first(L.axes[v]) in inds(o)]
    length(vs) == 2 || error("Gate being applied must act on two sites of the network.")
    a, b = vs
    haskey(L.eindex, (a, b)) || error("Vertices where the gates are being applied must be neighbors for now.")
    e = L.eindex[(a, b)]
    return L.edges[e + 1]
end

"""
    apply(o⃗, ψ, bpc::MI355XCache; maxdim, cutoff, normalize, callback)

`apply(o, ψ; envs, ...)` (src/apply.jl:97-160) with the product environment of
every gate taken from the device messages of `bpc`: each two-site gate runs
`simple_update_bp` on the device (batched FP64 Jacobi eigh / QR / SVD), gates on
disjoint bonds in one launch.  One-site gates act on the host.  Returns the new ψ.
"""
function ITensors.apply(os::Vector{ITensor}, psi::AbstractITensorNetwork, bpc::MI355XCache; maxdim = nothing,
        cutoff = nothing, normalize = false, callback = Returns(nothing))
    work = copy(bpc)
    L = work.layout
    check(work.ctx, ccall((:itnbp_set_sites, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}), work.net,
        flat_sites(psi, L)))
    psi = copy(psi)
    pending = collect(os)
    while !isempty(pending)
        used, batch, rest = Set(), ITensor[], ITensor[]
        for o in pending
            vs = [v for v in L.verts if first(L.axes[v]) in inds(o)]
            if length(vs) == 1
                isempty(batch) || (push!(rest, o); continue)
                t = ITensors.apply(o, psi[only(vs)])
                psi[only(vs)] = normalize ? t / ITensors.norm(t) : t
                continue
            end
            a, b = gate_sites(o, work)
            if a in used || b in used
                push!(rest, o)
            else
                push!(used, a, b)
                push!(batch, o)
            end
        end
        pending = rest
        isempty(batch) && continue
        n = length(batch)
        ids = Cint[]
        mats = Float64[]
        for o in batch
            a, b = gate_sites(o, work)
            push!(ids, L.eindex[(a, b)])
            append!(mats, vec(gate_matrix(o, work, a, b)))
        end
        max_keep = 64  # the largest kept rank (bonds grow to at most 64)
        te = Vector{Float64}(undef, n)
        sv = zeros(n * max_keep)
        kept = Vector{Cint}(undef, n)
        check(work.ctx, ccall((:itnbp_apply_gates, libitnbp), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Float64}, Cint, Cdouble, Cint, Ptr{Float64}, Ptr{Float64}, Cint,
                Ptr{Cint}),
            work.net, n, ids, mats, something(maxdim, 0), something(cutoff, -1.0), normalize ? 1 : 0, te, sv,
            max_keep, kept))
        pull_sites!(work)
        for (k, o) in enumerate(batch)
            a, b = gate_sites(o, work)
            for v in (a, b)
                psi[v] = permute(work.ptn[(v, "ket")], inds(psi[v])...)
            end
            callback(; singular_values = sv[((k - 1) * max_keep + 1):((k - 1) * max_keep + kept[k])],
                truncation_error = te[k])
        end
    end
    return psi
end

ITensors.apply(o::ITensor, psi::AbstractITensorNetwork, bpc::MI355XCache; kwargs...) =
    ITensors.apply(ITensor[o], psi, bpc; kwargs...)

# -- BP energy and its pullback ---------------------------------------------------

"""
    bp_energy(ψ, nsweeps, ops1, ops2; checkpoint = 1)

E = Σ_v ⟨ops1[v]⟩ + Σ_e ⟨ops2[e]⟩ from BP after `nsweeps` Jacobi sweeps from
identity messages.  `ops1[i]` is the d×d matrix O[s', s] of the i-th vertex
of `vertices(ψ)`, `ops2[j]` the (d_a d_b)² matrix of the j-th edge (a, b) of
`edges(ψ)`, first site fastest.  Differentiable through the rrule below.
"""
function bp_energy(psi::ITensorNetwork, nsweeps::Integer, ops1::Vector, ops2::Vector; checkpoint = 1)
    return first(bp_energy_and_gradient(psi, nsweeps, ops1, ops2; checkpoint))
end

function bp_energy_and_gradient(psi::ITensorNetwork, nsweeps::Integer, ops1::Vector, ops2::Vector;
        checkpoint = 1)
    bpc = MI355XCache(psi; identity_init = false)
    L = bpc.layout
    n = ccall((:itnbp_net_psi_numel, libitnbp), Int64, (Ptr{Cvoid},), bpc.net)
    grad = Vector{Float64}(undef, n)
    E = Ref(0.0)
    o1 = reduce(vcat, [vec(Float64.(m)) for m in ops1]; init = Float64[])
    # edge operators in the device edge order; reorder those named (b, a)
    ops2_dev = Dict{Int, Matrix{Float64}}()
    for (j, e) in enumerate(edges(psi))
        a, b = src(e), dst(e)
        m = Float64.(ops2[j])
        if L.edges[L.eindex[(a, b)] + 1] != (a, b)
            da, db = dim(first(L.axes[a])), dim(first(L.axes[b]))
            m = reshape(permutedims(reshape(m, da, db, da, db), (2, 1, 4, 3)), da * db, da * db)
        end
        ops2_dev[L.eindex[(a, b)]] = m
    end
    o2 = reduce(vcat, [vec(ops2_dev[e]) for e in 0:(length(L.edges) - 1) if haskey(ops2_dev, e)]; init = Float64[])
    check(bpc.ctx, ccall((:itnbp_energy_grad, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ref{Float64}, Ptr{Float64}),
        bpc.net, nsweeps, checkpoint, isempty(o1) ? C_NULL : o1, isempty(o2) ? C_NULL : o2, E, grad))
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

# The cotangent of ψ is returned as an ITensorNetwork on ψ's graph holding
# ∂E/∂ψ_v at every vertex (the natural tangent of a network of tensors).
function ChainRulesCore.rrule(::typeof(bp_energy), psi::ITensorNetwork, nsweeps::Integer, ops1, ops2; kwargs...)
    E, g = bp_energy_and_gradient(psi, nsweeps, ops1, ops2; kwargs...)
    function bp_energy_pullback(Ebar)
        dpsi = ITensorNetwork(Dict(v => Ebar * t for (v, t) in g))
        return (NoTangent(), dpsi, NoTangent(), NoTangent(), NoTangent())
    end
    return E, bp_energy_pullback
end

# -- inner products with alg = "bp" on the two-layer kernels (k_layer.hip) --------

# flat column-major data of every site tensor of `tn` permuted to `axes(v)`
# (real, or interleaved (re, im) pairs when `cplx`)
function flat_layer(tn, verts, axes, cplx::Bool)
    parts = [vec(array(permute(tn[v], axes(v)...))) for v in verts]
    if cplx
        z = reduce(vcat, [ComplexF64.(x) for x in parts]; init = ComplexF64[])
        return collect(reinterpret(Float64, z))
    end
    return reduce(vcat, [Float64.(x) for x in parts]; init = Float64[])
end

"""
    loginner_bp(ϕ, ψ; A = nothing, cache_update_kwargs = (;))

log of BP's estimate of ⟨ϕ|ψ⟩ (or ⟨ϕ|A|ψ⟩) on the device: the LinearFormNetwork
(BilinearFormNetwork with the operator's and the ket's link indices merged) with
one {bra, ket} partition per site, no default message (initialize_cache.jl:10-12),
`maxiter` 1 on a tree and required otherwise (abstractbeliefpropagationcache.jl:38,
315-317); then logscalar (:397-412).  `A[v]` carries the site indices of ψ (ket
side) and their primes (bra side).  `cache_update_kwargs` takes `maxiter`, `tol`
and `edge_sequence` ("forest_cover" or "parallel"), as logscalar's do
(contract.jl:41-58).
"""
function loginner_bp(phi::AbstractITensorNetwork, psi::AbstractITensorNetwork; A = nothing,
        cache_update_kwargs = (;), ctx::Context = default_context())
    maxiter = get(cache_update_kwargs, :maxiter, nothing)
    tol = get(cache_update_kwargs, :tol, nothing)
    edge_sequence = get(cache_update_kwargs, :edge_sequence, "forest_cover")
    edge_sequence in ("forest_cover", "parallel") ||
        error("loginner_bp: edge_sequence must be \"forest_cover\" or \"parallel\"")
    L = Layout(psi)
    verts = L.verts
    cplx = any(v -> eltype(psi[v]) <: Complex || eltype(phi[v]) <: Complex ||
                     (A !== nothing && eltype(A[v]) <: Complex), verts)
    blinks = Dict{Any, Index}()
    for (a, b) in L.edges
        l = only(linkinds(phi, a => b))
        blinks[(a, b)] = l
        blinks[(b, a)] = l
    end
    bra_axes(v) = [only(siteinds(phi, v)); [blinks[(v, w)] for w in neighbors(psi, v)]]
    ket_axes(v) = L.axes[v]
    ne = length(L.edges)
    edges_c = Cint[]
    for (a, b) in L.edges
        push!(edges_c, L.vindex[a], L.vindex[b])
    end
    inc_ptr = Cint[0]
    inc_edge = Cint[]
    for v in verts
        for w in neighbors(psi, v)
            push!(inc_edge, L.eindex[(v, w)])
        end
        push!(inc_ptr, length(inc_edge))
    end
    bchi = Cint[dim(blinks[(a, b)]) for (a, b) in L.edges]
    pchi = Cint[dim(L.links[(a, b)]) for (a, b) in L.edges]
    achi = A === nothing ? Cint[] : Cint[dim(only(linkinds(A, a => b))) for (a, b) in L.edges]
    kchi = A === nothing ? pchi : pchi .* achi
    phys = Cint[dim(only(siteinds(psi, v))) for v in verts]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:itnbp_lnet_create, libitnbp), Cint,
        (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Cint,
         Ref{Ptr{Cvoid}}),
        ctx.h, length(verts), ne, edges_c, inc_ptr, inc_edge, kchi, bchi, phys, Cint(cplx), h))
    lnet = h[]
    try
        bra = flat_layer(phi, verts, bra_axes, cplx)
        if A === nothing
            check(ctx, ccall((:itnbp_lnet_set_tensors, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                lnet, flat_layer(psi, verts, ket_axes, cplx), bra))
        else
            nk = ccall((:itnbp_lnet_ket_numel, libitnbp), Int64, (Ptr{Cvoid},), lnet)
            check(ctx, ccall((:itnbp_lnet_set_tensors, libitnbp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                lnet, zeros(Float64, (cplx ? 2 : 1) * nk), bra))
            # A_v[s', s, links of A in the ket's incident order]
            op_axes(v) = begin
                s = only(siteinds(psi, v))
                [ITensors.prime(s); s; [only(linkinds(A, v => w)) for w in neighbors(psi, v)]]
            end
            check(ctx, ccall((:itnbp_lnet_fold_operator, libitnbp), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cint}, Ptr{Float64}, Ptr{Cint}),
                lnet, flat_layer(psi, verts, ket_axes, cplx), pchi, flat_layer(A, verts, op_axes, cplx), achi))
        end
        if maxiter === nothing
            maxiter = is_tree(psi) ? 1 : error("You need to specify a number of iterations for BP!")
        end
        parallel = edge_sequence == "parallel"
        o = UpdateOpts(parallel ? ITNBP_SCHED_PARALLEL : ITNBP_SCHED_SEQUENTIAL, Cint(maxiter),
            tol === nothing ? -1.0 : Float64(tol), -1.0, PRECISIONS["f64"], 0.0, Cint(1), Cint(0), C_NULL,
            Cint(0))
        r = Ref{UpdateResult}()
        check(ctx, ccall((:itnbp_lnet_update, libitnbp), Cint, (Ptr{Cvoid}, Ref{UpdateOpts}, Ref{UpdateResult}),
            lnet, o, r))
        re, im = Ref(0.0), Ref(0.0)
        check(ctx, ccall((:itnbp_lnet_logscalar, libitnbp), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}),
            lnet, re, im))
        return (cplx || im[] != 0) ? complex(re[], im[]) : re[]
    finally
        ccall((:itnbp_lnet_destroy, libitnbp), Cint, (Ptr{Cvoid},), lnet)
    end
end

# the reference's inner / loginner with alg = "bp" (src/inner.jl:112-160); the
# inner network's link indices are never materialised, so dual_link_index_map has
# no effect, and a caller's own cache (`cache!`) stays on the reference's path
function bp_inner_kwargs(kw)
    haskey(kw, :cache!) && error("loginner_bp: a caller-supplied cache! is not supported on the device path")
    return (; cache_update_kwargs = get(kw, :cache_update_kwargs, (;)))
end
ITensorNetworks.loginner(::Algorithm"bp", phi::ITensorNetwork, psi::ITensorNetwork; kwargs...) =
    loginner_bp(phi, psi; bp_inner_kwargs(kwargs)...)
ITensorNetworks.loginner(::Algorithm"bp", phi::ITensorNetwork, A::ITensorNetwork, psi::ITensorNetwork; kwargs...) =
    loginner_bp(phi, psi; A, bp_inner_kwargs(kwargs)...)
ITensors.inner(::Algorithm"bp", phi::ITensorNetwork, psi::ITensorNetwork; kwargs...) =
    exp(loginner_bp(phi, psi; bp_inner_kwargs(kwargs)...))
ITensors.inner(::Algorithm"bp", phi::ITensorNetwork, A::ITensorNetwork, psi::ITensorNetwork; kwargs...) =
    exp(loginner_bp(phi, psi; A, bp_inner_kwargs(kwargs)...))

end # module
