"""itnbp: MI355X-native belief propagation on the norm network of a tensor
network state, with the BeliefPropagationCache / update / environment /
expect / apply API of ITensorNetworks.jl.

Every compute call goes through libitnbp.so (hand-written HIP kernels for
gfx950); there is no CPU fallback.
"""
from ._lib import ITNBPError, ITNBPLibraryMissing, Context, default_context, load  # noqa: F401
from .network import (ITensorNetwork, NamedGraph, comb_tree, heavy_hex, merge_link_indices,  # noqa: F401
                      named_grid, path_graph, productstate, random_instance_set, random_tensornetwork,
                      split_message)
from .cache import (BeliefPropagationCache, QuadraticFormNetwork, OPS, default_edge_sequence,  # noqa: F401
                    edge_scalars, environment, expect, expect2_on_cache, expect_on_cache,
                    identity_messages, initialize_cache, logscalar, message_diff, norm_sqr_network,
                    normalize, rdm, region_scalars, rescale, rescale_, scalar, update, update_,
                    updated_message, vertex_scalars, energy_and_gradient, sweeps_vjp, update_batched_)
from .apply import (Gate, apply, apply_gates_, apply_gates_vjp, edge_coloring, simple_update_round_,  # noqa: F401,E402
                    tebd_bp_)
from .io import load as load_network, save as save_network  # noqa: F401,E402
from .layer import LayeredBPCache, inner, inner_network_cache, loginner, norm_cache  # noqa: F401,E402
