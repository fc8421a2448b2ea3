"""Index helpers and worm algebra (host side).

split_idx/join_idx restate src/util.jl:3-23 (1-based, first index fastest);
worm_action/worm_inverse/worm_count restate src/worms.jl:4-7.
"""


def split_idx(dims, idx):
    """1-based compound index -> tuple of 1-based indices, first fastest (util.jl:3-14)."""
    idx -= 1
    r = []
    for d in dims:
        r.append(idx % d + 1)
        idx //= d
    return tuple(r)


def join_idx(dims, idxs):
    """Inverse of split_idx (util.jl:15-23)."""
    r = 0
    for idx, d in zip(reversed(tuple(idxs)), reversed(tuple(dims))):
        r *= d
        r += idx - 1
    return r + 1


def worm_action(worm, state, basis_size):
    """State after applying worm function `worm` (worms.jl:4-5)."""
    return (state + worm - 1) % basis_size + 1


def worm_inverse(worm, basis_size):
    """Inverse worm function (worms.jl:6)."""
    return basis_size - worm


def worm_count(basis_size):
    """Number of non-trivial worm functions (worms.jl:7)."""
    return basis_size - 1


def site_of_leg(leg, num_sites):
    """1-based leg -> 1-based bond site (vertex_data.jl:190)."""
    return leg - num_sites if leg > num_sites else leg
