// Compile-time dispatch over (dimension, block size, symmetric storage).
#pragma once

#include <type_traits>

#include "rg_internal.h"

namespace rg {

template <int V>
using IC = std::integral_constant<int, V>;
template <bool V>
using BC = std::integral_constant<bool, V>;

template <class F>
void dispatch_bs(int bs, F&& f)
{
    switch (bs) {
        case 1: f(IC<1>{}); break;
        case 2: f(IC<2>{}); break;
        case 3: f(IC<3>{}); break;
        case 4: f(IC<4>{}); break;
        default: throw Error(RG_VALUE_ERROR, "block size must be 1..4");
    }
}

// f(IC<DIM>, IC<BS>, BC<SYM>)
template <class F>
void dispatch(const Geom& g, bool sym, F&& f)
{
    dispatch_bs(g.bs, [&](auto B) {
        if (g.dim == 2) {
            if (sym) f(IC<2>{}, B, BC<true>{});
            else f(IC<2>{}, B, BC<false>{});
        } else {
            if (sym) f(IC<3>{}, B, BC<true>{});
            else f(IC<3>{}, B, BC<false>{});
        }
    });
}

// Grid of a chunked streaming kernel over n rows: a multiple of the XCD count,
// at most kMaxGrid blocks; chunk is a multiple of the block size.
struct Launch {
    int grid;
    int64_t chunk;
};
inline Launch chunked_launch(int64_t n, int max_grid = kMaxGrid)
{
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks < 1) blocks = 1;
    int64_t grid = blocks < max_grid ? blocks : max_grid;
    grid = ((grid + kNumXcd - 1) / kNumXcd) * kNumXcd;
    int64_t chunk = (n + grid - 1) / grid;
    chunk = ((chunk + kBlock - 1) / kBlock) * kBlock;
    if (chunk < kBlock) chunk = kBlock;
    return {(int)grid, chunk};
}

}  // namespace rg
