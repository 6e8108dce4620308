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

// Grid of a chunked streaming kernel over n rows: each block owns one
// contiguous chunk, a multiple of kBlock*rows_per_thread rows; the grid is a
// multiple of the XCD count and at most max_grid blocks.
struct Launch {
    int grid;
    int64_t chunk;
};
inline Launch chunked_launch(int64_t n, int rows_per_thread = 1, int max_grid = kMaxGrid)
{
    const int64_t unit = (int64_t)kBlock * rows_per_thread;
    const int64_t cap = (int64_t)max_grid - kNumXcd;
    int64_t k = (n + unit * cap - 1) / (unit * cap);
    if (k < 1) k = 1;
    const int64_t chunk = unit * k;
    int64_t grid = (n + chunk - 1) / chunk;
    if (grid < 1) grid = 1;
    grid = ((grid + kNumXcd - 1) / kNumXcd) * kNumXcd;
    return {(int)grid, chunk};
}

}  // namespace rg
