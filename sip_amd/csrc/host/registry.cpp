// Lookup of the compiled Stockham leaf kernels (generated instance tables).
#include <cstdlib>
#include <map>

#include "plan.h"

namespace sip {

#define SIP_STOCKHAM_PART(k)                                \
    extern const StockhamInstance g_stockham_part_##k[];    \
    extern const int g_stockham_part_##k##_count;
#include "stockham_parts.inc"
#undef SIP_STOCKHAM_PART

static const std::map<int64_t, const StockhamInstance *> &instances()
{
    static const std::map<int64_t, const StockhamInstance *> m = [] {
        std::map<int64_t, const StockhamInstance *> r;
#define SIP_STOCKHAM_PART(k)                                        \
    for (int i = 0; i < g_stockham_part_##k##_count; i++)           \
        r[g_stockham_part_##k[i].L] = &g_stockham_part_##k[i];
#include "stockham_parts.inc"
#undef SIP_STOCKHAM_PART
        return r;
    }();
    return m;
}

const StockhamInstance *find_instance(int64_t L)
{
    const auto &m = instances();
    auto it = m.find(L);
    return it == m.end() ? nullptr : it->second;
}

int64_t instance_count() { return (int64_t)instances().size(); }

extern const Stockham2dInstance g_stockham2d[9];
extern const Stockham2dInstance g_stockham2d_split[1];

const Stockham2dInstance *find_2d_instance(int64_t X, int64_t Y)
{
    // 128 x 128 slices take the single-plane tile by default (two workgroups per CU): 3D 128^3 x 64
    // 1035 -> 983 us, 2D 128^2 x 4096 338 -> 319 us on MI355X; SIP_FFT_TILE2D_SPLIT=0 selects
    // the two-plane tile
    if (X == 128 && Y == 128) {
        const char *e = std::getenv("SIP_FFT_TILE2D_SPLIT");
        if (!e || std::atoi(e) != 0) return &g_stockham2d_split[0];
    }
    for (const Stockham2dInstance &t : g_stockham2d)
        if (t.X == X && t.Y == Y) return &t;
    return nullptr;
}

}  // namespace sip
