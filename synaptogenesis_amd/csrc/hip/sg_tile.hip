// Tile kernels, full records: the base activator's launcher and the
// Izhikevich passes over records with every per-synapse field; compact
// records go to sg_tile_compact.hip.
#include "sg_tile_kernels.h"

namespace sg {

template <int WAVES, int CHUNK>
static hipError_t launch_tile_base_t(const sg_conn &c, hipStream_t st) {
    dim3 g((c.num_dest + 63) / 64), b(64 * WAVES);
#define SG_TB(T)                                                             \
    switch (c.out_type) {                                                    \
        case SG_OUT_BIT:                                                     \
            hipLaunchKernelGGL((tile_base_kernel<T, SG_OUT_BIT, WAVES,       \
                                CHUNK>), g, b, 0, st, c); break;             \
        case SG_OUT_INT:                                                     \
            hipLaunchKernelGGL((tile_base_kernel<T, SG_OUT_INT, WAVES,       \
                                CHUNK>), g, b, 0, st, c); break;             \
        default:                                                             \
            hipLaunchKernelGGL((tile_base_kernel<T, SG_OUT_FLOAT, WAVES,     \
                                CHUNK>), g, b, 0, st, c);                    \
    }
    switch (c.type) {
        case SG_FULLY_CONNECTED: SG_TB(SG_FULLY_CONNECTED); break;
        case SG_SUBSET: SG_TB(SG_SUBSET); break;
        case SG_CONVERGENT: SG_TB(SG_CONVERGENT); break;
        case SG_DIVERGENT: SG_TB(SG_DIVERGENT); break;
        case SG_SPARSE: SG_TB(SG_SPARSE); break;
        default: return hipErrorInvalidValue;
    }
#undef SG_TB
    return hipGetLastError();
}

// Fewer than two tiles per CU: sixteen wavefronts per tile.
hipError_t launch_tile_base(const sg_conn &c, hipStream_t st) {
    const int tiles = (c.num_dest + 63) / 64;
    if (tiles < 512) return launch_tile_base_t<16, 64>(c, st);
    return launch_tile_base_t<4, 32>(c, st);
}

hipError_t launch_iz_tile_compact(const sg_conn &c, const sg_iz_syn &s,
                                  const sg_iz_dest &d, int kind, int sub,
                                  int upd, int act, const float *st_,
                                  const float *dt_, hipStream_t st);

hipError_t launch_iz_tile(const sg_conn &c, const sg_iz_syn &s,
                          const sg_iz_dest &d, int kind, int sub, int upd,
                          int act, const float *st_, const float *dt_,
                          hipStream_t st) {
    if (s.src_stp)
        return launch_iz_tile_compact(c, s, d, kind, sub, upd, act, st_, dt_,
                                      st);
    if (upd && act)
        return launch_iz_tile_t<1, 1, 0>(c, s, d, kind, sub, st_, dt_, st);
    if (upd)
        return launch_iz_tile_t<1, 0, 0>(c, s, d, sub ? SG_IZ_SUB : SG_IZ_ADD,
                                         sub, st_, dt_, st);
    return launch_iz_tile_t<0, 1, 0>(c, s, d, kind, 0, st_, dt_, st);
}

}  // namespace sg
