// Tile kernels, compact records (sg_hip.h: SG_IZC_*): the Izhikevich
// passes whose presynaptic trace and short-term plasticity are kept per
// source.
#include "sg_tile_kernels.h"

namespace sg {

hipError_t launch_iz_tile_compact(const sg_conn &c, const sg_iz_syn &s,
                                  const sg_iz_dest &d, int kind, int sub,
                                  int upd, int act, const float *st_,
                                  const float *dt_, hipStream_t st) {
    if (upd && act)
        return launch_iz_tile_t<1, 1, 1>(c, s, d, kind, sub, st_, dt_, st);
    if (upd)
        return launch_iz_tile_t<1, 0, 1>(c, s, d, sub ? SG_IZ_SUB : SG_IZ_ADD,
                                         sub, st_, dt_, st);
    return launch_iz_tile_t<0, 1, 1>(c, s, d, kind, 0, st_, dt_, st);
}

}  // namespace sg
