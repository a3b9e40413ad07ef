// One instantiation of the tile-pass kernel (kernels_tile.hpp): complex128
// when PL_TILE_DOUBLE is 1 (else complex64), tiles of 2^(6 + PL_TILE_KB)
// amplitudes.  The build compiles this file once per pair.
#include "tiled_launch.hpp"

#if !defined(PL_TILE_DOUBLE) || !defined(PL_TILE_KB)
#error "PL_TILE_DOUBLE and PL_TILE_KB must be set"
#endif

namespace plamd {

using TileT = std::conditional_t<PL_TILE_DOUBLE, double, float>;
static_assert(!(PL_TILE_DOUBLE && PL_TILE_KB > 5), "complex128 tiles hold at most 2^11 amplitudes");

template <>
hipError_t launch_tile_pass_kb<TileT, PL_TILE_KB>(v2<TileT> *sv, uint64_t ntiles, const TpGeom &g,
                                                 const TrInst<TileT> *prog, int ninst,
                                                 const TpOp<TileT> *lops, bool uses_lds,
                                                 hipStream_t stream) {
    const size_t lds = uses_lds ? (size_t{1} << (6 + PL_TILE_KB)) * sizeof(v2<TileT>) : 0;
    hipLaunchKernelGGL((k_tile_pass<TileT, PL_TILE_KB>), dim3(static_cast<unsigned>(ntiles)), dim3(kTpLanes),
                       lds, stream, sv, ntiles, g, prog, ninst, lops);
    return hipGetLastError();
}

} // namespace plamd
