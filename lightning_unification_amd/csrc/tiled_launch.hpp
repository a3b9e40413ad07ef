// Launcher of the tile-pass kernel (kernels_tile.hpp).  Its register-
// resident variants are by far the largest device code of the engine, so
// each (precision, tile size) instantiation is compiled in a translation
// unit of its own: tiled.hip built once per pair with PL_TILE_DOUBLE and
// PL_TILE_KB set.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels_tile.hpp"

namespace plamd {

// One pass over every tile of 2^(6 + KB) amplitudes (KB = 3..6 for
// complex64, 3..5 for complex128: 64 registers of complex128 per lane do
// not fit a wavefront's register file): one wavefront per tile, LDS only
// when the pass has LDS operations.
template <class T, int KB>
hipError_t launch_tile_pass_kb(v2<T> *sv, uint64_t ntiles, const TpGeom &g, const TrInst<T> *prog, int ninst,
                               const TpOp<T> *lops, bool uses_lds, hipStream_t stream);

// each one is defined by the translation unit that instantiates it
#define PL_TILE_DECLARE(T, KB)                                                                             \
    template <>                                                                                            \
    hipError_t launch_tile_pass_kb<T, KB>(v2<T> *, uint64_t, const TpGeom &, const TrInst<T> *, int,       \
                                          const TpOp<T> *, bool, hipStream_t);
PL_TILE_DECLARE(float, 3)
PL_TILE_DECLARE(float, 4)
PL_TILE_DECLARE(float, 5)
PL_TILE_DECLARE(float, 6)
PL_TILE_DECLARE(double, 3)
PL_TILE_DECLARE(double, 4)
PL_TILE_DECLARE(double, 5)
#undef PL_TILE_DECLARE

template <class T>
inline hipError_t launch_tile_pass(int kb, v2<T> *sv, uint64_t ntiles, const TpGeom &g, const TrInst<T> *prog,
                                   int ninst, const TpOp<T> *lops, bool uses_lds, hipStream_t stream) {
    switch (kb) {
    case 3: return launch_tile_pass_kb<T, 3>(sv, ntiles, g, prog, ninst, lops, uses_lds, stream);
    case 4: return launch_tile_pass_kb<T, 4>(sv, ntiles, g, prog, ninst, lops, uses_lds, stream);
    case 5: return launch_tile_pass_kb<T, 5>(sv, ntiles, g, prog, ninst, lops, uses_lds, stream);
    default:
        if constexpr (std::is_same_v<T, float>) {
            return launch_tile_pass_kb<T, 6>(sv, ntiles, g, prog, ninst, lops, uses_lds, stream);
        } else {
            return hipErrorInvalidValue; // rejected by the caller before any launch
        }
    }
}

} // namespace plamd
