// Fast row-kernel instantiations: float bulk, ordinary stores (one
// translation unit per kind, precision and store kind, so the library builds
// in parallel).
#define PFC_INST_UNIT
#include "pfc_fast.hpp"

namespace pfc {

template bool row_fast_launch_impl<float, false>(int, int, dim3, dim3, size_t, hipStream_t, const V2<float>*, V2<float>*, int,
                                             const Lay&, const Nonlin&, const V2<double>*, const Slots&, RangePart*,
                                             const RangeCheck&);

}  // namespace pfc
