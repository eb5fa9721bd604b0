// Fast row-kernel instantiations: double bulk, ordinary stores (one
// translation unit per kind, precision and store kind, so the library builds
// in parallel).
#define PFC_INST_UNIT
#include "pfc_fast.hpp"

namespace pfc {

template bool row_fast_launch_impl<double, false>(int, int, dim3, dim3, size_t, hipStream_t, const V2<double>*, V2<double>*, int,
                                             const Lay&, const Nonlin&, const V2<double>*, const Slots&, RangePart*,
                                             const RangeCheck&);

}  // namespace pfc
