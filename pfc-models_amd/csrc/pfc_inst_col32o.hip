// Fast column-kernel instantiations: float bulk, ordinary stores.
#define PFC_INST_UNIT
#include "pfc_fast.hpp"

namespace pfc {

template bool col_fast_launch_impl<float, false>(int, int, int, dim3, dim3, size_t, hipStream_t, V2<float>*, const V2<float>*,
                                             V2<float>*, int, const Lay&, const ColTables&, const V2<double>*,
                                             const Slots&);

}  // namespace pfc
