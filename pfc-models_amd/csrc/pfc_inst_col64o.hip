// Fast column-kernel instantiations: double bulk, ordinary stores.
#define PFC_INST_UNIT
#include "pfc_fast.hpp"

namespace pfc {

template bool col_fast_launch_impl<double, false>(int, int, int, dim3, dim3, size_t, hipStream_t, V2<double>*, const V2<double>*,
                                             V2<double>*, int, const Lay&, const ColTables&, const V2<double>*,
                                             const Slots&);

}  // namespace pfc
