// Fast row-kernel instantiations: double bulk, non-temporal stores (one
// translation unit per kind, precision and store kind, so the library builds
// in parallel).
#define PFC_INST_UNIT
#include "pfc_fast.hpp"

namespace pfc {

#define PFC_X(LEN, VAR, PP, ...) static_assert(Sch<LEN, PP, __VA_ARGS__>::valid(), "bad schedule");
PFC_SCHEDULES(PFC_X)
#undef PFC_X

bool row_range_launch_f64(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<double>* S,
                          int Ny, const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part) {
#define PFC_X(LEN, VAR, PP, ...)                                                                             \
  if (M == LEN && var == VAR) {                                                                              \
    hipLaunchKernelGGL((k_row_range<double, Sch<LEN, PP, __VA_ARGS__>>), grid, block, lds, st, S, Ny, lay, tw, sl, part); \
    return true;                                                                                             \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

template bool row_fast_launch_impl<double, true>(int, int, dim3, dim3, size_t, hipStream_t, const V2<double>*, V2<double>*, int,
                                             const Lay&, const Nonlin&, const V2<double>*, const Slots&, RangePart*,
                                             const RangeCheck&);

}  // namespace pfc
