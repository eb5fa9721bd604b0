// Fast row-kernel instantiations, float bulk (one translation unit per kind and
// precision so the library builds in parallel).
#include "pfc_fast.hpp"

namespace pfc {

#define PFC_X(LEN, VAR, PP, ...) static_assert(Sch<LEN, PP, __VA_ARGS__>::valid(), "bad schedule");
PFC_SCHEDULES(PFC_X)
#undef PFC_X

bool row_fast_launch_f32(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<float>* S,
                         V2<float>* T, int Ny, const Lay& lay, const Nonlin& q, const V2<double>* tw,
                         const Slots& sl, RangePart* rpart) {
#define PFC_X(LEN, VAR, PP, ...)                                                                          \
  if (M == LEN && var == VAR) {                                                                           \
    if (rpart)                                                                                            \
      hipLaunchKernelGGL((k_row_fast<float, Sch<LEN, PP, __VA_ARGS__>, true>), grid, block, lds, st, S, T, Ny, lay, \
                         q, tw, sl, rpart);                                                               \
    else                                                                                                  \
      hipLaunchKernelGGL((k_row_fast<float, Sch<LEN, PP, __VA_ARGS__>, false>), grid, block, lds, st, S, T, Ny, lay, \
                         q, tw, sl, rpart);                                                               \
    return true;                                                                                          \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

bool row_range_launch_f32(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<float>* S,
                          int Ny, const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part) {
#define PFC_X(LEN, VAR, PP, ...)                                                                             \
  if (M == LEN && var == VAR) {                                                                              \
    hipLaunchKernelGGL((k_row_range<float, Sch<LEN, PP, __VA_ARGS__>>), grid, block, lds, st, S, Ny, lay, tw, sl, part); \
    return true;                                                                                             \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

}  // namespace pfc
