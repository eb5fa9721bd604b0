// Fast column-kernel instantiations, float bulk.
#include "pfc_fast.hpp"

namespace pfc {

bool col_fast_launch_f32(int Ny, int var, int mode, dim3 grid, dim3 block, size_t lds, hipStream_t st, V2<float>* S,
                         const V2<float>* T, V2<float>* D, int Nc, const Lay& lay, const ColTables& tb,
                         const V2<double>* tw, const Slots& sl) {
#define PFC_X(LEN, VAR, PP, ...)                                                                               \
  if (Ny == LEN && var == VAR) {                                                                               \
    using Sc = Sch<LEN, PP, __VA_ARGS__>;                                                                      \
    if (mode == COL_ETD1)                                                                                      \
      hipLaunchKernelGGL((k_col_fast<float, Sc, COL_ETD1>), grid, block, lds, st, S, T, D, Nc, lay, tb, tw, sl);        \
    else if (mode == COL_RK2_STAGE1)                                                                           \
      hipLaunchKernelGGL((k_col_fast<float, Sc, COL_RK2_STAGE1>), grid, block, lds, st, S, T, D, Nc, lay, tb, tw, sl);  \
    else                                                                                                       \
      hipLaunchKernelGGL((k_col_fast<float, Sc, COL_RK2_STAGE2>), grid, block, lds, st, S, T, D, Nc, lay, tb, tw, sl);  \
    return true;                                                                                               \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

}  // namespace pfc
