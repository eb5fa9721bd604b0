// Multi-shape step kernels (k_row_group, k_col_group), double bulk.
#include "pfc_fast.hpp"

namespace pfc {

#define PFC_X(CASE, CLS, LEN, PP, ...) static_assert(Sch<LEN, PP, __VA_ARGS__>::valid(), "bad group schedule");
PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X

template <int CLS>
static void launch_class(int kind, dim3 grid, dim3 block, size_t lds, hipStream_t st, const GroupMember<double>* mem,
                         int nmem) {
  if (kind == 0)
    hipLaunchKernelGGL((k_row_group<double, CLS>), grid, block, lds, st, mem, nmem);
  else if (kind == 1 + COL_ETD1)
    hipLaunchKernelGGL((k_col_group<double, COL_ETD1, CLS>), grid, block, lds, st, mem, nmem);
  else if (kind == 1 + COL_RK2_STAGE1)
    hipLaunchKernelGGL((k_col_group<double, COL_RK2_STAGE1, CLS>), grid, block, lds, st, mem, nmem);
  else
    hipLaunchKernelGGL((k_col_group<double, COL_RK2_STAGE2, CLS>), grid, block, lds, st, mem, nmem);
}

void group_launch_f64(int kind, int cls, int blocks, size_t lds, hipStream_t st, const GroupMember<double>* mem,
                      int nmem) {
  const dim3 grid(blocks), block(kGroupThreads);
  if (cls == 0)
    launch_class<0>(kind, grid, block, lds, st, mem, nmem);
  else
    launch_class<1>(kind, grid, block, lds, st, mem, nmem);
}

}  // namespace pfc
