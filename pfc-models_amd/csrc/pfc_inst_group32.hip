// Multi-shape step kernels (k_row_group, k_col_group, the persistent k_flow), float bulk.
#include "pfc_fast.hpp"

namespace pfc {

#define PFC_X(CASE, CLS, LEN, PP, ...) static_assert(Sch<LEN, PP, __VA_ARGS__>::valid(), "bad group schedule");
PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X

template <int CLS, bool NT>
static void launch_class(int kind, dim3 grid, dim3 block, size_t lds, hipStream_t st, const GroupMember<float>* mem,
                         int nmem) {
  if (kind == 0)
    hipLaunchKernelGGL((k_row_group<float, CLS, NT>), grid, block, lds, st, mem, nmem);
  else if (kind == 1 + COL_ETD1)
    hipLaunchKernelGGL((k_col_group<float, COL_ETD1, CLS, NT>), grid, block, lds, st, mem, nmem);
  else if (kind == 1 + COL_RK2_STAGE1)
    hipLaunchKernelGGL((k_col_group<float, COL_RK2_STAGE1, CLS, NT>), grid, block, lds, st, mem, nmem);
  else
    hipLaunchKernelGGL((k_col_group<float, COL_RK2_STAGE2, CLS, NT>), grid, block, lds, st, mem, nmem);
}

void group_launch_f32(int kind, int cls, bool nt, int blocks, size_t lds, hipStream_t st,
                      const GroupMember<float>* mem, int nmem) {
  const dim3 grid(blocks), block(kGroupThreads);
  if (cls == 0)
    nt ? launch_class<0, true>(kind, grid, block, lds, st, mem, nmem)
       : launch_class<0, false>(kind, grid, block, lds, st, mem, nmem);
  else
    nt ? launch_class<1, true>(kind, grid, block, lds, st, mem, nmem)
       : launch_class<1, false>(kind, grid, block, lds, st, mem, nmem);
}

template <int CLS, bool RK2>
static void flow_one(int blocks, size_t lds, hipStream_t st, const FlowPlan<float>& f) {
  hipLaunchKernelGGL(k_flow_reset<unsigned>, dim3((f.nwords + 255) / 256), dim3(256), 0, st, f.cnt, f.nwords);
  hipLaunchKernelGGL((k_flow<float, CLS, RK2>), dim3(blocks), dim3(kGroupThreads), lds, st, f);
  hipLaunchKernelGGL((k_flow_mark<float>), dim3(1), dim3(256), 0, st, f.rows, f.cnt0, f.nmem, f.abort);
}

void flow_launch_f32(int cls, bool rk2, int blocks, size_t lds, hipStream_t st, const FlowPlan<float>& f) {
  if (cls == 0)
    rk2 ? flow_one<0, true>(blocks, lds, st, f) : flow_one<0, false>(blocks, lds, st, f);
  else
    rk2 ? flow_one<1, true>(blocks, lds, st, f) : flow_one<1, false>(blocks, lds, st, f);
}

int flow_occupancy_f32(int cls, bool rk2, size_t lds) {
  int n = 0;
  hipError_t e;
  if (cls == 0)
    e = rk2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_flow<float, 0, true>, kGroupThreads, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_flow<float, 0, false>, kGroupThreads, lds);
  else
    e = rk2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_flow<float, 1, true>, kGroupThreads, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_flow<float, 1, false>, kGroupThreads, lds);
  return e == hipSuccess ? n : 0;
}

}  // namespace pfc
