// Instantiation of the extension DP kernel for C = 9..16 columns per lane per half (one
// translation unit per group so the unrolled variants compile in parallel).
#include "ext_dp_impl.h"

namespace nsh_ext {
void launch_group2(const ExtArgs &a, int grid, hipStream_t s)
{
	// the table path's per-column offsets would spill at C > 8: always the LDS profile here
	hipLaunchKernelGGL((ext_dp_kernel<9, 16, true>), dim3(grid), dim3(64), 0, s, a);
}
} // namespace nsh_ext
