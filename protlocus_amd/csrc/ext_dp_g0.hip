// Instantiation of the extension DP kernel for C = 1..4 columns per lane per half (one
// translation unit per group so the unrolled variants compile in parallel).
#include "ext_dp_impl.h"

namespace nsh_ext {
void launch_group0(const ExtArgs &a, int grid, hipStream_t s)
{
	if (a.pssm) hipLaunchKernelGGL((ext_dp_kernel<1, 4, true>), dim3(grid), dim3(64), 0, s, a);
	else hipLaunchKernelGGL((ext_dp_kernel<1, 4, false>), dim3(grid), dim3(64), 0, s, a);
}
} // namespace nsh_ext
