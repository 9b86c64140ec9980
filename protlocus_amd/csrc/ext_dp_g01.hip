// Instantiation of the extension DP kernel for the merged queue: C = 1..8 columns per lane per
// half in one kernel, for plans whose long problems must all start at once.
#include "ext_dp_impl.h"

namespace nsh_ext {
void launch_group01(const ExtArgs &a, int grid, hipStream_t s)
{
	hipLaunchKernelGGL((ext_dp_kernel<1, 8, false>), dim3(grid), dim3(64), 0, s, a);
}
} // namespace nsh_ext
