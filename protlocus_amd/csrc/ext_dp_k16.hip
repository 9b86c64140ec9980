// Instantiation of the extension DP kernel for one LDS variant (kept in its own translation
// unit so the two large unrolled variants compile in parallel).
#include "ext_dp_impl.h"

namespace nsh_ext {
void launch_large(const ExtArgs &a, int grid, hipStream_t s)
{
	hipLaunchKernelGGL(ext_dp_kernel<CMAX>, dim3(grid), dim3(64), 0, s, a);
}
} // namespace nsh_ext
