// Instantiation of the work-group-per-problem DP: int16 saturating arithmetic with CIGAR trace
// (ns_global_gs16b with NS_F_CIGAR).
#include <mutex>
#include "global_dp_impl.h"

namespace nsh_glob {
void launch_trace16(const GlobArgs &a, int n, int NW, int C, hipStream_t s)
{
	static std::once_flag once;
	std::call_once(once, [] { (void)set_lds_limit<Arith16, true, false>(); });
	launch_bucket<Arith16, true, false>(a, n, NW, C, s);
}
} // namespace nsh_glob
