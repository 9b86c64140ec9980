// Instantiation of the work-group-per-problem DP: int32 wrapping arithmetic, with CIGAR trace
// (ns_global_gs32b with NS_F_CIGAR) and score-only (ns_global_gs32b without it).
#include <mutex>
#include "global_dp_impl.h"

namespace nsh_glob {
void launch_trace32(const GlobArgs &a, int n, int NW, int C, hipStream_t s)
{
	static std::once_flag once;
	std::call_once(once, [] { (void)set_lds_limit<Arith32, true, false>(); });
	launch_bucket<Arith32, true, false>(a, n, NW, C, s);
}
void launch_score32(const GlobArgs &a, int n, int NW, int C, hipStream_t s)
{
	static std::once_flag once;
	std::call_once(once, [] { (void)set_lds_limit<Arith32, false, false>(); });
	launch_bucket<Arith32, false, false>(a, n, NW, C, s);
}
} // namespace nsh_glob
