// Instantiation of the work-group-per-problem DP in extension mode: the int16 score-only
// extension DP (ns_global_gs16b with NS_F_EXT_LEFT/RIGHT, or plain score-only) for proteins whose
// padded width exceeds the one-wave extension kernel's 2048 columns.
#include <mutex>
#include "global_dp_impl.h"

namespace nsh_glob {
void launch_ext16(const GlobArgs &a, int n, int NW, int C, hipStream_t s)
{
	static std::once_flag once;
	std::call_once(once, [] { (void)set_lds_limit<Arith16, false, true>(); });
	launch_bucket<Arith16, false, true>(a, n, NW, C, s);
}
} // namespace nsh_glob
