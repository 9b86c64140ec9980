// Global spliced DP with CIGAR on gfx950 (ns_global_gs16b/gs32b, nasw-sse.c:351-567, 779-973).
#include <hip/hip_runtime.h>
#include "nsh_internal.h"

int nsh_global_run(void *km, const char *ns, int32_t nl, const char *as, int32_t al,
		const ns_opt_t *opt, const uint8_t *ss, int32_t is32, ns_rst_t *r)
{
	(void)km; (void)ns; (void)nl; (void)as; (void)al; (void)opt; (void)ss; (void)is32; (void)r;
	nsh_set_error("global DP kernel not available in this build");
	return 0;
}
