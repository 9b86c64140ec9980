// Problem descriptor and kernel arguments of the work-group-per-problem spliced DP (host and
// device): the global DP with CIGAR trace (ns_global_gs16b / gs32b with NS_F_CIGAR), the int32
// score-only DP, and the extension mode for proteins longer than the one-wave extension kernel.
#ifndef NSH_GLOBAL_DP_TYPES_H
#define NSH_GLOBAL_DP_TYPES_H
#include <cstdint>
#include <hip/hip_runtime.h>

namespace nsh_glob {
constexpr int kCMaxG = 8, kNWMax = 16;   // up to 8 columns per lane, 16 waves: 8192 columns
constexpr int kLongRowsG = 2048;         // rows from which a problem starts first in its bucket, at issue priority

struct GlobProb {
	int64_t rec_off, aa_off, tr_off, pssm_off;   // pssm_off < 0: matrix scores
	int32_t nl, al, io, out;
	int32_t C, NW, slen, left;
	int32_t npad, mode;   // extension mode: the reference's padded width; mode 1 = plain score-only
};
struct GlobArgs {
	const GlobProb *prob;
	const uint32_t *rec; const uint8_t *aas; const int8_t *pssm; const int8_t *mat; int32_t asize;
	int32_t go, ge, fs;
	uint16_t *trace;        // per problem: nl rows x al columns (rows 0, 1 unused)
	int32_t *score;         // per problem: H(nl-1, al-1) (nasw-sse.c:559/962)
	// extension mode
	const int32_t *pen; int32_t pen_n, xdrop, eb;
	int32_t *ext_out;       // 4 ints per problem: nt_len, aa_len, score, rows evaluated
};

// columns per lane and waves of a problem of `cols` columns (al, or npad in extension mode):
// one wave per 256 columns up to 4096, then 16 waves with up to 8 columns per lane
inline void shape(int32_t cols, int32_t &C, int32_t &NW)
{
	if (cols <= 64 * 4 * kNWMax) {
		NW = (cols + 255) / 256;
		if (NW < 1) NW = 1;
		C = (cols + 64 * NW - 1) / (64 * NW);
	} else {
		NW = kNWMax;
		C = (cols + 64 * kNWMax - 1) / (64 * kNWMax);
	}
	if (C < 1) C = 1;
}

// one bucket of problems sharing NW and C (grid = n, 64 * NW threads), per kernel variant
void launch_trace16(const GlobArgs &a, int n, int NW, int C, hipStream_t s);
void launch_trace32(const GlobArgs &a, int n, int NW, int C, hipStream_t s);
void launch_score32(const GlobArgs &a, int n, int NW, int C, hipStream_t s);
void launch_ext16(const GlobArgs &a, int n, int NW, int C, hipStream_t s);
} // namespace nsh_glob
#endif
