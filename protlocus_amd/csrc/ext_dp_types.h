// Problem descriptor and kernel arguments of the batched extension DP (host and device).
#ifndef NSH_EXT_DP_TYPES_H
#define NSH_EXT_DP_TYPES_H
#include <cstdint>
#include <hip/hip_runtime.h>

namespace nsh_ext {
struct ExtProb {
	int64_t rec_off;   // first row record (uint32 units)
	int64_t aa_off;    // first DP-order residue code of the protein
	int64_t pssm_off;  // -1: score with the matrix; else byte offset of al x asize PSSM rows (query order)
	int32_t nl, al, npad, io;
	int32_t out, C, left;
	int32_t mode;      // 0: extension result; 1: plain score-only global DP, score H(last row, al-1)
	int32_t pair = -1; // >= 0: this problem and prob[pair] (from ExtArgs::pbase) share a wave, C per lane each
	int32_t pad_ = 0;
};

struct ExtArgs {
	const ExtProb *prob; int32_t n;
	const ExtProb *pbase;   // the whole descriptor array (pair partners)
	const uint32_t *rec; const uint8_t *aas; const int8_t *pssm; const int8_t *mat; int32_t asize;
	const int32_t *pen; int32_t pen_n;
	int32_t go, ge, fs, xdrop, eb;
	int32_t *counter;
	int32_t *out;      // 4 ints per problem: nt_len, aa_len, score, rows evaluated
};


// column-count groups, each a kernel with its own register and LDS footprint and occupancy; the
// groups run concurrently on separate streams
constexpr int kGroups = 3;
constexpr int kGroupLo[kGroups] = { 1, 5, 9 }, kGroupHi[kGroups] = { 4, 8, 16 };
// persistent waves per CU. With matrix scores the kernels are compiled for 6 / 3 / 1 waves per SIMD
// (the register budget that measured fastest: more waves hide the per-row dependency chain better
// than the few spilled registers cost); with per-problem profiles LDS (22 x 128C bytes per wave)
// also bounds residency.
#ifndef NSH_EXT_WPE0
#define NSH_EXT_WPE0 6
#endif
#ifndef NSH_EXT_WPE1
#define NSH_EXT_WPE1 3
#endif
#ifndef NSH_EXT_WPE01
#define NSH_EXT_WPE01 4
#endif
constexpr int kGroupWavesPerCU[kGroups] = { 4 * NSH_EXT_WPE0, 4 * NSH_EXT_WPE1, 3 };
constexpr int kGroupWavesPerCUPssm[kGroups] = { 12, 7, 3 };
// A long problem: rows x 128-column chunks at which one problem runs for a large part of a plan.
constexpr int64_t kLongWork = (int64_t)1 << 18;
// Merged queue: a matrix-scored plan that holds a long problem runs its C <= 8 problems as one
// queue, longest first, in one kernel (1..8 columns per lane, NSH_EXT_WPE01 waves per SIMD: on the
// map path's plan of 39,510 extensions, 4 ran it in 134-144 ms, 3 in 158 and 5 in 149). The groups'
// kernels run on separate streams, and two streams that share a hardware queue run one after the
// other, so a long problem in the second would start only when the first has drained.
constexpr int kMergedWavesPerCU = 4 * NSH_EXT_WPE01;
void launch_group01(const ExtArgs &a, int grid, hipStream_t s);
constexpr int kCmax = 16;
inline int group_of(int C) { return C <= 4? 0 : C <= 8? 1 : 2; }
void launch_group(int g, const ExtArgs &a, int grid, hipStream_t s);
} // namespace nsh_ext
#endif
