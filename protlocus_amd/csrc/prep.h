// Device-side input preparation shared by the DP engines: problem descriptor, tables, launcher.
#ifndef NSH_PREP_H
#define NSH_PREP_H
#include <cstdint>
#include <hip/hip_runtime.h>

namespace nsh_prep {
struct PrepProb {
	int64_t ns_off, as_off, rec_off, aa_off;   // ns_off: into the raw bases, or into PrepConst::genome when from_genome
	int32_t nl, al, left, io;
	int32_t has_ss, from_genome, pad_[2];
};
struct PrepConst {
	uint8_t nt4[256], aa20[256], codon[64];
	int32_t sp[6], sp_null;
	const uint8_t *genome;                     // an oriented genome in HBM that problems may read their bases from
};

// one work-group per problem; rec[i] = nas[i] | donor[i+1] << 8 | acceptor[i] << 16 (bytes)
void launch(const PrepProb *pp, int32_t n, const uint8_t *raw_ns, const uint8_t *raw_ss, const uint8_t *raw_as,
		uint32_t *rec, uint8_t *aas, const PrepConst *pc, hipStream_t s);
} // namespace nsh_prep

namespace nsh_feat {
// the 64 prescreen features of n candidates (features.hip); out: n x 64 floats, device
void launch(const nsh_prep::PrepProb *pp, int32_t n, const uint8_t *ns, const uint8_t *as,
		const nsh_prep::PrepConst *pc, const int8_t *mat, float *out, hipStream_t s);
// the same for candidates whose windows are slices of an oriented genome in HBM (d_g + g_off[k],
// nl[k] bases, at most 24,000) and whose proteins are host strings (al[k] at most 4,096);
// out: n x 64 floats on the host. 1 on success
int from_genome(const uint8_t *d_g, int32_t n, const int64_t *g_off, const int32_t *nl, const char *as,
		const int64_t *as_off, const int32_t *al, int64_t l_as, float *out);
constexpr int kMaxWindow = 24000, kMaxProtein = 4096;
} // namespace nsh_feat
#endif
