/* Test stand-in for the caller side of the GPU profile-HMM entry points: the profile object and
 * the four accessors that the reference's hmm.c defines for them (mp_hmm_length,
 * mp_hmm_emissions, mp_hmm_transitions, mp_hmm_gpu_slot; mppriv.h, hmm.c). Built into
 * tests/libhmm_caller.so and loaded with RTLD_GLOBAL by tests/test_boundary_gpu.py, so the
 * library's mp_hmm_dp_score_gpu_batch / mp_hmm_dp_score_gpu / mp_hmm_gpu_release are reached the
 * way the reference's map.c and hmm.c reach them. Emissions M x 20 and transitions M x 7 int16,
 * as mp_hmm_read1 stores them. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

struct mp_hmm_s {
	int32_t len;
	int16_t *emit, *trans;
	void *gpu[2];
	int32_t n_slot_calls;
};

int32_t mp_hmm_length(const struct mp_hmm_s *h) { return h->len; }
const int16_t *mp_hmm_emissions(const struct mp_hmm_s *h) { return h->emit; }
const int16_t *mp_hmm_transitions(const struct mp_hmm_s *h) { return h->trans; }
void **mp_hmm_gpu_slot(const struct mp_hmm_s *h, int which)
{
	struct mp_hmm_s *m = (struct mp_hmm_s*)h;
	++m->n_slot_calls;
	return which >= 0 && which < 2? &m->gpu[which] : 0;
}

struct mp_hmm_s *hc_make(int32_t len, const int16_t *emit, const int16_t *trans)
{
	struct mp_hmm_s *h = (struct mp_hmm_s*)calloc(1, sizeof(*h));
	h->len = len;
	h->emit = (int16_t*)malloc((size_t)len * 20 * 2);
	h->trans = (int16_t*)malloc((size_t)len * 7 * 2);
	memcpy(h->emit, emit, (size_t)len * 20 * 2);
	memcpy(h->trans, trans, (size_t)len * 7 * 2);
	return h;
}
void *hc_slot0(const struct mp_hmm_s *h) { return h->gpu[0]; }
int32_t hc_slot_calls(const struct mp_hmm_s *h) { return h->n_slot_calls; }
void hc_free(struct mp_hmm_s *h) { free(h->emit); free(h->trans); free(h); }
