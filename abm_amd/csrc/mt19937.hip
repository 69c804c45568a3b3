// Host-side MT19937 for the foraging random walk (no device code).
//
// The reference's explore branch draws np.random.uniform(lo, hi) from numpy's
// legacy global RandomState (abm/agent/supcalc.py:38-47), i.e. the published
// MT19937 generator of Matsumoto and Nishimura (1998) with 53-bit doubles made
// from two 32-bit outputs, (a >> 5) * 2^26 + (b >> 6), scaled by 2^-53.
// numpy keeps that generator in a plain struct of 624 key words and a
// position (numpy/random/src/mt19937/mt19937.h, mt19937_state); abm_mt19937_state
// has the same layout, so the host hands the address of a RandomState's state
// straight to these functions.  Peeking N doubles ahead and then skipping the
// number the kernel consumed leaves the generator exactly where the reference
// would, without copying RandomState objects.
#include <cstring>

#include "../../include/abm_amd.h"

namespace abm {
int set_error(int code, const char* fmt, ...);
}
using abm::set_error;

namespace {

constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;

void twist(uint32_t* mt) {
    for (int i = 0; i < kN; ++i) {
        const uint32_t y = (mt[i] & kUpper) | (mt[(i + 1) % kN] & kLower);
        mt[i] = mt[(i + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
    }
}

inline uint32_t next32(abm_mt19937_state* s) {
    if (s->pos >= kN) {
        twist(s->key);
        s->pos = 0;
    }
    uint32_t y = s->key[s->pos++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

inline double next_double(abm_mt19937_state* s) {
    const uint32_t a = next32(s) >> 5, b = next32(s) >> 6;
    return (a * 67108864.0 + b) / 9007199254740992.0;
}

int check_args(const void* states, int32_t n_states, const int64_t* counts, const char* who) {
    if (n_states < 0 || (n_states > 0 && (!states || !counts)))
        return set_error(ABM_ERR_INVALID, "%s: null states or counts", who);
    return ABM_OK;
}

}  // namespace

extern "C" {

int abm_mt19937_peek(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts,
                     double* out) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_peek")) return rc;
    int64_t total = 0;
    for (int32_t b = 0; b < n_states; ++b) {
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_peek: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
        total += counts[b];
    }
    if (total > 0 && !out) return set_error(ABM_ERR_INVALID, "abm_mt19937_peek: null output");
    abm_mt19937_state copy;
    for (int32_t b = 0; b < n_states; ++b) {
        std::memcpy(&copy, states[b], sizeof(copy));
        for (int64_t i = 0; i < counts[b]; ++i) *out++ = next_double(&copy);
    }
    return ABM_OK;
}

int abm_mt19937_skip(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_skip")) return rc;
    for (int32_t b = 0; b < n_states; ++b)
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_skip: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
    for (int32_t b = 0; b < n_states; ++b)
        for (int64_t i = 0; i < counts[b]; ++i) (void)next_double(states[b]);
    return ABM_OK;
}

int abm_mt19937_peek32(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts,
                       uint32_t* out) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_peek32")) return rc;
    int64_t total = 0;
    for (int32_t b = 0; b < n_states; ++b) {
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_peek32: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
        total += counts[b];
    }
    if (total > 0 && !out) return set_error(ABM_ERR_INVALID, "abm_mt19937_peek32: null output");
    abm_mt19937_state copy;
    for (int32_t b = 0; b < n_states; ++b) {
        if (counts[b] == 0) continue;
        std::memcpy(&copy, states[b], sizeof(copy));
        for (int64_t i = 0; i < counts[b]; ++i) *out++ = next32(&copy);
    }
    return ABM_OK;
}

int abm_mt19937_skip32(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_skip32")) return rc;
    for (int32_t b = 0; b < n_states; ++b)
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_skip32: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
    for (int32_t b = 0; b < n_states; ++b)
        for (int64_t i = 0; i < counts[b]; ++i) (void)next32(states[b]);
    return ABM_OK;
}

}  // extern "C"
