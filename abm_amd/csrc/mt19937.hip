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
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/abm_amd.h"

namespace abm {
int set_error(int code, const char* fmt, ...);
}
using abm::set_error;

namespace {

constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;

inline uint32_t twist_word(uint32_t cur, uint32_t nxt, uint32_t far) {
    const uint32_t y = (cur & kUpper) | (nxt & kLower);
    return far ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
}

// The recurrence over i = 0..623 in place, split where (i + 397) and
// (i + 1) wrap so that no index needs a modulo.
void twist(uint32_t* mt) {
    int i = 0;
    for (; i < kN - kM; ++i) mt[i] = twist_word(mt[i], mt[i + 1], mt[i + kM]);
    for (; i < kN - 1; ++i) mt[i] = twist_word(mt[i], mt[i + 1], mt[i + kM - kN]);
    mt[kN - 1] = twist_word(mt[kN - 1], mt[0], mt[kM - 1]);
}

inline uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// The next n outputs of s (into out, or dropped when out is null), advancing
// s: whole runs of the key at a time, a twist whenever the position reaches
// 624 and another output is needed (numpy twists lazily, so a state can end
// at position 624).
void draw_words(abm_mt19937_state* s, int64_t n, uint32_t* out) {
    uint32_t* key = s->key;
    int pos = s->pos;
    while (n > 0) {
        if (pos >= kN) {
            twist(key);
            pos = 0;
        }
        const int take = (int)std::min<int64_t>(n, kN - pos);
        if (out) {
            for (int j = 0; j < take; ++j) out[j] = temper(key[pos + j]);
            out += take;
        }
        pos += take;
        n -= take;
    }
    s->pos = pos;
}

// random_sample from pairs of outputs: (a >> 5) * 2^26 + (b >> 6), scaled by 2^-53.
void draw_doubles(abm_mt19937_state* s, int64_t n, double* out) {
    uint32_t buf[2 * 256];
    while (n > 0) {
        const int take = (int)std::min<int64_t>(n, 256);
        draw_words(s, 2 * take, out ? buf : nullptr);
        if (out) {
            for (int j = 0; j < take; ++j) {
                const uint32_t a = buf[2 * j] >> 5, b = buf[2 * j + 1] >> 6;
                out[j] = (a * 67108864.0 + b) / 9007199254740992.0;
            }
            out += take;
        }
        n -= take;
    }
}

// Run fn(b) for every state b; the generators are independent, so large
// requests are split over a few host threads (at most 8, each taking a
// contiguous share of the states).
template <typename Fn>
void for_each_state(int32_t n, int64_t work, Fn fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1u, 8u);
    nt = (int)std::min<int64_t>(nt, std::max<int64_t>(1, work / 32768));
    nt = std::min<int>(nt, std::max<int32_t>(1, n));
    if (nt <= 1) {
        for (int32_t b = 0; b < n; ++b) fn(b);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    auto share = [&](int t) {
        const int32_t lo = (int32_t)((int64_t)n * t / nt), hi = (int32_t)((int64_t)n * (t + 1) / nt);
        for (int32_t b = lo; b < hi; ++b) fn(b);
    };
    for (int t = 1; t < nt; ++t) pool.emplace_back(share, t);
    share(0);
    for (auto& th : pool) th.join();
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
    std::vector<int64_t> at(n_states + 1, 0);
    for (int32_t b = 0; b < n_states; ++b) at[b + 1] = at[b] + counts[b];
    for_each_state(n_states, total, [&](int32_t b) {
        if (counts[b] == 0) return;
        abm_mt19937_state copy;
        std::memcpy(&copy, states[b], sizeof(copy));
        draw_doubles(&copy, counts[b], out + at[b]);
    });
    return ABM_OK;
}

int abm_mt19937_skip(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_skip")) return rc;
    for (int32_t b = 0; b < n_states; ++b)
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_skip: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
    int64_t total = 0;
    for (int32_t b = 0; b < n_states; ++b) total += counts[b];
    for_each_state(n_states, total / 64, [&](int32_t b) { draw_words(states[b], 2 * counts[b], nullptr); });
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
    std::vector<int64_t> at(n_states + 1, 0);
    for (int32_t b = 0; b < n_states; ++b) at[b + 1] = at[b] + counts[b];
    for_each_state(n_states, total, [&](int32_t b) {
        if (counts[b] == 0) return;
        const abm_mt19937_state* st = states[b];
        if (st->pos + counts[b] <= kN) {   // no twist needed: read the key in place
            for (int64_t j = 0; j < counts[b]; ++j) out[at[b] + j] = temper(st->key[st->pos + j]);
            return;
        }
        abm_mt19937_state copy;
        std::memcpy(&copy, st, sizeof(copy));
        draw_words(&copy, counts[b], out + at[b]);
    });
    return ABM_OK;
}

int abm_mt19937_skip32(abm_mt19937_state* const* states, int32_t n_states, const int64_t* counts) {
    if (int rc = check_args(states, n_states, counts, "abm_mt19937_skip32")) return rc;
    for (int32_t b = 0; b < n_states; ++b)
        if (!states[b] || counts[b] < 0 || states[b]->pos < 0 || states[b]->pos > kN)
            return set_error(ABM_ERR_INVALID, "abm_mt19937_skip32: state %d is null, has a negative "
                             "count or a position outside [0, 624]", b);
    int64_t total = 0;
    for (int32_t b = 0; b < n_states; ++b) total += counts[b];
    for_each_state(n_states, total / 64, [&](int32_t b) { draw_words(states[b], counts[b], nullptr); });
    return ABM_OK;
}

}  // extern "C"
