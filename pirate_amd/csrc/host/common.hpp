// Small shared helpers for the pirate_amd host library: error type, power-of-two
// arithmetic, bit reversal. Host-only (no HIP includes).

#ifndef PIRATE_AMD_COMMON_HPP
#define PIRATE_AMD_COMMON_HPP

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace pirate_amd {

// All host-side failures are reported as pdd_error; the C-ABI layer converts them
// into status codes plus a thread-local message.
struct pdd_error : public std::runtime_error {
    explicit pdd_error(const std::string &s) : std::runtime_error(s) { }
};

#define PDD_CHECK(cond, msg)                                                     \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::ostringstream _pdd_ss;                                          \
            _pdd_ss << msg << " [" << #cond << " failed at " << __FILE__ << ":"  \
                    << __LINE__ << "]";                                          \
            throw ::pirate_amd::pdd_error(_pdd_ss.str());                        \
        }                                                                        \
    } while (0)

inline bool is_power_of_two(long n) { return (n > 0) && !(n & (n - 1)); }

inline long pow2(long n)
{
    PDD_CHECK((n >= 0) && (n <= 62), "pow2(): bad exponent " << n);
    return 1L << n;
}

// Exact log2 of a power of two (throws otherwise).
inline int integer_log2(long n)
{
    PDD_CHECK(is_power_of_two(n), "integer_log2(): " << n << " is not a power of two");
    int r = 0;
    while ((1L << r) < n)
        r++;
    return r;
}

// Reverses the low 'nbits' bits of i (reference: src_lib/utils.cpp:189-203).
inline long bit_reverse(long i, int nbits)
{
    long j = 0;
    for (int b = 0; b < nbits; b++) {
        j = (j << 1) | (i & 1);
        i >>= 1;
    }
    return j;
}

inline long xdiv(long m, long n)
{
    PDD_CHECK((n > 0) && (m >= 0) && (m % n == 0), "xdiv(): " << m << " is not divisible by " << n);
    return m / n;
}

inline long align_up(long n, long a) { return ((n + a - 1) / a) * a; }

template<typename T>
inline std::string vec_str(const std::vector<T> &v)
{
    std::ostringstream ss;
    ss << "[";
    for (size_t i = 0; i < v.size(); i++)
        ss << (i ? ", " : "") << v[i];
    ss << "]";
    return ss.str();
}

// Stage-1 -> stage-2 lag of ring row (freq_coarse, dm_brev), in tree time samples.
// lag = (2^stage2_rank - 1 - freq_coarse) * (bitrev(dm_brev) + [uflag] * 2^stage1_rank).
// Reference: src_lib/utils.cpp:218-239 (rb_lag).
inline long rb_lag(long freq_coarse, long dm_brev, int stage1_rank, int stage2_rank, bool uflag)
{
    long ndm = 1L << stage1_rank;
    long nfreq = 1L << stage2_rank;
    PDD_CHECK((freq_coarse >= 0) && (freq_coarse < nfreq), "rb_lag(): freq_coarse out of range");
    PDD_CHECK((dm_brev >= 0) && (dm_brev < ndm), "rb_lag(): dm_brev out of range");
    long dm = bit_reverse(dm_brev, stage1_rank);
    if (uflag)
        dm += ndm;
    return (nfreq - 1 - freq_coarse) * dm;
}

}  // namespace pirate_amd

#endif
