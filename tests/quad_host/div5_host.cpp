// div5_host.cpp -- TEST INFRASTRUCTURE: the compile-time truncating
// division by 10^N of dec_device.h (dshr10: a static shift by N bits, then
// floor division by 5^N by Barrett reduction or by one-word passes) built
// for the host CPU, so the CPU suite checks every form against exact
// integer division before any GPU run.  Each case (WI, N) is one template
// instantiation; div5_run(case, x, n, out) writes floor(x / 10^N) for n
// values of WI limbs each, and div5_run_passes the same through the pass
// form alone.
#define __HIPCC_RTC__ 1
#define __device__
#define __forceinline__ inline
#define __global__
#include <cstdint>
#include <cstring>
static inline unsigned upd_alignbyte(unsigned hi, unsigned lo, unsigned s)
{
    return (unsigned)((((unsigned long long)hi << 32) | lo) >> (8 * (s & 3)));
}
#define __builtin_amdgcn_alignbyte(a, b, c) upd_alignbyte((a), (b), (c))
#define __builtin_nontemporal_load(p) (*(p))
static inline int __shfl_xor(int v, int, int) { return v; }
#include "dec_device.h"
using namespace updec;

template <int WI, int N> static void run(const uint32_t* X, int n, uint32_t* out, bool passes)
{
    for (int i = 0; i < n; i++) {
        Dec<WI> a;
        for (int k = 0; k < WI; k++) a.m.w[k] = X[i * WI + k];
        a.neg = 0;
        Mag<WI> r;
        if (passes) {
            constexpr int BI = 32 * WI - N;
            constexpr int WM = BI <= 32 ? 1 : (BI + 31) / 32;
            const Mag<WM> t = resize<WM>(shr_static<WI, N>(a.m));
            r = resize<WI>(div_pow5_passes<WM, N>(t));
        } else {
            r = dshr10<WI, WI, N>(a).m;
        }
        for (int k = 0; k < WI; k++) out[i * WI + k] = r.w[k];
    }
}

#define CASES(X)                                                                                                   \
    X(1, 3) X(2, 9) X(8, 13) X(2, 14) X(2, 20) X(2, 27) X(3, 14) X(3, 30) X(4, 27) X(4, 38) X(5, 40) X(8, 50) X(8, 77) X(12, 100)          \
    X(16, 100) X(19, 100) X(21, 100) X(22, 100) X(32, 100) X(32, 200) X(32, 300) X(40, 133) X(48, 150) X(64, 14)   \
    X(64, 400) X(63, 600)

extern "C" int div5_ncases()
{
    int n = 0;
#define COUNT(w, k) n++;
    CASES(COUNT)
#undef COUNT
    return n;
}

extern "C" void div5_case(int c, int* wi, int* n, int* barrett)
{
    int i = 0;
#define INFO(w, k)                                                                                                 \
    if (i++ == c) {                                                                                                \
        *wi = w;                                                                                                   \
        *n = k;                                                                                                    \
        *barrett = barrett_pays(32 * w - k <= 32 ? 1 : (32 * w - k + 31) / 32, 32 * w - k < 1 ? 1 : 32 * w - k, k); \
    }
    CASES(INFO)
#undef INFO
}

extern "C" void div5_run(int c, const uint32_t* X, int n, uint32_t* out, int passes)
{
    int i = 0;
#define RUN(w, k) \
    if (i++ == c) run<w, k>(X, n, out, passes != 0);
    CASES(RUN)
#undef RUN
}

// recip_word against the 64-bit division it replaces, for every normalized
// d = 2^31 + start + k * step below 2^32; returns the number of mismatches
extern "C" long recip_check(uint32_t start, uint32_t step)
{
    long bad = 0;
    for (uint64_t d = (1ull << 31) + start; d < (1ull << 32); d += step)
        bad += recip_word((uint32_t)d) != (uint32_t)((~0ull) / d - (1ull << 32));
    return bad;
}

extern "C" long recip_check_range(uint32_t start, uint32_t count)
{
    long bad = 0;
    for (uint64_t d = (1ull << 31) + start; d < (1ull << 31) + start + count && d < (1ull << 32); d++)
        bad += recip_word((uint32_t)d) != (uint32_t)((~0ull) / d - (1ull << 32));
    return bad;
}
