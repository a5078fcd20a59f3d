// dmod_host.cpp -- TEST INFRASTRUCTURE: dmod (the device % of two decimals,
// dec_device.h) built for the host CPU so the CPU suite checks it against
// the closed form of the reference's ModCalc before any GPU run.  Operands
// are 32-limb magnitudes with a sign word; widths (WA, WB) = (32, 32) or
// (4, 8) (a narrower than b).
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

template <int WA, int WB, int FAITH>
static void run(const uint32_t* A, const uint32_t* B, const uint32_t* signs, int n, uint32_t* out)
{
    constexpr int WO = WA > WB ? WA : WB;
    for (int i = 0; i < n; i++) {
        Dec<WA> a;
        Dec<WB> b;
        for (int k = 0; k < WA; k++) a.m.w[k] = A[i * 32 + k];
        for (int k = 0; k < WB; k++) b.m.w[k] = B[i * 32 + k];
        a.neg = signs[2 * i];
        b.neg = signs[2 * i + 1];
        const Dec<WO> r = dmod<WO, WA, WB, FAITH>(a, b);
        for (int k = 0; k < 32; k++) out[i * 33 + k] = k < WO ? r.m.w[k] : 0u;
        out[i * 33 + 32] = r.neg;
    }
}

extern "C" void dmod_run(int narrow, int faith, const uint32_t* A, const uint32_t* B, const uint32_t* signs, int n,
                         uint32_t* out)
{
    if (narrow) {
        if (faith) run<4, 8, 1>(A, B, signs, n, out);
        else run<4, 8, 0>(A, B, signs, n, out);
    } else {
        if (faith) run<32, 32, 1>(A, B, signs, n, out);
        else run<32, 32, 0>(A, B, signs, n, out);
    }
}
