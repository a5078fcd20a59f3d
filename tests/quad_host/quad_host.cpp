// quad_host.cpp -- TEST INFRASTRUCTURE: the row-quad helpers of dec_device.h
// (load_quad, quad_row, quad_byte) built for the host CPU and checked
// against the per-row decoder load_compact, for every width 1..32 and every
// row count 1..13 (full and partial last quads).  Columns are exactly sized
// heap copies, so a build with AddressSanitizer also checks that no helper
// reads past a column.  Exit status 0 = every check passed.
#define __HIPCC_RTC__ 1
#define __device__
#define __forceinline__ inline
#define __global__
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>
static inline unsigned upd_alignbyte(unsigned hi, unsigned lo, unsigned s)
{
    return (unsigned)((((unsigned long long)hi << 32) | lo) >> (8 * (s & 3)));
}
#define __builtin_amdgcn_alignbyte(a, b, c) upd_alignbyte((a), (b), (c))
#define __builtin_nontemporal_load(p) (*(p))
static inline int __shfl_xor(int v, int, int) { return v; }
#include "dec_device.h"
using namespace updec;

static uint64_t rng = 88172645463325252ull;
static uint32_t next32()
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)rng;
}

static int failures = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            if (failures++ < 20) {             \
                fprintf(stderr, __VA_ARGS__);  \
                fputc('\n', stderr);           \
            }                                  \
        }                                      \
    } while (0)

template <int B> static void check_width()
{
    constexpr int NW = (B + 3) / 4;
    for (int rows = 1; rows <= 13; rows++) {
        const size_t nbytes = (size_t)rows * B;
        // column: random bytes (any bit pattern decodes: sign bit + magnitude)
        std::vector<uint8_t> img(nbytes);
        for (auto& b : img) b = (uint8_t)next32();
        // dword-aligned exact-size heap copy
        uint32_t* store = (uint32_t*)malloc((nbytes + 3) & ~(size_t)3);
        uint8_t* col = (uint8_t*)store;
        memcpy(col, img.data(), nbytes);
        const size_t nq = ((size_t)rows + 3) / 4;
        for (size_t q = 0; q < nq; q++) {
            const Quad<B> qd = load_quad<B>(col, q, nbytes);
            for (uint32_t k = 0; k < 4; k++) {
                const size_t row = q * 4 + k;
                if (row >= (size_t)rows) continue;
                const Dec<NW> a = quad_row<B>(qd, k);
                const Dec<NW> b = load_compact<B>(col, row, nbytes);
                bool same = a.neg == b.neg;
                for (int i = 0; i < NW; i++) same = same && a.m.w[i] == b.m.w[i];
                CHECK(same, "B=%d rows=%d row=%zu: quad_row differs from load_compact", B, rows, row);
            }
        }
        free(store);
    }
}

static void check_bytes()
{
    for (int rows = 1; rows <= 13; rows++) {
        std::vector<uint8_t> v(rows);
        for (auto& b : v) b = (uint8_t)next32();
        uint32_t* store = (uint32_t*)malloc(((size_t)rows + 3) & ~(size_t)3);
        memcpy(store, v.data(), rows);
        for (size_t q = 0; q * 4 < (size_t)rows; q++) {
            const Quad<1> kq = load_quad<1>((const uint8_t*)store, q, rows);
            for (uint32_t k = 0; k < 4 && q * 4 + k < (size_t)rows; k++)
                CHECK(quad_byte(kq, k) == v[q * 4 + k], "byte column rows=%d row=%zu", rows, q * 4 + k);
        }
        free(store);
        // int32 column through Quad<4>
        std::vector<int32_t> f(rows);
        for (auto& x : f) x = (int32_t)next32();
        int32_t* fs = (int32_t*)malloc((size_t)rows * 4);
        memcpy(fs, f.data(), (size_t)rows * 4);
        for (size_t q = 0; q * 4 < (size_t)rows; q++) {
            const Quad<4> fq = load_quad<4>((const uint8_t*)fs, q, (size_t)rows * 4);
            for (uint32_t k = 0; k < 4 && q * 4 + k < (size_t)rows; k++)
                CHECK((int32_t)fq.w[k] == f[q * 4 + k], "int32 column rows=%d row=%zu", rows, q * 4 + k);
        }
        free(fs);
    }
}

template <int... Bs> static void check_all(std::integer_sequence<int, Bs...>) { (check_width<Bs + 1>(), ...); }

int main()
{
    for (int rep = 0; rep < 20; rep++) check_all(std::make_integer_sequence<int, 32>{});
    check_bytes();
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("quad helpers: all checks passed\n");
    return 0;
}
