// segsum.hip -- segmented SUM of a compact decimal column
// (updec_segsum_compact, updec_segsum_compact_noalloc in include/updec.h).
//
// The reference sums a compact column per group with aggregate_column
// (src/CudaAcc/AriesEngineAlgorithm.h:832-857): rows are visited in
// `associated` order and segment s covers positions [groups[s], groups[s+1])
// (the last one up to `rows`); segreduce (kernel_segreduce.hxx:535-612)
// reduces each segment into a 133-byte nullable_type<Decimal>.
//
// Here the whole operation is three stream-ordered launches of kernels
// compiled with the library, with nothing read back to the host:
//   k_seg_zero    clears the per-segment accumulators;
//   k_seg_main<W> one block per tile of 4096 positions: a tile inside one
//                 segment (the common case for large groups) is summed in
//                 registers, reduced over the wave and the block, and added
//                 to the segment's accumulator once; in a tile that
//                 crosses segment starts each lane walks a run of 16
//                 consecutive positions, adding its running sum to a
//                 segment's accumulator when its segment changes;
//   k_seg_final<W> one lane per segment: carry-normalizes the accumulator
//                 and writes the image.
// An accumulator is AW = W + 2 words of 64 bits, one per 32-bit limb of a
// two's complement sum: every addition adds a partial sum's limbs
// (sign-extended to AW limbs) into the 64-bit words, so no carry crosses
// words until the final kernel, and a word cannot overflow before 2^32
// additions.  Integer addition is associative, so the result is the exact
// sum whatever order the atomics land in; the image holds its low 32 limbs,
// as the reference's 32-limb Decimal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/updec.h"
#include "api_internal.h"
#include "dec_device.h"
#include "planner.h"

using namespace updec;

namespace {

#define SEGCHK(expr)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return api_fail(UPDEC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kSegThreads = 256;
constexpr int kSegRowsPerThread = 16;
constexpr i64 kSegTile = (i64)kSegThreads * kSegRowsPerThread;

struct SegArgs {
    const u8* col;
    usize colbytes;
    u32 B;
    u32 pad;
    const i32* assoc;   // may be null: position i is row i
    const i32* groups;  // segment starts
    i32 ngroups;
    i32 pad2;
    i64 rows;
    u64* acc;  // ngroups * (W + 2) words
};

// largest s in [s0, s1] with groups[s] <= r; s0 - 1 when there is none
__device__ __forceinline__ i32 seg_search(const i32* g, i32 s0, i32 s1, i64 r)
{
    i32 lo = s0, hi = s1 + 1;
    while (lo < hi) {
        const i32 mid = (lo + hi) >> 1;
        if ((i64)g[mid] <= r) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

template <int AW> __device__ __forceinline__ void seg_flush(u64* acc, i32 s, const Mag<AW>& m)
{
    u64* p = acc + (usize)s * AW;
#pragma unroll
    for (int w = 0; w < AW; w++) atomicAdd(reinterpret_cast<unsigned long long*>(p + w), (unsigned long long)m.w[w]);
}

__global__ void __launch_bounds__(256) k_seg_zero(u64* p, usize n)
{
    for (usize i = (usize)blockIdx.x * 256u + threadIdx.x; i < n; i += (usize)gridDim.x * 256u) p[i] = 0u;
}

template <int W> __global__ void __launch_bounds__(kSegThreads) k_seg_main(SegArgs a)
{
    constexpr int AW = W + 2;
    __shared__ i32 s_ends[2];
    __shared__ u32 red[(kSegThreads / 64) * AW];
    const u32 tid = threadIdx.x;
    const i64 lo = (i64)blockIdx.x * kSegTile;
    const i64 hi = lo + kSegTile < a.rows ? lo + kSegTile : a.rows;
    const i32 last = a.ngroups - 1;
    if (tid == 0) s_ends[0] = seg_search(a.groups, 0, last, lo);
    if (tid == 64) s_ends[1] = seg_search(a.groups, 0, last, hi - 1);
    __syncthreads();
    const i32 sf = s_ends[0], sl = s_ends[1];
    if (sf == sl) {
        // the tile lies inside one segment (or before the first one)
        if (sf < 0) return;
        Mag<AW> acc = mzero<AW>();
        for (i64 i = lo + tid; i < hi; i += kSegThreads) {
            const usize row = a.assoc ? (usize)(u32)a.assoc[i] : (usize)i;
            acc_add<AW, W>(acc, load_compact_rt<W>(a.col, row, a.B, a.colbytes));
        }
        wave_reduce<AW>(acc);
        if ((tid & 63u) == 0)
#pragma unroll
            for (int w = 0; w < AW; w++) red[(tid >> 6) * AW + w] = acc.w[w];
        __syncthreads();
        if (tid == 0) {
            Mag<AW> t = mzero<AW>();
            for (int wv = 0; wv < kSegThreads / 64; wv++) {
                Mag<AW> o;
#pragma unroll
                for (int w = 0; w < AW; w++) o.w[w] = red[wv * AW + w];
                add_to<AW>(t, t, o);
            }
            seg_flush<AW>(a.acc, sf, t);
        }
        return;
    }
    // the tile crosses segment starts: every lane takes a run of
    // kSegRowsPerThread consecutive positions, so its segment changes only
    // at the starts inside its run (a forward walk over the starts) and it
    // adds its running sum to a segment once per segment it meets
    const i64 r0 = lo + (i64)tid * kSegRowsPerThread;
    const i64 r1 = r0 + kSegRowsPerThread < hi ? r0 + kSegRowsPerThread : hi;
    if (r0 >= hi) return;
    i32 s = seg_search(a.groups, sf < 0 ? 0 : sf, sl, r0);
    Mag<AW> acc = mzero<AW>();
    for (i64 i = r0; i < r1; i++) {
        i32 si = s;
        while (si < sl && (i64)a.groups[si + 1] <= i) si++;
        if (si != s) {
            if (s >= 0) seg_flush<AW>(a.acc, s, acc);
            acc = mzero<AW>();
            s = si;
        }
        if (s < 0) continue;
        const usize row = a.assoc ? (usize)(u32)a.assoc[i] : (usize)i;
        acc_add<AW, W>(acc, load_compact_rt<W>(a.col, row, a.B, a.colbytes));
    }
    if (s >= 0) seg_flush<AW>(a.acc, s, acc);
}

template <int W> __global__ void __launch_bounds__(64) k_seg_final(const u64* acc, i32 ngroups, u32 P, u32 S, u8* out)
{
    constexpr int AW = W + 2;
    const i32 g = (i32)(blockIdx.x * 64u + threadIdx.x);
    if (g >= ngroups) return;
    const u64* p = acc + (usize)g * AW;
    Mag<AW> m;
    u64 c = 0;
#pragma unroll
    for (int w = 0; w < AW; w++) {
        const u64 x = p[w];
        const u64 lo = (x & 0xffffffffull) + c;
        m.w[w] = (u32)lo;
        c = (x >> 32) + (lo >> 32);
    }
    write_sum_image<AW>(out + (usize)g * 133u, m, P & 0xffu, S & 0xffu, 0u);
}

// width class of a compact row of B bytes: limbs W >= ceil(B / 4)
int seg_class_words(int B)
{
    return B <= 8 ? 2 : B <= 16 ? 4 : B <= 32 ? 8 : B <= 64 ? 16 : B <= 128 ? 32 : 0;
}

int seg_check(const uint8_t* d_col, int precision, int64_t rows, const int32_t* d_groups, int32_t ngroups,
              uint8_t* d_out, int* B, int* W)
{
    if (rows < 0 || ngroups < 0 || (ngroups && (!d_groups || !d_out)) || (rows && !d_col) || precision <= 0)
        return api_fail(UPDEC_EINVAL, "bad argument");
    if (rows > (int64_t)INT32_MAX + 1)
        return api_fail(UPDEC_EINVAL, "positions are int32 (associated, groups): at most 2^31 rows");
    *B = compact_bytes(precision);
    *W = seg_class_words(*B);
    if (!*W) return api_fail(UPDEC_EUNSUPPORTED, "rows wider than 128 bytes (the 32-limb Decimal) are not summed");
    return UPDEC_OK;
}

template <int W>
int seg_launch_w(const uint8_t* d_col, int P, int S_, int B, int64_t rows, const int32_t* d_assoc, const int32_t* d_groups,
                 int32_t ngroups, uint8_t* d_out, u64* acc, hipStream_t st)
{
    constexpr int AW = W + 2;
    const usize nacc = (usize)ngroups * AW;
    const unsigned zb = (unsigned)std::min<usize>((nacc + 255) / 256, 4096);
    hipLaunchKernelGGL(k_seg_zero, dim3(zb), dim3(256), 0, st, acc, nacc);
    SEGCHK(hipGetLastError());
    if (rows > 0) {
        SegArgs a{d_col, (usize)rows * (usize)B, (u32)B, 0u, d_assoc, d_groups, ngroups, 0, rows, acc};
        const i64 grid = (rows + kSegTile - 1) / kSegTile;
        hipLaunchKernelGGL(k_seg_main<W>, dim3((unsigned)grid), dim3(kSegThreads), 0, st, a);
        SEGCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_seg_final<W>, dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0, st, (const u64*)acc, ngroups,
                       (u32)P, (u32)S_, d_out);
    SEGCHK(hipGetLastError());
    return UPDEC_OK;
}

int seg_launch(const uint8_t* d_col, int P, int S_, int B, int W, int64_t rows, const int32_t* d_assoc,
               const int32_t* d_groups, int32_t ngroups, uint8_t* d_out, u64* acc, hipStream_t st)
{
    switch (W) {
    case 2: return seg_launch_w<2>(d_col, P, S_, B, rows, d_assoc, d_groups, ngroups, d_out, acc, st);
    case 4: return seg_launch_w<4>(d_col, P, S_, B, rows, d_assoc, d_groups, ngroups, d_out, acc, st);
    case 8: return seg_launch_w<8>(d_col, P, S_, B, rows, d_assoc, d_groups, ngroups, d_out, acc, st);
    case 16: return seg_launch_w<16>(d_col, P, S_, B, rows, d_assoc, d_groups, ngroups, d_out, acc, st);
    default: return seg_launch_w<32>(d_col, P, S_, B, rows, d_assoc, d_groups, ngroups, d_out, acc, st);
    }
}

// ------------------------------------------------------------------ fixed scratch of the no-alloc entry point

// Accumulators of updec_segsum_compact_noalloc: 8 MiB on every device the
// library runs on, held until the process ends.  The array's name is a host
// stand-in; kernels get the device address from hipGetSymbolAddress.
constexpr usize kSegScratchWords = (usize)1 << 20;
__device__ __align__(256) u64 scratch[kSegScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
u64* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(u64** out)
{
    int dev = 0;
    SEGCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return api_fail(UPDEC_EUNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SEGCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<u64*>(p);
    }
    *out = g_scratch_ptr[dev];
    return UPDEC_OK;
}

}  // namespace

namespace updec {
// set-up of the no-alloc entry point's scratch on the current device (called
// by updec_init)
int segsum_scratch_setup()
{
    u64* p = nullptr;
    return scratch_for_current_device(&p);
}
}  // namespace updec

extern "C" {

int updec_segsum_compact(const uint8_t* d_col, int precision, int scale, int64_t rows, const int32_t* d_associated,
                         const int32_t* d_groups, int32_t ngroups, uint8_t* d_out, int mode, updec_stream_t stream)
{
    (void)mode;  // a sum is the same in both modes: the image keeps 32 limbs
    int B = 0, W = 0;
    int rc = seg_check(d_col, precision, rows, d_groups, ngroups, d_out, &B, &W);
    if (rc != UPDEC_OK || ngroups == 0) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    u64* acc = nullptr;
    SEGCHK(hipMallocAsync(reinterpret_cast<void**>(&acc), (usize)ngroups * (W + 2) * sizeof(u64), st));
    rc = seg_launch(d_col, precision, scale, B, W, rows, d_associated, d_groups, ngroups, d_out, acc, st);
    const hipError_t fe = hipFreeAsync(acc, st);
    if (rc != UPDEC_OK) return rc;
    if (fe != hipSuccess) return api_fail(UPDEC_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
    return UPDEC_OK;
}

int updec_segsum_compact_noalloc(const uint8_t* d_col, int precision, int scale, int64_t rows,
                                 const int32_t* d_associated, const int32_t* d_groups, int32_t ngroups, uint8_t* d_out,
                                 int mode, updec_stream_t stream)
{
    (void)mode;
    int B = 0, W = 0;
    int rc = seg_check(d_col, precision, rows, d_groups, ngroups, d_out, &B, &W);
    if (rc != UPDEC_OK || ngroups == 0) return rc;
    const usize need = (usize)ngroups * (W + 2);
    if (need > kSegScratchWords)
        return api_fail(UPDEC_EINVAL, "segments need " + std::to_string(need * 8) + " bytes of accumulators; the fixed "
                                      "scratch holds " + std::to_string(kSegScratchWords * 8) +
                                      " (at most " + std::to_string(kSegScratchWords / (W + 2)) +
                                      " segments at this precision)");
    u64* acc = nullptr;
    rc = scratch_for_current_device(&acc);
    if (rc != UPDEC_OK) return rc;
    return seg_launch(d_col, precision, scale, B, W, rows, d_associated, d_groups, ngroups, d_out, acc,
                      reinterpret_cast<hipStream_t>(stream));
}

int updec_segsum_noalloc_max_segments(int precision, int32_t* max_segments)
{
    if (precision <= 0 || !max_segments) return api_fail(UPDEC_EINVAL, "bad argument");
    const int W = seg_class_words(compact_bytes(precision));
    if (!W) return api_fail(UPDEC_EUNSUPPORTED, "rows wider than 128 bytes (the 32-limb Decimal) are not summed");
    *max_segments = (int32_t)(kSegScratchWords / (W + 2));
    return UPDEC_OK;
}

}  // extern "C"
