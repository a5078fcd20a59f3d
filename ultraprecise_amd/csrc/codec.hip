// codec.hip -- the compact codec at the edges of the path: bulk ingest of
// decimal text into compact columns, and bulk output formatting of compact
// columns (include/updec.h, "codec").
//
// Ingest, one lane per row, reproduces Decimal(P, S, text)
// (src/datatypes/decimal.cu:154-165) followed by ToCompactDecimal
// (decimal.cu:337-341), the per-value step of the CSV import
// (src/server/mysql/importCsvFile.cpp:640-660):
//   StringToDecimal (decimal.cu:2980-3025): after an optional leading '-',
//   every character c is accumulated as v = v*10 + (c - '0') modulo 2^1024,
//   (c - '0') signed, so a character that is not a digit still shifts and
//   offsets the value; a '.' sets the fraction flag and the character after
//   it is taken as a digit whatever it is; every character after the first
//   '.' counts one fractional digit.  A row ends at its length or at its
//   first NUL; a '.' that ends the row ends it (the reference goes on to
//   read its terminator as a digit and past it).
//   cast (decimal.cu:471-533): scale up multiplies by 10^(S - frac) modulo
//   2^1024; scale down is floor division by 10^(frac - S); the sign is kept.
//   ToCompactDecimal: the low B(P) bytes of the magnitude with the sign
//   OR-ed into bit 7 of the last one.
// Multiplication and addition modulo 2^(32 W) give the low 32 W bits of the
// reference's 32-limb result exactly, so a row whose cast only scales up is
// parsed at the output's width.  A row that is cast down divides the whole
// value: it is parsed at 32 limbs, or at 4, 8 or 16 when it is digits only
// and short enough that its value fits (10^38 < 2^127, 10^77 < 2^256,
// 10^154 < 2^512).
//
// Output, one lane per row, formats Decimal(compact, P, S)
// (decimal.cu:167-176) either as GetDecimal (decimal.cu:362-397, the form
// AriesDataBuffer::GetCompactDecimalAsString returns,
// src/CudaAcc/AriesEngineDef.cu:817-829): '+' or '-', eight upper-case hex
// digits per limb from the top nonzero limb down, "(" three digits of the
// scale ")"; or as decimal text with exactly `scale` fractional digits
// ("-123.4500"), the form the reference's decimal tests print.  Each row
// goes to a fixed stride-byte slot, NUL-terminated and zero-filled; rows are
// assembled in LDS and written out with coalesced dword stores.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/updec.h"
#include "api_internal.h"
#include "dec_device.h"
#include "planner.h"

using namespace updec;

namespace {

#define CODECCHK(expr)                                                                                       \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return api_fail(UPDEC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline hipStream_t st_of(updec_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// the reference's Decimal holds 32 limbs; B(P) above 128 bytes would not fit it
constexpr int kMaxParseBytes = 128;

// limb class (2, 4, 8, 16, 32) that holds B bytes
int limb_class(int B)
{
    const int w = (B + 3) / 4;
    return w <= 2 ? 2 : w <= 4 ? 4 : w <= 8 ? 8 : w <= 16 ? 16 : 32;
}

// ---------------------------------------------------------------- ingest

// scan of one row: characters taken, fractional digits, sign, and whether
// every character taken is a decimal digit
struct RowScan {
    u32 ndig, frac, neg, plain;
};

__device__ RowScan scan_row(const char* s, u32 len)
{
    RowScan r{0u, 0u, 0u, 1u};
    u32 i = 0;
    if (len > 0 && s[0] == '-') {
        r.neg = 1u;
        i = 1;
    }
    bool dot = false;
    for (; i < len; i++) {
        char c = s[i];
        if (c == 0) break;
        if (c == '.') {
            dot = true;
            i++;
            if (i >= len || s[i] == 0) break;
            c = s[i];
        }
        if (dot) r.frac++;
        r.plain &= ((u32)((int)(signed char)c - '0') <= 9u) ? 1u : 0u;
        r.ndig++;
    }
    return r;
}

// v = v * 10^k + chunk (chunk signed) modulo 2^(32 W)
template <int W> __device__ void fold_chunk(Mag<W>& v, u32 k, i64 chunk)
{
    mul_word<W>(v, pow10_word(k));
    const u32 ext = chunk < 0 ? 0xffffffffu : 0u;
    const u64 lo = (u64)chunk;
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const u32 a = i == 0 ? (u32)lo : (i == 1 ? (u32)(lo >> 32) : ext);
        const u64 t = (u64)v.w[i] + a + c;
        v.w[i] = (u32)t;
        c = t >> 32;
    }
}

// StringToDecimal modulo 2^(32 W), nine characters per fold
template <int W> __device__ Mag<W> parse_value(const char* s, u32 len)
{
    Mag<W> v = mzero<W>();
    u32 i = (len > 0 && s[0] == '-') ? 1u : 0u;
    i64 chunk = 0;
    u32 k = 0;
    for (; i < len; i++) {
        char c = s[i];
        if (c == 0) break;
        if (c == '.') {
            i++;
            if (i >= len || s[i] == 0) break;
            c = s[i];
        }
        chunk = chunk * 10 + ((i64)(signed char)c - '0');
        if (++k == 9u) {
            fold_chunk<W>(v, k, chunk);
            chunk = 0;
            k = 0;
        }
    }
    if (k) fold_chunk<W>(v, k, chunk);
    return v;
}

// parse, cast to scale S, low bytes as WO words
template <int W, int WO> __device__ Mag<WO> parse_cast(const char* s, u32 len, const RowScan& sc, u32 S)
{
    Mag<W> v = parse_value<W>(s, len);
    if (S >= sc.frac) v = mul_p10_rt<W>(v, S - sc.frac);
    else v = div_p10_rt<W>(v, sc.frac - S);
    return resize<WO>(v);
}

// rows per block of the ingest and format kernels
constexpr int kCodecBlock = 64;

// Rows of at most 16 bytes (every DECIMAL(15,2) text, for one) take one
// pass over registers: the dwords covering the row are loaded (only those
// holding a byte of it, so nothing past the text is touched beyond its last
// dword), realigned to the row's first byte, and the same character rules
// run over the 16 byte positions with static indices.  Sixteen characters
// keep the accumulated value within an int64 (|v| < 208 * 10^16 / 9).
struct ShortRow {
    i64 v;
    u32 frac, neg, plain;
};

__device__ ShortRow scan_short(const char* s, u32 len)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(s);
    const u32* p = reinterpret_cast<const u32*>(a & ~(uintptr_t)3);
    const u32 sh = (u32)(a & 3u);
    const u32 nd = (sh + len + 3u) >> 2;
    u32 w[6];
#pragma unroll
    for (int k = 0; k < 5; k++) w[k] = (u32)k < nd ? p[k] : 0u;
    w[5] = 0u;
    u32 t[4];
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = __builtin_amdgcn_alignbyte(w[j + 1], w[j], sh);
    ShortRow r{0, 0u, 0u, 1u};
    bool dot = false, force = false, stop = false;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const u32 c = (t[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        if ((u32)i >= len || stop) continue;
        if (i == 0 && c == (u32)'-') {
            r.neg = 1u;
            continue;
        }
        if (c == 0u) {  // the row ends at its first NUL, also right after a '.'
            stop = true;
            continue;
        }
        if (!force && c == (u32)'.') {
            dot = true;
            force = true;  // the next character is taken as a digit whatever it is
            continue;
        }
        force = false;
        if (dot) r.frac++;
        const i32 d = (i32)(signed char)(char)c - '0';
        r.plain &= ((u32)d <= 9u) ? 1u : 0u;
        r.v = r.v * 10 + d;
    }
    return r;
}

template <int WO>
__global__ void __launch_bounds__(kCodecBlock) k_parse_compact(const char* __restrict__ chars,
                                                               const int64_t* __restrict__ offsets, int64_t rows,
                                                               u32 B, u32 S, u8* __restrict__ out)
{
    extern __shared__ u32 tile[];  // kCodecBlock * B bytes, rounded up to dwords
    u8* t8 = reinterpret_cast<u8*>(tile);
    const int64_t base = (int64_t)blockIdx.x * kCodecBlock;
    const int64_t row = base + threadIdx.x;
    if (row < rows) {
        const int64_t o0 = offsets[row], o1 = offsets[row + 1];
        const char* s = chars + o0;
        const u32 len = o1 > o0 ? (u32)(o1 - o0) : 0u;
        Mag<WO> m;
        u32 neg;
        bool done = false;
        if (len <= 16u) {
            const ShortRow sr = scan_short(s, len);
            neg = sr.neg;
            if (S >= sr.frac || sr.plain) {
                // v * 10^(S - frac) modulo 2^(32 WO), or floor(v / 10^(frac - S)) for v >= 0
                u64 u = (u64)sr.v;
                if (S < sr.frac) {
                    const u32 n = sr.frac - S;  // <= 16
                    u64 d = 1;
                    for (u32 k = 0; k < n; k++) d *= 10u;
                    u /= d;
                }
                const u32 ext = (S >= sr.frac && sr.v < 0) ? 0xffffffffu : 0u;
#pragma unroll
                for (int i = 0; i < WO; i++) m.w[i] = i == 0 ? (u32)u : (i == 1 ? (u32)(u >> 32) : ext);
                if (S > sr.frac) m = mul_p10_rt<WO>(m, S - sr.frac);
                done = true;
            }
        }
        if (!done) {
        const RowScan sc = scan_row(s, len);
        neg = sc.neg;
        const bool down = S < sc.frac;
        if (!down) {
            m = parse_cast<WO, WO>(s, len, sc, S);
        } else if (sc.plain && sc.ndig <= 38u) {
            m = parse_cast<(WO > 4 ? WO : 4), WO>(s, len, sc, S);
        } else if (sc.plain && sc.ndig <= 77u) {
            m = parse_cast<(WO > 8 ? WO : 8), WO>(s, len, sc, S);
        } else if (sc.plain && sc.ndig <= 154u) {
            m = parse_cast<(WO > 16 ? WO : 16), WO>(s, len, sc, S);
        } else {
            m = parse_cast<32, WO>(s, len, sc, S);
        }
        }
        u8* dst = t8 + (u32)threadIdx.x * B;
#pragma unroll
        for (int b = 0; b < 4 * WO; b++)
            if ((u32)b < B) dst[b] = (u8)(m.w[b >> 2] >> (8 * (b & 3)));
        dst[B - 1] |= (u8)(neg << 7);
    }
    __syncthreads();
    const int64_t nrows = rows - base < kCodecBlock ? rows - base : kCodecBlock;
    const u32 n = (u32)nrows * B;
    u8* d = out + base * (int64_t)B;
    // the block's span starts at base * B: dword stores when that is aligned
    if (((uintptr_t)d & 3u) == 0) {
        for (u32 i = threadIdx.x; i < n / 4; i += kCodecBlock) reinterpret_cast<u32*>(d)[i] = tile[i];
        for (u32 i = (n / 4) * 4 + threadIdx.x; i < n; i += kCodecBlock) d[i] = t8[i];
    } else {
        for (u32 i = threadIdx.x; i < n; i += kCodecBlock) d[i] = t8[i];
    }
}

// ---------------------------------------------------------------- output

template <int W> __device__ u32 div_1e9_rem(Mag<W>& x)
{
    u32 r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) x.w[i] = div_word_const<1000000000u>(r, x.w[i]);
    return r;
}

// LDS per row: the formatted slot (`slot` bytes, the longest form rounded
// to dwords) and, for text, a digit area of slot + 12 bytes (digits come in
// groups of nine, up to eight past the longest form)
template <int W>
__global__ void __launch_bounds__(kCodecBlock) k_format_compact(const u8* __restrict__ col, int64_t rows, u32 B,
                                                                u32 S, u32 hex, u32 slot, u32 stride,
                                                                char* __restrict__ out)
{
    extern __shared__ u32 tile[];
    u8* t8 = reinterpret_cast<u8*>(tile);
    const int64_t base = (int64_t)blockIdx.x * kCodecBlock;
    const int64_t row = base + threadIdx.x;
    u8* o = t8 + (u32)threadIdx.x * slot;
    for (u32 i = 0; i < slot; i++) o[i] = 0;
    if (row < rows) {
        const Dec<W> v = load_compact_rt<W>(col, (usize)row, B, (usize)rows * B);
        u32 p = 0;
        if (hex) {
            o[p++] = v.neg ? '-' : '+';
            bool lead = false;
#pragma unroll
            for (int i = W - 1; i >= 0; i--) {
                lead = lead || v.m.w[i] != 0u;
                if (lead) {
#pragma unroll
                    for (int j = 7; j >= 0; j--) {
                        const u32 nib = (v.m.w[i] >> (4 * j)) & 0xFu;
                        o[p++] = (u8)(nib < 10u ? '0' + nib : 'A' + nib - 10u);
                    }
                }
            }
            o[p++] = '(';
            o[p++] = (u8)('0' + S / 100u);
            o[p++] = (u8)('0' + S / 10u % 10u);
            o[p++] = (u8)('0' + S % 10u);
            o[p++] = ')';
        } else {
            // decimal digits, least significant first, into the digit area
            u8* dig = t8 + (u32)kCodecBlock * slot + (u32)threadIdx.x * (slot + 12u);
            Mag<W> x = v.m;
            u32 n = 0;
            while (!is_zero(x)) {
                u32 r = div_1e9_rem<W>(x);
                for (int k = 0; k < 9; k++) {
                    dig[n++] = (u8)('0' + r % 10u);
                    r /= 10u;
                }
            }
            while (n > 1 && dig[n - 1] == '0') n--;
            if (n == 0) dig[n++] = '0';
            while (n <= S) dig[n++] = '0';
            if (v.neg) o[p++] = '-';
            for (int i = (int)n - 1; i >= 0; i--) {
                o[p++] = dig[i];
                if ((u32)i == S && S > 0) o[p++] = '.';
            }
        }
    }
    __syncthreads();
    // the block's rows are consecutive stride-byte slots: dword stores,
    // zeros past each row's formatted slot
    const int64_t nrows = rows - base < kCodecBlock ? rows - base : kCodecBlock;
    const u32 sw = stride / 4, lw = slot / 4;
    const u32 nw = (u32)nrows * sw;
    u32* d = reinterpret_cast<u32*>(out + base * (int64_t)stride);
    for (u32 i = threadIdx.x; i < nw; i += kCodecBlock) {
        const u32 r = i / sw, w = i - r * sw;
        d[i] = w < lw ? tile[r * lw + w] : 0u;
    }
}

// characters of the longest decimal text of DECIMAL(P, S): sign, the
// digits of 2^(8B - 1) - 1 (or S + 1 when more), the point, NUL
int64_t text_slot(int P, int S)
{
    const int B = compact_bytes(P);
    const int bits = 8 * B - 1;
    const int digits = (int)std::floor(bits * 0.30102999566398119521) + 1;
    const int body = digits > S + 1 ? digits : S + 1;
    return 1 + body + (S > 0 ? 1 : 0) + 1;
}

// GetDecimal's longest form: sign, 8 hex digits per limb of a magnitude
// below 2^(8B - 1), "(SSS)", NUL
int64_t hex_slot(int P)
{
    const int B = compact_bytes(P);
    const int limbs = (8 * B - 1 + 31) / 32;
    return 1 + 8 * (int64_t)limbs + 5 + 1;
}

}  // namespace

extern "C" {

int updec_parse_compact(const char* d_chars, const int64_t* d_offsets, int64_t rows, int32_t precision,
                        int32_t scale, uint8_t* d_out, updec_stream_t stream)
{
    if (rows < 0 || (rows && (!d_chars || !d_offsets || !d_out))) return api_fail(UPDEC_EINVAL, "bad argument");
    if (precision < 1 || scale < 0 || scale > precision) return api_fail(UPDEC_EINVAL, "bad precision or scale");
    const int B = compact_bytes(precision);
    if (B > kMaxParseBytes)
        return api_fail(UPDEC_EUNSUPPORTED, "B(P) above 128 bytes: the reference's 32-limb Decimal cannot hold it");
    if (rows == 0) return UPDEC_OK;
    const unsigned grid = (unsigned)((rows + kCodecBlock - 1) / kCodecBlock);
    const size_t lds = ((size_t)kCodecBlock * B + 3) / 4 * 4;
    hipStream_t st = st_of(stream);
    switch (limb_class(B)) {
    case 2: hipLaunchKernelGGL(k_parse_compact<2>, dim3(grid), dim3(kCodecBlock), lds, st, d_chars, d_offsets, rows, (u32)B, (u32)scale, d_out); break;
    case 4: hipLaunchKernelGGL(k_parse_compact<4>, dim3(grid), dim3(kCodecBlock), lds, st, d_chars, d_offsets, rows, (u32)B, (u32)scale, d_out); break;
    case 8: hipLaunchKernelGGL(k_parse_compact<8>, dim3(grid), dim3(kCodecBlock), lds, st, d_chars, d_offsets, rows, (u32)B, (u32)scale, d_out); break;
    case 16: hipLaunchKernelGGL(k_parse_compact<16>, dim3(grid), dim3(kCodecBlock), lds, st, d_chars, d_offsets, rows, (u32)B, (u32)scale, d_out); break;
    default: hipLaunchKernelGGL(k_parse_compact<32>, dim3(grid), dim3(kCodecBlock), lds, st, d_chars, d_offsets, rows, (u32)B, (u32)scale, d_out); break;
    }
    CODECCHK(hipGetLastError());
    return UPDEC_OK;
}

int updec_format_stride(int32_t precision, int32_t scale, int32_t form, int64_t* stride)
{
    if (!stride || precision < 1 || scale < 0 || scale > precision) return api_fail(UPDEC_EINVAL, "bad argument");
    if (form != UPDEC_FORMAT_HEX && form != UPDEC_FORMAT_TEXT) return api_fail(UPDEC_EINVAL, "unknown format");
    if (compact_bytes(precision) > kMaxParseBytes) return api_fail(UPDEC_EUNSUPPORTED, "B(P) above 128 bytes");
    const int64_t n = form == UPDEC_FORMAT_HEX ? hex_slot(precision) : text_slot(precision, scale);
    *stride = (n + 3) / 4 * 4;
    return UPDEC_OK;
}

int updec_format_compact(const uint8_t* d_col, int64_t rows, int32_t precision, int32_t scale, int32_t form,
                         char* d_out, int64_t stride, updec_stream_t stream)
{
    int64_t need = 0;
    const int rc = updec_format_stride(precision, scale, form, &need);
    if (rc != UPDEC_OK) return rc;
    if (rows < 0 || (rows && (!d_col || !d_out))) return api_fail(UPDEC_EINVAL, "bad argument");
    if (stride < need || (stride & 3) || stride > (1 << 20))
        return api_fail(UPDEC_EINVAL, "stride below updec_format_stride, not a multiple of 4, or above 1 MiB");
    if (reinterpret_cast<uintptr_t>(d_out) & 3u) return api_fail(UPDEC_EINVAL, "output must be 4-byte aligned");
    if (rows == 0) return UPDEC_OK;
    const int B = compact_bytes(precision);
    const unsigned grid = (unsigned)((rows + kCodecBlock - 1) / kCodecBlock);
    const u32 hex = form == UPDEC_FORMAT_HEX ? 1u : 0u;
    const u32 slot = (u32)need;
    const size_t lds = (size_t)kCodecBlock * slot + (hex ? 0 : (size_t)kCodecBlock * (slot + 12u));
    hipStream_t st = st_of(stream);
    const u32 sb = (u32)B, ss = (u32)scale, sd = (u32)stride;
    switch (limb_class(B)) {
    case 2: hipLaunchKernelGGL(k_format_compact<2>, dim3(grid), dim3(kCodecBlock), lds, st, d_col, rows, sb, ss, hex, slot, sd, d_out); break;
    case 4: hipLaunchKernelGGL(k_format_compact<4>, dim3(grid), dim3(kCodecBlock), lds, st, d_col, rows, sb, ss, hex, slot, sd, d_out); break;
    case 8: hipLaunchKernelGGL(k_format_compact<8>, dim3(grid), dim3(kCodecBlock), lds, st, d_col, rows, sb, ss, hex, slot, sd, d_out); break;
    case 16: hipLaunchKernelGGL(k_format_compact<16>, dim3(grid), dim3(kCodecBlock), lds, st, d_col, rows, sb, ss, hex, slot, sd, d_out); break;
    default: hipLaunchKernelGGL(k_format_compact<32>, dim3(grid), dim3(kCodecBlock), lds, st, d_col, rows, sb, ss, hex, slot, sd, d_out); break;
    }
    CODECCHK(hipGetLastError());
    return UPDEC_OK;
}

}  // extern "C"
