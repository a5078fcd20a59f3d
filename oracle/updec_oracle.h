/*
 * updec_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's host `Decimal` arithmetic
 * (src/datatypes/decimal.hxx:130-141, src/datatypes/decimal.cu) used as the
 * parity oracle for the HIP path and as the CPU baseline timed beside it.
 * Nothing in the product (ultraprecise_amd/) may link, load or call this.
 *
 * Two evaluation modes:
 *   OD_FAITHFUL (0): 32 x 32-bit limbs, magnitudes wrap mod 2^1024, the
 *       reference's DivOrMod path quirks (uint64 wrap in path (i), uint32
 *       round-compare wrap in path (ii), the single-bit-short rule in path (iii)).
 *   OD_EXTENDED (1): same rules, exact (OD_MAXW limbs, no wrap, no quirks).
 */
#ifndef UPDEC_ORACLE_H
#define UPDEC_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OD_FAITHFUL 0
#define OD_EXTENDED 1
#ifndef OD_MAXW
#define OD_MAXW 192
#endif

typedef struct od_dec {
    int32_t sign;   /* 0 = +, 1 = - (sign-magnitude; -0 is representable) */
    int32_t prec;   /* informational, as in the reference */
    int32_t frac;   /* scale: value = (-1)^sign * M / 10^frac */
    int32_t error;  /* 0 ok, 2 = ERR_DIV_BY_ZERO (decimal.hxx:27) */
    uint32_t v[OD_MAXW]; /* little-endian base-2^32 magnitude */
} od_dec;

/* layout helpers */
int  od_compact_bytes(int P);                       /* B(P): decimal.cu:92, closed form past 299 */
void od_zero(od_dec* d);
void od_from_compact(od_dec* d, const uint8_t* buf, int P, int S);   /* decimal.cu:168-176 */
void od_to_compact(const od_dec* d, uint8_t* buf, int len);          /* decimal.cu:337-341 */
int  od_from_string(od_dec* d, const char* s, int mode);             /* decimal.cu:2980-3025 */
void od_cast(od_dec* d, int P, int S, int mode);                     /* decimal.cu:471-533 */
void od_truncate(od_dec* d, int p, int mode);                        /* decimal.cu:541-570 */
void od_from_int32_scale(od_dec* d, int32_t t, int32_t scale, int mode); /* decimal.cu:230-260 */
void od_from_int64(od_dec* d, int64_t t);                            /* decimal.cu:301-309 */
void od_rescale_up(od_dec* d, int scale, int mode);                  /* decimal.cu:263-268 */

/* arithmetic (in place on a) */
void od_shl10(od_dec* d, int n, int mode);                           /* decimal.cu:2028-2058 */
void od_shr10(od_dec* d, int n, int mode);                           /* decimal.cu:2003-2026 */
void od_neg(od_dec* d);                                              /* decimal.cu:578-586 */
void od_add(od_dec* a, const od_dec* b, int mode);                   /* decimal.cu:1284-1323 */
void od_sub(od_dec* a, const od_dec* b, int mode);                   /* decimal.cu:1575-1613 */
void od_mul(od_dec* a, const od_dec* b, int mode);                   /* decimal.cu:1325-1351 */
void od_div(od_dec* a, const od_dec* b, int mode);                   /* decimal.cu:2201-2333 */
void od_mod(od_dec* a, const od_dec* b, int mode);                   /* decimal.cu:2621-2716 */
int  od_cmp(const od_dec* a, const od_dec* b, int mode);             /* decimal.cu:638-745 */
int  od_is_zero(const od_dec* d);
int  od_bitlen(const od_dec* d);

/* 133-byte nullable_type<Decimal> image: flag, sign, prec, frac, error, v[32] (aries_types.hxx:298-310) */
void od_to_nullable_image(const od_dec* d, uint8_t* out133);
/* decimal string of value (for tests): "-123.4500" */
int  od_to_string(const od_dec* d, char* buf, int cap);
/* GetDecimal hex form (decimal.cu:362-397): "+0000000C(002)" */
int  od_format_hex(const od_dec* d, char* out);
/* bulk text -> compact (importCsvFile.cpp:640-660 via decimal.cu:154-165, 337-341) */
void od_parse_compact(const char* chars, const int64_t* offsets, int64_t rows, int P, int S, uint8_t* out);
/* bulk compact -> hex (hex = 1) or decimal text (hex = 0), stride-byte NUL-padded slots */
void od_format_compact(const uint8_t* col, int64_t rows, int P, int S, int hex, char* out, int64_t stride);

/* overflow accounting for OD_EXTENDED (set when a result would exceed OD_MAXW limbs) */
int  od_overflow_flag(void);
void od_clear_overflow(void);

/* ---- plan evaluation (the rewritten expression tree, text form) ----
 * plan := '(' OP ' ' P ' ' S ' ' plan ' ' plan ')'  OP in ADD SUB MUL DIV
 *       | '(' 'NEG' ' ' plan ')'
 *       | '(' 'CAST' ' ' P ' ' S ' ' plan ')'   od_cast to DECIMAL(P,S)
 *       | '(' 'TRUNC' ' ' P ' ' S ' ' p ' ' plan ')'   od_truncate(p) with the
 *                                              argument's plan type as prec/frac
 *       | 'c' IDX ':' P ':' S          compact column operand
 *       | 'k' ('+'|'-') HEX ':' S      constant: big-endian hex magnitude, scale
 */
typedef struct od_plan od_plan;
od_plan* od_plan_parse(const char* text);
void     od_plan_free(od_plan* p);
int      od_plan_result(const od_plan* p, int* P, int* S);

/* evaluate rows [0,rows) into compact output rows of `outlen` bytes */
int od_plan_eval(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                 int64_t rows, uint8_t* out, int outlen, int mode, int threads);
/* SUM of the plan value over rows (optionally only rows with mask[i] != 0).
 * Writes the 133-byte nullable image and returns the row count summed. */
int64_t od_plan_sum(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                    int64_t rows, const uint8_t* mask, uint8_t* out133, int mode, int threads);
/* grouped SUM: gid[i] in [0,ngroups) or <0 to skip; out133 = ngroups * 133 bytes; counts = ngroups int64 */
int od_plan_group_sum(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                      int64_t rows, const int32_t* gid, int ngroups, uint8_t* out133, int64_t* counts,
                      int mode, int threads);
/* segmented SUM over a compact column in `associated` order with segment starts `groups`
 * (AriesEngineAlgorithm.h:832-857, kernel_segreduce.hxx:609-612) */
int od_segsum_compact(const uint8_t* col, int P, int S, int64_t rows, const int32_t* associated,
                      const int32_t* groups, int ngroups, uint8_t* out133, int mode, int threads);
/* AVG = SUM image / Decimal((int64)count) (AriesSqlOperator_helpers.cu:717-764, kernel_util.hxx:427-438) */
void od_avg_from_sum(const uint8_t* sum133, int64_t count, uint8_t* out133, int mode);

#ifdef __cplusplus
}
#endif
#endif
