// planner.h -- host side of the decimal expression path: parse, the
// reference's alignment scheduling / constant pre-calculation rewrite,
// per-node result types, limb-width analysis, and HIP source generation.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "vm.h"

namespace updec {

enum class Mode { Faithful = 0, Extended = 1 };

// Host decimal used for constant pre-calculation (the reference folds
// constants with Decimal::AddDecimalOnCpu / MulDecimalOnCpu on the host,
// AriesEngineShell.cpp:360-597).  Exact limbs; faithful mode reduces every
// result mod 2^1024 like the reference's 32-limb Decimal.
struct HostDec {
    static constexpr int NW = 96;
    uint32_t sign = 0;
    int32_t prec = 0;
    int32_t frac = 0;
    uint32_t v[NW] = {0};

    static HostDec from_string(const std::string& s, Mode m);     // decimal.cu:2980-3025
    static HostDec from_int(int64_t t, int prec);                  // decimal.cu:281-309
    static HostDec from_int_scale(int32_t t, int32_t scale, Mode m);// decimal.cu:230-260
    HostDec rescaled_up(int scale, Mode m) const;                  // decimal.cu:263-268
    void add(const HostDec& d, Mode m);                            // decimal.cu:1284-1323
    void mul(const HostDec& d, Mode m);                            // decimal.cu:1325-1351
    bool is_zero() const;
    bool equals_one() const;                                        // operator!=(Decimal, int)
    void check_and_set_real_precision();                           // decimal.cu:2060-2088
    int bitlen() const;
    std::string hex() const;
};

struct ColumnType {
    int P = 0, S = 0;
    // Optional statistic: every stored magnitude is < 2^bits (0 = unknown,
    // then the layout bound 8*B(P)-1 is used).  Narrower bounds give
    // narrower limb arrays; a wrong bound gives wrong results.
    int bits = 0;
};

// Expression tree, mirroring the reference's BinaryTree/MultipleTree
// (AriesEngineShell.cpp:19-95) closely enough to apply the same rewrite.
// Cast is cast(expr, P, S): Decimal(P, S).cast(value) (decimal.cu:471-533),
// multiply up or truncate down to scale S.  Trunc is truncate(expr, p):
// Decimal::truncate(p) (decimal.cu:541-570) with the argument's type as its
// prec and frac; p (in ival) may be negative.  Both are function nodes: the
// rewrite treats them as operands and rewrites their argument on its own.
enum class NK { Column, Int, Dec, Calc, Cast, Trunc };
enum class Op { Add = 0, Sub = 1, Mul = 2, Div = 3, Mod = 4 };

struct Expr {
    NK kind = NK::Column;
    Op op = Op::Add;
    int col = -1;
    int64_t ival = 0;
    HostDec dval;
    int P = 0, S = 0;       // value_type.DataType.{Precision,Scale}
    bool reverse = false;    // pending negation (value_reverse)
    std::vector<std::shared_ptr<Expr>> ch;
};
using ExprP = std::shared_ptr<Expr>;

// Parse infix text: identifiers (column names), integer and decimal
// literals, + - * / %, parentheses, unary minus.
ExprP parse_expr(const std::string& text, const std::vector<std::string>& names,
                 const std::vector<ColumnType>& types, Mode mode, std::string* err);

// Frontend result types (CommonBiaodashi.cpp:3338-3377 with
// Calc*TargetPrecision, decimal.cu:1353-1358,1568-1572,1783-1786,1997-2000).
void infer_frontend_types(Expr& e, Mode mode);

// Type of truncate(x, p) for x of type (P, S): the prec and frac that
// Decimal::truncate leaves (decimal.cu:541-570).
void trunc_type(int P, int S, int p, int* outP, int* outS);

// The reference's AlignOptimize (AriesEngineShell.cpp:646-685) applied to
// one calculation tree.  Returns the rewritten tree whose nodes carry the
// per-node (P,S) of ListToBinTree (AriesEngineShell.cpp:90-144).
ExprP align_optimize(const ExprP& e, Mode mode);

// Typed evaluation plan
enum class PK { Col, Const, Neg, Add, Sub, Mul, Div, Cast, Mod, Trunc };
struct PNode {
    PK kind;
    int P = 0, S = 0;
    int tp = 0;              // Trunc: the truncate argument p
    int col = -1;
    HostDec k;               // Const
    int a = -1, b = -1;      // children
    int bits = 0;            // magnitude bound: value < 2^bits (faithful mode: <= 1024, values wrap)
    int ubits = 0;           // the same bound without the 32-limb wrap
    int W = 1;               // limbs
    int koff = -1;           // Const: word offset of {sign, W limbs} in the plan's constant block
};
struct Plan {
    std::vector<PNode> n;
    int root = -1;
    Mode mode = Mode::Faithful;
    int outP = 0, outS = 0, outB = 0;   // materialized result type and compact bytes
    // Constants are kernel data, as the reference's constValues
    // (AriesExprCalcNode.cpp:346-349): the generated source reads constant
    // k's sign and limbs from the launch's constant array at kbase + koff,
    // so plans that differ only in constant values of the same widths share
    // one source text (one hipRTC compile).
    int kwords = 0;                     // words of this plan's constant block
    int kbase = 0;                      // offset of the block in the launch's constant array
    std::vector<uint32_t> const_data() const;   // the block: per constant, sign then W limbs
    // canonical plan text (see oracle's od_plan_parse); with values = false
    // constants print as their slot and scale ("k@3:2"), the source's shape
    std::string text(bool values = true) const;
};

int compact_bytes(int P);   // __PREC_BYTE_ARRAY (decimal.cu:92), closed form past 299

// Build a plan from a rewritten tree.  Column types give the operand
// layout.  Returns false with *err set when the tree uses something the
// HIP path does not implement.
bool build_plan(const ExprP& rewritten, const std::vector<ColumnType>& cols, Mode requested, Plan* out,
                std::string* err);

// parse -> AlignOptimize -> typed plan.  mode: 0 = pick faithful when the
// plan stays inside the reference's envelope, else extended; 1 = faithful
// (an error outside the envelope); 2 = extended.
bool plan_expression(const std::string& text, const std::vector<std::string>& names,
                     const std::vector<ColumnType>& types, int mode, Plan* out, std::string* err);

// Classification: true when the plan stays inside the reference's defined
// envelope (every shift <= 38, scales <= 255, precisions <= 299, divisions
// finish in one chunk), i.e. the reference-faithful contract applies.
bool plan_is_faithful_envelope(const Plan& p, const std::vector<ColumnType>& cols);

// ---- code generation ----
struct GenOptions {
    int block = 256;
    // grouped aggregation: slot copies per wave for the LDS atomic
    // accumulators (1 = one set shared by the 64 lanes, 64 = per lane)
    int slot_split = 4;
    // aggregation row loops take a lane's four consecutive rows per step
    // ("row quads") when the columns they load are narrow enough
    bool row_quads = true;
    // block-column inputs (IterArgs: blocks, block start rows, optional row
    // indices with -1 for NULL) instead of one contiguous array per column
    bool blocks = false;
};

// Row quads (aggregation row loops): a lane owns rows [4q, 4q + 4) and
// loads each narrow column as the 4 B whole dwords those rows occupy.  Used
// when every column the row loop loads inline is at most kQuadMaxBytes wide
// and their quads together (with the filter and key quads) take at most
// kQuadMaxWords registers.  The calc kernel keeps one row per lane: its
// rows of 8 and 16 bytes are already whole-dword loads and stores, and
// output quads measured slower there.
constexpr int kQuadMaxBytes = 16;
constexpr int kQuadMaxWords = 40;

// one-pass aggregation: threads per block of the final kernel, and the
// words of state the main kernel accumulates before it (overflow flag,
// most slots used, 2048-word key bitmap, 2 words padding the cleared span
// to a multiple of 16 bytes), which the host clears per run
constexpr int kAggFinalThreads = 512;
constexpr int kAggOnepassStatWords = 2 + 2048 + 2;

// The fused aggregation launch keeps every lane at <= 2^kAggRowsPerLaneLog2
// rows; the generated accumulator widths rely on it.
constexpr int kAggRowsPerLaneLog2 = 14;

// algorithmic 32 x 32 -> 64 limb products per row of a plan
int64_t plan_limb_products(const Plan& p);

// true when the calc kernel of this plan takes one row per thread with no
// row loop (the launch needs ceil(rows / block) blocks)
bool calc_one_row(const Plan& p);
// threads (and rows) per block of the calc kernel of a plan
int calc_block(const Plan& p);

// calc kernel: one compact output row per input row (the reference's
// calc_expr_<id>, AriesExprCalcNode.cpp:1285-1318)
std::string gen_calc_kernel(const Plan& p, const std::vector<ColumnType>& cols, const std::string& name,
                            const GenOptions& o);

// fused aggregation kernel for several plans over the same rows:
//   optional int32 filter column (value <= / < / >= / > / == / != constant),
//   optional group id from up to two 1-byte key columns through a LUT,
//   per plan a SUM accumulator per group, plus a row count per group.
struct AggSpec {
    std::vector<Plan> plans;
    int filter_col = -1;      // index into the int32 column list
    int filter_op = 0;        // 0 <=, 1 <, 2 >=, 3 >, 4 ==, 5 !=
    int64_t filter_val = 0;
    int nkeys = 0;            // 0, 1 or 2 key byte columns
    int ngroups = 1;          // groups addressed by the LUT (max)
    // one-pass form: every block numbers the keys it meets in an LDS hash
    // table of at most `maxslots` entries; the final kernel merges the
    // per-block tables by key.  Needs nkeys > 0 and 64-bit accumulators.
    bool onepass = false;
    int maxslots = 16;
    // AVG(expr_k) = SUM / COUNT for every bit k set, written by the final
    // kernel of the one-pass form (DivisionInt64 rule, faithful or exact)
    uint32_t avg_mask = 0;
    bool avg_faithful = true;
};

// true when the one-pass grouped form applies to this aggregation
bool agg_onepass_ok(const AggSpec& s, const GenOptions& o);
// words per block record of the one-pass form: MAXS entries of
// {key+1, NP x AW2 sums, 2 count words}
int agg_onepass_record_words(const AggSpec& s);
// Returns an empty string when the accumulators cannot fit in LDS;
// *block_out receives the block size the kernels were generated for.
std::string gen_agg_kernels(const AggSpec& s, const std::vector<ColumnType>& cols, const std::string& name,
                            const GenOptions& o, int* acc_words, int* block_out);
// limbs of one block partial of the fused aggregation
int agg_acc_words(const AggSpec& s);
// rows a lane takes per step of the aggregation row loop (1 or 4)
int agg_rows_per_step(const AggSpec& s, const std::vector<ColumnType>& cols, const GenOptions& o);

// ---- the precompiled interpreter (vm.h) ----
// Register program for one plan: the plan tree is expanded (shared nodes
// are recomputed, so no value stays live longer than in the tree), operands
// are ordered by register need (Sethi-Ullman) and every register holds W
// limbs, W the smallest class that holds every node's exact bound.
struct VmProgram {
    std::vector<updec_vm::Ins> ins;
    std::vector<uint32_t> consts;   // W + 1 words per constant: sign, limbs
    int wclass = 0, W = 0, nregs = 0, root = 0;
    int outB = 0, outP = 0, outS = 0;
    bool faithful = false;
    std::string text() const;       // one line per instruction
};
bool build_vm_program(const Plan& p, const std::vector<ColumnType>& cols, VmProgram* out, std::string* err);
// multi-lane calc kernel `name` of a program of 64 or more limbs, L limbs per lane
std::string gen_calc_ml(const VmProgram& vp, const std::vector<ColumnType>& cols, int L, const std::string& name);

}  // namespace updec
