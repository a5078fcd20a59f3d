// vm.h -- program format of the precompiled decimal VM, shared by the host
// program builder (planner.cpp) and the device interpreter (vm.hip).
//
// The per-query JIT path (hipRTC) specializes every limb count and scale at
// compile time, which costs 0.3-5 s of compile per new expression.  A
// serving loop that sees a new random expression every few milliseconds
// cannot wait for that, so the same plans also run on an interpreter that
// is compiled once with the library: one kernel per register width class
// W (2, 4, 8, 16, 32 limbs), evaluating a short register program per row.
// Scales become run-time shift counts; everything else is the same limb
// code as the JIT kernels, so results are identical.
#ifndef UPDEC_VM_H
#define UPDEC_VM_H
#include <stdint.h>

namespace updec_vm {

enum Op : uint8_t {
    OP_MOV = 0,   // dst = a
    OP_NEG = 1,   // dst = -a (sign flip, also of zero)
    OP_ADD = 2,   // dst = a + b   (operands pre-scaled by their shift)
    OP_SUB = 3,   // dst = a - b
    OP_MUL = 4,   // dst = a * b
    OP_DIV = 5,   // dst = a / b with S = arg (= scale(b) + 4)
    OP_UP = 6,    // dst = a * 10^arg (cast to a larger scale)
    OP_DOWN = 7,  // dst = floor(|a| / 10^arg), sign kept (cast to a smaller scale)
};

// K_LAST: the previous instruction's result, kept in VGPRs (the builder
// forwards a value that only the next instruction reads and then does not
// store it: dst = kNoStore; root = kNoStore reads the last result).
enum Kind : uint8_t { K_REG = 0, K_COL = 1, K_CONST = 2, K_LAST = 3 };
constexpr uint8_t kNoStore = 0xff;

constexpr int kMaxRegs = 8;
constexpr int kMaxCols = 8;
constexpr int kMaxIns = 512;
// W = 2 .. 32: one value per lane; W = 64 .. 512: one value per group of
// lanes (dec_mlane.h), L limbs per lane, for extended-mode programs whose
// exact intermediates need more than 32 limbs
constexpr int kNumClasses = 9;
constexpr int kFirstMultiLane = 5;
constexpr int kClassW[kNumClasses] = {2, 4, 8, 16, 32, 64, 128, 256, 512};
constexpr int kMaxAccWords = 514;               // accumulator limbs of the widest class (W + 2)
// registers per class: the register file is held in LDS, so the wide
// one-value-per-lane classes get fewer registers to keep four blocks
// resident per CU; a multi-lane register is L + 1 words per lane
constexpr int kClassRegs[kNumClasses] = {8, 8, 8, 6, 4, 8, 8, 8, 8};

struct Operand {
    uint8_t kind;    // Kind
    uint8_t idx;     // register, column or constant index
    uint16_t shift;  // multiply by 10^shift before use (ADD/SUB alignment)
};

struct Ins {
    uint8_t op;
    uint8_t dst;
    uint16_t arg;
    Operand a, b;
};

// One query of a batch.  Constants are stored at the query's class width as
// W + 1 words {sign, limbs...} starting at const_off.
struct Query {
    const uint8_t* col[kMaxCols];
    uint32_t colbytes[kMaxCols];
    uint64_t rows;
    uint32_t ins_off, nins;
    uint32_t const_off;
    uint32_t root;            // register holding the result
    uint32_t outB;            // compact bytes of the materialized result (SUM input)
    uint32_t outP, outS;      // image type
    uint32_t faithful;        // DivOrMod path quirks
    uint32_t wclass;          // index into kClassW
    uint32_t item0, nitems;   // this query's work items (contiguous)
    uint8_t* image;           // 133-byte nullable_type<Decimal> of SUM
    int64_t* count;           // COUNT(*)
};

// A contiguous row range of one query.  Items of one width class are
// contiguous in the batch's item array.
struct Item {
    uint32_t query;
    uint32_t pad;
    uint64_t row0, row1;
};

// Device-resident batch header: per class item ranges and the work counters
// the persistent kernels take items from.
struct Header {
    uint32_t nqueries;
    uint32_t class_item0[kNumClasses];
    uint32_t class_nitems[kNumClasses];
    uint32_t counter[kNumClasses];
};

}  // namespace updec_vm
#endif
