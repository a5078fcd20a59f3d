// vm_host.cpp -- TEST INFRASTRUCTURE: the interpreter's per-row code
// (vm_device.h over dec_device.h) built for the host CPU, so the CPU test
// suite can check interpreter programs row by row against the oracle
// without a GPU.  The few GPU builtins the header uses get host stand-ins
// here; the product only ever runs this code on the GPU.
#define __HIPCC_RTC__ 1
#define __device__
#define __forceinline__ inline
#define __global__
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <string>
static inline unsigned upd_alignbyte(unsigned hi, unsigned lo, unsigned s) { return (unsigned)((((unsigned long long)hi << 32) | lo) >> (8 * (s & 3))); }
#define __builtin_amdgcn_alignbyte(a, b, c) upd_alignbyte((a), (b), (c))
#define __builtin_nontemporal_load(p) (*(p))
// wave_reduce is not used on the host (rows are summed sequentially); the
// declaration only satisfies name lookup in the header's template
static inline int __shfl_xor(int v, int, int) { return v; }
#include "dec_device.h"
#include "vm_device.h"
#include "planner.h"
using namespace updec;
using namespace updec_vm;

template <int W> void run(const VmProgram& pr, Query& Q, std::vector<uint32_t>& prog, int64_t rows, uint8_t* img)
{
    constexpr int AW = W + 2;
    Mag<AW> acc = mzero<AW>();
    std::vector<u32> regs(kMaxRegs * (W + 1), 0u);
    RegFile<W, 1> R{regs.data()};
    for (int64_t r = 0; r < rows; r++) {
        Dec<W> v = vm_eval<W>(&Q, prog.data(), pr.consts.data(), (usize)r, R);
        acc_add<AW, W>(acc, stored_rt<W>(v, Q.outB));
    }
    write_sum_image<AW>(img, acc, Q.outP & 0xff, Q.outS & 0xff, 0);
}

extern "C" int vmh_sum(const char* text, const char** names, const int* P, const int* S, const int* bits, int ncols,
                       int mode, const uint8_t* const* data, int64_t rows, uint8_t* img, char* err, int cap)
{
    std::vector<std::string> nm;
    std::vector<ColumnType> ty(ncols);
    for (int i = 0; i < ncols; i++) { nm.push_back(names[i]); ty[i].P = P[i]; ty[i].S = S[i]; ty[i].bits = bits[i]; }
    Plan plan;
    std::string e;
    if (!plan_expression(text, nm, ty, mode, &plan, &e)) { snprintf(err, cap, "%s", e.c_str()); return 1; }
    VmProgram pr;
    if (!build_vm_program(plan, ty, &pr, &e)) { snprintf(err, cap, "%s", e.c_str()); return 2; }
    std::vector<uint32_t> prog;
    for (auto& I : pr.ins) {
        prog.push_back((uint32_t)I.op | ((uint32_t)I.dst << 8) | ((uint32_t)I.arg << 16));
        prog.push_back((uint32_t)I.a.kind | ((uint32_t)I.a.idx << 8) | ((uint32_t)I.a.shift << 16));
        prog.push_back((uint32_t)I.b.kind | ((uint32_t)I.b.idx << 8) | ((uint32_t)I.b.shift << 16));
    }
    Query Q{};
    for (int c = 0; c < ncols; c++) { Q.col[c] = data[c]; Q.colbytes[c] = compact_bytes(P[c]); }
    Q.rows = rows; Q.nins = pr.ins.size(); Q.root = pr.root; Q.outB = pr.outB; Q.outP = pr.outP; Q.outS = pr.outS;
    Q.faithful = pr.faithful; Q.wclass = pr.wclass;
    switch (pr.W) {
    case 2: run<2>(pr, Q, prog, rows, img); break;
    case 4: run<4>(pr, Q, prog, rows, img); break;
    case 8: run<8>(pr, Q, prog, rows, img); break;
    case 16: run<16>(pr, Q, prog, rows, img); break;
    case 32: run<32>(pr, Q, prog, rows, img); break;
    // the multi-lane classes' programs, evaluated one value per thread at
    // the same width: the reference their group evaluation must match
    case 64: run<64>(pr, Q, prog, rows, img); break;
    case 128: run<128>(pr, Q, prog, rows, img); break;
    default: run<256>(pr, Q, prog, rows, img); break;
    }
    return 0;
}
