// vm_device.h -- the interpreter's per-row evaluation: one register
// program over values of W limbs (see vm.h for the program format).
// Included by vm.hip; header-only so a host build of the same code can
// check it against the oracle on the CPU.
#ifndef UPDEC_VM_DEVICE_H
#define UPDEC_VM_DEVICE_H
#include "dec_device.h"
#include "vm.h"

namespace updec_vm {

using namespace updec;

template <int W> UPD_DEV Dec<W> reg_get(const Dec<W> (&R)[kMaxRegs], u32 i)
{
    Dec<W> r = R[0];
#pragma unroll
    for (int k = 1; k < kMaxRegs; k++)
        if (i == (u32)k) r = R[k];
    return r;
}

template <int W> UPD_DEV void reg_set(Dec<W> (&R)[kMaxRegs], u32 i, const Dec<W>& v)
{
#pragma unroll
    for (int k = 0; k < kMaxRegs; k++)
        if (i == (u32)k) R[k] = v;
}

// operand word: kind | idx << 8 | shift << 16
template <int W>
UPD_DEV Dec<W> fetch(u32 ow, const Query* q, const u32* cst, const Dec<W> (&R)[kMaxRegs], usize row)
{
    const u32 kind = ow & 0xffu, idx = (ow >> 8) & 0xffu, shift = ow >> 16;
    Dec<W> v;
    if (kind == K_REG) {
        v = reg_get<W>(R, idx);
    } else if (kind == K_COL) {
        const u32 B = q->colbytes[idx];
        v = load_compact_rt<W>(q->col[idx], row, B, (usize)q->rows * B);
    } else {
        const u32* k = cst + idx * (u32)(W + 1);
        v.neg = k[0];
#pragma unroll
        for (int w = 0; w < W; w++) v.m.w[w] = k[1 + w];
    }
    if (shift) v.m = mul_p10_rt<W>(v.m, shift);
    return v;
}

template <int W> UPD_DEV Dec<W> vm_eval(const Query* q, const u32* prog, const u32* cst, usize row)
{
    Dec<W> R[kMaxRegs];
#pragma unroll
    for (int k = 0; k < kMaxRegs; k++) {
        R[k].m = mzero<W>();
        R[k].neg = 0u;
    }
    const u32 n = q->nins;
    const bool faith = q->faithful != 0u;
    for (u32 pc = 0; pc < n; pc++) {
        const u32 w0 = prog[3 * pc], w1 = prog[3 * pc + 1], w2 = prog[3 * pc + 2];
        const u32 op = w0 & 0xffu, dst = (w0 >> 8) & 0xffu, arg = w0 >> 16;
        const Dec<W> a = fetch<W>(w1, q, cst, R, row);
        Dec<W> r;
        if (op == OP_MOV) {
            r = a;
        } else if (op == OP_NEG) {
            r = dneg<W>(a);
        } else if (op == OP_UP) {
            r.m = mul_p10_rt<W>(a.m, arg);
            r.neg = a.neg;
        } else if (op == OP_DOWN) {
            r.m = div_p10_rt<W>(a.m, arg);
            r.neg = a.neg;
        } else {
            const Dec<W> b = fetch<W>(w2, q, cst, R, row);
            if (op == OP_ADD) r = dadd<W, W, W>(a, b);
            else if (op == OP_SUB) r = dsub<W, W, W>(a, b);
            else if (op == OP_MUL) r = dmul<W, W, W>(a, b);
            else r = ddiv_rt<W>(a, b, arg, faith);
        }
        reg_set<W>(R, dst, r);
    }
    return reg_get<W>(R, q->root);
}

}  // namespace updec_vm
#endif
