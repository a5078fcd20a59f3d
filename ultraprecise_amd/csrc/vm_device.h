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

// The register file lives in LDS, laid out [register][limb][lane] so the
// lanes of a wave touch consecutive dwords (no bank conflicts).  In VGPRs
// it would need a run-time index, which the compiler turns into scratch
// memory.  `base` points at this lane's column of the block's array; STRIDE
// is the block size (lanes between consecutive limbs).
template <int W, int STRIDE> struct RegFile {
    u32* base;

    UPD_DEV Dec<W> get(u32 i) const
    {
        const u32* p = base + i * (u32)((W + 1) * STRIDE);
        Dec<W> r;
#pragma unroll
        for (int w = 0; w < W; w++) r.m.w[w] = p[w * STRIDE];
        r.neg = p[W * STRIDE];
        return r;
    }
    UPD_DEV void set(u32 i, const Dec<W>& v)
    {
        u32* p = base + i * (u32)((W + 1) * STRIDE);
#pragma unroll
        for (int w = 0; w < W; w++) p[w * STRIDE] = v.m.w[w];
        p[W * STRIDE] = v.neg;
    }
};

// operand word: kind | idx << 8 | shift << 16
template <int W, class RF>
UPD_DEV Dec<W> fetch(u32 ow, const Query* q, const u32* cst, const RF& R, const Dec<W>& last, usize row)
{
    const u32 kind = ow & 0xffu, idx = (ow >> 8) & 0xffu, shift = ow >> 16;
    Dec<W> v;
    if (kind == K_LAST) {
        v = last;
    } else if (kind == K_REG) {
        v = R.get(idx);
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

// evaluate the query's program for one row; R: the lane's register file
template <int W, class RF> UPD_DEV Dec<W> vm_eval(const Query* q, const u32* prog, const u32* cst, usize row, RF& R)
{
    const u32 n = q->nins;
    const bool faith = q->faithful != 0u;
    Dec<W> last;
    last.m = mzero<W>();
    last.neg = 0u;
    for (u32 pc = 0; pc < n; pc++) {
        const u32 w0 = prog[3 * pc], w1 = prog[3 * pc + 1], w2 = prog[3 * pc + 2];
        const u32 op = w0 & 0xffu, dst = (w0 >> 8) & 0xffu, arg = w0 >> 16;
        const Dec<W> a = fetch<W, RF>(w1, q, cst, R, last, row);
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
            const Dec<W> b = fetch<W, RF>(w2, q, cst, R, last, row);
            if (op == OP_ADD) r = dadd<W, W, W>(a, b);
            else if (op == OP_SUB) r = dsub<W, W, W>(a, b);
            else if (op == OP_MUL) r = dmul<W, W, W>(a, b);
            else r = ddiv_rt<W>(a, b, arg, faith);
        }
        if (dst != kNoStore) R.set(dst, r);
        last = r;
    }
    return q->root == kNoStore ? last : R.get(q->root);
}

}  // namespace updec_vm
#endif
