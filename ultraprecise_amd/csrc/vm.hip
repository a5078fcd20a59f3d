// vm.hip -- the precompiled decimal interpreter and its batch launcher
// (include/updec.h, "interpreter and batches").
//
// Kernels: one persistent k_vm_sum<W> per register width class.  Its
// blocks take work items (a row range of one query) from a per-class
// counter, evaluate the query's register program for every row, fold the
// materialized result into a per-lane two's complement accumulator and
// write one partial per item; k_vm_final adds a query's partials in item
// order and writes the SUM image and COUNT.  A batch of many queries is one
// set of launches (captured once into a HIP graph, the classes on parallel
// branches), a single query is the same launches with one query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/updec.h"
#include "api_internal.h"
#include "dec_device.h"
#include "dec_mlane.h"
#include "planner.h"
#include "vm.h"
#include "vm_device.h"

using namespace updec;
using namespace updec_vm;

namespace {

#define VMCHK(expr)                                                                                          \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return api_fail(UPDEC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock[kNumClasses] = {256, 256, 128, 64, 64, 128, 128, 128, 128};

template <int W> constexpr int class_of()
{
    return W == 2 ? 0 : W == 4 ? 1 : W == 8 ? 2 : W == 16 ? 3 : W == 32 ? 4 : W == 64 ? 5 : W == 128 ? 6 : W == 256 ? 7 : 8;
}

template <int W, int T>
__global__ void __launch_bounds__(T)
    k_vm_sum(Header* h, const Query* qs, const u32* prog, const u32* cst, const Item* items, u32* partials)
{
    constexpr int CLS = class_of<W>();
    constexpr int AW = W + 2;
    constexpr int NWAVE = T / 64;
    __shared__ u32 s_item;
    __shared__ u32 red[NWAVE * AW];
    __shared__ u32 regs[kClassRegs[CLS] * (W + 1) * T];
    const u32 tid = threadIdx.x;
    RegFile<W, T> R{regs + tid};
    const u32 n = h->class_nitems[CLS];
    const u32 i0 = h->class_item0[CLS];
    for (;;) {
        // The barrier first: the thread-0 work that ends one iteration (the
        // partial) and the thread-0 atomic that starts the next must stay
        // separated by a barrier.  Without one the compiler may merge the
        // two thread-0 regions into a divergent exit of the item loop, and
        // the other lanes of wave 0 re-enter the loop (and its barriers)
        // ahead of thread 0, re-reading an item number only thread 0 can
        // advance.  It also orders the reads of s_item and red of the last
        // iteration before they are written again.
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(&h->counter[CLS], 1u);
        __syncthreads();
        const u32 it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n) break;
        const Item I = items[i0 + it];
        const Query* q = qs + I.query;
        const u32* qp = prog + 3u * q->ins_off;
        const u32* qc = cst + q->const_off;
        const u32 ob = q->outB;
        Mag<AW> acc = mzero<AW>();
        for (usize row = I.row0 + tid; row < I.row1; row += T) {
            const Dec<W> v = vm_eval<W>(q, qp, qc, row, R);
            acc_add<AW, W>(acc, stored_rt<W>(v, ob));
        }
        wave_reduce<AW>(acc);
        if ((tid & 63u) == 0) {
#pragma unroll
            for (int w = 0; w < AW; w++) red[(tid >> 6) * AW + w] = acc.w[w];
        }
        __syncthreads();
        if (tid == 0) {
            Mag<AW> t = mzero<AW>();
            for (int wv = 0; wv < NWAVE; wv++) {
                Mag<AW> o;
#pragma unroll
                for (int w = 0; w < AW; w++) o.w[w] = red[wv * AW + w];
                add_to<AW>(t, t, o);
            }
            u32* dst = partials + (usize)(i0 + it) * kMaxAccWords;
#pragma unroll
            for (int w = 0; w < AW; w++) dst[w] = t.w[w];
        }
    }
}

// ---------------------------------------------------------------- multi-lane classes

// the lane's slice of register i: L limbs and the sign, in LDS at
// [register][word][lane] like RegFile
template <int L, int STRIDE> struct MLRegFile {
    u32* base;
    UPD_DEV ml::MDec<L> get(u32 i) const
    {
        const u32* p = base + i * (u32)((L + 1) * STRIDE);
        ml::MDec<L> r;
#pragma unroll
        for (int w = 0; w < L; w++) r.m.w[w] = p[w * STRIDE];
        r.neg = p[L * STRIDE];
        return r;
    }
    UPD_DEV void set(u32 i, const ml::MDec<L>& v)
    {
        u32* p = base + i * (u32)((L + 1) * STRIDE);
#pragma unroll
        for (int w = 0; w < L; w++) p[w * STRIDE] = v.m.w[w];
        p[L * STRIDE] = v.neg;
    }
};

template <int L, int TPI, class RF>
UPD_DEV ml::MDec<L> fetch_ml(u32 ow, const Query* q, const u32* cst, const RF& R, const ml::MDec<L>& last, usize row,
                             const ml::Grp<TPI>& g)
{
    constexpr int W = L * TPI;
    const u32 kind = ow & 0xffu, idx = (ow >> 8) & 0xffu, shift = ow >> 16;
    ml::MDec<L> v;
    if (kind == K_LAST) {
        v = last;
    } else if (kind == K_REG) {
        v = R.get(idx);
    } else if (kind == K_COL) {
        const u32 B = q->colbytes[idx];
        v = ml::load_compact<L, TPI>(q->col[idx], row, B, (usize)q->rows * B, g);
    } else {
        // constants at the class width: {sign, W limbs}
        const u32* k = cst + idx * (u32)(W + 1);
        v.neg = k[0];
#pragma unroll
        for (int w = 0; w < L; w++) v.m.w[w] = k[1 + g.rank * L + w];
    }
    if (shift) ml::mul_p10<L, TPI>(v.m, shift, g);
    return v;
}

// the register program of one row, evaluated by the row's group of TPI lanes
// (extended mode: multi-lane classes hold no faithful programs)
template <int L, int TPI, class RF>
UPD_DEV ml::MDec<L> vm_eval_ml(const Query* q, const u32* prog, const u32* cst, usize row, RF& R, const ml::Grp<TPI>& g)
{
    const u32 n = q->nins;
    ml::MDec<L> last;
    last.m = ml::mzero<L>();
    last.neg = 0u;
    for (u32 pc = 0; pc < n; pc++) {
        const u32 w0 = prog[3 * pc], w1 = prog[3 * pc + 1], w2 = prog[3 * pc + 2];
        const u32 op = w0 & 0xffu, dst = (w0 >> 8) & 0xffu, arg = w0 >> 16;
        const ml::MDec<L> a = fetch_ml<L, TPI, RF>(w1, q, cst, R, last, row, g);
        ml::MDec<L> r;
        if (op == OP_MOV) {
            r = a;
        } else if (op == OP_NEG) {
            r = a;
            r.neg ^= 1u;
        } else if (op == OP_UP) {
            r = a;
            ml::mul_p10<L, TPI>(r.m, arg, g);
        } else if (op == OP_DOWN) {
            r = a;
            ml::div_p10<L, TPI>(r.m, arg, g);
        } else {
            const ml::MDec<L> b = fetch_ml<L, TPI, RF>(w2, q, cst, R, last, row, g);
            if (op == OP_ADD) {
                r = ml::add<L, TPI>(a, b, g);
            } else if (op == OP_SUB) {
                r = ml::sub<L, TPI>(a, b, g);
            } else if (op == OP_MUL) {
                r.m = ml::mul<L, TPI>(a.m, b.m, g);
                r.neg = a.neg ^ b.neg;
            } else {
                r = ml::div<L, TPI>(a, b, arg, g);
            }
        }
        if (dst != kNoStore) R.set(dst, r);
        last = r;
    }
    return q->root == kNoStore ? last : R.get(q->root);
}

// Persistent kernel of a multi-lane class: blocks take work items as
// k_vm_sum does; within an item the block's T / TPI groups take rows in
// turn, each group evaluating its row's program and folding the stored
// value into its two's complement accumulator; the block then adds its
// groups' accumulators column by column and writes the item's partial.
template <int L, int TPI, int T>
__global__ void __launch_bounds__(T)
    k_vm_sum_ml(Header* h, const Query* qs, const u32* prog, const u32* cst, const Item* items, u32* partials)
{
    constexpr int W = L * TPI;
    constexpr int CLS = class_of<W>();
    constexpr int AW = W + 2;
    constexpr int NG = T / TPI;
    __shared__ u32 s_item;
    __shared__ u32 regs[kClassRegs[CLS] * (L + 1) * T];
    __shared__ u32 gacc[NG * AW];
    __shared__ u64 csum[AW];
    const u32 tid = threadIdx.x;
    const ml::Grp<TPI> g;
    const u32 grp = tid / (u32)TPI;
    MLRegFile<L, T> R{regs + tid};
    const u32 n = h->class_nitems[CLS];
    const u32 i0 = h->class_item0[CLS];
    for (;;) {
        // as in k_vm_sum: the barrier keeps thread 0's item fetch apart from
        // its partial write of the previous iteration
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(&h->counter[CLS], 1u);
        __syncthreads();
        const u32 it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n) break;
        const Item I = items[i0 + it];
        const Query* q = qs + I.query;
        const u32* qp = prog + 3u * q->ins_off;
        const u32* qc = cst + q->const_off;
        const u32 ob = q->outB;
        ml::Acc<L> acc = ml::acc_zero<L>();
        for (usize row = I.row0 + grp; row < I.row1; row += NG) {
            const ml::MDec<L> v = vm_eval_ml<L, TPI>(q, qp, qc, row, R, g);
            ml::acc_add<L, TPI>(acc, ml::stored<L, TPI>(v, ob, g), g);
        }
#pragma unroll
        for (int w = 0; w < L; w++) gacc[grp * AW + g.rank * L + w] = acc.m.w[w];
        if (g.rank == TPI - 1) {
            gacc[grp * AW + W] = acc.x0;
            gacc[grp * AW + W + 1] = acc.x1;
        }
        __syncthreads();
        for (u32 w = tid; w < (u32)AW; w += T) {
            u64 s = 0;
            for (u32 k = 0; k < (u32)NG; k++) s += gacc[k * AW + w];
            csum[w] = s;
        }
        __syncthreads();
        if (tid == 0) {
            u32* dst = partials + (usize)(i0 + it) * kMaxAccWords;
            u64 c = 0;
            for (int w = 0; w < AW; w++) {
                const u64 t = csum[w] + c;
                dst[w] = (u32)t;
                c = t >> 32;
            }
        }
    }
}

// SUM image and COUNT of a query of a multi-lane class: its items'
// partials added column by column (64-bit words, no carries), then one
// carry pass; the image keeps the low 32 limbs of the magnitude
template <int AW> __device__ void final_query_wide(const Query* q, const u32* partials)
{
    __shared__ u64 cs[AW];
    __shared__ u32 mag[AW];
    for (u32 w = threadIdx.x; w < (u32)AW; w += 64) {
        u64 s = 0;
        for (u32 k = 0; k < q->nitems; k++) s += partials[(usize)(q->item0 + k) * kMaxAccWords + w];
        cs[w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 c = 0;
        for (int w = 0; w < AW; w++) {
            const u64 t = cs[w] + c;
            mag[w] = (u32)t;
            c = t >> 32;
        }
        const bool neg = (mag[AW - 1] >> 31) != 0;
        // magnitude of the two's complement value
        u64 cc = neg ? 1u : 0u;
        const u32 fm = neg ? 0xffffffffu : 0u;
        u8* o = q->image;
        o[0] = 1;
        o[1] = neg ? 1 : 0;
        o[2] = (u8)(q->outP & 0xffu);
        o[3] = (u8)(q->outS & 0xffu);
        o[4] = 0;
        for (int w = 0; w < 32; w++) {
            const u64 t = (u64)(mag[w] ^ fm) + cc;
            const u32 x = (u32)t;
            cc = t >> 32;
            o[5 + 4 * w + 0] = (u8)x;
            o[5 + 4 * w + 1] = (u8)(x >> 8);
            o[5 + 4 * w + 2] = (u8)(x >> 16);
            o[5 + 4 * w + 3] = (u8)(x >> 24);
        }
        *q->count = (int64_t)q->rows;
    }
}

template <int AW> __device__ void final_query(const Query* q, const u32* partials)
{
    Mag<AW> acc = mzero<AW>();
    for (u32 k = threadIdx.x; k < q->nitems; k += 64) {
        const u32* p = partials + (usize)(q->item0 + k) * kMaxAccWords;
        Mag<AW> o;
#pragma unroll
        for (int w = 0; w < AW; w++) o.w[w] = p[w];
        add_to<AW>(acc, acc, o);
    }
    wave_reduce<AW>(acc);
    if (threadIdx.x == 0) {
        write_sum_image<AW>(q->image, acc, q->outP & 0xffu, q->outS & 0xffu, 0u);
        *q->count = (int64_t)q->rows;
    }
}

__global__ void __launch_bounds__(64) k_vm_final(const Header* h, const Query* qs, const u32* partials)
{
    const u32 qi = blockIdx.x;
    if (qi >= h->nqueries) return;
    const Query* q = qs + qi;
    switch (q->wclass) {
    case 0: final_query<4>(q, partials); break;
    case 1: final_query<6>(q, partials); break;
    case 2: final_query<10>(q, partials); break;
    case 3: final_query<18>(q, partials); break;
    case 4: final_query<34>(q, partials); break;
    case 5: final_query_wide<66>(q, partials); break;
    case 6: final_query_wide<130>(q, partials); break;
    case 7: final_query_wide<258>(q, partials); break;
    default: final_query_wide<514>(q, partials); break;
    }
}

// ---------------------------------------------------------------- host

struct Pending {
    std::shared_ptr<const VmProgram> prog;
    const uint8_t* cols[kMaxCols];
    uint32_t colbytes[kMaxCols];
    int64_t rows;
    int64_t chunk;  // rows per work item of this query
    uint8_t* image;
    int64_t* count;
};

template <typename T> int host_alloc(T** p, size_t n)
{
    VMCHK(hipHostMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(1, n), hipHostMallocDefault));
    return UPDEC_OK;
}
template <typename T> int dev_alloc(T** p, size_t n)
{
    VMCHK(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(1, n)));
    return UPDEC_OK;
}

}  // namespace

struct updec_vm_s {
    std::shared_ptr<const VmProgram> prog;
    std::vector<uint32_t> colbytes;
};

struct updec_batch_s {
    int max_queries = 0, max_items = 0, max_prog_words = 0, max_const_words = 0;
    // defaults sized so one 10^7-row query alone makes >= 4 items per CU
    int64_t rows_per_item[kNumClasses] = {8192, 8192, 4096, 2048, 1024, 512, 256, 128, 64};
    int blocks_per_cu[kNumClasses] = {4, 4, 4, 4, 4, 4, 4, 4, 4};
    // lanes per value of the multi-lane classes: W / 4 (4 limbs per lane) or W / 8 (8 limbs)
    int lanes[kNumClasses] = {1, 1, 1, 1, 1, 16, 32, 64, 64};
    Header* h_hdr = nullptr;
    Query* h_q = nullptr;
    Item* h_items = nullptr;
    uint32_t* h_prog = nullptr;
    uint32_t* h_cst = nullptr;
    Header* d_hdr = nullptr;
    Query* d_q = nullptr;
    Item* d_items = nullptr;
    uint32_t* d_prog = nullptr;
    uint32_t* d_cst = nullptr;
    uint32_t* d_partials = nullptr;
    hipEvent_t staged = nullptr;  // the last launch's host-to-device copies are done
    bool staged_pending = false;
    hipStream_t side[kNumClasses] = {};
    hipEvent_t fork = nullptr, join[kNumClasses] = {};
    hipGraphExec_t graph = nullptr;
    std::vector<Pending> pending;
    // space the pending queries take at launch
    uint32_t pend_items = 0, pend_prog_words = 0, pend_const_words = 0;
    std::map<const VmProgram*, int> pend_programs;
    std::mutex mu;

    ~updec_batch_s()
    {
        if (staged_pending && staged) (void)hipEventSynchronize(staged);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (auto& s : side)
            if (s) (void)hipStreamDestroy(s);
        for (auto& e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (staged) (void)hipEventDestroy(staged);
        (void)hipHostFree(h_hdr);
        (void)hipHostFree(h_q);
        (void)hipHostFree(h_items);
        (void)hipHostFree(h_prog);
        (void)hipHostFree(h_cst);
        (void)hipFree(d_hdr);
        (void)hipFree(d_q);
        (void)hipFree(d_items);
        (void)hipFree(d_prog);
        (void)hipFree(d_cst);
        (void)hipFree(d_partials);
    }

    void clear_pending()
    {
        pending.clear();
        pend_items = pend_prog_words = pend_const_words = 0;
        pend_programs.clear();
    }

    int grid(int cls, uint32_t nitems) const
    {
        const int g = api_device_cus() * blocks_per_cu[cls];
        return std::max(1, std::min<int>(g, (int)nitems));
    }

    // launch the class kernels (on `s` or on the side streams) and the final kernel
    int launch_kernels(hipStream_t s, bool fork_join, const uint32_t* nitems_per_class)
    {
        auto launch_class = [&](int c, hipStream_t st) -> int {
            const int g = nitems_per_class ? grid(c, nitems_per_class[c]) : api_device_cus() * blocks_per_cu[c];
            switch (c) {
            case 0: hipLaunchKernelGGL((k_vm_sum<2, kBlock[0]>), dim3(g), dim3(kBlock[0]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials); break;
            case 1: hipLaunchKernelGGL((k_vm_sum<4, kBlock[1]>), dim3(g), dim3(kBlock[1]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials); break;
            case 2: hipLaunchKernelGGL((k_vm_sum<8, kBlock[2]>), dim3(g), dim3(kBlock[2]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials); break;
            case 3: hipLaunchKernelGGL((k_vm_sum<16, kBlock[3]>), dim3(g), dim3(kBlock[3]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials); break;
            case 4: hipLaunchKernelGGL((k_vm_sum<32, kBlock[4]>), dim3(g), dim3(kBlock[4]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials); break;
#define UPD_ML(Lx, TPIx, C) hipLaunchKernelGGL((k_vm_sum_ml<Lx, TPIx, kBlock[C]>), dim3(g), dim3(kBlock[C]), 0, st, d_hdr, d_q, d_prog, d_cst, d_items, d_partials)
            case 5: if (lanes[5] == 8) UPD_ML(8, 8, 5); else UPD_ML(4, 16, 5); break;
            case 6: if (lanes[6] == 16) UPD_ML(8, 16, 6); else UPD_ML(4, 32, 6); break;
            case 7: if (lanes[7] == 32) UPD_ML(8, 32, 7); else UPD_ML(4, 64, 7); break;
            default: UPD_ML(8, 64, 8); break;   // 512 limbs: a whole wave, 8 limbs per lane
#undef UPD_ML
            }
            VMCHK(hipGetLastError());
            return UPDEC_OK;
        };
        if (fork_join) {
            VMCHK(hipEventRecord(fork, s));
            for (int c = 0; c < kNumClasses; c++) {
                VMCHK(hipStreamWaitEvent(side[c], fork, 0));
                int rc = launch_class(c, side[c]);
                if (rc) return rc;
                VMCHK(hipEventRecord(join[c], side[c]));
                VMCHK(hipStreamWaitEvent(s, join[c], 0));
            }
        } else {
            for (int c = 0; c < kNumClasses; c++) {
                if (nitems_per_class && nitems_per_class[c] == 0) continue;
                int rc = launch_class(c, s);
                if (rc) return rc;
            }
        }
        const uint32_t nq = fork_join ? (uint32_t)max_queries : h_hdr->nqueries;
        if (nq) {
            hipLaunchKernelGGL(k_vm_final, dim3(nq), dim3(64), 0, s, d_hdr, d_q, d_partials);
            VMCHK(hipGetLastError());
        }
        return UPDEC_OK;
    }
};

// ---------------------------------------------------------------- C ABI

int updec_vm_compile(const char* text, const updec_column_desc* cols, int ncols, int mode, updec_vm_t* out)
{
    if (!text || !out || ncols < 0 || ncols > kMaxCols || (ncols && !cols)) return api_fail(UPDEC_EINVAL, "bad argument");
    std::vector<std::string> names;
    auto types = api_types(cols, ncols, &names);
    Plan plan;
    std::string err;
    if (!api_make_plan(text, names, types, mode, &plan, &err)) return api_fail(UPDEC_ECOMPILE, err);
    auto prog = std::make_shared<VmProgram>();
    if (!build_vm_program(plan, types, prog.get(), &err)) return api_fail(UPDEC_EUNSUPPORTED, err);
    std::unique_ptr<updec_vm_s> v(new updec_vm_s());
    v->prog = prog;
    for (auto& t : types) v->colbytes.push_back((uint32_t)compact_bytes(t.P));
    *out = v.release();
    return UPDEC_OK;
}

int updec_vm_info(updec_vm_t v, int32_t* width, int32_t* nins, int32_t* nregs, int32_t* precision, int32_t* scale,
                  int32_t* faithful)
{
    if (!v) return api_fail(UPDEC_EINVAL, "null program");
    const VmProgram& p = *v->prog;
    if (width) *width = p.W;
    if (nins) *nins = (int32_t)p.ins.size();
    if (nregs) *nregs = p.nregs;
    if (precision) *precision = p.outP;
    if (scale) *scale = p.outS;
    if (faithful) *faithful = p.faithful ? 1 : 0;
    return UPDEC_OK;
}

int updec_vm_listing(updec_vm_t v, char* buf, size_t cap, size_t* needed)
{
    if (!v) return api_fail(UPDEC_EINVAL, "null program");
    std::string s = v->prog->text();
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return UPDEC_OK;
}

void updec_vm_free(updec_vm_t v) { delete v; }

int updec_batch_create(int32_t max_queries, int32_t max_items, updec_batch_t* out)
{
    if (!out || max_queries <= 0 || max_items < max_queries) return api_fail(UPDEC_EINVAL, "bad argument");
    std::unique_ptr<updec_batch_s> b(new updec_batch_s());
    b->max_queries = max_queries;
    b->max_items = max_items;
    b->max_prog_words = max_queries * 3 * 64;
    b->max_const_words = max_queries * 8 * (512 + 1);
    int rc;
    if ((rc = host_alloc(&b->h_hdr, 1)) || (rc = host_alloc(&b->h_q, max_queries)) || (rc = host_alloc(&b->h_items, max_items)) ||
        (rc = host_alloc(&b->h_prog, b->max_prog_words)) || (rc = host_alloc(&b->h_cst, b->max_const_words)) ||
        (rc = dev_alloc(&b->d_hdr, 1)) || (rc = dev_alloc(&b->d_q, max_queries)) || (rc = dev_alloc(&b->d_items, max_items)) ||
        (rc = dev_alloc(&b->d_prog, b->max_prog_words)) || (rc = dev_alloc(&b->d_cst, b->max_const_words)) ||
        (rc = dev_alloc(&b->d_partials, (size_t)max_items * kMaxAccWords)))
        return rc;
    memset(b->h_hdr, 0, sizeof(Header));
    VMCHK(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    VMCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (int c = 0; c < kNumClasses; c++) {
        VMCHK(hipStreamCreateWithFlags(&b->side[c], hipStreamNonBlocking));
        VMCHK(hipEventCreateWithFlags(&b->join[c], hipEventDisableTiming));
    }
    *out = b.release();
    return UPDEC_OK;
}

int updec_batch_tuning(updec_batch_t b, int32_t width_class, int64_t rows_per_item, int32_t blocks_per_cu)
{
    if (!b || width_class < 0 || width_class >= kNumClasses || rows_per_item < 64 || blocks_per_cu < 1 ||
        blocks_per_cu > 64)
        return api_fail(UPDEC_EINVAL, "bad argument");
    std::lock_guard<std::mutex> g(b->mu);
    if (!b->pending.empty()) return api_fail(UPDEC_EINVAL, "tuning changes while queries are staged");
    b->rows_per_item[width_class] = rows_per_item;
    b->blocks_per_cu[width_class] = blocks_per_cu;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return UPDEC_OK;
}

int updec_batch_lanes(updec_batch_t b, int32_t width_class, int32_t lanes)
{
    if (!b || width_class < kFirstMultiLane || width_class >= kNumClasses)
        return api_fail(UPDEC_EINVAL, "lanes per value apply to the multi-lane classes (W = 64 .. 512)");
    const int W = kClassW[width_class];
    if ((lanes != W / 4 && lanes != W / 8) || lanes > 64)
        return api_fail(UPDEC_EINVAL, "lanes per value must be W / 4 or W / 8, at most 64");
    std::lock_guard<std::mutex> g(b->mu);
    if (!b->pending.empty()) return api_fail(UPDEC_EINVAL, "tuning changes while queries are staged");
    b->lanes[width_class] = lanes;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return UPDEC_OK;
}

int updec_batch_add(updec_batch_t b, updec_vm_t v, const uint8_t* const* d_cols, int64_t rows, uint8_t* d_image,
                    int64_t* d_count, int32_t* index_out)
{
    if (!b || !v || rows < 0 || !d_image || !d_count || (!d_cols && !v->colbytes.empty()))
        return api_fail(UPDEC_EINVAL, "bad argument");
    std::lock_guard<std::mutex> g(b->mu);
    const VmProgram& pr = *v->prog;
    // the class's tuned rows per item, raised (in steps of 64 rows) for a
    // query that would otherwise need more items than the whole batch holds
    int64_t ch = b->rows_per_item[pr.wclass];
    // a large query takes items of at least rows / (4 x the class's resident
    // blocks): four items per block keep the load balanced while the
    // per-item work (fetch, block reduction, partial write) stays small next
    // to the item's rows
    const int64_t slots = 4 * (int64_t)api_device_cus() * b->blocks_per_cu[pr.wclass];
    const int64_t big = (rows + slots - 1) / slots;
    if (big > ch) ch = (big + 63) / 64 * 64;
    if ((rows + ch - 1) / ch > b->max_items) ch = ((rows + b->max_items - 1) / b->max_items + 63) / 64 * 64;
    const uint32_t items = (uint32_t)std::max<int64_t>(1, (rows + ch - 1) / ch);
    const bool new_prog = b->pend_programs.find(&pr) == b->pend_programs.end();
    const uint32_t pw = new_prog ? 3u * (uint32_t)pr.ins.size() : 0u, cw = new_prog ? (uint32_t)pr.consts.size() : 0u;
    if ((int)b->pending.size() >= b->max_queries || b->pend_items + items > (uint32_t)b->max_items ||
        b->pend_prog_words + pw > (uint32_t)b->max_prog_words || b->pend_const_words + cw > (uint32_t)b->max_const_words) {
        if (b->pending.empty()) return api_fail(UPDEC_EINVAL, "query does not fit an empty batch");
        return api_fail(UPDEC_EFULL, "batch is full");
    }
    Pending p{};
    p.prog = v->prog;
    for (size_t c = 0; c < v->colbytes.size(); c++) {
        if (!d_cols[c] && rows > 0) return api_fail(UPDEC_EINVAL, "null column");
        p.cols[c] = d_cols[c];
        p.colbytes[c] = v->colbytes[c];
    }
    p.rows = rows;
    p.chunk = ch;
    p.image = d_image;
    p.count = d_count;
    b->pending.push_back(p);
    b->pend_items += items;
    b->pend_prog_words += pw;
    b->pend_const_words += cw;
    b->pend_programs[&pr] = 1;
    if (index_out) *index_out = (int32_t)b->pending.size() - 1;
    return UPDEC_OK;
}

int updec_batch_pending(updec_batch_t b, int32_t* n)
{
    if (!b || !n) return api_fail(UPDEC_EINVAL, "bad argument");
    std::lock_guard<std::mutex> g(b->mu);
    *n = (int32_t)b->pending.size();
    return UPDEC_OK;
}

int updec_batch_launch(updec_batch_t b, int use_graph, updec_stream_t stream)
{
    if (!b) return api_fail(UPDEC_EINVAL, "null batch");
    std::lock_guard<std::mutex> g(b->mu);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (b->pending.empty()) return UPDEC_OK;
    // the staging buffers are rewritten below: the previous copies must be done
    if (b->staged_pending) {
        VMCHK(hipEventSynchronize(b->staged));
        b->staged_pending = false;
    }
    // items per class, in class order; programs and constants deduplicated
    uint32_t nitems[kNumClasses] = {};
    for (const Pending& p : b->pending) {
        const int c = p.prog->wclass;
        nitems[c] += (uint32_t)std::max<int64_t>(1, (p.rows + p.chunk - 1) / p.chunk);
    }
    uint32_t total = 0, first[kNumClasses];
    for (int c = 0; c < kNumClasses; c++) {
        first[c] = total;
        total += nitems[c];
    }
    if (total > (uint32_t)b->max_items) {
        b->clear_pending();
        return api_fail(UPDEC_EINVAL, "batch needs more work items than it was created for");
    }
    std::map<const VmProgram*, std::pair<uint32_t, uint32_t>> placed;  // program -> (ins_off, const_off)
    uint32_t prog_words = 0, const_words = 0, fill[kNumClasses];
    memcpy(fill, first, sizeof fill);
    Header& H = *b->h_hdr;
    memset(&H, 0, sizeof H);
    H.nqueries = (uint32_t)b->pending.size();
    for (int c = 0; c < kNumClasses; c++) {
        H.class_item0[c] = first[c];
        H.class_nitems[c] = nitems[c];
    }
    for (size_t qi = 0; qi < b->pending.size(); qi++) {
        const Pending& p = b->pending[qi];
        const VmProgram& pr = *p.prog;
        auto it = placed.find(&pr);
        if (it == placed.end()) {
            const uint32_t pw = 3u * (uint32_t)pr.ins.size(), cw = (uint32_t)pr.consts.size();
            if (prog_words + pw > (uint32_t)b->max_prog_words || const_words + cw > (uint32_t)b->max_const_words) {
                b->clear_pending();
                return api_fail(UPDEC_EINVAL, "batch programs exceed the staging space");
            }
            for (size_t k = 0; k < pr.ins.size(); k++) {
                const Ins& I = pr.ins[k];
                uint32_t* w = b->h_prog + prog_words + 3 * k;
                w[0] = (uint32_t)I.op | ((uint32_t)I.dst << 8) | ((uint32_t)I.arg << 16);
                w[1] = (uint32_t)I.a.kind | ((uint32_t)I.a.idx << 8) | ((uint32_t)I.a.shift << 16);
                w[2] = (uint32_t)I.b.kind | ((uint32_t)I.b.idx << 8) | ((uint32_t)I.b.shift << 16);
            }
            if (cw) memcpy(b->h_cst + const_words, pr.consts.data(), 4u * cw);
            it = placed.emplace(&pr, std::make_pair(prog_words / 3u, const_words)).first;
            prog_words += pw;
            const_words += cw;
        }
        Query& Q = b->h_q[qi];
        memset(&Q, 0, sizeof Q);
        for (int c = 0; c < kMaxCols; c++) {
            Q.col[c] = p.cols[c];
            Q.colbytes[c] = p.colbytes[c];
        }
        Q.rows = (uint64_t)p.rows;
        Q.ins_off = it->second.first;
        Q.nins = (uint32_t)pr.ins.size();
        Q.const_off = it->second.second;
        Q.root = (uint32_t)pr.root;
        Q.outB = (uint32_t)pr.outB;
        Q.outP = (uint32_t)pr.outP;
        Q.outS = (uint32_t)pr.outS;
        Q.faithful = pr.faithful ? 1u : 0u;
        Q.wclass = (uint32_t)pr.wclass;
        Q.image = p.image;
        Q.count = p.count;
        const int c = pr.wclass;
        const int64_t ch = p.chunk;
        const uint32_t n = (uint32_t)std::max<int64_t>(1, (p.rows + ch - 1) / ch);
        Q.item0 = fill[c];
        Q.nitems = n;
        for (uint32_t k = 0; k < n; k++) {
            Item& I = b->h_items[fill[c] + k];
            I.query = (uint32_t)qi;
            I.pad = 0;
            I.row0 = (uint64_t)std::min<int64_t>(p.rows, (int64_t)k * ch);
            I.row1 = (uint64_t)std::min<int64_t>(p.rows, (int64_t)(k + 1) * ch);
        }
        fill[c] += n;
    }
    const size_t nq = b->pending.size();
    b->clear_pending();
    VMCHK(hipMemcpyAsync(b->d_hdr, b->h_hdr, sizeof(Header), hipMemcpyHostToDevice, s));
    VMCHK(hipMemcpyAsync(b->d_q, b->h_q, sizeof(Query) * nq, hipMemcpyHostToDevice, s));
    VMCHK(hipMemcpyAsync(b->d_items, b->h_items, sizeof(Item) * total, hipMemcpyHostToDevice, s));
    if (prog_words) VMCHK(hipMemcpyAsync(b->d_prog, b->h_prog, 4u * prog_words, hipMemcpyHostToDevice, s));
    if (const_words) VMCHK(hipMemcpyAsync(b->d_cst, b->h_cst, 4u * const_words, hipMemcpyHostToDevice, s));
    VMCHK(hipEventRecord(b->staged, s));
    b->staged_pending = true;
    if (!use_graph) return b->launch_kernels(s, false, nitems);
    if (!b->graph) {
        // capture once: every class kernel on its own branch, sized for a
        // full device; kernels with no items exit at once
        hipStream_t cap;
        VMCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        int rc = UPDEC_OK;
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = b->launch_kernels(cap, true, nullptr);
            hipError_t e2 = hipStreamEndCapture(cap, &graph);
            if (rc == UPDEC_OK && e2 != hipSuccess) rc = api_fail(UPDEC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
        } else {
            rc = api_fail(UPDEC_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        }
        if (rc == UPDEC_OK) {
            e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = api_fail(UPDEC_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
    }
    VMCHK(hipGraphLaunch(b->graph, s));
    return UPDEC_OK;
}

void updec_batch_free(updec_batch_t b) { delete b; }
