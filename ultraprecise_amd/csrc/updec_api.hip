// updec_api.hip -- the C ABI (include/updec.h): hipRTC module cache,
// expression / aggregation launchers, and the static kernels (generator,
// statistics, group discovery, AVG division).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/updec.h"
#include "api_internal.h"
#include "dec_device.h"
#include "gen_common.h"
#include "planner.h"

// text of dec_device.h, generated at build time, handed to hipRTC
#include "dec_device_embed.inc"

using namespace updec;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(UPDEC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipStream_t S(updec_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ hipRTC module cache

struct Module {
    hipModule_t mod = nullptr;
    std::map<std::string, hipFunction_t> fns;
    std::mutex mu;
    ~Module()
    {
        if (mod) (void)hipModuleUnload(mod);
    }
    hipFunction_t fn(const std::string& name, std::string* err)
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = fns.find(name);
        if (it != fns.end()) return it->second;
        hipFunction_t f = nullptr;
        hipError_t e = hipModuleGetFunction(&f, mod, name.c_str());
        if (e != hipSuccess) {
            *err = "hipModuleGetFunction(" + name + "): " + hipGetErrorString(e);
            return nullptr;
        }
        fns[name] = f;
        return f;
    }
};
using ModuleP = std::shared_ptr<Module>;

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// On-disk store of compiled code objects, so a kernel compiled by one
// process is loaded by the next without hipRTC: one file per source, named
// by a hash of the source and of the device library text, holding both so a
// hash collision or a changed library is detected and recompiled.
// Directory: $UPDEC_KERNEL_CACHE ("off" disables), else
// $XDG_CACHE_HOME/updec, else $HOME/.cache/updec.  Failures to read or
// write it only cost a compile.
class DiskCache {
  public:
    DiskCache()
    {
        const char* env = getenv("UPDEC_KERNEL_CACHE");
        if (env && *env) {
            if (strcmp(env, "off") != 0) dir_ = env;
        } else if (const char* x = getenv("XDG_CACHE_HOME"); x && *x) {
            dir_ = std::string(x) + "/updec";
        } else if (const char* h = getenv("HOME"); h && *h) {
            dir_ = std::string(h) + "/.cache/updec";
        }
        lib_hash_ = fnv1a(kDecMlaneSource, sizeof(kDecMlaneSource) - 1,
                          fnv1a(kDecDeviceSource, sizeof(kDecDeviceSource) - 1));
    }
    bool load(const std::string& src, std::vector<char>* code)
    {
        if (dir_.empty()) return false;
        FILE* f = fopen(path(src).c_str(), "rb");
        if (!f) return false;
        bool ok = false;
        uint64_t hdr[3];
        if (fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == kMagic && hdr[1] == lib_hash_ && hdr[2] == src.size()) {
            std::string s(src.size(), '\0');
            if (fread(&s[0], 1, s.size(), f) == s.size() && s == src) {
                long pos = ftell(f);
                fseek(f, 0, SEEK_END);
                long end = ftell(f);
                if (pos > 0 && end > pos) {
                    code->resize((size_t)(end - pos));
                    fseek(f, pos, SEEK_SET);
                    ok = fread(code->data(), 1, code->size(), f) == code->size();
                }
            }
        }
        fclose(f);
        return ok;
    }
    void store(const std::string& src, const std::vector<char>& code)
    {
        if (dir_.empty()) return;
        for (size_t i = 0; i <= dir_.size(); i++) {   // mkdir -p
            if (i == dir_.size() || (dir_[i] == '/' && i > 0)) (void)mkdir(dir_.substr(0, i).c_str(), 0755);
        }
        const std::string fin = path(src);
        const std::string tmp = fin + ".tmp" + std::to_string((unsigned long)getpid());
        FILE* f = fopen(tmp.c_str(), "wb");
        if (!f) return;
        const uint64_t hdr[3] = {kMagic, lib_hash_, (uint64_t)src.size()};
        bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(src.data(), 1, src.size(), f) == src.size() &&
                  fwrite(code.data(), 1, code.size(), f) == code.size();
        ok = (fclose(f) == 0) && ok;
        if (!ok || rename(tmp.c_str(), fin.c_str()) != 0) (void)remove(tmp.c_str());
    }

  private:
    static constexpr uint64_t kMagic = 0x314b504455ull;   // "UPDK1"
    std::string path(const std::string& src) const
    {
        char name[48];
        snprintf(name, sizeof name, "/%016llx.hsaco",
                 (unsigned long long)fnv1a(src.data(), src.size(), lib_hash_ ^ 0x9e3779b97f4a7c15ull));
        return dir_ + name;
    }
    std::string dir_;
    uint64_t lib_hash_ = 0;
};

std::atomic<int64_t> g_rtc_compiles{0}, g_mem_hits{0}, g_disk_hits{0};

// LRU of compiled modules keyed by the full source text, as the
// reference's dynamic kernel manager (DynamicKernel.cu:10-58, LRU_COUNT 100).
// Generated sources carry no constant values (they are kernel data), so
// queries that differ only in their literals share a module.
class ModuleCache {
  public:
    static ModuleCache& get()
    {
        static ModuleCache c;
        return c;
    }
    ModuleP compile(const std::string& src, std::string* err)
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = map_.find(src);
            if (it != map_.end()) {
                lru_.splice(lru_.begin(), lru_, it->second.second);
                g_mem_hits++;
                return it->second.first;
            }
        }
        std::vector<char> code;
        if (disk_.load(src, &code)) {
            g_disk_hits++;
        } else {
            if (!rtc(src, &code, err)) return nullptr;
            g_rtc_compiles++;
            disk_.store(src, code);
        }
        auto m = std::make_shared<Module>();
        hipError_t e = hipModuleLoadData(&m->mod, code.data());
        if (e != hipSuccess) {
            *err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
            m->mod = nullptr;
            return nullptr;
        }
        std::lock_guard<std::mutex> g(mu_);
        lru_.push_front(src);
        map_[src] = {m, lru_.begin()};
        while (map_.size() > kCap) {
            map_.erase(lru_.back());
            lru_.pop_back();
        }
        return m;
    }
    // compile only (no device needed): used to check generated sources
    static bool rtc(const std::string& src, std::vector<char>* code, std::string* err)
    {
        hiprtcProgram prog;
        const char* hdr_src[] = {kDecDeviceSource, kDecMlaneSource};
        const char* hdr_name[] = {"dec_device.h", "dec_mlane.h"};
        if (hiprtcCreateProgram(&prog, src.c_str(), "updec_kernel.hip", 2, hdr_src, hdr_name) != HIPRTC_SUCCESS) {
            *err = "hiprtcCreateProgram failed";
            return false;
        }
        const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                              "-packed-fp32-ops"};
        hiprtcResult r = hiprtcCompileProgram(prog, sizeof(opts) / sizeof(opts[0]), opts);
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        if (r != HIPRTC_SUCCESS) {
            *err = "hipRTC compile failed: " + log;
            hiprtcDestroyProgram(&prog);
            return false;
        }
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code->resize(cs);
        hiprtcGetCode(prog, code->data());
        hiprtcDestroyProgram(&prog);
        return true;
    }

  private:
    static constexpr size_t kCap = 100;
    DiskCache disk_;
    std::mutex mu_;
    std::list<std::string> lru_;
    std::unordered_map<std::string, std::pair<ModuleP, std::list<std::string>::iterator>> map_;
};

int device_cus()
{
    static std::atomic<int> cus{0};
    int c = cus.load();
    if (c) return c;
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus.store(v);
    return v;
}

std::vector<ColumnType> to_types(const updec_column_desc* cols, int n, std::vector<std::string>* names)
{
    std::vector<ColumnType> t(n);
    for (int i = 0; i < n; i++) {
        t[i].P = cols[i].precision;
        t[i].S = cols[i].scale;
        t[i].bits = cols[i].bits;
        if (names) names->push_back(cols[i].name ? cols[i].name : ("c" + std::to_string(i)));
    }
    return t;
}

// parse -> AlignOptimize -> plan; mode AUTO picks faithful inside the envelope
bool make_plan(const std::string& text, const std::vector<std::string>& names, const std::vector<ColumnType>& types,
               int mode, Plan* out, std::string* err)
{
    return plan_expression(text, names, types, mode, out, err);
}

int copy_text(const std::string& s, char* buf, size_t cap, size_t* needed)
{
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return UPDEC_OK;
}

std::atomic<uint64_t> g_kernel_id{0};

}  // namespace

namespace updec {
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
int api_device_cus() { return device_cus(); }
std::vector<ColumnType> api_types(const updec_column_desc* cols, int n, std::vector<std::string>* names)
{
    return to_types(cols, n, names);
}
bool api_make_plan(const std::string& text, const std::vector<std::string>& names,
                   const std::vector<ColumnType>& types, int mode, Plan* out, std::string* err)
{
    return make_plan(text, names, types, mode, out, err);
}
}  // namespace updec

// ------------------------------------------------------------------ static kernels

#define UPD_MAXCOLS 16
struct ColArgs {
    const u8* col[UPD_MAXCOLS];
    usize colbytes[UPD_MAXCOLS];
    usize rows;
    u8* out;
    const u32* kc;  // the plans' constants (kernel data)
};
struct AggArgs {
    const u8* col[UPD_MAXCOLS];
    usize colbytes[UPD_MAXCOLS];
    usize rows;
    const i32* fcol;
    const u8* key0;
    const u8* key1;
    const u8* lut;
    u32* partials;
    const u32* kc;
};
// column-iterator kernel arguments (the generated kernels' ColIter and
// IterArgs)
struct ColIter {
    const u8* const* blocks;
    const i64* starts;
    usize positions;
    i32 nblocks;
    u32 stride;  // bytes per row: B, or B + 1 for nullable rows
    const void* idx;
    const void* const* iblocks;
    const i64* istarts;
    i32 niblocks;
    u32 flags;  // bit 0 nullable rows; bits 1-2 index kind (0 none, 1 int32, 2 nullable int32)
};
static_assert(sizeof(ColIter) == 64, "ColIter layout");
struct IterArgs {
    ColIter it[UPD_MAXCOLS];
    usize rows;
    const i32* fcol;
    const u8* key0;
    const u8* key1;
    const u8* lut;
    u32* partials;
    u8* out;
    u8* valid;
    const u32* kc;
    u32 outnull;
    u32 nullcols;  // bit c: column c can be NULL (divisions by zero under it give NULL)
    u32 skipnull;  // count only rows whose values are not NULL
    u32 pad;
};
static_assert(sizeof(IterArgs) == 1112, "IterArgs layout");

// validate block-column descriptors (one shared int32 index) and fill the
// iterator arguments: every column gets the shared index and, so that a
// position is NULL in all columns or in none, the smallest column's row
// count; no division-by-zero NULLs (nullcols = 0), as this form defines
static int fill_blocks(IterArgs* a, const std::vector<ColumnType>& types, const updec_column_blocks* cols,
                       const int32_t* d_indices, int64_t rows)
{
    *a = IterArgs{};
    usize limit = ~(usize)0;
    for (size_t i = 0; i < types.size(); i++) {
        const updec_column_blocks& c = cols[i];
        if (c.nblocks < 1 || !c.d_blocks || !c.d_starts || c.rows < 0)
            return api_fail(UPDEC_EINVAL, "column " + std::to_string(i) + ": no blocks");
        limit = std::min(limit, (usize)c.rows);
    }
    if (types.empty()) limit = 0;
    // without indices row i reads position i: every column must hold the rows
    if (!d_indices && (usize)rows > limit) return api_fail(UPDEC_EINVAL, "rows exceed the columns' rows");
    for (size_t i = 0; i < types.size(); i++) {
        ColIter& t = a->it[i];
        t.blocks = cols[i].d_blocks;
        t.starts = reinterpret_cast<const i64*>(cols[i].d_starts);
        t.nblocks = cols[i].nblocks;
        t.positions = limit;
        t.stride = (u32)compact_bytes(types[i].P);
        t.idx = d_indices;
        t.niblocks = d_indices ? 1 : 0;
        t.flags = d_indices ? (UPDEC_INDEX_INT32 << 1) : 0u;
    }
    a->rows = (usize)rows;
    a->skipnull = 1;
    return UPDEC_OK;
}

// validate column iterators (updec_column_iter) and fill the arguments
static int fill_iters(IterArgs* a, const std::vector<ColumnType>& types, const updec_column_iter* cols, int64_t rows)
{
    *a = IterArgs{};
    if (!types.empty() && !cols) return api_fail(UPDEC_EINVAL, "null column iterator array");
    for (size_t i = 0; i < types.size(); i++) {
        const updec_column_iter& c = cols[i];
        const std::string col = "column " + std::to_string(i) + ": ";
        if (c.nblocks < 1 || !c.d_blocks || !c.d_starts || c.positions < 0) return api_fail(UPDEC_EINVAL, col + "no blocks");
        if (c.index_kind < UPDEC_INDEX_NONE || c.index_kind > UPDEC_INDEX_NULLABLE_INT32)
            return api_fail(UPDEC_EINVAL, col + "unknown index kind");
        if (c.index_kind == UPDEC_INDEX_NONE && rows > c.positions)
            return api_fail(UPDEC_EINVAL, col + "rows exceed the positions of a column without index");
        if (c.index_kind != UPDEC_INDEX_NONE) {
            if (c.nindex_blocks < 1) return api_fail(UPDEC_EINVAL, col + "index without blocks");
            if (c.nindex_blocks == 1 && !c.d_index) return api_fail(UPDEC_EINVAL, col + "null index");
            if (c.nindex_blocks > 1 && (!c.d_index_blocks || !c.d_index_starts))
                return api_fail(UPDEC_EINVAL, col + "null index block arrays");
        }
        ColIter& t = a->it[i];
        t.blocks = c.d_blocks;
        t.starts = reinterpret_cast<const i64*>(c.d_starts);
        t.nblocks = c.nblocks;
        t.positions = (usize)c.positions;
        t.stride = (u32)compact_bytes(types[i].P) + (c.nullable ? 1u : 0u);
        t.idx = c.d_index;
        t.iblocks = c.d_index_blocks;
        t.istarts = reinterpret_cast<const i64*>(c.d_index_starts);
        t.niblocks = c.index_kind == UPDEC_INDEX_NONE ? 0 : c.nindex_blocks;
        t.flags = (c.nullable ? 1u : 0u) | ((u32)c.index_kind << 1);
        if (c.nullable || c.index_kind != UPDEC_INDEX_NONE) a->nullcols |= 1u << i;
    }
    a->rows = (usize)rows;
    return UPDEC_OK;
}

__global__ void __launch_bounds__(256) k_gen_compact(updec_gen::Desc d, uint64_t key, int64_t row0, int64_t rows,
                                                     u8* out, int nbytes)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    u8 buf[128];
    updec_gen::gen_row(d, key, (uint64_t)(row0 + i), buf, nbytes);
    for (int b = 0; b < nbytes; b++) out[i * nbytes + b] = buf[b];
}

__global__ void __launch_bounds__(256) k_gen_int32(uint64_t key, int32_t lo, int32_t hi, int64_t row0, int64_t rows,
                                                   int32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) out[i] = updec_gen::gen_int32(key, (uint64_t)(row0 + i), lo, hi);
}

struct ByteChoices {
    u8 c[64];
    int n;
};
__global__ void __launch_bounds__(256) k_gen_byte(uint64_t key, ByteChoices ch, int64_t row0, int64_t rows, u8* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) out[i] = updec_gen::gen_byte(key, (uint64_t)(row0 + i), ch.c, ch.n);
}

typedef unsigned int v4u __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(v4u* __restrict__ dst, const v4u* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Issue-rate probe of the limb multiply: every lane runs 8 independent
// chains of 32 x 32 -> 64 multiply-adds (v_mad_u64_u32), `iters` steps
// each, and stores the sums so nothing is folded away.  8 * iters * lanes
// multiply-adds per launch.
__global__ void __launch_bounds__(256) k_mad_peak(u64* out, u32 iters, u32 seed)
{
    const u32 t = blockIdx.x * 256u + threadIdx.x;
    u32 m[8];
    u64 acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        m[k] = (t + 1u) * 2654435761u + (u32)k * 40503u + seed;
        acc[k] = (u64)k;
    }
    for (u32 i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] = (u64)m[k] * (u32)(acc[k] >> 32 | 1u) + (u32)acc[k];
    }
    u64 s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) s ^= acc[k];
    out[t] = s;
}

// largest magnitude bit length of a compact column (sign bit excluded)
__global__ void __launch_bounds__(256) k_maxbits(const u8* col, int nbytes, int64_t rows, u32* out)
{
    u32 best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
        const u8* r = col + i * nbytes;
        u32 bl = 0;
        for (int b = nbytes - 1; b >= 0; b--) {
            u32 x = r[b];
            if (b == nbytes - 1) x &= 0x7fu;
            if (x) { bl = 8u * (u32)b + 32u - (u32)__builtin_clz(x); break; }
        }
        best = bl > best ? bl : best;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        u32 o = (u32)__shfl_xor((int)best, off, 64);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63u) == 0) atomicMax(out, best);
}

__device__ __forceinline__ bool filter_pass(i32 v, int op, i64 c)
{
    switch (op) {
    case 0: return v <= c;
    case 1: return v < c;
    case 2: return v >= c;
    case 3: return v > c;
    case 4: return v == c;
    default: return v != c;
    }
}

// group discovery: presence bitmap of (key0 << 8 | key1) over filtered rows.
// Each block builds its bitmap in LDS and merges the nonzero words into the
// global bitmap once, so global atomics are per block, not per row.
__global__ void __launch_bounds__(256) k_key_presence(const u8* key0, const u8* key1, int nkeys, const i32* fcol,
                                                      int fop, i64 fval, int64_t rows, u32* bitmap)
{
    __shared__ u32 bm[2048];
    for (int i = threadIdx.x; i < 2048; i += 256) bm[i] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
        if (fcol && !filter_pass(fcol[i], fop, fval)) continue;
        u32 k = nkeys == 2 ? (((u32)key0[i] << 8) | key1[i]) : key0[i];
        u32 bit = 1u << (k & 31u);
        if (!(bm[k >> 5] & bit)) atomicOr(&bm[k >> 5], bit);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2048; i += 256)
        if (bm[i]) atomicOr(&bitmap[i], bm[i]);
}

// AVG: Decimal(sum) /= Decimal((int64)count) per group (kernel_util.hxx:427-438)
template <int FAITH>
__global__ void __launch_bounds__(64) k_div_int64(const u8* sums, int64_t stride, const int64_t* counts, u8* out, int64_t n,
                                                  int64_t out_stride)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    div_image_int64<FAITH>(sums + i * stride, counts[i], out + i * out_stride);
}

// ------------------------------------------------------------------ objects

// The constants of an expression or aggregation, uploaded to the device at
// the first launch and passed to every launch (the reference's constValues,
// AriesExprCalcNode.cpp:346-349).
struct ConstBlock {
    std::vector<u32> host;
    u32* dev = nullptr;
    int upload(hipStream_t st)
    {
        if (dev || host.empty()) return UPDEC_OK;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&dev), host.size() * sizeof(u32)));
        HIPCHK(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(u32), hipMemcpyHostToDevice, st));
        return UPDEC_OK;
    }
    ~ConstBlock()
    {
        if (dev) (void)hipFree(dev);
    }
};

struct updec_expr_s {
    Plan plan;
    std::vector<ColumnType> types;
    ConstBlock consts;
    std::string name, src;
    ModuleP mod;
    hipFunction_t fn = nullptr;
    // the block-column variant, generated and compiled at first use
    ModuleP bmod;
    hipFunction_t bfn = nullptr;
    // multi-lane form (extended-mode plans whose exact intermediates need 64
    // or more limbs): the register program, lanes per row (1 = one row per
    // lane, the kernel above), the kernels compiled per lane count
    std::unique_ptr<VmProgram> vp;
    int lanes = 1;
    ConstBlock ml_consts;
    std::map<int, std::pair<ModuleP, hipFunction_t>> ml_fn;
};

struct updec_agg_s {
    AggSpec spec;
    std::vector<ColumnType> types;
    ConstBlock consts;   // every plan's block, at the plan's kbase
    int nkeys = 0, has_filter = 0;
    std::string last_src;
    struct GenMod {
        ModuleP mod;
        std::string src;
        int block = 256;
    };
    std::map<int, GenMod> by_ngroups;
    u32* d_bitmap = nullptr;
    u8* d_lut = nullptr;
    u32* d_partials = nullptr;
    size_t partial_words = 0;
    // one-pass grouped form
    bool onepass = false;
    ModuleP op_mod;
    std::string op_src;
    int op_block = 256;
    // one-pass state, two regions of kAggOnepassStatWords used by turns:
    // [0] a block's key table overflowed, [1] most slots a block used,
    // [2..2050) bitmap of the keys seen.  A run's final kernel clears the
    // other region for the next run; stat_dirty asks the host to clear both
    // (first run, or a run that stopped between its two launches).
    u32* d_meta = nullptr;
    int stat_phase = 0;
    bool stat_dirty = true;
    // the final kernel's output, written by it into pinned host memory:
    // [0] ngroups, [1] too many groups, [2] copy of the overflow flag,
    // [4..] keys
    u32* h_meta = nullptr;
    u32* h_meta_dev = nullptr;  // its device address
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    ~updec_agg_s()
    {
        if (d_bitmap) (void)hipFree(d_bitmap);
        if (d_lut) (void)hipFree(d_lut);
        if (d_partials) (void)hipFree(d_partials);
        if (d_meta) (void)hipFree(d_meta);
        if (h_meta) (void)hipHostFree(h_meta);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

struct updec_module_s {
    ModuleP mod;
};

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* updec_last_error(void) { return g_err.c_str(); }
int updec_version(void) { return 1; }

int updec_init(int device)
{
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipFree(nullptr));
    return segsum_scratch_setup();
}

int updec_malloc(void** d_ptr, size_t bytes)
{
    if (!d_ptr) return fail(UPDEC_EINVAL, "bad argument");
    HIPCHK(hipMalloc(d_ptr, bytes ? bytes : 4));
    return UPDEC_OK;
}
int updec_free(void* d_ptr)
{
    if (d_ptr) HIPCHK(hipFree(d_ptr));
    return UPDEC_OK;
}
int updec_malloc_async(void** d_ptr, size_t bytes, updec_stream_t stream)
{
    if (!d_ptr) return fail(UPDEC_EINVAL, "null pointer");
    HIPCHK(hipMallocAsync(d_ptr, bytes ? bytes : 4, S(stream)));
    return UPDEC_OK;
}
int updec_free_async(void* d_ptr, updec_stream_t stream)
{
    if (d_ptr) HIPCHK(hipFreeAsync(d_ptr, S(stream)));
    return UPDEC_OK;
}
int updec_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, updec_stream_t stream)
{
    if (bytes) HIPCHK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return UPDEC_OK;
}
int updec_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, updec_stream_t stream)
{
    if (bytes) HIPCHK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return UPDEC_OK;
}
int updec_memset(void* d_ptr, int value, size_t bytes, updec_stream_t stream)
{
    if (bytes) HIPCHK(hipMemsetAsync(d_ptr, value, bytes, S(stream)));
    return UPDEC_OK;
}
int updec_stream_create(updec_stream_t* stream)
{
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreate(&s));
    *stream = s;
    return UPDEC_OK;
}
int updec_stream_destroy(updec_stream_t stream)
{
    if (stream) HIPCHK(hipStreamDestroy(S(stream)));
    return UPDEC_OK;
}
int updec_stream_sync(updec_stream_t stream)
{
    HIPCHK(hipStreamSynchronize(S(stream)));
    return UPDEC_OK;
}
int updec_device_sync(void)
{
    HIPCHK(hipDeviceSynchronize());
    return UPDEC_OK;
}
int updec_device_info(int32_t* cus, int32_t* clock_khz, int64_t* mem_bytes, char* name, size_t name_cap)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    if (cus) *cus = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (mem_bytes) *mem_bytes = (int64_t)p.totalGlobalMem;
    if (name && name_cap) {
        snprintf(name, name_cap, "%s %s", p.name, p.gcnArchName);
    }
    return UPDEC_OK;
}
int updec_mem_info(int64_t* free_bytes, int64_t* total_bytes)
{
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return UPDEC_OK;
}
int updec_event_create(void** ev)
{
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    *ev = e;
    return UPDEC_OK;
}
int updec_event_record(void* ev, updec_stream_t stream)
{
    HIPCHK(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), S(stream)));
    return UPDEC_OK;
}
int updec_event_elapsed(void* ev0, void* ev1, float* ms)
{
    HIPCHK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(ev1)));
    HIPCHK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(ev0), reinterpret_cast<hipEvent_t>(ev1)));
    return UPDEC_OK;
}
int updec_event_query(void* ev, int32_t* done)
{
    if (!ev || !done) return fail(UPDEC_EINVAL, "bad argument");
    hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(ev));
    if (e == hipSuccess) {
        *done = 1;
        return UPDEC_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return UPDEC_OK;
    }
    return fail(UPDEC_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
}
int updec_event_sync(void* ev)
{
    if (!ev) return fail(UPDEC_EINVAL, "bad argument");
    HIPCHK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(ev)));
    return UPDEC_OK;
}
int updec_event_destroy(void* ev)
{
    if (ev) HIPCHK(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
    return UPDEC_OK;
}

int updec_copy_kernel(void* d_dst, const void* d_src, size_t bytes, updec_stream_t stream)
{
    if ((bytes & 15u) || (reinterpret_cast<uintptr_t>(d_dst) & 15u) || (reinterpret_cast<uintptr_t>(d_src) & 15u))
        return fail(UPDEC_EINVAL, "copy needs 16-byte aligned pointers and size");
    size_t n = bytes / 16;
    if (!n) return UPDEC_OK;
    unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)device_cus() * 16);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, S(stream), reinterpret_cast<v4u*>(d_dst),
                       reinterpret_cast<const v4u*>(d_src), n);
    HIPCHK(hipGetLastError());
    return UPDEC_OK;
}

int updec_mad_peak(int32_t blocks, int32_t iters, double* ms_out, double* macs_out)
{
    if (blocks < 1 || iters < 1 || !ms_out || !macs_out) return fail(UPDEC_EINVAL, "bad argument");
    u64* d = nullptr;
    HIPCHK(hipMalloc(reinterpret_cast<void**>(&d), (size_t)blocks * 256 * sizeof(u64)));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_mad_peak, dim3(blocks), dim3(256), 0, nullptr, d, (u32)iters, 1u);   // warm-up
    HIPCHK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_mad_peak, dim3(blocks), dim3(256), 0, nullptr, d, (u32)iters, 2u);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipFree(d));
    *ms_out = ms;
    *macs_out = 8.0 * (double)iters * (double)blocks * 256.0;
    return UPDEC_OK;
}

int updec_compact_bytes(int precision) { return compact_bytes(precision); }

int updec_column_maxbits(const uint8_t* d_col, int precision, int64_t rows, int32_t* bits_out, updec_stream_t stream)
{
    if (!bits_out || rows < 0) return fail(UPDEC_EINVAL, "bad argument");
    int nb = compact_bytes(precision);
    if (nb <= 0) return fail(UPDEC_EINVAL, "bad precision");
    u32* d = nullptr;
    HIPCHK(hipMallocAsync((void**)&d, sizeof(u32), S(stream)));
    HIPCHK(hipMemsetAsync(d, 0, sizeof(u32), S(stream)));
    if (rows > 0) {
        int64_t blocks = std::min<int64_t>((rows + 255) / 256, (int64_t)device_cus() * 8);
        hipLaunchKernelGGL(k_maxbits, dim3((unsigned)blocks), dim3(256), 0, S(stream), d_col, nb, rows, d);
        HIPCHK(hipGetLastError());
    }
    u32 h = 0;
    HIPCHK(hipMemcpyAsync(&h, d, sizeof(u32), hipMemcpyDeviceToHost, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    HIPCHK(hipFreeAsync(d, S(stream)));
    *bits_out = (int32_t)h;
    return UPDEC_OK;
}

int updec_expr_compile(const char* text, const updec_column_desc* cols, int ncols, int mode, updec_expr_t* out)
{
    if (!text || !out || ncols < 0 || ncols > UPD_MAXCOLS || (ncols && !cols)) return fail(UPDEC_EINVAL, "bad argument");
    std::vector<std::string> names;
    auto types = to_types(cols, ncols, &names);
    std::unique_ptr<updec_expr_s> e(new updec_expr_s());
    std::string err;
    if (!make_plan(text, names, types, mode, &e->plan, &err)) return fail(UPDEC_ECOMPILE, err);
    e->types = types;
    e->consts.host = e->plan.const_data();
    e->name = "calc_expr_" + std::to_string(g_kernel_id.fetch_add(1));
    GenOptions o;
    o.block = calc_block(e->plan);
    e->src = gen_calc_kernel(e->plan, types, "calc_expr", o);
    // a plan too wide for one row per lane also gets the multi-lane form
    // (updec_expr_lanes selects it)
    if (e->plan.mode != Mode::Faithful && ncols <= updec_vm::kMaxCols) {
        std::unique_ptr<VmProgram> vp(new VmProgram());
        std::string verr;
        if (build_vm_program(e->plan, types, vp.get(), &verr) && vp->wclass >= updec_vm::kFirstMultiLane &&
            vp->outB == e->plan.outB) {
            e->ml_consts.host = vp->consts;
            e->vp = std::move(vp);
        }
    }
    *out = e.release();
    return UPDEC_OK;
}

int updec_expr_lanes(updec_expr_t e, int32_t lanes_per_row, int32_t* width_out)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    const int W = e->vp ? e->vp->W : 0;
    if (width_out) *width_out = W;
    if (lanes_per_row == 0) return UPDEC_OK;
    if (lanes_per_row == 1 || (W && (lanes_per_row == W / 4 || lanes_per_row == W / 8) && lanes_per_row <= 64)) {
        e->lanes = lanes_per_row;
        return UPDEC_OK;
    }
    return fail(UPDEC_EUNSUPPORTED, W ? "lanes per row must be 1, W/4 or W/8 (at most 64)"
                                      : "this expression has only the one-row-per-lane form");
}

int updec_expr_source_ml(updec_expr_t e, int32_t lanes_per_row, char* buf, size_t cap, size_t* needed)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    if (!e->vp || (lanes_per_row != e->vp->W / 4 && lanes_per_row != e->vp->W / 8))
        return fail(UPDEC_EUNSUPPORTED, "no multi-lane form at this lane count");
    return copy_text(gen_calc_ml(*e->vp, e->types, e->vp->W / lanes_per_row, "calc_ml"), buf, cap, needed);
}

int updec_expr_info(updec_expr_t e, int32_t* P, int32_t* S_, int32_t* out_bytes, int32_t* faithful)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    if (P) *P = e->plan.outP;
    if (S_) *S_ = e->plan.outS;
    if (out_bytes) *out_bytes = e->plan.outB;
    if (faithful) *faithful = e->plan.mode == Mode::Faithful ? 1 : 0;
    return UPDEC_OK;
}

int updec_expr_work(updec_expr_t e, int64_t* limb_products_per_row)
{
    if (!e || !limb_products_per_row) return fail(UPDEC_EINVAL, "bad argument");
    *limb_products_per_row = plan_limb_products(e->plan);
    return UPDEC_OK;
}

int updec_expr_plan(updec_expr_t e, char* buf, size_t cap, size_t* needed)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    char hdr[64];
    snprintf(hdr, sizeof hdr, "%s %d %d %d ", e->plan.mode == Mode::Faithful ? "F" : "X", e->plan.outP, e->plan.outS,
             e->plan.outB);
    return copy_text(hdr + e->plan.text(), buf, cap, needed);
}

int updec_expr_source(updec_expr_t e, char* buf, size_t cap, size_t* needed)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    return copy_text(e->src, buf, cap, needed);
}

int updec_expr_source_iter(updec_expr_t e, char* buf, size_t cap, size_t* needed)
{
    if (!e) return fail(UPDEC_EINVAL, "null expression");
    GenOptions o;
    o.blocks = true;
    o.block = calc_block(e->plan);
    return copy_text(gen_calc_kernel(e->plan, e->types, "calc_iter", o), buf, cap, needed);
}

// every column pointer of a launch over rows > 0 must be set: a NULL column
// is refused on the host instead of faulting the device
static int check_cols(const uint8_t* const* d_cols, size_t ncols, int64_t rows)
{
    if (rows <= 0 || ncols == 0) return UPDEC_OK;
    if (!d_cols) return fail(UPDEC_EINVAL, "null column array");
    for (size_t i = 0; i < ncols; i++)
        if (!d_cols[i]) return fail(UPDEC_EINVAL, "null column " + std::to_string(i));
    return UPDEC_OK;
}

int updec_expr_calc(updec_expr_t e, const uint8_t* const* d_cols, int64_t rows, uint8_t* d_out, updec_stream_t stream)
{
    if (!e || rows < 0 || (rows && (!d_cols || !d_out))) return fail(UPDEC_EINVAL, "bad argument");
    if (int rc = check_cols(d_cols, e->types.size(), rows)) return rc;
    if (reinterpret_cast<uintptr_t>(d_out) & 3u) return fail(UPDEC_EINVAL, "output must be 4-byte aligned");
    if (rows == 0) return UPDEC_OK;
    std::string err;
    if (e->lanes > 1) {
        // multi-lane form: a group of `lanes` lanes per row, 256-thread
        // blocks striding over the rows
        auto& mf = e->ml_fn[e->lanes];
        if (!mf.first) {
            mf.first = ModuleCache::get().compile(gen_calc_ml(*e->vp, e->types, e->vp->W / e->lanes, "calc_ml"), &err);
            if (!mf.first) return fail(UPDEC_ECOMPILE, err);
            mf.second = mf.first->fn("calc_ml", &err);
            if (!mf.second) return fail(UPDEC_ECOMPILE, err);
        }
        if (int rc = e->ml_consts.upload(S(stream))) return rc;
        struct MlArgs {
            const uint8_t* col[updec_vm::kMaxCols];
            usize rows;
            uint8_t* out;
            const u32* kc;
        } a{};
        for (size_t i = 0; i < e->types.size(); i++) a.col[i] = d_cols[i];
        a.rows = (usize)rows;
        a.out = d_out;
        a.kc = e->ml_consts.dev;
        const int64_t per_block = 256 / e->lanes;
        const int64_t blocks = std::min<int64_t>((rows + per_block - 1) / per_block, (int64_t)device_cus() * 16);
        void* params[] = {&a};
        HIPCHK(hipModuleLaunchKernel(mf.second, (unsigned)blocks, 1, 1, 256, 1, 1, 0, S(stream), params, nullptr));
        return UPDEC_OK;
    }
    if (!e->mod) {
        e->mod = ModuleCache::get().compile(e->src, &err);
        if (!e->mod) return fail(UPDEC_ECOMPILE, err);
        e->fn = e->mod->fn("calc_expr", &err);
        if (!e->fn) return fail(UPDEC_ECOMPILE, err);
    }
    if (int rc = e->consts.upload(S(stream))) return rc;
    ColArgs a{};
    for (size_t i = 0; i < e->types.size(); i++) {
        a.col[i] = d_cols[i];
        a.colbytes[i] = (usize)rows * (usize)compact_bytes(e->types[i].P);
    }
    a.rows = (usize)rows;
    a.out = d_out;
    a.kc = e->consts.dev;
    const int T = calc_block(e->plan);
    int64_t tiles = (rows + T - 1) / T;
    if (tiles > INT32_MAX) return fail(UPDEC_EINVAL, "too many rows for one launch");
    unsigned grid = calc_one_row(e->plan) ? (unsigned)tiles
                                          : (unsigned)std::min<int64_t>(tiles, (int64_t)device_cus() * 32 * (256 / T));
    void* params[] = {&a};
    HIPCHK(hipModuleLaunchKernel(e->fn, grid, 1, 1, T, 1, 1, 0, S(stream), params, nullptr));
    return UPDEC_OK;
}

// launch the iterator variant of an expression's calc kernel over filled arguments
static int calc_iter_launch(updec_expr_t e, IterArgs* a, int64_t rows, uint8_t* d_out, uint8_t* d_valid, int out_nullable,
                            updec_stream_t stream)
{
    if (rows == 0) return UPDEC_OK;
    std::string err;
    if (!e->bmod) {
        GenOptions o;
        o.blocks = true;
        o.block = calc_block(e->plan);
        e->bmod = ModuleCache::get().compile(gen_calc_kernel(e->plan, e->types, "calc_iter", o), &err);
        if (!e->bmod) return fail(UPDEC_ECOMPILE, err);
        e->bfn = e->bmod->fn("calc_iter", &err);
        if (!e->bfn) return fail(UPDEC_ECOMPILE, err);
    }
    if (int rc2 = e->consts.upload(S(stream))) return rc2;
    a->out = d_out;
    a->valid = d_valid;
    a->kc = e->consts.dev;
    a->outnull = out_nullable ? 1u : 0u;
    const int T = calc_block(e->plan);
    int64_t tiles = (rows + T - 1) / T;
    if (tiles > INT32_MAX) return fail(UPDEC_EINVAL, "too many rows for one launch");
    unsigned grid = calc_one_row(e->plan) ? (unsigned)tiles
                                          : (unsigned)std::min<int64_t>(tiles, (int64_t)device_cus() * 32 * (256 / T));
    void* params[] = {a};
    HIPCHK(hipModuleLaunchKernel(e->bfn, grid, 1, 1, T, 1, 1, 0, S(stream), params, nullptr));
    return UPDEC_OK;
}

int updec_expr_calc_blocks(updec_expr_t e, const updec_column_blocks* cols, const int32_t* d_indices, int64_t rows,
                           uint8_t* d_out, uint8_t* d_valid, updec_stream_t stream)
{
    if (!e || rows < 0 || (rows && (!d_out || (!cols && !e->types.empty())))) return fail(UPDEC_EINVAL, "bad argument");
    if (reinterpret_cast<uintptr_t>(d_out) & 3u) return fail(UPDEC_EINVAL, "output must be 4-byte aligned");
    IterArgs a;
    int rc = fill_blocks(&a, e->types, cols, d_indices, rows);
    if (rc != UPDEC_OK) return rc;
    return calc_iter_launch(e, &a, rows, d_out, d_valid, 0, stream);
}

int updec_expr_calc_iter(updec_expr_t e, const updec_column_iter* cols, int64_t rows, uint8_t* d_out,
                         int32_t out_nullable, uint8_t* d_valid, updec_stream_t stream)
{
    if (!e || rows < 0 || (rows && !d_out)) return fail(UPDEC_EINVAL, "bad argument");
    if (reinterpret_cast<uintptr_t>(d_out) & 3u) return fail(UPDEC_EINVAL, "output must be 4-byte aligned");
    IterArgs a;
    int rc = fill_iters(&a, e->types, cols, rows);
    if (rc != UPDEC_OK) return rc;
    return calc_iter_launch(e, &a, rows, d_out, d_valid, out_nullable, stream);
}

void updec_expr_free(updec_expr_t e) { delete e; }

int updec_agg_compile(const updec_agg_desc* d, updec_agg_t* out)
{
    if (!d || !out || d->ncols < 0 || d->ncols > UPD_MAXCOLS || d->nexprs <= 0 || d->nkeys < 0 || d->nkeys > 2)
        return fail(UPDEC_EINVAL, "bad argument");
    std::vector<std::string> names;
    std::unique_ptr<updec_agg_s> a(new updec_agg_s());
    a->types = to_types(d->cols, d->ncols, &names);
    std::string err;
    for (int k = 0; k < d->nexprs; k++) {
        Plan p;
        if (!make_plan(d->exprs[k], names, a->types, d->mode, &p, &err))
            return fail(UPDEC_ECOMPILE, "expr " + std::to_string(k) + ": " + err);
        // constants of plan k follow those of plans 0..k-1
        p.kbase = (int)a->consts.host.size();
        const std::vector<u32> kd = p.const_data();
        a->consts.host.insert(a->consts.host.end(), kd.begin(), kd.end());
        a->spec.plans.push_back(std::move(p));
    }
    a->has_filter = d->has_filter;
    a->spec.filter_col = d->has_filter ? 0 : -1;
    a->spec.filter_op = d->filter_op;
    a->spec.filter_val = d->filter_value;
    a->nkeys = d->nkeys;
    a->spec.nkeys = d->nkeys;
    if (d->nexprs < 32 && (d->avg_mask >> d->nexprs)) return fail(UPDEC_EINVAL, "avg_mask names an expression past nexprs");
    a->spec.avg_mask = d->avg_mask;
    a->spec.avg_faithful = d->mode != UPDEC_MODE_EXTENDED;
    a->onepass = agg_onepass_ok(a->spec, GenOptions());
    *out = a.release();
    return UPDEC_OK;
}

// kernel names by form, so profiles tell the forms apart: one-pass keyed
// (agg_keys_*), two-pass through the key table (agg_lut_*), no grouping
// (agg_sum_*)
static std::string form_name(const updec_agg_s* a, bool onepass)
{
    return onepass ? "agg_keys" : (a->nkeys == 0 ? "agg_sum" : "agg_lut");
}

static int agg_module(updec_agg_t a, int ngroups, bool iter, ModuleP* mod, int* block, std::string* err)
{
    const int mkey = ngroups | (iter ? 1 << 20 : 0);
    auto it = a->by_ngroups.find(mkey);
    if (it != a->by_ngroups.end()) {
        *mod = it->second.mod;
        *block = it->second.block;
        a->last_src = it->second.src;
        return UPDEC_OK;
    }
    AggSpec s = a->spec;
    s.ngroups = ngroups;
    GenOptions o;
    o.blocks = iter;
    int aw = 0, blk = 256;
    std::string src = gen_agg_kernels(s, a->types, form_name(a, false), o, &aw, &blk);
    if (src.empty()) {
        *err = "the accumulators of " + std::to_string(ngroups) + " groups do not fit in LDS";
        return UPDEC_EUNSUPPORTED;
    }
    ModuleP m = ModuleCache::get().compile(src, err);
    if (!m) return UPDEC_ECOMPILE;
    a->by_ngroups[mkey] = {m, src, blk};
    a->last_src = src;
    *mod = m;
    *block = blk;
    return UPDEC_OK;
}

static void fill_args(updec_agg_t a, AggArgs* args, const uint8_t* const* d_cols, const int32_t* d_filter,
                      const uint8_t* d_key0, const uint8_t* d_key1, int64_t rows)
{
    *args = AggArgs{};
    for (size_t i = 0; i < a->types.size(); i++) {
        args->col[i] = d_cols ? d_cols[i] : nullptr;
        args->colbytes[i] = (usize)rows * (usize)compact_bytes(a->types[i].P);
    }
    args->rows = (usize)rows;
    args->fcol = d_filter;
    args->key0 = d_key0;
    args->key1 = d_key1;
    args->lut = a->d_lut;
    args->partials = a->d_partials;
    args->kc = a->consts.dev;
}

// rps: rows a lane takes per loop step (1, or 4 for row quads)
// AVG images of the two-pass forms: one k_div_int64 launch per AVG
// expression over the groups' SUM images and counts
static int launch_avgs(updec_agg_t a, uint8_t* d_images, const int64_t* d_counts, int ngroups, int max_groups,
                       hipStream_t st)
{
    const uint32_t mask = a->spec.avg_mask;
    if (!mask || ngroups <= 0) return UPDEC_OK;
    const int NP = (int)a->spec.plans.size();
    int navg = 0;
    for (int k = 0; k < NP; k++) navg += (mask >> k) & 1u;
    uint8_t* avgs = d_images + (size_t)max_groups * NP * 133;
    const unsigned blocks = (unsigned)((ngroups + 63) / 64);
    for (int k = 0, j = 0; k < NP; k++) {
        if (!((mask >> k) & 1u)) continue;
        const u8* sums = d_images + (size_t)k * 133;
        u8* out = avgs + (size_t)j * 133;
        if (a->spec.avg_faithful)
            hipLaunchKernelGGL(k_div_int64<1>, dim3(blocks), dim3(64), 0, st, sums, (int64_t)NP * 133, d_counts, out,
                               (int64_t)ngroups, (int64_t)navg * 133);
        else
            hipLaunchKernelGGL(k_div_int64<0>, dim3(blocks), dim3(64), 0, st, sums, (int64_t)NP * 133, d_counts, out,
                               (int64_t)ngroups, (int64_t)navg * 133);
        HIPCHK(hipGetLastError());
        j++;
    }
    return UPDEC_OK;
}

static int64_t agg_grid(hipFunction_t f, int64_t rows, int T, int rps)
{
    const int64_t steps = (rows + rps - 1) / rps;
    int64_t tiles = std::max<int64_t>(1, (steps + T - 1) / T);
    // the occupancy query is made once per kernel and block size (a
    // repeated run would otherwise pay it on every call); a handle reused by
    // a later module can only make the grid a poor size, never wrong, since
    // the per-lane row bound below does not depend on it
    static std::mutex occ_mu;
    static std::map<std::pair<hipFunction_t, int>, int> occ;
    int blocks_per_cu = 0;
    {
        std::lock_guard<std::mutex> g(occ_mu);
        auto it = occ.find({f, T});
        if (it != occ.end()) {
            blocks_per_cu = it->second;
        } else {
            if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, f, T, 0) != hipSuccess ||
                blocks_per_cu < 1)
                blocks_per_cu = 2;
            occ[{f, T}] = blocks_per_cu;
        }
    }
    int64_t grid = std::min<int64_t>(tiles, (int64_t)device_cus() * blocks_per_cu);
    // every lane <= 2^14 rows: at most 2^12 quads (or 2^14 single rows)
    const int64_t per_block = (int64_t)T << (kAggRowsPerLaneLog2 - 2);
    const int64_t min_grid = ((rows + 3) / 4 + per_block - 1) / per_block;
    return std::max(grid, std::max<int64_t>(min_grid, 1));
}

// One pass: keys numbered per block in LDS, merged on the device; one
// host synchronization at the end to return the group count and keys.
// Sets *overflowed when some block met more distinct keys than its table
// holds; the caller then runs the two-pass form.
static int agg_run_onepass(updec_agg_t a, const uint8_t* const* d_cols, const int32_t* d_filter, const uint8_t* d_key0,
                           const uint8_t* d_key1, int64_t rows, uint8_t* d_images, int64_t* d_counts, int32_t max_groups,
                           int32_t* ngroups_out, uint16_t* keys_out, hipStream_t st, bool* overflowed)
{
    *overflowed = false;
    std::string err;
    if (!a->op_mod) {
        AggSpec s = a->spec;
        s.onepass = true;
        GenOptions o;
        int aw = 0;
        a->op_src = gen_agg_kernels(s, a->types, form_name(a, true), o, &aw, &a->op_block);
        if (a->op_src.empty()) return fail(UPDEC_EUNSUPPORTED, "one-pass accumulators do not fit in LDS");
        a->op_mod = ModuleCache::get().compile(a->op_src, &err);
        if (!a->op_mod) return fail(UPDEC_ECOMPILE, err);
    }
    a->last_src = a->op_src;
    hipFunction_t fmain = a->op_mod->fn(form_name(a, true) + "_main", &err);
    hipFunction_t ffin = fmain ? a->op_mod->fn(form_name(a, true) + "_final", &err) : nullptr;
    if (!fmain || !ffin) return fail(UPDEC_ECOMPILE, err);
    const int T = a->op_block;
    const int64_t grid = agg_grid(fmain, rows, T, agg_rows_per_step(a->spec, a->types, GenOptions()));
    const size_t need = (size_t)grid * (size_t)agg_onepass_record_words(a->spec);
    if (need > a->partial_words) {
        if (a->d_partials) HIPCHK(hipFree(a->d_partials));
        a->d_partials = nullptr;
        HIPCHK(hipMalloc((void**)&a->d_partials, need * sizeof(u32)));
        a->partial_words = need;
    }
    if (!a->d_meta) {
        HIPCHK(hipMalloc((void**)&a->d_meta, 2 * kAggOnepassStatWords * sizeof(u32)));
        a->stat_dirty = true;
    }
    if (!a->h_meta) {
        HIPCHK(hipHostMalloc((void**)&a->h_meta, (4 + 256) * sizeof(u32), hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void**)&a->h_meta_dev, a->h_meta, 0));
    }
    if (a->stat_dirty) HIPCHK(hipMemsetAsync(a->d_meta, 0, 2 * kAggOnepassStatWords * sizeof(u32), st));
    a->stat_dirty = true;
    u32* stat = a->d_meta + (size_t)a->stat_phase * kAggOnepassStatWords;
    u32* next = a->d_meta + (size_t)(1 - a->stat_phase) * kAggOnepassStatWords;
    if (int krc = a->consts.upload(st)) return krc;
    AggArgs args;
    fill_args(a, &args, d_cols, d_filter, d_key0, d_key1, rows);
    u32* ovf = stat;
    if (!a->ev0) {
        HIPCHK(hipEventCreate(&a->ev0));
        HIPCHK(hipEventCreate(&a->ev1));
    }
    // events recorded around the launch (events attached to the dispatch
    // with hipExtModuleLaunchKernel cost about 16 us more host time per
    // launch); the span also takes in the few microseconds between the first
    // marker and the dispatch
    void* pm[] = {&args, &ovf};
    HIPCHK(hipEventRecord(a->ev0, st));
    HIPCHK(hipModuleLaunchKernel(fmain, (uint32_t)grid, 1, 1, T, 1, 1, 0, st, pm, nullptr));
    HIPCHK(hipEventRecord(a->ev1, st));
    a->timed = true;
    const u32* rec = a->d_partials;
    u32 nb = (u32)grid, maxg = (u32)max_groups;
    uint8_t* img = d_images;
    uint64_t* cnt = reinterpret_cast<uint64_t*>(d_counts);
    u32* meta = a->h_meta_dev;
    uint8_t* avgs = a->spec.avg_mask ? d_images + (size_t)max_groups * a->spec.plans.size() * 133 : nullptr;
    void* pf[] = {&rec, &nb, &stat, &next, &maxg, &img, &cnt, &meta, &avgs};
    const int NP = (int)a->spec.plans.size();
    HIPCHK(hipModuleLaunchKernel(ffin, (unsigned)(max_groups * (NP + 1)), 1, 1, kAggFinalThreads, 1, 1, 0, st, pf, nullptr));
    a->stat_dirty = false;
    a->stat_phase = 1 - a->stat_phase;
    HIPCHK(hipStreamSynchronize(st));
    if (a->h_meta[2]) {
        *overflowed = true;
        return UPDEC_OK;
    }
    if (a->h_meta[1]) return fail(UPDEC_EINVAL, "more groups than max_groups");
    const int ng = (int)a->h_meta[0];
    *ngroups_out = ng;
    if (keys_out)
        for (int g = 0; g < ng; g++) {
            const u32 k = a->h_meta[4 + g];
            keys_out[g] = (uint16_t)(a->nkeys == 2 ? (((k >> 8) & 0xff) | ((k & 0xff) << 8)) : k);
        }
    return UPDEC_OK;
}

// Two-pass form: key discovery into a host-numbered table, then one
// generated kernel pass with a group LUT (or no keys: one group).  With
// `iter` the columns come through column iterators (the kernel variant
// generated with iterator loads); else the contiguous columns d_cols.
static int agg_run_twopass(updec_agg_t a, const IterArgs* iter, const uint8_t* const* d_cols, const int32_t* d_filter,
                           const uint8_t* d_key0, const uint8_t* d_key1, int64_t rows, uint8_t* d_images,
                           int64_t* d_counts, int32_t max_groups, int32_t* ngroups_out, uint16_t* keys_out,
                           hipStream_t st)
{
    int ngroups = 1;
    std::vector<uint16_t> keys;
    if (a->nkeys > 0) {
        if (!a->d_bitmap) HIPCHK(hipMalloc((void**)&a->d_bitmap, 65536 / 8));
        if (!a->d_lut) HIPCHK(hipMalloc((void**)&a->d_lut, 65536));
        HIPCHK(hipMemsetAsync(a->d_bitmap, 0, 65536 / 8, st));
        if (rows > 0) {
            int64_t blocks = std::min<int64_t>((rows + 255) / 256, (int64_t)device_cus() * 8);
            hipLaunchKernelGGL(k_key_presence, dim3((unsigned)blocks), dim3(256), 0, st, d_key0, d_key1, a->nkeys,
                               a->has_filter ? d_filter : nullptr, a->spec.filter_op, (i64)a->spec.filter_val, rows,
                               a->d_bitmap);
            HIPCHK(hipGetLastError());
        }
        std::vector<u32> bm(65536 / 32);
        HIPCHK(hipMemcpyAsync(bm.data(), a->d_bitmap, 65536 / 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<u8> lut(65536, 0xff);
        for (u32 k = 0; k < 65536; k++)
            if (bm[k >> 5] & (1u << (k & 31))) {
                if ((int)keys.size() >= 255) return fail(UPDEC_EUNSUPPORTED, "more than 255 groups");
                lut[k] = (u8)keys.size();
                keys.push_back((uint16_t)(a->nkeys == 2 ? (((k >> 8) & 0xff) | ((k & 0xff) << 8)) : k));
            }
        ngroups = (int)keys.size();
        if (ngroups > max_groups) return fail(UPDEC_EINVAL, "more groups than max_groups");
        *ngroups_out = ngroups;
        if (keys_out)
            for (int g = 0; g < ngroups; g++) keys_out[g] = keys[g];
        if (ngroups == 0) return UPDEC_OK;
        HIPCHK(hipMemcpyAsync(a->d_lut, lut.data(), 65536, hipMemcpyHostToDevice, st));
    } else {
        if (max_groups < 1) return fail(UPDEC_EINVAL, "max_groups < 1");
        *ngroups_out = 1;
        if (keys_out) keys_out[0] = 0;
    }
    std::string err;
    ModuleP mod;
    int T = 256;
    int mrc = agg_module(a, ngroups, iter != nullptr, &mod, &T, &err);
    if (mrc != UPDEC_OK) return fail(mrc, err);
    hipFunction_t fmain = mod->fn(form_name(a, false) + "_main", &err);
    hipFunction_t ffin = fmain ? mod->fn(form_name(a, false) + "_final", &err) : nullptr;
    if (!fmain || !ffin) return fail(UPDEC_ECOMPILE, err);

    const int NP = (int)a->spec.plans.size();
    const int aw = agg_acc_words(a->spec);
    const size_t PW = (size_t)NP * ngroups * aw + 2 * (size_t)ngroups;
    GenOptions go;
    go.blocks = iter != nullptr;
    const int64_t grid = agg_grid(fmain, rows, T, agg_rows_per_step(a->spec, a->types, go));
    if ((size_t)grid * PW > a->partial_words) {
        if (a->d_partials) HIPCHK(hipFree(a->d_partials));
        a->d_partials = nullptr;
        HIPCHK(hipMalloc((void**)&a->d_partials, (size_t)grid * PW * sizeof(u32)));
        a->partial_words = (size_t)grid * PW;
    }
    if (int krc = a->consts.upload(st)) return krc;
    AggArgs args;
    IterArgs iargs;
    void* pm[1];
    if (iter) {
        iargs = *iter;
        iargs.fcol = d_filter;
        iargs.key0 = d_key0;
        iargs.key1 = d_key1;
        iargs.lut = a->d_lut;
        iargs.partials = a->d_partials;
        iargs.kc = a->consts.dev;
        pm[0] = &iargs;
    } else {
        fill_args(a, &args, d_cols, d_filter, d_key0, d_key1, rows);
        pm[0] = &args;
    }
    if (!a->ev0) {
        HIPCHK(hipEventCreate(&a->ev0));
        HIPCHK(hipEventCreate(&a->ev1));
    }
    HIPCHK(hipEventRecord(a->ev0, st));
    HIPCHK(hipModuleLaunchKernel(fmain, (uint32_t)grid, 1, 1, T, 1, 1, 0, st, pm, nullptr));
    HIPCHK(hipEventRecord(a->ev1, st));
    a->timed = true;
    const u32* cp = a->d_partials;
    u32 nb = (u32)grid;
    uint8_t* img = d_images;
    uint64_t* cnt = reinterpret_cast<uint64_t*>(d_counts);
    void* pf[] = {&cp, &nb, &img, &cnt};
    HIPCHK(hipModuleLaunchKernel(ffin, (unsigned)(NP * ngroups + ngroups), 1, 1, 256, 1, 1, 0, st, pf, nullptr));
    return launch_avgs(a, d_images, d_counts, ngroups, max_groups, st);
}

int updec_agg_run(updec_agg_t a, const uint8_t* const* d_cols, const int32_t* d_filter, const uint8_t* d_key0,
                  const uint8_t* d_key1, int64_t rows, uint8_t* d_images, int64_t* d_counts, int32_t max_groups,
                  int32_t* ngroups_out, uint16_t* keys_out, updec_stream_t stream)
{
    if (!a || rows < 0 || !d_images || !d_counts || !ngroups_out) return fail(UPDEC_EINVAL, "bad argument");
    if (int rc = check_cols(d_cols, a->types.size(), rows)) return rc;
    if (max_groups < 1 || max_groups > 255) return fail(UPDEC_EINVAL, "max_groups must be in [1, 255]");
    if (a->has_filter && !d_filter) return fail(UPDEC_EINVAL, "filter column missing");
    if (a->nkeys >= 1 && !d_key0) return fail(UPDEC_EINVAL, "key column missing");
    if (a->nkeys == 2 && !d_key1) return fail(UPDEC_EINVAL, "second key column missing");
    hipStream_t st = S(stream);
    if (a->onepass) {
        bool overflowed = false;
        int rc = agg_run_onepass(a, d_cols, d_filter, d_key0, d_key1, rows, d_images, d_counts, max_groups, ngroups_out,
                                 keys_out, st, &overflowed);
        if (rc != UPDEC_OK || !overflowed) return rc;
        // more distinct keys in a block than the one-pass table holds: two-pass form
    }
    return agg_run_twopass(a, nullptr, d_cols, d_filter, d_key0, d_key1, rows, d_images, d_counts, max_groups,
                           ngroups_out, keys_out, st);
}

int updec_agg_run_blocks(updec_agg_t a, const updec_column_blocks* cols, const int32_t* d_indices, int64_t rows,
                         uint8_t* d_images, int64_t* d_counts, updec_stream_t stream)
{
    if (!a || rows < 0 || !d_images || !d_counts || (!cols && !a->types.empty())) return fail(UPDEC_EINVAL, "bad argument");
    if (a->has_filter || a->nkeys > 0)
        return fail(UPDEC_EUNSUPPORTED, "block-column inputs take ungrouped, unfiltered aggregations (see updec_agg_run_iter)");
    IterArgs args;
    int rc = fill_blocks(&args, a->types, cols, d_indices, rows);
    if (rc != UPDEC_OK) return rc;
    int32_t ng = 0;
    return agg_run_twopass(a, &args, nullptr, nullptr, nullptr, nullptr, rows, d_images, d_counts, 1, &ng, nullptr,
                           S(stream));
}

int updec_agg_run_iter(updec_agg_t a, const updec_column_iter* cols, int64_t rows, const int32_t* d_filter,
                       const uint8_t* d_key0, const uint8_t* d_key1, uint8_t* d_images, int64_t* d_counts,
                       int32_t max_groups, int32_t* ngroups_out, uint16_t* keys_out, updec_stream_t stream)
{
    if (!a || rows < 0 || !d_images || !d_counts || !ngroups_out) return fail(UPDEC_EINVAL, "bad argument");
    if (max_groups < 1 || max_groups > 255) return fail(UPDEC_EINVAL, "max_groups must be in [1, 255]");
    if (a->has_filter && !d_filter) return fail(UPDEC_EINVAL, "filter column missing");
    if (a->nkeys >= 1 && !d_key0) return fail(UPDEC_EINVAL, "key column missing");
    if (a->nkeys == 2 && !d_key1) return fail(UPDEC_EINVAL, "second key column missing");
    IterArgs args;
    int rc = fill_iters(&args, a->types, cols, rows);
    if (rc != UPDEC_OK) return rc;
    if (args.nullcols && a->spec.avg_mask)
        return fail(UPDEC_EUNSUPPORTED, "AVG over columns that can be NULL needs the non-NULL count: divide the SUM "
                                        "images with updec_div_int64 instead of avg_mask");
    return agg_run_twopass(a, &args, nullptr, d_filter, d_key0, d_key1, rows, d_images, d_counts, max_groups,
                           ngroups_out, keys_out, S(stream));
}

int updec_agg_last_kernel_ms(updec_agg_t a, float* ms)
{
    if (!a || !ms || !a->timed) return fail(UPDEC_EINVAL, "no timed run");
    HIPCHK(hipEventSynchronize(a->ev1));
    HIPCHK(hipEventElapsedTime(ms, a->ev0, a->ev1));
    return UPDEC_OK;
}

int updec_agg_source(updec_agg_t a, int32_t ngroups, char* buf, size_t cap, size_t* needed)
{
    if (!a || ngroups < 0 || ngroups > 255) return fail(UPDEC_EINVAL, "bad argument");
    if (ngroups == 0 && !a->onepass) return fail(UPDEC_EINVAL, "no one-pass form for this aggregation");
    AggSpec s = a->spec;
    s.onepass = ngroups == 0;
    s.ngroups = ngroups;
    GenOptions o;
    int aw = 0, blk = 0;
    std::string src = gen_agg_kernels(s, a->types, form_name(a, s.onepass), o, &aw, &blk);
    if (src.empty()) return fail(UPDEC_EUNSUPPORTED, "accumulators do not fit in LDS");
    return copy_text(src, buf, cap, needed);
}

int updec_agg_source_iter(updec_agg_t a, int32_t ngroups, char* buf, size_t cap, size_t* needed)
{
    if (!a || ngroups < 1 || ngroups > 255) return fail(UPDEC_EINVAL, "bad argument");
    AggSpec s = a->spec;
    s.onepass = false;
    s.ngroups = ngroups;
    GenOptions o;
    o.blocks = true;
    int aw = 0, blk = 0;
    std::string src = gen_agg_kernels(s, a->types, form_name(a, false), o, &aw, &blk);
    if (src.empty()) return fail(UPDEC_EUNSUPPORTED, "accumulators do not fit in LDS");
    return copy_text(src, buf, cap, needed);
}

void updec_agg_free(updec_agg_t a) { delete a; }

int updec_div_int64(const uint8_t* d_sums, int64_t stride_bytes, const int64_t* d_counts, uint8_t* d_out, int64_t n,
                    int mode, updec_stream_t stream)
{
    if (n < 0 || (n && (!d_sums || !d_counts || !d_out)) || stride_bytes < 133) return fail(UPDEC_EINVAL, "bad argument");
    if (n == 0) return UPDEC_OK;
    unsigned blocks = (unsigned)((n + 63) / 64);
    if (mode == UPDEC_MODE_EXTENDED)
        hipLaunchKernelGGL(k_div_int64<0>, dim3(blocks), dim3(64), 0, S(stream), d_sums, stride_bytes, d_counts, d_out, n,
                           (int64_t)133);
    else
        hipLaunchKernelGGL(k_div_int64<1>, dim3(blocks), dim3(64), 0, S(stream), d_sums, stride_bytes, d_counts, d_out, n,
                           (int64_t)133);
    HIPCHK(hipGetLastError());
    return UPDEC_OK;
}

int updec_module_cache_stats(int64_t* rtc_compiles, int64_t* memory_hits, int64_t* disk_hits)
{
    if (rtc_compiles) *rtc_compiles = g_rtc_compiles.load();
    if (memory_hits) *memory_hits = g_mem_hits.load();
    if (disk_hits) *disk_hits = g_disk_hits.load();
    return UPDEC_OK;
}

int updec_compile(const char* src, size_t len, updec_module_t* out)
{
    if (!src || !out) return fail(UPDEC_EINVAL, "bad argument");
    std::string err;
    ModuleP m = ModuleCache::get().compile(std::string(src, len), &err);
    if (!m) return fail(UPDEC_ECOMPILE, err);
    *out = new updec_module_s{m};
    return UPDEC_OK;
}

int updec_launch(updec_module_t m, const char* fn, uint32_t grid, uint32_t block, const void* args, size_t args_size,
                 updec_stream_t stream)
{
    if (!m || !fn || grid == 0 || block == 0 || block > 1024) return fail(UPDEC_EINVAL, "bad argument");
    std::string err;
    hipFunction_t f = m->mod->fn(fn, &err);
    if (!f) return fail(UPDEC_ECOMPILE, err);
    size_t sz = args_size;
    std::vector<uint8_t> buf((const uint8_t*)args, (const uint8_t*)args + args_size);
    void* extra[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, buf.data(), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
    HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, S(stream), nullptr, extra));
    return UPDEC_OK;
}

void updec_module_free(updec_module_t m) { delete m; }

int updec_check_compile(const char* src, size_t len)
{
    if (!src) return fail(UPDEC_EINVAL, "bad argument");
    std::vector<char> code;
    std::string err;
    if (!ModuleCache::rtc(std::string(src, len), &code, &err)) return fail(UPDEC_ECOMPILE, err);
    return UPDEC_OK;
}

int updec_gen_compact(const updec_gen_desc* g, int64_t row0, int64_t rows, uint8_t* out, int on_device,
                      updec_stream_t stream)
{
    if (!g || !out || rows < 0) return fail(UPDEC_EINVAL, "bad argument");
    int nb = compact_bytes(g->precision);
    if (nb <= 0 || nb > 128) return fail(UPDEC_EINVAL, "precision out of generator range");
    if (g->kind == UPDEC_GEN_DIGITS && (g->digits < 1 || g->digits > 460)) return fail(UPDEC_EINVAL, "digits out of range");
    if (g->kind == UPDEC_GEN_BELOW && (g->bound_words < 1 || g->bound_words > updec_gen::GEN_MAX_BOUND_WORDS))
        return fail(UPDEC_EINVAL, "bound_words out of range");
    if (g->kind < UPDEC_GEN_RANGE || g->kind > UPDEC_GEN_BELOW) return fail(UPDEC_EINVAL, "unknown generator kind");
    updec_gen::Desc d{};
    d.kind = g->kind; d.precision = g->precision; d.digits = g->digits; d.neg_permille = g->neg_permille;
    d.lo = g->lo; d.hi = g->hi; d.step = g->step; d.seed = g->seed; d.column_id = g->column_id;
    d.bound_words = g->kind == UPDEC_GEN_BELOW ? g->bound_words : 0;
    for (int i = 0; i < d.bound_words; i++) d.bound[i] = g->bound[i];
    uint64_t key = updec_gen::column_key(g->seed, g->column_id);
    if (on_device) {
        if (rows == 0) return UPDEC_OK;
        hipLaunchKernelGGL(k_gen_compact, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, S(stream), d, key, row0, rows,
                           out, nb);
        HIPCHK(hipGetLastError());
        return UPDEC_OK;
    }
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (rows < 65536) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        int64_t a = rows * t / nt, b = rows * (t + 1) / nt;
        th.emplace_back([=]() {
            for (int64_t i = a; i < b; i++) updec_gen::gen_row(d, key, (uint64_t)(row0 + i), out + i * nb, nb);
        });
    }
    for (auto& x : th) x.join();
    return UPDEC_OK;
}

int updec_gen_int32(uint64_t seed, uint32_t column_id, int32_t lo, int32_t hi, int64_t row0, int64_t rows, int32_t* out,
                    int on_device, updec_stream_t stream)
{
    if (!out || rows < 0 || hi < lo) return fail(UPDEC_EINVAL, "bad argument");
    uint64_t key = updec_gen::column_key(seed, column_id);
    if (on_device) {
        if (rows == 0) return UPDEC_OK;
        hipLaunchKernelGGL(k_gen_int32, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, S(stream), key, lo, hi, row0, rows,
                           out);
        HIPCHK(hipGetLastError());
        return UPDEC_OK;
    }
    for (int64_t i = 0; i < rows; i++) out[i] = updec_gen::gen_int32(key, (uint64_t)(row0 + i), lo, hi);
    return UPDEC_OK;
}

int updec_gen_byte(uint64_t seed, uint32_t column_id, const uint8_t* choices, int32_t nchoices, int64_t row0,
                   int64_t rows, uint8_t* out, int on_device, updec_stream_t stream)
{
    if (!out || !choices || nchoices <= 0 || nchoices > 64 || rows < 0) return fail(UPDEC_EINVAL, "bad argument");
    uint64_t key = updec_gen::column_key(seed, column_id);
    if (on_device) {
        if (rows == 0) return UPDEC_OK;
        ByteChoices ch{};
        memcpy(ch.c, choices, nchoices);
        ch.n = nchoices;
        hipLaunchKernelGGL(k_gen_byte, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, S(stream), key, ch, row0, rows, out);
        HIPCHK(hipGetLastError());
        return UPDEC_OK;
    }
    for (int64_t i = 0; i < rows; i++) out[i] = updec_gen::gen_byte(key, (uint64_t)(row0 + i), choices, nchoices);
    return UPDEC_OK;
}

}  // extern "C"
