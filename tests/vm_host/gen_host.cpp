// gen_host.cpp -- TEST INFRASTRUCTURE: the kernel generators of planner.cpp
// built for the host, exporting the generated source of a variant the C ABI
// builds only at first use (block-column inputs), so the CPU suite can
// compile it with hipRTC and read it.
#include <cstring>
#include <string>
#include <vector>

#include "planner.h"

using namespace updec;

static std::string g_src;

extern "C" const char* gen_blocks_source(const char* text, const char* const* names, const int* P, const int* S, int n,
                                         int agg)
{
    std::vector<std::string> nm(names, names + n);
    std::vector<ColumnType> types(n);
    for (int i = 0; i < n; i++) types[i] = ColumnType{P[i], S[i], 0};
    Plan p;
    std::string err;
    if (!plan_expression(text, nm, types, 0, &p, &err)) {
        g_src = "error: " + err;
        return g_src.c_str();
    }
    GenOptions o;
    o.blocks = true;
    if (agg) {
        AggSpec s;
        s.plans = {p};
        s.ngroups = 1;
        int aw = 0, blk = 0;
        g_src = gen_agg_kernels(s, types, "agg_blocks", o, &aw, &blk);
    } else {
        g_src = gen_calc_kernel(p, types, "calc_blocks", o);
    }
    return g_src.c_str();
}
