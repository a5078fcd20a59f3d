// api_internal.h -- helpers shared by the C ABI translation units
// (updec_api.hip, vm.hip, segsum.hip).
#pragma once
#include <string>
#include <vector>

#include "../../include/updec.h"
#include "planner.h"

namespace updec {

// record the calling thread's error message, return `code`
int api_fail(int code, const std::string& msg);
// compute units of the current device (cached)
int api_device_cus();
// column descriptors -> planner types (and names)
std::vector<ColumnType> api_types(const updec_column_desc* cols, int n, std::vector<std::string>* names);
// parse -> AlignOptimize -> plan; mode AUTO picks faithful inside the envelope
bool api_make_plan(const std::string& text, const std::vector<std::string>& names,
                   const std::vector<ColumnType>& types, int mode, Plan* out, std::string* err);
// device address of the segmented SUM's fixed scratch on the current device
// (segsum.hip), looked up once per device
int segsum_scratch_setup();

}  // namespace updec
