// Error reporting for a host-only build of planner.cpp: the planner's entry
// points report failures through nf::fail, which the product library defines
// beside its HIP runtime.  This definition keeps the message for
// nf_last_error, so that planner.cpp links and runs without HIP under the
// address and undefined-behaviour sanitizers.
#include <string>

namespace nf {
static thread_local std::string last_error;
int fail(int code, const std::string &msg) {
    last_error = msg;
    return code;
}
}  // namespace nf

extern "C" const char *nf_last_error() { return nf::last_error.c_str(); }
