#include "common.hpp"

namespace cfdx {
std::string &last_error_slot()
{
    static thread_local std::string msg;
    return msg;
}
}  // namespace cfdx

extern "C" const char *cfdx_last_error(void) { return cfdx::last_error_slot().c_str(); }
