// Library-wide C entry points: error text and ABI version.
#include <string>

#include "dpmi_internal.h"

namespace dpmi {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace dpmi

extern "C" const char* dpmi_last_error(void) { return dpmi::g_last_error.c_str(); }

extern "C" int dpmi_abi_version(void) { return 1; }
