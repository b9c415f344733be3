// Error reporting shared by the library's source files: the message is
// returned by fvdbm_last_error() of the calling thread.
#pragma once

#include <string>

namespace fvdbm {
int set_error(int code, const std::string &msg);
}  // namespace fvdbm
