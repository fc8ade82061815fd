// Library-internal helpers shared between translation units.
#pragma once
#include <string>

namespace inlm {

// message returned by inlm_last_error(NULL)
void set_global_error(const std::string &msg);

}  // namespace inlm
