/* params_io.h -- params.kv reader and runout writer (drop-in for the
 * reference's include/params_io.h:8-9). */
#pragma once

#include <string>

#include "types.h"

namespace mpcd {

Params read_params_kv(const std::string& filepath);
void write_runout_kv(const std::string& filepath, const RunOutStep& runout);

}  // namespace mpcd
