/* dump_io.h -- cell-field dump (drop-in for the reference's
 * include/dump_io.h:8). */
#pragma once

#include <string>

#include "types.h"

namespace mpcd {

void write_cell_fields_bin(const std::string& filepath, const CellFields& G);

}  // namespace mpcd
