// Grid HDF5 writer (placeholder until the writer lands).
#include "tc3_ctx.hpp"

namespace tc3 {
int write_grid_hdf5(const tc3_ctx *, int, const char *, std::string &err) {
    err = "grid HDF5 output is not available in this build";
    return 6;
}
}  // namespace tc3
