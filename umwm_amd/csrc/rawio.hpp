// Self-describing raw binary container for the model's gridded output and
// restart files (the roles of the reference's umwmout_*.nc and
// umwmrst_*.nc, src/umwm_io.F90:317-854, src/umwm_restart.F90:18-145):
// the same variable names, dimension names, shapes (Fortran order, fastest
// first) and units, without NetCDF.  Layout, little-endian:
//   "UMWMRAW1"  int32 nvars
//   per variable: str name, str units, str description, int32 type
//   (1 float32, 2 float64, 3 int32), int32 ndims, ndims x (str dim, int64
//   size), then prod(size) values
// where str is int32 length + bytes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace umwm {

enum RawType : int32_t { RAW_F32 = 1, RAW_F64 = 2, RAW_I32 = 3 };

struct RawVar {
  std::string name, units, description;
  int32_t type = RAW_F32;
  std::vector<std::string> dims;  // fastest first
  std::vector<int64_t> shape;
  std::vector<double> data;       // prod(shape) values, Fortran order
};

// 0 on success, else a message in err.
int raw_write(const std::string& path, const std::vector<RawVar>& vars, std::string& err);
int raw_read(const std::string& path, std::vector<RawVar>& vars, std::string& err);

}  // namespace umwm
