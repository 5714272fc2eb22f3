// Definitions shared by the C-ABI translation units (not part of the ABI).
#pragma once
#include "../host/fs_setup.hpp"

struct fs_mesh {
  fs::Mesh m;
  bool prepared = false;
  int order_prepared = -1;
};

namespace fse {
// text of a per-instance device error flag
const char* device_error_message(int e);
// the calling thread's last error text
const char* last_error_text();
}  // namespace fse
