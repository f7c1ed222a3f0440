// Drives the library through include/fs3d.hpp the way a C++ host would.
// Without a GPU it checks that world creation throws std::runtime_error; with
// one it runs the 64^3 rain scene for 256 ticks and prints the final hash.
#include <cstdio>
#include <stdexcept>

#include "fs3d.hpp"

int main() {
  std::printf("abi %d devices %d\n", fs3d_abi_version(), fs3d::device_count());
  if (fs3d::device_count() == 0) {
    try {
      fs3d::World w({8, 8, 8});
      std::printf("unexpected: world created without a device\n");
      return 1;
    } catch (const std::runtime_error &e) {
      std::printf("no-device error: %s\n", e.what());
      return 0;
    }
  }
  fs3d::World w({64, 64, 64});
  w.fill_rain(0, fs3d::threshold(3, 10));
  const std::vector<uint64_t> h = w.run(256, 256);
  std::printf("hash %llu\n", (unsigned long long)h[0]);
  try {
    w.set_voxel(64, 0, 0, 1);
    return 1;
  } catch (const std::runtime_error &) {
  }
  return 0;
}
