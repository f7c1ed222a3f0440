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
  w.prepare_run(256, 256);
  const std::vector<uint64_t> h = w.run(256, 256);
  std::printf("hash %llu\n", (unsigned long long)h[0]);
  // the same ticks with skipping off give the same state
  fs3d::World d({64, 64, 64}, FS3D_WORLD_DENSE);
  d.fill_rain(0, fs3d::threshold(3, 10));
  d.step(256);
  // fused pairs need whole rows per wave: nx > 256
  fs3d::World f({320, 16, 16}, FS3D_WORLD_DENSE);
  const bool small_fused = d.fuse();
  f.set_fuse(true);
  std::printf("skip %d dense-skip %d fuse %d/%d same %d ms %d\n", w.skip(), d.skip(), (int)small_fused,
              (int)f.fuse(), (int)(d.hash() == w.hash()), (int)(w.last_run_ms_per_tick() > 0.0));
  const float pos[3] = {0.0f, 0.0f, -5.0f}, rot[3] = {0.0f, 0.0f, 0.0f};
  const fs3d_camera cam = w.key_camera(pos, rot, 1.5f);
  // + 0.0f turns a negative zero positive
  std::printf("camera forward %.1f %.1f %.1f\n", cam.basis[6] + 0.0f, cam.basis[7] + 0.0f, cam.basis[8] + 0.0f);
  try {
    w.set_voxel(64, 0, 0, 1);
    return 1;
  } catch (const std::runtime_error &) {
  }
  return 0;
}
