// fs3d.hpp -- header-only C++ wrapper over the fs3d C ABI (include/fs3d.h).
//
// RAII handles for worlds, batches and servers.  A failing call is reported the
// way the reference reports errors (util::displayError, src/util/debug.cpp:23-27):
// "ERROR: <message>" on stderr, then std::runtime_error(message) is thrown.
//
// Typical use at the reference's engine-loop seam (src/engine/engine.cpp:59-69):
//
//   fs3d::World world({256, 256, 256});
//   world.fill_rain(seed, fs3d::threshold(3, 10));
//   while (running) {
//     handle_events();
//     world.step(1);                        // the tick, asynchronous
//     world.render(camera, W, H, rgb, hit, face);  // hand the grid to the screen
//   }
#ifndef FS3D_HPP
#define FS3D_HPP

#include <array>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fs3d.h"

namespace fs3d {

[[noreturn]] inline void fail(const std::string &msg) {
  std::cerr << "ERROR: " << msg << std::endl;
  throw std::runtime_error(msg);
}

inline void check(int rc) {
  if (rc != FS3D_OK) fail(fs3d_last_error());
}

inline uint64_t threshold(uint64_t num, uint64_t den) {
  uint64_t t = 0;
  check(fs3d_threshold(num, den, &t));
  return t;
}

inline int device_count() {
  int n = 0;
  check(fs3d_device_count(&n));
  return n;
}

class World {
 public:
  explicit World(std::array<int32_t, 3> dims, uint32_t flags = 0) : dims_(dims) {
    check(fs3d_world_create(dims_.data(), flags, &w_));
  }
  ~World() { fs3d_world_destroy(w_); }
  World(const World &) = delete;
  World &operator=(const World &) = delete;
  World(World &&o) noexcept : w_(o.w_), dims_(o.dims_) { o.w_ = nullptr; }

  fs3d_world *handle() const { return w_; }
  const std::array<int32_t, 3> &dims() const { return dims_; }
  size_t size() const { return (size_t)dims_[0] * dims_[1] * dims_[2]; }

  void set_voxel(int32_t x, int32_t y, int32_t z, uint8_t v) { check(fs3d_world_set_voxel(w_, x, y, z, v)); }
  uint8_t get_voxel(int32_t x, int32_t y, int32_t z) const {
    uint8_t v = 0;
    check(fs3d_world_get_voxel(w_, x, y, z, &v));
    return v;
  }
  void upload(const std::vector<uint8_t> &g) { check(fs3d_world_upload(w_, g.data(), g.size())); }
  std::vector<uint8_t> download() const {
    std::vector<uint8_t> g(size());
    check(fs3d_world_download(w_, g.data(), g.size()));
    return g;
  }
  uint64_t tick_count() const {
    uint64_t t = 0;
    check(fs3d_world_get_tick(w_, &t));
    return t;
  }
  void set_tick(uint64_t t) { check(fs3d_world_set_tick(w_, t)); }

  void step(int32_t nticks = 1) { check(fs3d_world_step(w_, nticks)); }
  uint64_t tick() {
    uint64_t h = 0;
    check(fs3d_world_tick(w_, &h));
    return h;
  }
  uint64_t hash() {
    uint64_t h = 0;
    check(fs3d_world_hash(w_, &h));
    return h;
  }
  std::vector<uint64_t> run(int32_t nticks, int32_t hash_every = 0) {
    std::vector<uint64_t> h(hash_every > 0 ? nticks / hash_every : 0);
    check(fs3d_world_run(w_, nticks, hash_every, h.empty() ? nullptr : h.data()));
    return h;
  }
  void prepare_run(int32_t nticks, int32_t hash_every = 0) {
    check(fs3d_world_prepare_run(w_, nticks, hash_every));
  }
  double last_run_ms_per_tick() const {
    double ms = 0.0;
    check(fs3d_world_last_run_ms(w_, &ms));
    return ms;
  }
  void sync() { check(fs3d_world_sync(w_)); }

  void save(const std::string &path) { check(fs3d_world_save(w_, path.c_str())); }
  void load(const std::string &path) { check(fs3d_world_load(w_, path.c_str())); }
  void set_skip(int32_t mode) { check(fs3d_world_set_skip(w_, mode)); }
  int32_t skip() const {
    int32_t m = 0;
    check(fs3d_world_get_skip(w_, &m));
    return m;
  }
  void set_fuse(bool on) { check(fs3d_world_set_fuse(w_, on ? 1 : 0)); }
  // FS3D_TRACK_AUTO, FS3D_TRACK_ON or FS3D_TRACK_OFF: renderer occupancy tracking between frames
  void set_render_track(int mode) { check(fs3d_world_set_render_track(w_, mode)); }
  bool fuse() const {
    int32_t on = 0;
    check(fs3d_world_get_fuse(w_, &on));
    return on != 0;
  }
  void set_segment(int32_t seg_bytes) { check(fs3d_world_set_segment(w_, seg_bytes)); }

  void fill_rain(uint64_t seed, uint64_t thr) { check(fs3d_world_fill_rain(w_, seed, thr)); }
  int64_t fill_boxes(uint64_t seed, const fs3d_box_params &p) {
    int64_t n = 0;
    check(fs3d_world_fill_boxes(w_, seed, &p, &n));
    return n;
  }
  void fill_pillars(uint64_t seed, const fs3d_pillar_params &p) { check(fs3d_world_fill_pillars(w_, seed, &p)); }

  void render(const fs3d_camera &cam, int32_t W, int32_t H, std::vector<float> &rgb, std::vector<uint64_t> &hit,
              std::vector<uint8_t> &face) {
    rgb.resize((size_t)W * H * 3);
    hit.resize((size_t)W * H);
    face.resize((size_t)W * H);
    check(fs3d_world_render(w_, &cam, W, H, rgb.data(), hit.data(), face.data()));
  }
  // The voxel camera for the reference's key-driven camera state.
  fs3d_camera key_camera(const float cam_pos[3], const float cam_rot[3], float aspect) const {
    fs3d_camera cam;
    check(fs3d_key_camera(dims_.data(), cam_pos, cam_rot, aspect, &cam));
    return cam;
  }

 private:
  fs3d_world *w_ = nullptr;
  std::array<int32_t, 3> dims_;
};

class Batch {
 public:
  explicit Batch(const std::vector<std::array<int32_t, 3>> &dims) : n_((int32_t)dims.size()) {
    std::vector<int32_t> flat;
    for (const auto &d : dims) flat.insert(flat.end(), d.begin(), d.end());
    check(fs3d_batch_create(n_, flat.data(), &b_));
  }
  ~Batch() { fs3d_batch_destroy(b_); }
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  fs3d_batch *handle() const { return b_; }
  void fill_rain(const std::vector<uint64_t> &seeds, const std::vector<uint64_t> &thr) {
    check(fs3d_batch_fill_rain(b_, seeds.data(), thr.data()));
  }
  std::vector<uint64_t> run(int32_t nticks, int32_t mode = FS3D_BATCH_TUNED) {
    std::vector<uint64_t> h(n_);
    check(fs3d_batch_run(b_, nticks, mode, h.data()));
    return h;
  }
  void autotune(int32_t trial_ticks = 10) {
    int32_t s, t, g, m;
    double ms;
    check(fs3d_batch_autotune(b_, trial_ticks, &s, &t, &g, &m, &ms));
  }

 private:
  fs3d_batch *b_ = nullptr;
  int32_t n_;
};

class Server {
 public:
  explicit Server(const fs3d_server_config &cfg) { check(fs3d_server_create(&cfg, &s_)); }
  ~Server() { fs3d_server_destroy(s_); }
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  int64_t submit(std::array<int32_t, 3> dims, uint64_t seed, uint64_t thr, int32_t ticks) {
    int64_t id = -1;
    check(fs3d_server_submit(s_, dims.data(), seed, thr, ticks, &id));
    return id;
  }
  void poll() { check(fs3d_server_poll(s_, nullptr, nullptr, nullptr)); }
  void drain() { check(fs3d_server_drain(s_)); }
  uint64_t result_hash(int64_t id) const {
    int32_t st = 0;
    uint64_t h = 0;
    double lat = 0;
    check(fs3d_server_result(s_, id, &st, &h, &lat));
    if (st != FS3D_REQ_DONE) fail("request " + std::to_string(id) + " is not done");
    return h;
  }
  fs3d_server_stats stats() const {
    fs3d_server_stats st;
    check(fs3d_server_get_stats(s_, &st));
    return st;
  }

 private:
  fs3d_server *s_ = nullptr;
};

}  // namespace fs3d

#endif
