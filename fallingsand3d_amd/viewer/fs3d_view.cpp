// fs3d_view -- a headless viewer shell over the fs3d C ABI.
//
// It mirrors the reference's interactive loop:
//   - VulkanEngine::run (src/engine/engine.cpp:59-69): every frame handles the
//     pending window events, then (here) ticks the world, then draws unless the
//     window is minimised;
//   - Window::handleEvents (src/engine/window.cpp:34-107): key-down / key-up
//     events set and clear "holding" flags for W A S D, left CTRL, SPACE and
//     the LEFT / RIGHT arrows; a quit event ends the loop; minimise / restore
//     toggle drawing;
//   - Renderer::draw (src/engine/rendering/renderer.cpp:437-467): each drawn
//     frame first advances the camera by one fixed 0.016 s step for the held
//     keys (fs3d_camera_integrate), then renders (here: the voxel ray caster
//     from fs3d_key_camera, written as a binary PPM).
//
// The viewer opens no window, so events come from a script (a
// file, or "-" for stdin), one per line; '#' starts a comment:
//   down KEY | up KEY    KEY: W A S D CTRL SPACE LEFT RIGHT
//   frames N             run N frames of the loop with the current key state
//   minimize | restore   window minimised / restored (no drawing while minimised)
//   set X Y Z V          a brush edit: voxel (X, Y, Z) becomes material V
//   snap                 write the next drawn frame even off the --every cadence
//   quit                 end the loop (the rest of the script is not read)
// Events take effect at the start of the next frame, as SDL_PollEvent delivers
// them.  The whole script is parsed before the GPU is touched, so a malformed
// line is reported without running anything.
//
// Output: frame_NNNNNN.ppm every --every drawn frames (0: only on snap) in
// --out, and one summary line on stdout:
//   frames F drawn D ticks T hash H cam X Y Z yaw R tick_ms A render_ms B
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "fs3d.hpp"

namespace {

struct Event {
  enum Kind { Down, Up, Frames, Minimize, Restore, Set, Snap, Quit } kind;
  uint32_t key = 0;
  long long n = 0;
  int x = 0, y = 0, z = 0, v = 0;
  int line = 0;
};

bool key_bit(const std::string &name, uint32_t *bit) {
  static const struct {
    const char *name;
    uint32_t bit;
  } keys[] = {{"W", FS3D_KEY_W},         {"A", FS3D_KEY_A},         {"S", FS3D_KEY_S},
              {"D", FS3D_KEY_D},         {"CTRL", FS3D_KEY_CTRL},   {"SPACE", FS3D_KEY_SPACE},
              {"LEFT", FS3D_KEY_LEFT},   {"RIGHT", FS3D_KEY_RIGHT}};
  for (const auto &k : keys)
    if (name == k.name) {
      *bit = k.bit;
      return true;
    }
  return false;
}

std::vector<Event> parse_script(std::istream &in, const std::string &src) {
  std::vector<Event> ev;
  std::string raw;
  int line = 0;
  auto bad = [&](const std::string &why) { fs3d::fail(src + ":" + std::to_string(line) + ": " + why); };
  while (std::getline(in, raw)) {
    line++;
    const size_t hash = raw.find('#');
    if (hash != std::string::npos) raw.resize(hash);
    std::istringstream ls(raw);
    std::string word;
    if (!(ls >> word)) continue;
    Event e{};
    e.line = line;
    if (word == "down" || word == "up") {
      std::string k;
      if (!(ls >> k) || !key_bit(k, &e.key)) bad("expected a key (W A S D CTRL SPACE LEFT RIGHT)");
      e.kind = word == "down" ? Event::Down : Event::Up;
    } else if (word == "frames") {
      if (!(ls >> e.n) || e.n < 0) bad("frames needs a count >= 0");
      e.kind = Event::Frames;
    } else if (word == "minimize") {
      e.kind = Event::Minimize;
    } else if (word == "restore") {
      e.kind = Event::Restore;
    } else if (word == "set") {
      if (!(ls >> e.x >> e.y >> e.z >> e.v) || e.v < 0 || e.v > 255) bad("set needs X Y Z V (V in 0..255)");
      e.kind = Event::Set;
    } else if (word == "snap") {
      e.kind = Event::Snap;
    } else if (word == "quit") {
      e.kind = Event::Quit;
    } else {
      bad("unknown event '" + word + "'");
    }
    std::string extra;
    if (ls >> extra) bad("unexpected '" + extra + "'");
    ev.push_back(e);
  }
  return ev;
}

void write_ppm(const std::string &path, int W, int H, const std::vector<float> &rgb) {
  std::ofstream f(path, std::ios::binary);
  if (!f) fs3d::fail("cannot write " + path + ": " + std::strerror(errno));
  f << "P6\n" << W << " " << H << "\n255\n";
  std::vector<unsigned char> row((size_t)W * 3);
  for (int y = 0; y < H; y++) {
    for (int i = 0; i < W * 3; i++) {
      float v = rgb[(size_t)y * W * 3 + i];
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      row[i] = (unsigned char)(v * 255.0f + 0.5f);
    }
    f.write(reinterpret_cast<const char *>(row.data()), (std::streamsize)row.size());
  }
  if (!f) fs3d::fail("write failed: " + path);
}

void usage() {
  std::cerr << "usage: fs3d_view --script FILE|- [--dims X Y Z] [--scene rain|boxes|pillars] [--seed N]\n"
               "                 [--ticks-per-frame K] [--size W H] [--every N] [--out DIR]\n";
}

}  // namespace

int main(int argc, char **argv) {
  int32_t dims[3] = {128, 128, 128};
  std::string scene = "boxes", script, out = ".";
  uint64_t seed = 1;
  int tpf = 1, W = 320, H = 180, every = 10;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) {
        usage();
        std::exit(2);
      }
    };
    if (a == "--dims") {
      need(3);
      for (int k = 0; k < 3; k++) dims[k] = std::atoi(argv[++i]);
    } else if (a == "--scene") {
      need(1);
      scene = argv[++i];
    } else if (a == "--seed") {
      need(1);
      seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (a == "--ticks-per-frame") {
      need(1);
      tpf = std::atoi(argv[++i]);
    } else if (a == "--size") {
      need(2);
      W = std::atoi(argv[++i]);
      H = std::atoi(argv[++i]);
    } else if (a == "--every") {
      need(1);
      every = std::atoi(argv[++i]);
    } else if (a == "--out") {
      need(1);
      out = argv[++i];
    } else if (a == "--script") {
      need(1);
      script = argv[++i];
    } else {
      usage();
      return 2;
    }
  }
  if (script.empty() || tpf < 0 || W < 1 || H < 1 || every < 0 ||
      (scene != "rain" && scene != "boxes" && scene != "pillars")) {
    usage();
    return 2;
  }
  try {
    std::vector<Event> events;
    if (script == "-") {
      events = parse_script(std::cin, "<stdin>");
    } else {
      std::ifstream f(script);
      if (!f) fs3d::fail("cannot read " + script);
      events = parse_script(f, script);
    }

    fs3d::World world({dims[0], dims[1], dims[2]});
    if (scene == "rain") {
      world.fill_rain(seed, fs3d::threshold(3, 10));
    } else if (scene == "boxes") {
      const fs3d_box_params bp{8, 32, 20, 10, 5, 1 << 20};
      world.fill_boxes(seed, bp);
    } else {
      // the 2048^3 pillar scene's parameters scaled to this grid
      const double sx = std::min(dims[0], dims[1]) / 2048.0, sz = dims[2] / 2048.0;
      auto at_least = [](double v, int lo) { return std::max(lo, (int)(v + 0.5)); };
      fs3d_pillar_params pp;
      pp.n_pillars = at_least(256.0 * sx * sx, 1);
      pp.r_min = at_least(4.0 * sx, 1);
      pp.r_max = std::max(pp.r_min, at_least(32.0 * sx, 1));
      pp.h_min = at_least(256.0 * sz, 1);
      pp.h_max = std::max(pp.h_min, at_least(1800.0 * sz, 1));
      pp.pool_w_min = at_least(256.0 * sx, 1);
      pp.pool_w_max = std::max(pp.pool_w_min, at_least(682.0 * sx, 1));
      pp.pool_d_min = pp.pool_w_min;
      pp.pool_d_max = pp.pool_w_max;
      pp.pool_h_min = at_least(32.0 * sz, 1);
      pp.pool_h_max = std::max(pp.pool_h_min, at_least(128.0 * sz, 1));
      world.fill_pillars(seed, pp);
    }

    // the reference's initial camera (src/engine/rendering/renderer.h:148-149)
    float cam_pos[3] = {0.0f, 0.0f, -5.0f}, cam_rot[3] = {0.0f, 0.0f, 0.0f};
    uint32_t held = 0;
    bool minimized = false, quit = false, snap = false;
    long long frames = 0, drawn = 0, ticks = 0;
    double tick_ms = 0.0, render_ms = 0.0;
    std::vector<float> rgb;
    std::vector<uint64_t> hit;
    std::vector<uint8_t> face;
    std::vector<Event> pending;
    const float aspect = (float)W / (float)H;

    auto run_frame = [&]() {
      // handleEvents: deliver what arrived since the last frame
      for (const Event &e : pending) {
        switch (e.kind) {
          case Event::Down: held |= e.key; break;
          case Event::Up: held &= ~e.key; break;
          case Event::Minimize: minimized = true; break;
          case Event::Restore: minimized = false; break;
          case Event::Set: world.set_voxel(e.x, e.y, e.z, (uint8_t)e.v); break;
          case Event::Snap: snap = true; break;
          default: break;
        }
      }
      pending.clear();
      // the tick at the engine-loop seam
      auto t0 = std::chrono::steady_clock::now();
      if (tpf > 0) {
        world.step(tpf);
        world.sync();
      }
      tick_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      ticks += tpf;
      frames++;
      if (minimized) return;
      // draw: the camera step for the held keys, then the frame
      fs3d::check(fs3d_camera_integrate(cam_pos, cam_rot, held));
      fs3d_camera cam;
      fs3d::check(fs3d_key_camera(dims, cam_pos, cam_rot, aspect, &cam));
      world.render(cam, W, H, rgb, hit, face);
      double ms = 0.0;
      fs3d::check(fs3d_world_last_render_ms(world.handle(), &ms));
      render_ms += ms;
      if (snap || (every > 0 && drawn % every == 0)) {
        char name[64];
        std::snprintf(name, sizeof name, "/frame_%06lld.ppm", drawn);
        write_ppm(out + name, W, H, rgb);
        snap = false;
      }
      drawn++;
    };

    for (const Event &e : events) {
      if (e.kind == Event::Quit) {
        quit = true;
        break;
      }
      if (e.kind == Event::Frames) {
        for (long long k = 0; k < e.n; k++) run_frame();
      } else {
        pending.push_back(e);
      }
    }
    (void)quit;
    std::cout << "frames " << frames << " drawn " << drawn << " ticks " << ticks << " hash " << world.hash()
              << " cam " << cam_pos[0] << " " << cam_pos[1] << " " << cam_pos[2] << " yaw " << cam_rot[1]
              << " tick_ms " << (frames ? tick_ms / frames : 0.0) << " render_ms "
              << (drawn ? render_ms / drawn : 0.0) << std::endl;
  } catch (const std::exception &) {
    return 1;  // the message is already on stderr
  }
  return 0;
}
