// C ABI of the renderers and the camera helpers.
//
// Camera block: fs3d_camera_globaldata fills the reference's per-frame
// GlobalData (src/engine/rendering/materials.h:19-28, written in
// Material::writeGlobalAndObjectData, src/engine/rendering/materials.cpp:531-575):
// camPos (x, y, z, 1); aspect (1700/900, 850, 450, 0); camDir (0, 0, 1, 1);
// view = Rx(0) * Ry(camRot.y degrees) * T(camPos); proj = a right-handed
// perspective (70 degrees, 1700/900, near 0.01, far 200, clip z in [-1, 1]) with
// proj[1][1] negated; viewproj = proj * view; invViewProj = its inverse.
// Matrices are column-major, 16 floats each (76 floats in all).
//
// fs3d_camera_integrate mirrors the reference's fixed-step camera motion
// (src/engine/rendering/renderer.cpp:437-467): 1.5 units and 10 degrees per
// 0.016 s frame for each held key.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fs3d_render.h"

using namespace fs3d;

namespace {

// 4x4 column-major, c[col][row]
struct M4 {
  float c[4][4];
};

M4 identity() {
  M4 m{};
  for (int i = 0; i < 4; i++) m.c[i][i] = 1.0f;
  return m;
}

// r.col(i) = a.col(0)*b[i][0] + a.col(1)*b[i][1] + a.col(2)*b[i][2] + a.col(3)*b[i][3]
M4 mul(const M4 &a, const M4 &b) {
  M4 r{};
  for (int i = 0; i < 4; i++)
    for (int row = 0; row < 4; row++)
      r.c[i][row] = ((a.c[0][row] * b.c[i][0] + a.c[1][row] * b.c[i][1]) + a.c[2][row] * b.c[i][2]) +
                    a.c[3][row] * b.c[i][3];
  return r;
}

// Rotation by `angle` radians about the unit axis (Rodrigues' formula).
M4 rotation(float angle, float ax, float ay, float az) {
  const float c = std::cos(angle), s = std::sin(angle);
  const float tx = (1.0f - c) * ax, ty = (1.0f - c) * ay, tz = (1.0f - c) * az;
  M4 r = identity();
  r.c[0][0] = c + tx * ax;
  r.c[0][1] = tx * ay + s * az;
  r.c[0][2] = tx * az - s * ay;
  r.c[1][0] = ty * ax - s * az;
  r.c[1][1] = c + ty * ay;
  r.c[1][2] = ty * az + s * ax;
  r.c[2][0] = tz * ax + s * ay;
  r.c[2][1] = tz * ay - s * ax;
  r.c[2][2] = c + tz * az;
  return r;
}

M4 translation(float x, float y, float z) {
  M4 t = identity();
  t.c[3][0] = x;
  t.c[3][1] = y;
  t.c[3][2] = z;
  return t;
}

M4 perspective_rh(float fovy, float aspect, float znear, float zfar) {
  const float th = std::tan(fovy / 2.0f);
  M4 p{};
  p.c[0][0] = 1.0f / (aspect * th);
  p.c[1][1] = 1.0f / th;
  p.c[2][2] = -(zfar + znear) / (zfar - znear);
  p.c[2][3] = -1.0f;
  p.c[3][2] = -(2.0f * zfar * znear) / (zfar - znear);
  return p;
}

// Inverse by Gauss-Jordan elimination with partial pivoting in double precision.
M4 inverse(const M4 &m) {
  double a[4][8];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      a[r][c] = m.c[c][r];
      a[r][4 + c] = r == c ? 1.0 : 0.0;
    }
  for (int col = 0; col < 4; col++) {
    int piv = col;
    for (int r = col + 1; r < 4; r++)
      if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
    for (int k = 0; k < 8; k++) std::swap(a[col][k], a[piv][k]);
    const double d = a[col][col];
    for (int k = 0; k < 8; k++) a[col][k] /= d;
    for (int r = 0; r < 4; r++)
      if (r != col) {
        const double f = a[r][col];
        for (int k = 0; k < 8; k++) a[r][k] -= f * a[col][k];
      }
  }
  M4 inv{};
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) inv.c[c][r] = (float)a[r][4 + c];
  return inv;
}

float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

constexpr int kMaxImageSide = 1 << 16;  // 16x16-pixel tiles: at most 4096 tiles per side

// Grids from this size up track changed bricks between frames by default: a
// full read of the grid costs more than about ten ticks of recording them.
constexpr size_t kTrackMinVoxels = (size_t)1 << 29;

bool track_wanted(const fs3d_world *w) {
  return w->track_mode == FS3D_TRACK_ON || (w->track_mode == FS3D_TRACK_AUTO && w->n >= kTrackMinVoxels);
}

int check_cam(const fs3d_camera *cam) {
  if (!cam) return set_error(FS3D_ERR_INVALID, "camera is NULL");
  if (!(cam->aspect > 0.0f)) return set_error(FS3D_ERR_INVALID, "camera aspect must be > 0");
  return FS3D_OK;
}

}  // namespace

extern "C" {

int fs3d_palette(float *rgba) {
  if (!rgba) return set_error(FS3D_ERR_INVALID, "rgba is NULL");
  palette_host(rgba);
  return clear_error();
}

int fs3d_camera_globaldata(const float cam_pos[3], const float cam_rot[3], float *block) {
  if (!cam_pos || !cam_rot || !block) return set_error(FS3D_ERR_INVALID, "NULL argument");
  const M4 view = mul(mul(rotation(radians(0.0f), 1.0f, 0.0f, 0.0f), rotation(radians(cam_rot[1]), 0.0f, 1.0f, 0.0f)),
                      translation(cam_pos[0], cam_pos[1], cam_pos[2]));
  M4 proj = perspective_rh(radians(70.0f), 1700.0f / 900.0f, 0.01f, 200.0f);
  proj.c[1][1] *= -1.0f;
  const M4 vp = mul(proj, view);
  const M4 ivp = inverse(vp);
  const float head[12] = {cam_pos[0], cam_pos[1], cam_pos[2], 1.0f, 1700.0f / 900.0f, 850.0f, 450.0f, 0.0f,
                          0.0f,       0.0f,       1.0f,       1.0f};
  std::memcpy(block, head, sizeof head);
  std::memcpy(block + 12, &view, 64);
  std::memcpy(block + 28, &proj, 64);
  std::memcpy(block + 44, &vp, 64);
  std::memcpy(block + 60, &ivp, 64);
  return clear_error();
}

int fs3d_camera_integrate(float cam_pos[3], float cam_rot[3], uint32_t keys) {
  if (!cam_pos || !cam_rot) return set_error(FS3D_ERR_INVALID, "NULL argument");
  const float move = 1.5f, turn = 10.0f;
  if (keys & FS3D_KEY_W)
    cam_pos[2] += move * 0.016f;
  else if (keys & FS3D_KEY_S)
    cam_pos[2] -= move * 0.016f;
  if (keys & FS3D_KEY_A)
    cam_pos[0] += move * 0.016f;
  else if (keys & FS3D_KEY_D)
    cam_pos[0] -= move * 0.016f;
  if (keys & FS3D_KEY_SPACE)
    cam_pos[1] -= move * 0.016f;
  else if (keys & FS3D_KEY_CTRL)
    cam_pos[1] += move * 0.016f;
  if (keys & FS3D_KEY_RIGHT)
    cam_rot[1] += turn * 0.016f;
  else if (keys & FS3D_KEY_LEFT)
    cam_rot[1] -= turn * 0.016f;
  return clear_error();
}

// Orbit camera around the grid centre c at distance dist_scale * max(n):
// eye = c + R (cos a cos e, sin a cos e, sin e); forward = normalize(c - eye);
// screen-left = normalize(+z x forward); screen-down = left x forward; light at
// c - max(n) (2, 5, 3).  Double precision, stored as float.
int fs3d_orbit_camera(const int32_t dims[3], double angle, double dist_scale, double elev, float aspect,
                      fs3d_camera *out) {
  if (!dims || !out || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || !(aspect > 0.0f))
    return set_error(FS3D_ERR_INVALID, "orbit camera: bad arguments");
  const double c[3] = {dims[0] * 0.5, dims[1] * 0.5, dims[2] * 0.5};
  const double m = std::max(std::max((double)dims[0], (double)dims[1]), (double)dims[2]);
  const double R = dist_scale * m;
  const double eye[3] = {c[0] + R * std::cos(angle) * std::cos(elev), c[1] + R * std::sin(angle) * std::cos(elev),
                         c[2] + R * std::sin(elev)};
  double f[3] = {c[0] - eye[0], c[1] - eye[1], c[2] - eye[2]};
  const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double &v : f) v /= fl;
  double l[3] = {-f[1], f[0], 0.0};
  const double ll = std::sqrt(l[0] * l[0] + l[1] * l[1]);
  for (double &v : l) v /= ll;
  const double dn[3] = {l[1] * f[2] - l[2] * f[1], l[2] * f[0] - l[0] * f[2], l[0] * f[1] - l[1] * f[0]};
  for (int a = 0; a < 3; a++) {
    out->origin[a] = (float)eye[a];
    out->basis[a] = (float)l[a];
    out->basis[3 + a] = (float)dn[a];
    out->basis[6 + a] = (float)f[a];
  }
  out->light[0] = (float)(c[0] - 2.0 * m);
  out->light[1] = (float)(c[1] - 5.0 * m);
  out->light[2] = (float)(c[2] - 3.0 * m);
  out->aspect = aspect;
  return clear_error();
}

int fs3d_key_camera(const int32_t dims[3], const float cam_pos[3], const float cam_rot[3], float aspect,
                    fs3d_camera *out) {
  if (!dims || !cam_pos || !cam_rot || !out || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || !(aspect > 0.0f))
    return set_error(FS3D_ERR_INVALID, "key camera: bad arguments");
  // the reference's view matrix, as in fs3d_camera_globaldata
  const M4 view = mul(mul(rotation(radians(0.0f), 1.0f, 0.0f, 0.0f), rotation(radians(cam_rot[1]), 0.0f, 1.0f, 0.0f)),
                      translation(cam_pos[0], cam_pos[1], cam_pos[2]));
  // Camera axes in world space are the rows of the rotation part; the eye is
  // -R^T t.  The camera looks down -z with +y up on screen (the projection's
  // flipped y turns Vulkan's downward NDC y back up), so screen-left = -row 0,
  // screen-down = -row 1 and forward = -row 2.
  double eye[3], axis[3][3];
  for (int j = 0; j < 3; j++) {
    eye[j] = 0.0;
    for (int i = 0; i < 3; i++) eye[j] -= (double)view.c[j][i] * view.c[3][i];
    for (int k = 0; k < 3; k++) axis[k][j] = -(double)view.c[j][k];
  }
  // World -> grid: the grid spans [-1, 1] along its longest axis around the
  // origin; world +X = grid +x, world +Y = grid +z (up), world +Z = grid -y.
  const double half = 0.5 * std::max(std::max((double)dims[0], (double)dims[1]), (double)dims[2]);
  const double c[3] = {dims[0] * 0.5, dims[1] * 0.5, dims[2] * 0.5};
  auto to_grid = [](const double v[3], double g[3]) {
    g[0] = v[0];
    g[1] = -v[2];
    g[2] = v[1];
  };
  double g[3];
  to_grid(eye, g);
  for (int a = 0; a < 3; a++) out->origin[a] = (float)(c[a] + half * g[a]);
  for (int k = 0; k < 3; k++) {
    to_grid(axis[k], g);
    for (int a = 0; a < 3; a++) out->basis[3 * k + a] = (float)g[a];
  }
  const double light[3] = {2.0, 5.0, 3.0};  // fs_raymarch's light position (shaders/fs_raymarch.frag:51)
  to_grid(light, g);
  for (int a = 0; a < 3; a++) out->light[a] = (float)(c[a] + half * g[a]);
  out->aspect = aspect;
  return clear_error();
}

int fs3d_render_sdf(const float cam_pos[3], float aspect, int32_t W, int32_t H, float *rgba) {
  if (!cam_pos || !rgba || W < 1 || H < 1 || W > kMaxImageSide || H > kMaxImageSide || !(aspect > 0.0f))
    return set_error(FS3D_ERR_INVALID, "render_sdf: bad arguments");
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  const size_t bytes = (size_t)W * H * 4 * sizeof(float);
  float *d = nullptr;
  FS3D_HIP(hipMalloc(&d, bytes));
  hipError_t e = launch_sdf(cam_pos, aspect, W, H, d, nullptr);
  if (e == hipSuccess) e = hipMemcpy(rgba, d, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "render_sdf failed: %s", hipGetErrorString(e));
  return clear_error();
}

int fs3d_box_vertex(const float cam_pos[3], const float inv_model[16], float origin[3]) {
  if (!cam_pos || !inv_model || !origin) return set_error(FS3D_ERR_INVALID, "box_vertex: NULL argument");
  const float v[4] = {cam_pos[0] * -1.0f, cam_pos[1] * -1.0f, cam_pos[2] * -1.0f, 1.0f};
  for (int i = 0; i < 3; i++)
    origin[i] = ((inv_model[0 * 4 + i] * v[0] + inv_model[1 * 4 + i] * v[1]) + inv_model[2 * 4 + i] * v[2]) +
                inv_model[3 * 4 + i] * v[3];
  return clear_error();
}

int fs3d_render_box_march(const float *origin, const float *local, int32_t n, float *rgba) {
  if (!origin || !local || !rgba || n < 0) return set_error(FS3D_ERR_INVALID, "box_march: bad arguments");
  if (n == 0) return clear_error();
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  const size_t in_bytes = (size_t)n * 3 * sizeof(float), out_bytes = (size_t)n * 4 * sizeof(float);
  float *d = nullptr;
  FS3D_HIP(hipMalloc(&d, 2 * in_bytes + out_bytes));
  hipError_t e = hipMemcpy(d, origin, in_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 3 * (size_t)n, local, in_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_box_march(d, d + 3 * (size_t)n, n, d + 6 * (size_t)n, nullptr);
  if (e == hipSuccess) e = hipMemcpy(rgba, d + 6 * (size_t)n, out_bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "box_march failed: %s", hipGetErrorString(e));
  return clear_error();
}

int fs3d_world_render(fs3d_world *w, const fs3d_camera *cam, int32_t W, int32_t H, float *rgb, uint64_t *hit,
                      uint8_t *face) {
  if (!w) return set_error(FS3D_ERR_INVALID, "world handle is NULL");
  if (int rc = check_cam(cam)) return rc;
  if (W < 1 || H < 1 || W > kMaxImageSide || H > kMaxImageSide || !rgb || !hit || !face)
    return set_error(FS3D_ERR_INVALID, "render: bad image arguments (1 <= W, H <= %d)", kMaxImageSide);
  FS3D_HIP(hipSetDevice(w->dev));
  const size_t px = (size_t)W * H;
  if (px > w->render_px) {
    FS3D_HIP(hipStreamSynchronize(w->stream));
    if (w->d_rgb) FS3D_HIP(hipFree(w->d_rgb));
    if (w->d_hit) FS3D_HIP(hipFree(w->d_hit));
    if (w->d_face) FS3D_HIP(hipFree(w->d_face));
    w->d_rgb = nullptr;
    w->d_hit = nullptr;
    w->d_face = nullptr;
    w->render_px = 0;
    FS3D_HIP(hipMalloc(&w->d_rgb, px * 3 * sizeof(float)));
    FS3D_HIP(hipMalloc(&w->d_hit, px * sizeof(u64)));
    FS3D_HIP(hipMalloc(&w->d_face, px));
    w->render_px = px;
  }
  if (!w->d_occ) FS3D_HIP(hipMalloc(&w->d_occ, occ_words(w->nx, w->ny, w->nz) * sizeof(u64)));
  if (!w->d_dirty) {
    const size_t bytes = dirty_words(w->nx, w->ny, w->nz) * sizeof(u64);
    FS3D_HIP(hipMalloc(&w->d_dirty, bytes));
    FS3D_HIP(hipMemsetAsync(w->d_dirty, 0, bytes, w->stream));
  }
  if (!w->occ_track && track_wanted(w)) {
    // from now on skipping ticks record the bricks they change, so the next
    // render reads only those; graphs captured before lack that step
    FS3D_HIP(hipStreamSynchronize(w->stream));
    world_drop_graphs(w);
    w->occ_track = true;
  }
  RenderCam rc_cam;
  std::memcpy(rc_cam.origin, cam->origin, sizeof rc_cam.origin);
  std::memcpy(rc_cam.basis, cam->basis, sizeof rc_cam.basis);
  std::memcpy(rc_cam.light, cam->light, sizeof rc_cam.light);
  rc_cam.aspect = cam->aspect;
  FS3D_HIP(hipEventRecord(w->ev0, w->stream));
  FS3D_HIP(launch_dda(w->buf[w->cur], w->nx, w->ny, w->nz, rc_cam, W, H, w->d_rgb, w->d_hit, w->d_face, w->d_occ,
                      w->d_dirty, w->occ_valid, w->stream));
  w->occ_valid = true;
  FS3D_HIP(hipEventRecord(w->ev1, w->stream));
  FS3D_HIP(hipMemcpyAsync(rgb, w->d_rgb, px * 3 * sizeof(float), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipMemcpyAsync(hit, w->d_hit, px * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipMemcpyAsync(face, w->d_face, px, hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  float ms = 0.f;
  FS3D_HIP(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  w->last_render_ms = ms;
  return clear_error();
}

int fs3d_world_set_render_track(fs3d_world *w, int32_t mode) {
  if (!w) return set_error(FS3D_ERR_INVALID, "world handle is NULL");
  if (mode < FS3D_TRACK_AUTO || mode > FS3D_TRACK_OFF) return set_error(FS3D_ERR_INVALID, "track mode %d", mode);
  w->track_mode = mode;
  if (w->occ_track && !track_wanted(w)) {
    // ticks stop recording changed bricks: the next render reads the whole grid
    FS3D_HIP(hipSetDevice(w->dev));
    FS3D_HIP(hipStreamSynchronize(w->stream));
    world_drop_graphs(w);
    w->occ_track = false;
    w->occ_valid = false;
  }
  return clear_error();
}

int fs3d_world_get_render_track(const fs3d_world *w, int32_t *mode) {
  if (!w || !mode) return set_error(FS3D_ERR_INVALID, "NULL argument");
  *mode = w->track_mode;
  return clear_error();
}

int fs3d_world_render_maps(fs3d_world *w, uint64_t *coarse, size_t n_coarse, uint64_t *fine, size_t n_fine,
                           uint64_t *region, size_t n_region) {
  if (!w || !coarse || !fine || !region) return set_error(FS3D_ERR_INVALID, "NULL argument");
  if (!w->d_occ) return set_error(FS3D_ERR_STATE, "the world has not been rendered");
  const size_t nb = ((size_t)w->nx + 7) / 8 * (((size_t)w->ny + 7) / 8) * (((size_t)w->nz + 7) / 8);
  const size_t nc = dirty_words(w->nx, w->ny, w->nz), nr = occ_words(w->nx, w->ny, w->nz) - nc - nb;
  if (n_coarse < nc || n_fine < nb || n_region < nr)
    return set_error(FS3D_ERR_INVALID, "maps need %zu coarse, %zu fine and %zu region words", nc, nb, nr);
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipMemcpyAsync(coarse, w->d_occ, nc * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipMemcpyAsync(fine, w->d_occ + nc, nb * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipMemcpyAsync(region, w->d_occ + nc + nb, nr * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_last_render_ms(const fs3d_world *w, double *ms) {
  if (!w || !ms) return set_error(FS3D_ERR_INVALID, "NULL argument");
  *ms = w->last_render_ms;
  return clear_error();
}

}  // extern "C"
