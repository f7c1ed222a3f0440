/*
 * fs3d oracle (TEST INFRASTRUCTURE ONLY) -- CPU transcription of the renderers.
 *
 * Every float operation below is one IEEE-754 binary32 operation in the order
 * written (built with -ffp-contract=off); the HIP kernels are checked against
 * these functions.  Conventions follow the reference's full-screen ray marcher:
 *
 *   pixel -> ray (shaders/fs_raymarch.frag:67-75, shaders/fs_raymarch.vert:30-36
 *   and the quad's UVs, src/engine/rendering/renderer.cpp:1253-1267): pixel
 *   (px, py), py = 0 at the top, has inUV = ((W - px - 0.5)/W, (py + 0.5)/H);
 *   uv = inUV*2 - 1; uv.y *= 1/aspect; dir = normalize(uv.x, uv.y, 1), so the
 *   camera's +x is screen-left, +y screen-down and +z forward.
 *   normalize(v) = v / sqrt((x*x + y*y) + z*z); dot(a,b) = (ax*bx + ay*by) + az*bz.
 *
 *   fso_raymarch_sdf restates shaders/fs_raymarch.frag:10-81: a sphere of radius
 *   0.5 at the origin, at most 64 steps, hit when the distance is below 0.001,
 *   stop when t > 1000 (tested after the hit test), central-difference normal
 *   with h = 0.001, colour (1,0,0) * max(0.05, dot(n, normalize(p - (2,5,3))))
 *   (the vector points from the light to p, as in the shader), miss = black.
 *
 *   fso_render_dda is the project's voxel renderer: the same pixel mapping with a
 *   camera basis (world dir = dx*bx + dy*by + dz*bz), a slab test against the
 *   grid box [0,nx]x[0,ny]x[0,nz], then Amanatides & Woo's voxel traversal
 *   ("A Fast Voxel Traversal Algorithm for Ray Tracing", Eurographics 1987) to
 *   the first non-EMPTY cell, ties between axes broken x, then y, then z.  The
 *   face is the one the ray entered the cell through (0 -x, 1 +x, 2 -y, 3 +y,
 *   4 -z, 5 +z; 6 when the ray starts inside a non-empty cell).  Colour = palette
 *   entry (STATIC 127, SAND 162, FLUID 137, any other byte its own index) times
 *   max(0.05, dot(n, normalize(p - light))), the fs_raymarch lighting with the
 *   face normal (-dir when inside).  A miss gives black, hit = UINT64_MAX,
 *   face = 255.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fs3d_oracle.h"

static void nrm3(float x, float y, float z, float out[3]) {
  const float l = sqrtf((x * x + y * y) + z * z);
  out[0] = x / l;
  out[1] = y / l;
  out[2] = z / l;
}

static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

void fso_pixel_dir(int px, int py, int W, int H, float aspect, float d[3]) {
  const float u = ((float)(W - px) - 0.5f) / (float)W;
  const float v = ((float)py + 0.5f) / (float)H;
  const float ux = u * 2.0f - 1.0f;
  float uy = v * 2.0f - 1.0f;
  const float inv = 1.0f / aspect;
  uy = uy * inv;
  nrm3(ux, uy, 1.0f, d);
}

static float sdf_map(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z) - 0.5f; }

void fso_raymarch_sdf(const float cam_pos[3], float aspect, int W, int H, float *rgba) {
#ifdef _OPENMP
  int nt = fso_get_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
#endif
  for (int py = 0; py < H; py++)
    for (int px = 0; px < W; px++) {
      float d[3];
      fso_pixel_dir(px, py, W, H, aspect, d);
      const float *o = cam_pos;
      float col[3] = {0.0f, 0.0f, 0.0f};
      float t = 0.0f;
      for (int i = 0; i < 64; i++) {
        const float p[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
        const float dist = sdf_map(p[0], p[1], p[2]);
        if (dist < 0.001f) {
          const float h = 0.001f;
          const float gx = sdf_map(p[0] + h, p[1] + 0.0f, p[2] + 0.0f) - sdf_map(p[0] - h, p[1] - 0.0f, p[2] - 0.0f);
          const float gy = sdf_map(p[0] + 0.0f, p[1] + h, p[2] + 0.0f) - sdf_map(p[0] - 0.0f, p[1] - h, p[2] - 0.0f);
          const float gz = sdf_map(p[0] + 0.0f, p[1] + 0.0f, p[2] + h) - sdf_map(p[0] - 0.0f, p[1] - 0.0f, p[2] - h);
          float n[3], dl[3];
          nrm3(gx, gy, gz, n);
          nrm3(p[0] - 2.0f, p[1] - 5.0f, p[2] - 3.0f, dl);
          const float diff = fmaxf(0.05f, dot3(n, dl));
          col[0] = 1.0f * diff;
          col[1] = 0.0f * diff;
          col[2] = 0.0f * diff;
          break;
        } else if (t > 1000.0f) {
          break;
        }
        t = t + dist;
      }
      float *out = rgba + ((size_t)py * W + px) * 4;
      out[0] = col[0];
      out[1] = col[1];
      out[2] = col[2];
      out[3] = 1.0f;
    }
}

static int palette_index(uint8_t m) {
  switch (m) {
    case FSO_STATIC: return 127;
    case FSO_SAND: return 162;
    case FSO_FLUID: return 137;
    default: return m;
  }
}

void fso_render_dda(const uint8_t *g, int nx, int ny, int nz, const fso_camera *cam, const float *palette, int W,
                    int H, float *rgb, uint64_t *hit, uint8_t *face) {
  const float nd[3] = {(float)nx, (float)ny, (float)nz};
  const int ni[3] = {nx, ny, nz};
#ifdef _OPENMP
  int nt = fso_get_threads();
#pragma omp parallel for schedule(dynamic, 2) num_threads(nt)
#endif
  for (int py = 0; py < H; py++)
    for (int px = 0; px < W; px++) {
      const size_t pix = (size_t)py * W + px;
      float dc[3], d[3];
      fso_pixel_dir(px, py, W, H, cam->aspect, dc);
      const float *bx = cam->basis, *by = cam->basis + 3, *bz = cam->basis + 6;
      nrm3((dc[0] * bx[0] + dc[1] * by[0]) + dc[2] * bz[0], (dc[0] * bx[1] + dc[1] * by[1]) + dc[2] * bz[1],
           (dc[0] * bx[2] + dc[1] * by[2]) + dc[2] * bz[2], d);
      const float *o = cam->origin;
      rgb[pix * 3 + 0] = 0.0f;
      rgb[pix * 3 + 1] = 0.0f;
      rgb[pix * 3 + 2] = 0.0f;
      hit[pix] = UINT64_MAX;
      face[pix] = 255;
      /* slab test */
      float tmin = 0.0f, tmax = INFINITY;
      int enter = -1, miss = 0;
      for (int a = 0; a < 3; a++) {
        if (d[a] == 0.0f) {
          if (o[a] < 0.0f || o[a] > nd[a]) miss = 1;
          continue;
        }
        const float t1 = (0.0f - o[a]) / d[a];
        const float t2 = (nd[a] - o[a]) / d[a];
        const float lo = fminf(t1, t2), hi = fmaxf(t1, t2);
        if (lo > tmin) {
          tmin = lo;
          enter = a;
        }
        if (hi < tmax) tmax = hi;
      }
      if (miss || tmin > tmax) continue;
      /* first cell */
      int idx[3], step[3];
      float tnext[3], tdelta[3];
      for (int a = 0; a < 3; a++) {
        const float p = o[a] + tmin * d[a];
        int i = (int)floorf(p);
        if (i < 0) i = 0;
        if (i > ni[a] - 1) i = ni[a] - 1;
        idx[a] = i;
        if (d[a] > 0.0f) {
          step[a] = 1;
          tdelta[a] = fabsf(1.0f / d[a]);
          tnext[a] = ((float)(i + 1) - o[a]) / d[a];
        } else if (d[a] < 0.0f) {
          step[a] = -1;
          tdelta[a] = fabsf(1.0f / d[a]);
          tnext[a] = ((float)i - o[a]) / d[a];
        } else {
          step[a] = 0;
          tdelta[a] = INFINITY;
          tnext[a] = INFINITY;
        }
      }
      int f = enter >= 0 ? 2 * enter + (d[enter] > 0.0f ? 0 : 1) : 6;
      float t = tmin;
      const int maxit = nx + ny + nz + 3;
      for (int it = 0; it < maxit; it++) {
        const size_t li = ((size_t)idx[2] * ny + idx[1]) * nx + idx[0];
        const uint8_t m = g[li];
        if (m != 0) {
          const float *c = palette + 4 * palette_index(m);
          float n[3] = {0.0f, 0.0f, 0.0f};
          if (f == 6) {
            n[0] = -d[0];
            n[1] = -d[1];
            n[2] = -d[2];
          } else {
            n[f >> 1] = (f & 1) ? 1.0f : -1.0f;
          }
          const float hp[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
          float dl[3];
          nrm3(hp[0] - cam->light[0], hp[1] - cam->light[1], hp[2] - cam->light[2], dl);
          const float diff = fmaxf(0.05f, dot3(n, dl));
          rgb[pix * 3 + 0] = c[0] * diff;
          rgb[pix * 3 + 1] = c[1] * diff;
          rgb[pix * 3 + 2] = c[2] * diff;
          hit[pix] = (uint64_t)li;
          face[pix] = (uint8_t)f;
          break;
        }
        int a;
        if (tnext[0] <= tnext[1] && tnext[0] <= tnext[2])
          a = 0;
        else if (tnext[1] <= tnext[2])
          a = 1;
        else
          a = 2;
        if (step[a] == 0) break;
        t = tnext[a];
        idx[a] += step[a];
        if (idx[a] < 0 || idx[a] >= ni[a]) break;
        tnext[a] = tnext[a] + tdelta[a];
        f = 2 * a + (step[a] > 0 ? 0 : 1);
      }
    }
}

/* Orbit camera around the grid centre c at distance dist_scale * max(n):
 * eye = c + R*(cos(th)cos(el), sin(th)cos(el), sin(el)); forward = normalize(c -
 * eye); screen-left = normalize(up x forward) with up = +z; screen-down = left x
 * forward.  Light at c - max(n) * (2, 5, 3): with the fs_raymarch lighting (the
 * vector runs from the light to the point) faces pointing away from the light
 * are lit, so the light sits below and behind.  Computed in double, stored as
 * float. */
void fso_orbit_camera(int nx, int ny, int nz, double angle, double dist_scale, double elev, float aspect,
                      fso_camera *out) {
  const double c[3] = {nx * 0.5, ny * 0.5, nz * 0.5};
  double m = nx;
  if (ny > m) m = ny;
  if (nz > m) m = nz;
  const double R = dist_scale * m;
  const double eye[3] = {c[0] + R * cos(angle) * cos(elev), c[1] + R * sin(angle) * cos(elev), c[2] + R * sin(elev)};
  double f[3] = {c[0] - eye[0], c[1] - eye[1], c[2] - eye[2]};
  double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int a = 0; a < 3; a++) f[a] /= fl;
  /* left = up x f, up = (0,0,1) */
  double l[3] = {-f[1], f[0], 0.0};
  double ll = sqrt(l[0] * l[0] + l[1] * l[1]);
  for (int a = 0; a < 3; a++) l[a] /= ll;
  /* down = left x f */
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
}

/*
 * fso_box_vertex / fso_box_march restate the scene's active ray marcher
 * (shaders/raymarch_sphere.vert:30-45 and shaders/raymarch_sphere.frag:9-53).
 * Vertex stage: the fragment's ray origin is the camera in object space,
 * invModel * vec4(-camPos, 1) (column-major; each output component summed
 * ((m0*v0 + m1*v1) + m2*v2) + m3*v3).  Fragment stage: dir = normalize(local -
 * origin); march t from 0 for at most 64 steps through the box SDF of half-size
 * 0.5, d = length(max(q,0)) + min(max(q.x, max(q.y, q.z)), 0) with q = |p| - 0.5;
 * a hit (d < 0.001) gives (max(0.05, density), 0, 0) where density starts at 1
 * and drops by 0.05 per step taken; t > 1000 (tested after the hit test) or 64
 * steps give black; alpha 1.
 */
void fso_box_vertex(const float cam_pos[3], const float inv_model[16], float origin[3]) {
  const float v[4] = {cam_pos[0] * -1.0f, cam_pos[1] * -1.0f, cam_pos[2] * -1.0f, 1.0f};
  for (int i = 0; i < 3; i++)
    origin[i] = ((inv_model[0 * 4 + i] * v[0] + inv_model[1 * 4 + i] * v[1]) + inv_model[2 * 4 + i] * v[2]) +
                inv_model[3 * 4 + i] * v[3];
}

static float box_sdf(float x, float y, float z) {
  const float qx = fabsf(x) - 0.5f, qy = fabsf(y) - 0.5f, qz = fabsf(z) - 0.5f;
  const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
  const float len = sqrtf((mx * mx + my * my) + mz * mz);
  return len + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
}

void fso_box_march(const float *origin, const float *local, int n, float *rgba) {
  for (int k = 0; k < n; k++) {
    const float *o = origin + 3 * k, *l = local + 3 * k;
    float d[3];
    nrm3(l[0] - o[0], l[1] - o[1], l[2] - o[2], d);
    float r = 0.0f, t = 0.0f, density = 1.0f;
    for (int i = 0; i < 64; i++) {
      const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
      const float dist = box_sdf(px, py, pz);
      if (dist < 0.001f) {
        r = fmaxf(0.05f, density);
        break;
      } else if (t > 1000.0f) {
        break;
      }
      t = t + dist;
      density = density - 0.05f;
    }
    rgba[4 * k + 0] = r;
    rgba[4 * k + 1] = 0.0f;
    rgba[4 * k + 2] = 0.0f;
    rgba[4 * k + 3] = 1.0f;
  }
}
