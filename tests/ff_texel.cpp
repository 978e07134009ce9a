// Checks the final gather's float-float NEAREST texel decision against an
// FP64 evaluation of the same formula. Built and run by
// tests/test_float_float.py (host build of the kernel's header).
#include <math.h>
#include <stdint.h>

#include "../many-lights-gi_amd/csrc/float_float.hpp"

namespace {
uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uni(uint64_t* s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// column-major 4x4 helpers in double
void mul4(const double* a, const double* b, double* o) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double v = 0;
      for (int k = 0; k < 4; ++k) v += a[k * 4 + r] * b[c * 4 + k];
      o[c * 4 + r] = v;
    }
}
bool inv4(const double* m, double* o) {  // Gauss-Jordan
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][c + 4] = r == c;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
    if (a[p][c] == 0) return false;
    for (int k = 0; k < 8; ++k) {
      double t = a[c][k];
      a[c][k] = a[p][k];
      a[p][k] = t;
    }
    const double d = a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] /= d;
    for (int r = 0; r < 4; ++r)
      if (r != c) {
        const double f = a[r][c];
        for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
      }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[c * 4 + r] = a[r][c + 4];
  return true;
}

// FP64 evaluation of the texel coordinate (the formula texel_exact follows);
// up_z lets a test plant a wrong basis.
void texel_f64(const float* m, int w, int h, int x, int y, float depth, const float* vp, const float* vn, int S,
               double up_z, double* cx, double* cy) {
  const double c0 = (double)x / w * 2.0 - 1.0, c1 = (double)y / h * 2.0 - 1.0, c2 = (double)depth * 2.0 - 1.0;
  double r[4];
  for (int row = 0; row < 4; ++row)
    r[row] = (double)m[row] * c0 + (double)m[4 + row] * c1 + ((double)m[8 + row] * c2 + (double)m[12 + row]);
  const double dx = r[0] / r[3] - vp[0], dy = r[1] / r[3] - vp[1], dz = r[2] / r[3] - vp[2];
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  const double nx = vn[0], ny = vn[1], nz = vn[2];
  const double qx = ny * up_z - nz, qy = -nx * up_z, qz = nx;
  const double ql = sqrt(qx * qx + qy * qy + qz * qz);
  const double sx = qx / ql, sy = qy / ql, sz = qz / ql;
  const double ux = sy * nz - sz * ny, uy = sz * nx - sx * nz, uz = sx * ny - sy * nx;
  const double den = 1.0 + (nx * dx + ny * dy + nz * dz) / dist;
  *cx = ((sx * dx + sy * dy + sz * dz) / dist / den + 1.0) * 0.5 * S;
  *cy = ((ux * dx + uy * dy + uz * dz) / dist / den + 1.0) * 0.5 * S;
}
}  // namespace

extern "C" {
// n random lit pairs seen by a 1920 x 1080 camera in the 10^3 room, ISM side
// S, with texel estimates off by up to 0.45 texel (odd pairs) or 4.5e-4
// texel (even pairs). Returns the decisions that differ (cases within 1e-10 texel of an edge
// are skipped); *near counts the cases within 2.5e-4 texel of an edge.
long check_ff_texel(long n, uint64_t seed, int S, double up_z_ref, long* near) {
  const int w = 1920, h = 1080;
  // camera: eye (5, 5, 9), looking at (5, 4, 0), fovy 60, near 0.05, far 50
  const double ex = 5, ey = 5, ez = 9, cxw = 5, cyw = 4, czw = 0;
  double f[3] = {cxw - ex, cyw - ey, czw - ez};
  double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& v : f) v /= fl;
  const double up[3] = {0, 1, 0};
  double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (double& v : s) v /= sl;
  const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  double V[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                  -(s[0] * ex + s[1] * ey + s[2] * ez), -(u[0] * ex + u[1] * ey + u[2] * ez),
                  f[0] * ex + f[1] * ey + f[2] * ez, 1};
  const double th = tan(60.0 * M_PI / 360.0), zn = 0.05, zf = 50.0;
  double P[16] = {0};
  P[0] = 1.0 / (w / (double)h * th);
  P[5] = 1.0 / th;
  P[10] = -(zf + zn) / (zf - zn);
  P[11] = -1;
  P[14] = -2 * zf * zn / (zf - zn);
  double VP[16], IV[16];
  mul4(P, V, VP);
  inv4(VP, IV);
  float m[16];
  for (int i = 0; i < 16; ++i) m[i] = (float)IV[i];
  long bad = 0;
  *near = 0;
  for (long i = 0; i < n; ++i) {
    // a world point in the room, projected to a pixel and its FP32 depth
    const double px = 0.5 + 9 * uni(&seed), py = 0.5 + 9 * uni(&seed), pz = 0.5 + 8 * uni(&seed);
    double c[4];
    for (int r = 0; r < 4; ++r) c[r] = VP[r] * px + VP[4 + r] * py + VP[8 + r] * pz + VP[12 + r];
    if (c[3] <= 0) continue;
    const double nxd = c[0] / c[3], nyd = c[1] / c[3], nzd = c[2] / c[3];
    if (fabs(nxd) >= 1 || fabs(nyd) >= 1 || fabs(nzd) >= 1) continue;
    const int x = (int)((nxd + 1) * 0.5 * w), y = (int)((nyd + 1) * 0.5 * h);
    const float depth = (float)(nzd * 0.5 + 0.5);
    // a VPL in the room with a unit normal towards the point (lit)
    const float vp[3] = {(float)(0.5 + 9 * uni(&seed)), (float)(0.5 + 9 * uni(&seed)), (float)(0.5 + 9 * uni(&seed))};
    double nv[3] = {px - vp[0] + 2 * uni(&seed) - 1, py - vp[1] + 2 * uni(&seed) - 1, pz - vp[2] + 2 * uni(&seed) - 1};
    const double nl = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    const float vn[3] = {(float)(nv[0] / nl), (float)(nv[1] / nl), (float)(nv[2] / nl)};
    double rx, ry;
    texel_f64(m, w, h, x, y, depth, vp, vn, S, (double)0.01f, &rx, &ry);
    if (!(rx >= 0 && rx < S && ry >= 0 && ry < S)) continue;  // outside the VPL's hemisphere image
    const double ex_ = fmin(rx - floor(rx), ceil(rx) - rx), ey_ = fmin(ry - floor(ry), ceil(ry) - ry);
    if (fmin(ex_, ey_) < 2.5e-4) ++*near;
    if (fmin(ex_, ey_) < 1e-10) continue;
    double qx, qy;
    texel_f64(m, w, h, x, y, depth, vp, vn, S, up_z_ref, &qx, &qy);
    // the kernel's estimates are within ~1e-4 texel; any within 0.5 must do
    const double jx = 0.9 * uni(&seed) - 0.45, jy = 0.9 * uni(&seed) - 0.45;
    const double sx = (i & 1) ? jx : jx * 1e-3, sy = (i & 1) ? jy : jy * 1e-3;
    int tx, ty;
    mlgi::ff::texel_exact(m, w, h, x, y, depth, vp, vn, S, (float)(rx + sx), (float)(ry + sy), &tx, &ty);
    if (tx != (int)floor(qx) || ty != (int)floor(qy)) ++bad;
  }
  return bad;
}
}
