// gs_project.hip -- projections of a density map at Euler angles, the
// templates a picking run correlates (the reference takes them from EMAN's
// project3d, protocol step 5, which it does not ship).
//
// out[j](row, col) = sum over w of the trilinear value of the volume at
//   R_j^T (u, v, w) + c,   u = col - c, v = row - c, w = -c .. box-1-c,
// c = box / 2, R_j = Rz(phi) Rx(alt) Rz(az) (EMAN convention), voxels outside
// the box count as 0: the volume rotated into its own box with trilinear
// interpolation, then summed along z.
//
// One thread per output pixel walks its ray; the walk is clipped to the
// part of the ray whose trilinear cell touches the box.  The volume (23 MB at
// box 180) is read through the caches: consecutive lanes hold consecutive
// pixels of a row, whose rays are one voxel apart, so the gathers of a wave
// fall into a few cache lines.  The bound is the gather rate of L1/L2, not
// HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/gisspa.h"
#include "gs_host.hpp"

namespace gs {
namespace {

constexpr int PWG = 256;

__device__ __forceinline__ float voxel(const float* __restrict__ vol, int n, int x, int y, int z) {
  return ((unsigned)x < (unsigned)n && (unsigned)y < (unsigned)n && (unsigned)z < (unsigned)n)
             ? vol[((long)z * n + y) * n + x]
             : 0.0f;
}

// w range of the ray base + w dir (one coordinate) with the coordinate in
// (-1, n): trilinear cells outside that interval hold only zero voxels
__device__ __forceinline__ void clip_axis(float b, float d, int n, float& lo, float& hi) {
  if (fabsf(d) < 1e-12f) {
    if (!(b > -1.0f && b < (float)n)) { lo = 1.0f; hi = 0.0f; }
    return;
  }
  float t0 = (-1.0f - b) / d, t1 = ((float)n - b) / d;
  if (t0 > t1) { const float t = t0; t0 = t1; t1 = t; }
  lo = fmaxf(lo, t0);
  hi = fminf(hi, t1);
}

__global__ __launch_bounds__(PWG) void k_project(const float* __restrict__ vol, int n,
                                                 const float* __restrict__ rot,
                                                 float* __restrict__ out) {
  const int j = blockIdx.y;
  const int pix = blockIdx.x * PWG + threadIdx.x;
  if (pix >= n * n) return;
  const int row = pix / n, col = pix % n;
  const int c = n / 2;
  const float* R = rot + 9 * j;  // row-major R_j
  const float u = (float)(col - c), v = (float)(row - c);
  // p(w) = R^T (u, v, w) + c = u R[0,:] + v R[1,:] + w R[2,:] + c
  const float bx = u * R[0] + v * R[3] + (float)c, by = u * R[1] + v * R[4] + (float)c,
              bz = u * R[2] + v * R[5] + (float)c;
  const float dx = R[6], dy = R[7], dz = R[8];
  float lo = (float)(-c), hi = (float)(n - 1 - c);
  clip_axis(bx, dx, n, lo, hi);
  clip_axis(by, dy, n, lo, hi);
  clip_axis(bz, dz, n, lo, hi);
  const int w0 = max(-c, (int)floorf(lo)), w1 = min(n - 1 - c, (int)ceilf(hi));
  float sum = 0.0f;
  for (int w = w0; w <= w1; ++w) {
    const float x = bx + (float)w * dx, y = by + (float)w * dy, z = bz + (float)w * dz;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
    const float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    const float c00 = voxel(vol, n, x0, y0, z0) * (1.0f - fx) + voxel(vol, n, x0 + 1, y0, z0) * fx;
    const float c10 = voxel(vol, n, x0, y0 + 1, z0) * (1.0f - fx) +
                      voxel(vol, n, x0 + 1, y0 + 1, z0) * fx;
    const float c01 = voxel(vol, n, x0, y0, z0 + 1) * (1.0f - fx) +
                      voxel(vol, n, x0 + 1, y0, z0 + 1) * fx;
    const float c11 = voxel(vol, n, x0, y0 + 1, z0 + 1) * (1.0f - fx) +
                      voxel(vol, n, x0 + 1, y0 + 1, z0 + 1) * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy, c1 = c01 * (1.0f - fy) + c11 * fy;
    sum += c0 * (1.0f - fz) + c1 * fz;
  }
  out[(long)j * n * n + pix] = sum;
}

// R = Rz(phi) Rx(alt) Rz(az), row-major, angles in degrees
void euler_rotation(double alt, double az, double phi, float* r) {
  const double d = M_PI / 180.0;
  const double ca = std::cos(alt * d), sa = std::sin(alt * d), cz = std::cos(az * d),
               sz = std::sin(az * d), cp = std::cos(phi * d), sp = std::sin(phi * d);
  const double m[9] = {cp * cz - ca * sz * sp,  cp * sz + ca * cz * sp,  sp * sa,
                       -sp * cz - ca * sz * cp, -sp * sz + ca * cz * cp, cp * sa,
                       sa * sz,                 -sa * cz,                ca};
  for (int k = 0; k < 9; ++k) r[k] = (float)m[k];
}

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};

int project(int device, const float* vol, int box, const float* euler, int nviews, float* out) {
  if (!vol || !euler || !out || box < 2 || box > 1024 || nviews < 0) return GS_ERR_ARG;
  if (nviews == 0) return GS_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return GS_ERR_DEVICE;
  }
  HIPCHK(hipSetDevice(device));
  std::vector<float> rot((size_t)nviews * 9);
  for (int j = 0; j < nviews; ++j)
    euler_rotation(euler[3 * j], euler[3 * j + 1], euler[3 * j + 2], &rot[(size_t)j * 9]);
  StreamGuard sg;
  HIPCHK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  const size_t nvox = (size_t)box * box * box, npix = (size_t)box * box;
  DevBuf dvol, drot, dout;
  CHK(dvol.ensure(nvox * sizeof(float)));
  CHK(drot.ensure(rot.size() * sizeof(float)));
  // views in chunks (grid y), so the output buffer stays below 1 GiB
  const int chunk = (int)std::max<size_t>(
      1, std::min<size_t>(std::min(nviews, 65535), (1u << 28) / npix));
  CHK(dout.ensure((size_t)chunk * npix * sizeof(float)));
  HIPCHK(hipMemcpyAsync(dvol.p, vol, nvox * sizeof(float), hipMemcpyHostToDevice, sg.s));
  HIPCHK(hipMemcpyAsync(drot.p, rot.data(), rot.size() * sizeof(float), hipMemcpyHostToDevice,
                        sg.s));
  for (int j0 = 0; j0 < nviews; j0 += chunk) {
    const int nv = std::min(chunk, nviews - j0);
    const dim3 grid((unsigned)((npix + PWG - 1) / PWG), (unsigned)nv);
    k_project<<<grid, PWG, 0, sg.s>>>(dvol.as<float>(), box, drot.as<float>() + 9 * (size_t)j0,
                                      dout.as<float>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out + (size_t)j0 * npix, dout.p, (size_t)nv * npix * sizeof(float),
                          hipMemcpyDeviceToHost, sg.s));
  }
  HIPCHK(hipStreamSynchronize(sg.s));
  return GS_OK;
}

}  // namespace
}  // namespace gs

extern "C" int gs_project(int device, const float* vol, int box, const float* euler, int nviews,
                          float* out) {
  return gs::project(device, vol, box, euler, nviews, out);
}
