// gs_dedup.hip -- duplicate removal of a pick list on the GPU
// (python_scripts/remove_repeat_particles_from_list.py:9-145).
//
// The script walks the picks of each micrograph in list order.  A pick i
// that is still unmarked when its turn comes compares itself with every later
// unmarked pick j whose centre lies closer than `center` pixels and whose
// orientation is within `euler_thres` degrees (of 0 or of a 180-degree turn);
// the lower-scored of the two is marked (i on a tie), and i goes on marking
// later picks even once it has been marked itself.  Only "unmarked when its
// turn comes" depends on the walk, so the walk is split into
//   1. the pair relation (every pair of a micrograph, in parallel): sorted
//      neighbour lists of every pick;
//   2. active(i) = unmarked at its turn = no active earlier neighbour j with
//      score_j > score_i, settled in rounds (each round settles every pick
//      whose earlier neighbours are settled, so at least the first open one);
//   3. the final mark of an active pick i: a later neighbour j with
//      score_j >= score_i that no active pick before i had marked.
// The pair test repeats the script's double-precision operations in order
// (no FMA contraction; the Euler matrices come from the host's C library, as
// the script's come from CPython's math module).  The one device
// transcendental is acos: pairs whose Euler distance lies within 1e-9 degree
// of a breakpoint of the threshold test are settled on the host with the C
// library's acos, so the result equals the script's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/gisspa.h"
#include "gs_host.hpp"

namespace gs {
namespace {

constexpr double kScriptPi = 3.14159265359;  // remove_repeat_particles_from_list.py:13
constexpr int DWG = 256;
constexpr double kAmbiguity = 1e-9;          // degrees
constexpr int kRoundsPerCheck = 4;

enum : int { ST_OPEN = 0, ST_ACTIVE = 1, ST_MARKED = 2 };

// Threshold test on the quaternion angle (:132-136): the distance itself or
// its folding about 180 degrees.
__host__ __device__ inline bool euler_close(double ed, double thr) {
#pragma clang fp contract(off)
  double d2 = fmod(ed, 180.0);
  if (d2 > thr + 1) d2 = 180 - d2;
  return ed < thr || d2 < thr;
}

__host__ __device__ inline double euler_angle(double q0) {
#pragma clang fp contract(off)
  return fabs(acos(q0) * 2 * 180 / kScriptPi);
}

// Orientation test of outer pick a (matrix ea) and inner pick b (matrix eb),
// :62-136, matrices in the script's m_o layout.  Returns 0 (not a duplicate,
// or a pair the script skips), 1 (duplicate) or 2 (undecided within the
// acos accuracy; *q0 receives the quaternion component for the host).
__device__ int pair_orientation(const double* __restrict__ ea, const double* __restrict__ eb,
                                double thr, double* q0out) {
#pragma clang fp contract(off)
  double m[9], mo[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r * 3 + c] = ea[c * 3 + r];  // the script's m = m_o^T
  for (int k = 0; k < 9; ++k) mo[k] = eb[k];
  double rm[9];
  rm[0] = m[0] * mo[0] + m[3] * mo[1] + m[6] * mo[2];
  rm[1] = mo[0] * m[1] + mo[1] * m[4] + mo[2] * m[7];
  rm[2] = mo[0] * m[2] + mo[1] * m[5] + mo[2] * m[8];
  rm[3] = mo[3] * m[0] + mo[4] * m[3] + mo[5] * m[6];
  rm[4] = mo[3] * m[1] + mo[4] * m[4] + mo[5] * m[7];
  rm[5] = mo[3] * m[2] + mo[4] * m[5] + mo[5] * m[8];
  rm[6] = mo[6] * m[0] + mo[7] * m[3] + mo[8] * m[6];
  rm[7] = mo[6] * m[1] + mo[7] * m[4] + mo[8] * m[7];
  rm[8] = mo[6] * m[2] + mo[7] * m[5] + mo[8] * m[8];
  double q0 = 0, t;
  const double tr = rm[0] + rm[4] + rm[8];
  if (tr > 0) {
    t = sqrt(1 + rm[0] + rm[4] + rm[8]) * 2;  // left to right, as :94
    if (t == 0) return 0;
    q0 = 0.25 * t;
  } else {
    // the x-dominant case (:102-109) is overwritten by the y/z case below
    // it in the script; only its t == 0 exit remains
    if (rm[0] >= rm[4] && rm[0] >= rm[8]) {
      t = sqrt(1 + rm[0] - rm[4] - rm[8]) * 2;
      if (t == 0) return 0;
    }
    if (rm[4] > rm[8]) {
      if (1 - rm[0] + rm[4] - rm[8] < 0) return 0;
      t = sqrt(1 - rm[0] + rm[4] - rm[8]) * 2;
      if (t == 0) return 0;
      q0 = (rm[2] - rm[6]) / t;
    } else {
      if (1 - rm[0] - rm[4] + rm[8] < 0) return 0;
      t = sqrt(1 - rm[0] - rm[4] + rm[8]) * 2;
      if (t == 0) return 0;
      q0 = (rm[3] - rm[1]) / t;
    }
  }
  if (q0 > 1 || q0 < -1) return 0;
  const double ed = euler_angle(q0);
  const bool c = euler_close(ed, thr);
  if (euler_close(ed - kAmbiguity, thr) != c || euler_close(ed + kAmbiguity, thr) != c) {
    *q0out = q0;
    return 2;
  }
  return c ? 1 : 0;
}

struct DedupArgs {
  const double* x;
  const double* y;
  const double* e;     // n x 9 Euler matrices
  const int* glo;      // first pick of each pick's group
  const int* ghi;      // one past its last
  int n;
  double center, center2, thr;
};

__device__ bool resolved_duplicate(const unsigned long long* keys, int nk, unsigned long long key) {
  int lo = 0, hi = nk;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < nk && keys[lo] == key;
}

// Neighbour relation.  Workgroup (bx, by) takes picks i of block bx and the
// candidates j of chunk by (ch picks wide) of the block's group range, staged
// through LDS in tiles of DWG (x, y).  FILL = false: count the duplicates per
// (chunk, pick) and list the undecided pairs (once, from their outer pick);
// FILL = true: write them at off[i] + off2[by * n + i] (chunks in order, so
// every pick's list is in index order), the undecided pairs taken from the
// host's sorted keys.
template <bool FILL>
__global__ __launch_bounds__(DWG) void k_dd_pairs(DedupArgs d, int ch, int* __restrict__ cnt,
                                                  const int* __restrict__ off,
                                                  const int* __restrict__ off2,
                                                  int* __restrict__ nbr,
                                                  int* __restrict__ amb_n, int amb_cap,
                                                  int2* __restrict__ amb_pair,
                                                  double* __restrict__ amb_q0,
                                                  const unsigned long long* __restrict__ keys,
                                                  int nkeys) {
#pragma clang fp contract(off)
  __shared__ double sx[DWG], sy[DWG];
  const int i0 = blockIdx.x * DWG;
  const int i = i0 + threadIdx.x;
  const bool valid = i < d.n;
  const int glo = d.glo[i0];
  const int jlo = glo + blockIdx.y * ch;
  const int jhi = min(jlo + ch, d.ghi[min(i0 + DWG, d.n) - 1]);
  const int lo = valid ? d.glo[i] : 0, hi = valid ? d.ghi[i] : 0;
  const double xi = valid ? d.x[i] : 0.0, yi = valid ? d.y[i] : 0.0;
  int count = 0;
  const long slot = (long)blockIdx.y * d.n + i;
  const int w0 = (FILL && valid) ? off[i] + off2[slot] : 0;
  const int wend = (FILL && valid) ? w0 + cnt[slot] : 0;
  for (int t0 = jlo; t0 < jhi; t0 += DWG) {
    __syncthreads();
    if (t0 + (int)threadIdx.x < jhi) {
      sx[threadIdx.x] = d.x[t0 + threadIdx.x];
      sy[threadIdx.x] = d.y[t0 + threadIdx.x];
    }
    __syncthreads();
    const int k0 = max(lo, t0) - t0, k1 = min(hi, min(t0 + DWG, jhi)) - t0;
    for (int k = k0; k < k1; ++k) {
      const int j = t0 + k;
      if (j == i) continue;
      // (x2 - x1), (y2 - y1) with pick 1 the outer (lower-index) one, :54
      const double dx = j > i ? sx[k] - xi : xi - sx[k];
      const double dy = j > i ? sy[k] - yi : yi - sy[k];
      const double r2 = dx * dx + dy * dy;
      if (r2 > d.center2) continue;  // conservative pre-test, the exact test follows
      if (!(fabs(sqrt(r2)) < d.center)) continue;
      const int a = min(i, j), b = max(i, j);
      double q0 = 0;
      int r = pair_orientation(d.e + 9 * (size_t)a, d.e + 9 * (size_t)b, d.thr, &q0);
      if (r == 2) {
        if (FILL) {
          const unsigned long long key = (unsigned long long)a * (unsigned)d.n + (unsigned)b;
          r = resolved_duplicate(keys, nkeys, key) ? 1 : 0;
        } else {
          if (i == a) {
            const int slot = atomicAdd(amb_n, 1);
            if (slot < amb_cap) {
              amb_pair[slot] = make_int2(a, b);
              amb_q0[slot] = q0;
            }
          }
          r = 0;  // the host adds the resolved duplicates to the counts
        }
      }
      if (r == 1) {
        // both passes classify every pair alike, so the list has room; the
        // bound only keeps a write inside it
        if (FILL && w0 + count < wend) nbr[w0 + count] = j;
        ++count;
      }
    }
  }
  if (!FILL && valid) cnt[slot] = count;
}

// Resolved undecided pairs join the counts (one slot per side).
__global__ void k_dd_bump(int* __restrict__ cnt, const long* __restrict__ slots, int ns) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns) atomicAdd(cnt + slots[k], 1);
}

// Per pick: chunk offsets relative to the pick's list (off2) and its total.
__global__ __launch_bounds__(DWG) void k_dd_chunk_offsets(const int* __restrict__ cnt, int n,
                                                          int nch, int* __restrict__ off2,
                                                          int* __restrict__ tot) {
  const int i = blockIdx.x * DWG + threadIdx.x;
  if (i >= n) return;
  int r = 0;
  for (int c = 0; c < nch; ++c) {
    off2[(long)c * n + i] = r;
    r += cnt[(long)c * n + i];
  }
  tot[i] = r;
}

// Exclusive prefix sum of tot[0..n) into off[0..n], one workgroup of SWG
// threads (each a contiguous run), the grand total also in *total.
constexpr int SWG = 1024;
__global__ __launch_bounds__(SWG) void k_dd_scan(const int* __restrict__ tot, int n,
                                                 int* __restrict__ off,
                                                 long long* __restrict__ total) {
  __shared__ long long part[SWG];
  const int per = (n + SWG - 1) / SWG;
  const int a = min(n, threadIdx.x * per), b = min(n, a + per);
  long long sum = 0;
  for (int i = a; i < b; ++i) sum += tot[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < SWG; d <<= 1) {  // inclusive scan of the run sums
    const long long v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  long long r = part[threadIdx.x] - sum;
  for (int i = a; i < b; ++i) {
    off[i] = (int)min(r, (long long)INT_MAX);
    r += tot[i];
  }
  if (threadIdx.x == SWG - 1) {
    off[n] = (int)min(part[SWG - 1], (long long)INT_MAX);
    *total = part[SWG - 1];
  }
}

__device__ inline int load_state(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_state(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One settling round: pick i is marked at its turn iff an active earlier
// neighbour outscores it; it is active iff every earlier neighbour that
// outscores it is settled and none is active.  States only move from open to
// settled, so reading a neighbour's old or new state gives the same outcome
// (at worst one round later).
__global__ __launch_bounds__(DWG) void k_dd_round(const int* __restrict__ off,
                                                  const int* __restrict__ nbr,
                                                  const double* __restrict__ s, int n,
                                                  int* state, int* __restrict__ open_count) {
  const int i = blockIdx.x * DWG + threadIdx.x;
  if (i >= n || load_state(state + i) != ST_OPEN) return;
  const double si = s[i];
  bool marked = false, pending = false;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const int j = nbr[k];
    if (j > i) break;  // lists are in index order
    if (s[j] > si) {
      const int st = load_state(state + j);
      if (st == ST_ACTIVE) {
        marked = true;
        break;
      }
      if (st == ST_OPEN) pending = true;
    }
  }
  if (marked) store_state(state + i, ST_MARKED);
  else if (!pending) store_state(state + i, ST_ACTIVE);
  else atomicAdd(open_count, 1);
}

// first_mark[j] = the first active earlier neighbour that outscores j (the
// pick whose turn marks j), n if none.
__global__ __launch_bounds__(DWG) void k_dd_first_mark(const int* __restrict__ off,
                                                       const int* __restrict__ nbr,
                                                       const double* __restrict__ s, int n,
                                                       const int* __restrict__ state,
                                                       int* __restrict__ first_mark) {
  const int j = blockIdx.x * DWG + threadIdx.x;
  if (j >= n) return;
  int f = n;
  for (int k = off[j]; k < off[j + 1]; ++k) {
    const int i = nbr[k];
    if (i > j) break;
    if (state[i] == ST_ACTIVE && s[i] > s[j]) {
      f = i;
      break;
    }
  }
  first_mark[j] = f;
}

// keep[i]: active at its turn and not marked during it, i.e. no later
// neighbour j with score_j >= score_i that was still unmarked when i's turn
// reached it (first_mark[j] > i; first_mark[j] == i cannot hold for such j).
__global__ __launch_bounds__(DWG) void k_dd_keep(const int* __restrict__ off,
                                                 const int* __restrict__ nbr,
                                                 const double* __restrict__ s, int n,
                                                 const int* __restrict__ state,
                                                 const int* __restrict__ first_mark,
                                                 unsigned char* __restrict__ keep) {
  const int i = blockIdx.x * DWG + threadIdx.x;
  if (i >= n) return;
  bool k = state[i] == ST_ACTIVE;
  if (k) {
    const double si = s[i];
    for (int q = off[i + 1] - 1; q >= off[i]; --q) {
      const int j = nbr[q];
      if (j < i) break;
      if (si <= s[j] && first_mark[j] > i) {
        k = false;
        break;
      }
    }
  }
  keep[i] = k ? 1 : 0;
}

// Euler matrix in the script's m_o layout (:63-71), angles converted with
// the script's pi (:43-45).
void euler_matrix(double alt_deg, double az_deg, double phi_deg, double* e) {
#pragma clang fp contract(off)
  const double alt = alt_deg * kScriptPi / 180, az = az_deg * kScriptPi / 180,
               phi = phi_deg * kScriptPi / 180;
  // each function value once (the script calls them again per element; the
  // values are the same)
  const double ca = std::cos(alt), sa = std::sin(alt), cz = std::cos(az), sz = std::sin(az),
               cp = std::cos(phi), sp = std::sin(phi);
  e[0] = cp * cz - sp * ca * sz;
  e[1] = cp * sz + sp * ca * cz;
  e[2] = sp * sa;
  e[3] = (-1) * sp * cz - cp * ca * sz;
  e[4] = (-1) * sp * sz + cp * ca * cz;
  e[5] = cp * sa;
  e[6] = sa * sz;
  e[7] = (-1) * sa * cz;
  e[8] = ca;
}

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};

__global__ void k_dd_fill(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill_int(int* p, int n, int v, hipStream_t st) {
  if (n > 0) k_dd_fill<<<(n + 255) / 256, 256, 0, st>>>(p, n, v);
}

// Device memory of one call.  Owning: every request is a fresh allocation,
// freed when the arena goes.  Scratch: requests are 256-byte aligned slices
// of a caller-free region (nothing is allocated); a request beyond its end
// is GS_ERR_CAPACITY.
class DevArena {
 public:
  DevArena() = default;
  DevArena(void* base, size_t cap) : base_(static_cast<char*>(base)), cap_(cap) {}
  template <class T>
  int get(size_t count, T** out) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (!base_) {
      owned_.emplace_back(new DevBuf);
      CHK(owned_.back()->ensure(bytes));
      *out = static_cast<T*>(owned_.back()->p);
      return GS_OK;
    }
    const size_t at = (used_ + 255) & ~size_t(255);
    if (at > cap_ || bytes > cap_ - at) return GS_ERR_CAPACITY;
    *out = reinterpret_cast<T*>(base_ + at);
    used_ = at + bytes;
    return GS_OK;
  }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
  std::vector<std::unique_ptr<DevBuf>> owned_;
};

// Host staging of one call: std::vector storage, or slices of a static
// region (the allocation-free entry point, which holds its lock meanwhile).
class HostArena {
 public:
  HostArena() = default;
  HostArena(void* base, size_t cap) : base_(static_cast<char*>(base)), cap_(cap) {}
  template <class T>
  int get(size_t count, T** out) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (!base_) {
      owned_.emplace_back(bytes);
      *out = reinterpret_cast<T*>(owned_.back().data());
      return GS_OK;
    }
    const size_t at = (used_ + 63) & ~size_t(63);
    if (at > cap_ || bytes > cap_ - at) return GS_ERR_CAPACITY;
    *out = reinterpret_cast<T*>(base_ + at);
    used_ = at + bytes;
    return GS_OK;
  }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
  std::vector<std::vector<unsigned char>> owned_;
};

int check_args(const gs_pick_rec* recs, size_t n_, double center, double thr, unsigned char* keep,
               size_t* n_kept) {
  if (n_kept) *n_kept = 0;
  if ((!recs || !keep) && n_ > 0) return GS_ERR_ARG;
  if (n_ > (size_t)(INT_MAX / 16)) return GS_ERR_ARG;
  if (!(center == center) || !(thr == thr)) return GS_ERR_ARG;  // NaN
  return GS_OK;
}

int select_device(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return GS_ERR_DEVICE;  // as gs_create
  }
  HIPCHK(hipSetDevice(device));
  return GS_OK;
}

// The duplicate removal of n (>= 1) picks on stream st, device buffers from
// `dev`, host staging from `host`; amb_cap0 is the first capacity of the
// list of undecided pairs (a larger count repeats the counting pass with the
// exact capacity, within max_amb).
int remove_repeats_run(DevArena& dev, HostArena& host, hipStream_t st, const gs_pick_rec* recs,
                       int n, double center, double thr, unsigned char* keep, size_t* n_kept,
                       int amb_cap0, long long max_amb) {
  // host side: structure of arrays, Euler matrices, group ranges
  double *x, *y, *s, *e;
  int *glo, *ghi;
  CHK(host.get(n, &x));
  CHK(host.get(n, &y));
  CHK(host.get(n, &s));
  CHK(host.get((size_t)n * 9, &e));
  CHK(host.get(n, &glo));
  CHK(host.get(n, &ghi));
  for (int i = 0; i < n; ++i) {
    x[i] = recs[i].x;
    y[i] = recs[i].y;
    s[i] = recs[i].score;
    euler_matrix(recs[i].alt, recs[i].az, recs[i].phi, &e[(size_t)i * 9]);
  }
  for (int lo = 0; lo < n;) {
    int hi = lo + 1;
    while (hi < n && recs[hi].group == recs[lo].group) ++hi;
    for (int i = lo; i < hi; ++i) {
      glo[i] = lo;
      ghi[i] = hi;
    }
    lo = hi;
  }

  double *dx, *dy, *ds, *de;
  int *dglo, *dghi, *doff, *damb_n;
  const size_t nd = (size_t)n * sizeof(double), ni = (size_t)n * sizeof(int);
  CHK(dev.get(n, &dx));
  CHK(dev.get(n, &dy));
  CHK(dev.get(n, &ds));
  CHK(dev.get((size_t)n * 9, &de));
  CHK(dev.get(n, &dglo));
  CHK(dev.get(n, &dghi));
  CHK(dev.get((size_t)n + 1, &doff));
  CHK(dev.get(1, &damb_n));
  HIPCHK(hipMemcpyAsync(dx, x, nd, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dy, y, nd, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ds, s, nd, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(de, e, nd * 9, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dglo, glo, ni, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dghi, ghi, ni, hipMemcpyHostToDevice, st));

  DedupArgs a;
  a.x = dx;
  a.y = dy;
  a.e = de;
  a.glo = dglo;
  a.ghi = dghi;
  a.n = n;
  a.center = center;
  // every r2 with sqrt(r2) < center satisfies r2 <= center^2 (1 + 1e-12)
  a.center2 = center > 0 ? center * center * (1 + 1e-12) : -1.0;
  a.thr = thr;
  const int blocks = (n + DWG - 1) / DWG;
  // j chunks: about 4096 workgroups in all, chunks a multiple of DWG wide
  int maxrange = 0;
  for (int bx = 0; bx < blocks; ++bx)
    maxrange = std::max(maxrange, ghi[std::min(n, (bx + 1) * DWG) - 1] - glo[bx * DWG]);
  const int want = std::max(1, 4096 / blocks);
  const int ch = std::max(DWG, (((maxrange + want - 1) / want) + DWG - 1) / DWG * DWG);
  const int nch = (maxrange + ch - 1) / ch;
  const dim3 grid2(blocks, nch);
  const size_t nslot = (size_t)n * nch;
  int *dcnt, *doff2, *dtot;
  long long* dtotal;
  CHK(dev.get(nslot, &dcnt));
  CHK(dev.get(nslot, &doff2));
  CHK(dev.get(n, &dtot));
  CHK(dev.get(1, &dtotal));

  // 1. counts and undecided pairs (a second pass if the pair list overflowed)
  int amb_cap = amb_cap0, amb_n = 0;
  int2* damb_pair;
  double* damb_q0;
  for (;;) {
    CHK(dev.get(amb_cap, &damb_pair));
    CHK(dev.get(amb_cap, &damb_q0));
    fill_int(damb_n, 1, 0, st);
    k_dd_pairs<false><<<grid2, DWG, 0, st>>>(a, ch, dcnt, nullptr, nullptr, nullptr, damb_n,
                                              amb_cap, damb_pair, damb_q0, nullptr, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&amb_n, damb_n, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (amb_n <= amb_cap) break;
    if (amb_n > max_amb) return GS_ERR_CAPACITY;
    amb_cap = amb_n;
  }
  int2* amb_pair;
  double* amb_q0;
  CHK(host.get(amb_n, &amb_pair));
  CHK(host.get(amb_n, &amb_q0));
  if (amb_n > 0) {
    HIPCHK(hipMemcpyAsync(amb_pair, damb_pair, amb_n * sizeof(int2), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(amb_q0, damb_q0, amb_n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // undecided pairs: the C library's acos decides; a duplicate joins the
  // counts of both picks, in the chunk that holds the other pick
  unsigned long long* keys;
  long* bumps;
  CHK(host.get(amb_n, &keys));
  CHK(host.get((size_t)2 * amb_n, &bumps));
  int nkeys = 0, nbumps = 0;
  auto chunk_slot = [&](int i, int j) -> long {
    return (long)((j - glo[(i / DWG) * DWG]) / ch) * n + i;
  };
  for (int k = 0; k < amb_n; ++k) {
    if (!euler_close(euler_angle(amb_q0[k]), thr)) continue;
    const int pa = amb_pair[k].x, pb = amb_pair[k].y;
    keys[nkeys++] = (unsigned long long)pa * (unsigned)n + (unsigned)pb;
    bumps[nbumps++] = chunk_slot(pa, pb);
    bumps[nbumps++] = chunk_slot(pb, pa);
  }
  std::sort(keys, keys + nkeys);
  unsigned long long* dkeys;
  long* dbump;
  CHK(dev.get(nkeys, &dkeys));
  CHK(dev.get(nbumps, &dbump));
  if (nkeys > 0) {
    HIPCHK(hipMemcpyAsync(dkeys, keys, nkeys * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          st));
    HIPCHK(hipMemcpyAsync(dbump, bumps, nbumps * sizeof(long), hipMemcpyHostToDevice, st));
    k_dd_bump<<<(nbumps + 255) / 256, 256, 0, st>>>(dcnt, dbump, nbumps);
  }
  k_dd_chunk_offsets<<<blocks, DWG, 0, st>>>(dcnt, n, nch, doff2, dtot);
  k_dd_scan<<<1, SWG, 0, st>>>(dtot, n, doff, dtotal);
  HIPCHK(hipGetLastError());
  long long total = 0;
  HIPCHK(hipMemcpyAsync(&total, dtotal, sizeof(long long), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // (keeps the host copies of the keys and bumps alive until here)
  if (total > INT_MAX) return GS_ERR_CAPACITY;
  int* dnbr;
  CHK(dev.get((size_t)total, &dnbr));

  // 2. neighbour lists
  k_dd_pairs<true><<<grid2, DWG, 0, st>>>(a, ch, dcnt, doff, doff2, dnbr, nullptr, 0, nullptr,
                                           nullptr, dkeys, nkeys);
  HIPCHK(hipGetLastError());

  // 3. settle the picks that are unmarked at their turn, in rounds
  int *dstate, *dopen;
  CHK(dev.get(n, &dstate));
  CHK(dev.get(kRoundsPerCheck, &dopen));
  fill_int(dstate, n, ST_OPEN, st);
  int rounds = 0;
  for (;;) {
    fill_int(dopen, kRoundsPerCheck, 0, st);
    for (int r = 0; r < kRoundsPerCheck; ++r)
      k_dd_round<<<blocks, DWG, 0, st>>>(doff, dnbr, ds, n, dstate, dopen + r);
    HIPCHK(hipGetLastError());
    int open = 0;
    HIPCHK(hipMemcpyAsync(&open, dopen + kRoundsPerCheck - 1, sizeof(int), hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
    rounds += kRoundsPerCheck;
    if (open == 0) break;
    if (rounds > n + kRoundsPerCheck) return GS_ERR_DEVICE;  // every round settles a pick
  }

  // 4. final marks
  int* dfirst;
  unsigned char* dkeep;
  CHK(dev.get(n, &dfirst));
  CHK(dev.get(n, &dkeep));
  k_dd_first_mark<<<blocks, DWG, 0, st>>>(doff, dnbr, ds, n, dstate, dfirst);
  k_dd_keep<<<blocks, DWG, 0, st>>>(doff, dnbr, ds, n, dstate, dfirst, dkeep);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(keep, dkeep, n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (n_kept) {
    size_t kept = 0;
    for (int i = 0; i < n; ++i) kept += keep[i] ? 1 : 0;
    *n_kept = kept;
  }
  return GS_OK;
}

int remove_repeats(int device, const gs_pick_rec* recs, size_t n_, double center, double thr,
                   unsigned char* keep, size_t* n_kept) {
  CHK(check_args(recs, n_, center, thr, keep, n_kept));
  if (n_ == 0) return GS_OK;
  CHK(select_device(device));
  StreamGuard sg;
  HIPCHK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  DevArena dev;
  HostArena host;
  return remove_repeats_run(dev, host, sg.s, recs, (int)n_, center, thr, keep, n_kept, 1 << 16,
                            LLONG_MAX);
}

// ---------------------------------------------------------------------------
// Allocation-free variant: every device buffer of the call is a slice of
// `scratch`, one static array per device (each GPU the library runs on holds
// it until the process ends); the host staging is a static region.  Calls
// are serialised by a lock (they share the staging); calls from other
// processes on the same GPU have their own copy of the array.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)GS_NOALLOC_SCRATCH_MIB << 20;
constexpr size_t kHostStageBytes = (size_t)12 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
alignas(64) static unsigned char host_stage[kHostStageBytes];
std::mutex noalloc_lock;
constexpr int kMaxDevices = 64;
void* scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

int remove_repeats_noalloc(int device, const gs_pick_rec* recs, size_t n_, double center,
                           double thr, unsigned char* keep, size_t* n_kept) {
  CHK(check_args(recs, n_, center, thr, keep, n_kept));
  if (n_ > (size_t)GS_NOALLOC_MAX_PICKS) return GS_ERR_CAPACITY;
  if (n_ == 0) return GS_OK;
  if (device >= kMaxDevices) return GS_ERR_DEVICE;
  CHK(select_device(device));
  std::lock_guard<std::mutex> g(noalloc_lock);
  if (!scratch_ptr[device]) {
    // the array's name is a host-side stand-in: its device address on the
    // current device comes from the runtime, once per device
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return GS_ERR_DEVICE;
    scratch_ptr[device] = p;
  }
  DevArena dev(scratch_ptr[device], kScratchBytes);
  HostArena host(host_stage, kHostStageBytes);
  // the per-thread default stream: no stream is created
  return remove_repeats_run(dev, host, hipStreamPerThread, recs, (int)n_, center, thr, keep,
                            n_kept, GS_NOALLOC_MAX_UNDECIDED, GS_NOALLOC_MAX_UNDECIDED);
}

}  // namespace
}  // namespace gs

extern "C" int gs_remove_repeats(int device, const gs_pick_rec* recs, size_t n, double center,
                                 double euler_thres, unsigned char* keep, size_t* n_kept) {
  return gs::remove_repeats(device, recs, n, center, euler_thres, keep, n_kept);
}

extern "C" int gs_remove_repeats_noalloc(int device, const gs_pick_rec* recs, size_t n,
                                         double center, double euler_thres, unsigned char* keep,
                                         size_t* n_kept) {
  return gs::remove_repeats_noalloc(device, recs, n, center, euler_thres, keep, n_kept);
}
