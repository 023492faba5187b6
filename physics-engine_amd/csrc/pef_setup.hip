// Scene construction: Akinci boundary volumes on the GPU and the lattice
// builders of particles.rs. This translation unit is compiled with
// -ffp-contract=off: psi must equal the reference's FP64 sums bit for bit, so
// every expression is evaluated as written, one rounding per operation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pef_common.h"
#include "pef_launch.h"

namespace pef {

// Cubic spline W in FP64, operation for operation as kernels.rs:40-100.
__device__ __host__ inline double cubic_w_f64(double r, double h, int dim) {
  if (h <= 0.0 || r < 0.0) return 0.0;
  double q = r / h;
  if (q >= 2.0) return 0.0;
  const double pi = 3.14159265358979323846;
  double a = (dim == 2) ? 10.0 / (7.0 * pi * h * h) : 1.0 / (pi * h * h * h);
  if (q < 1.0) return a * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
  double t = 2.0 - q;
  return a * 0.25 * t * t * t;
}

constexpr int kPsiCap = 1024;  // partners within 2h of one boundary sample

struct PsiGrid {
  double inv_cell;
  int32_t ix0, iy0, iz0, nx, ny, nz;
  int32_t dim;
};

__global__ void k_psi_count(uint32_t nb, PsiGrid g, const double* __restrict__ bx, uint32_t* __restrict__ cell_of,
                            uint32_t* __restrict__ rank_of, uint32_t* __restrict__ counts) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  int32_t kx = floor_i32(bx[3ull * k] * g.inv_cell) - g.ix0;
  int32_t ky = floor_i32(bx[3ull * k + 1] * g.inv_cell) - g.iy0;
  int32_t kz = (g.dim == 2) ? 0 : floor_i32(bx[3ull * k + 2] * g.inv_cell) - g.iz0;
  kx = kx < 0 ? 0 : (kx >= g.nx ? g.nx - 1 : kx);  // bounds come from the same keys; clamp is defensive
  ky = ky < 0 ? 0 : (ky >= g.ny ? g.ny - 1 : ky);
  kz = kz < 0 ? 0 : (kz >= g.nz ? g.nz - 1 : kz);
  uint32_t c = ((uint32_t)kx * (uint32_t)g.ny + (uint32_t)ky) * (uint32_t)g.nz + (uint32_t)kz;
  cell_of[k] = c;
  rank_of[k] = atomicAdd(&counts[c], 1u);
}

__global__ void k_psi_scatter(uint32_t nb, const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ rank_of,
                              const uint32_t* __restrict__ start, uint32_t* __restrict__ members) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  members[start[cell_of[k]] + rank_of[k]] = k;
}

// psi_b = 1 / (W(0) + sum over other samples with r < 2h, ascending index)
// (particles.rs:252-279).
__global__ void k_psi(uint32_t nb, PsiGrid g, double h, const double* __restrict__ bx,
                      const uint32_t* __restrict__ cell_of, const uint32_t* __restrict__ start,
                      const uint32_t* __restrict__ members, double* __restrict__ psi, uint32_t* __restrict__ overflow) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const double support = 2.0 * h;
  const double xb = bx[3ull * k], yb = bx[3ull * k + 1], zb = bx[3ull * k + 2];
  uint32_t c = cell_of[k];
  int32_t lz = (int32_t)(c % (uint32_t)g.nz);
  uint32_t t = c / (uint32_t)g.nz;
  int32_t ly = (int32_t)(t % (uint32_t)g.ny);
  int32_t lx = (int32_t)(t / (uint32_t)g.ny);
  uint32_t part[kPsiCap];
  int n = 0;
  const int zlo = (g.dim == 2) ? 0 : -1, zhi = (g.dim == 2) ? 0 : 1;
  for (int dz = zlo; dz <= zhi; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int32_t x = lx + dx, y = ly + dy, z = lz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
        uint32_t cc = ((uint32_t)x * (uint32_t)g.ny + (uint32_t)y) * (uint32_t)g.nz + (uint32_t)z;
        for (uint32_t p = start[cc]; p < start[cc + 1]; ++p) {
          uint32_t j = members[p];
          if (j == k) continue;
          double ddx = bx[3ull * j] - xb, ddy = bx[3ull * j + 1] - yb, ddz = bx[3ull * j + 2] - zb;
          double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
          if (!(r < support)) continue;
          if (n < kPsiCap) {
            // insertion into ascending order
            int q = n;
            while (q > 0 && part[q - 1] > j) {
              part[q] = part[q - 1];
              --q;
            }
            part[q] = j;
          } else {
            atomicOr(overflow, 1u);
          }
          ++n;
        }
      }
  double delta = cubic_w_f64(0.0, h, g.dim);
  int m = n < kPsiCap ? n : kPsiCap;
  for (int q = 0; q < m; ++q) {
    uint32_t j = part[q];
    double ddx = bx[3ull * j] - xb, ddy = bx[3ull * j + 1] - yb, ddz = bx[3ull * j + 2] - zb;
    double r = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    delta += cubic_w_f64(r, h, g.dim);
  }
  psi[k] = (delta > 1e-30) ? 1.0 / delta : 0.0;
}

// Exclusive scan on the host side of a small device array is avoided: the
// counts are scanned on the device by one block per 8192 cells.
__global__ void k_psi_scan_local(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                 uint32_t* __restrict__ sums, uint32_t C) {
  __shared__ uint32_t sh[1024];
  uint32_t base = blockIdx.x * 8192u + threadIdx.x * 8u;
  uint32_t v[8], s = 0;
  for (int q = 0; q < 8; ++q) {
    v[q] = (base + q < C) ? in[base + q] : 0u;
    s += v[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele over the 1024 thread sums
    uint32_t x = (threadIdx.x >= (uint32_t)o) ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t pre = sh[threadIdx.x] - s;
  for (int q = 0; q < 8; ++q) {
    if (base + q < C) out[base + q] = pre;
    pre += v[q];
  }
  if (threadIdx.x == 1023) sums[blockIdx.x] = sh[1023];
}

__global__ void k_psi_scan_fix(uint32_t* __restrict__ out, const uint32_t* __restrict__ sums, uint32_t C,
                               uint32_t nblk) {
  // block offsets are summed serially per thread: nblk is small for scene setup
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > C) return;
  uint32_t b = i / 8192u, off = 0;
  for (uint32_t q = 0; q < b && q < nblk; ++q) off += sums[q];
  if (i < C) {
    out[i] += off;
  } else {
    uint32_t tot = 0;
    for (uint32_t q = 0; q < nblk; ++q) tot += sums[q];
    out[C] = tot;
  }
}

#define PEF_CHECK(x)                   \
  do {                                 \
    if ((x) != hipSuccess) goto fail;  \
  } while (0)

int akinci_psi_gpu(hipStream_t st, int dim, double h, uint64_t nb, const double* bx_host, double* psi_host) {
  if (nb == 0) return 0;
  const double support = 2.0 * h;
  PsiGrid g;
  g.dim = dim;
  g.inv_cell = 1.0 / support;
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (uint64_t k = 0; k < nb; ++k)
    for (int a = 0; a < 3; ++a) {
      int32_t key = (a == 2 && dim == 2) ? 0 : floor_i32(bx_host[3 * k + a] * g.inv_cell);
      mn[a] = std::min(mn[a], key);
      mx[a] = std::max(mx[a], key);
    }
  g.ix0 = mn[0];
  g.iy0 = mn[1];
  g.iz0 = mn[2];
  g.nx = mx[0] - mn[0] + 1;
  g.ny = mx[1] - mn[1] + 1;
  g.nz = mx[2] - mn[2] + 1;
  uint64_t C64 = (uint64_t)g.nx * (uint64_t)g.ny * (uint64_t)g.nz;
  if (C64 > (1ull << 30) || nb > (1ull << 31)) return 2;
  uint32_t C = (uint32_t)C64;
  uint32_t nblk = (C + 8191u) / 8192u;
  double *d_bx = nullptr, *d_psi = nullptr;
  uint32_t *d_cell = nullptr, *d_rank = nullptr, *d_counts = nullptr, *d_start = nullptr, *d_members = nullptr,
           *d_sums = nullptr, *d_ovf = nullptr;
  uint32_t ovf = 0;
  int rc = 4;
  const uint32_t n32 = (uint32_t)nb;
  const uint32_t bs = 256;
  const uint32_t gb = (n32 + bs - 1) / bs;
  PEF_CHECK(hipMalloc(&d_bx, 3 * nb * sizeof(double)));
  PEF_CHECK(hipMalloc(&d_psi, nb * sizeof(double)));
  PEF_CHECK(hipMalloc(&d_cell, nb * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_rank, nb * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_members, nb * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_counts, (size_t)C * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_start, ((size_t)C + 1) * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_sums, ((size_t)nblk + 1) * sizeof(uint32_t)));
  PEF_CHECK(hipMalloc(&d_ovf, sizeof(uint32_t)));
  PEF_CHECK(hipMemcpyAsync(d_bx, bx_host, 3 * nb * sizeof(double), hipMemcpyHostToDevice, st));
  PEF_CHECK(hipMemsetAsync(d_counts, 0, (size_t)C * sizeof(uint32_t), st));
  PEF_CHECK(hipMemsetAsync(d_ovf, 0, sizeof(uint32_t), st));
  hipLaunchKernelGGL(k_psi_count, dim3(gb), dim3(bs), 0, st, n32, g, d_bx, d_cell, d_rank, d_counts);
  hipLaunchKernelGGL(k_psi_scan_local, dim3(nblk), dim3(1024), 0, st, d_counts, d_start, d_sums, C);
  hipLaunchKernelGGL(k_psi_scan_fix, dim3((C + 1 + 255) / 256), dim3(256), 0, st, d_start, d_sums, C, nblk);
  hipLaunchKernelGGL(k_psi_scatter, dim3(gb), dim3(bs), 0, st, n32, d_cell, d_rank, d_start, d_members);
  hipLaunchKernelGGL(k_psi, dim3(gb), dim3(bs), 0, st, n32, g, h, d_bx, d_cell, d_start, d_members, d_psi, d_ovf);
  PEF_CHECK(hipGetLastError());
  PEF_CHECK(hipMemcpyAsync(psi_host, d_psi, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  PEF_CHECK(hipMemcpyAsync(&ovf, d_ovf, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  PEF_CHECK(hipStreamSynchronize(st));
  rc = ovf ? 8 : 0;
fail:
  (void)hipFree(d_bx);
  (void)hipFree(d_psi);
  (void)hipFree(d_cell);
  (void)hipFree(d_rank);
  (void)hipFree(d_members);
  (void)hipFree(d_counts);
  (void)hipFree(d_start);
  (void)hipFree(d_sums);
  (void)hipFree(d_ovf);
  return rc;
}

}  // namespace pef

// ---------------------------------------------------------------------------
// Host scene builders (particles.rs:164-347), FP64, evaluated as written.
// ---------------------------------------------------------------------------
extern "C" {

int32_t pef_mass_from_spacing(double rest_density, double dx, int32_t dim, double* out) {
  if (!out) return 1;
  if (dim != 2 && dim != 3) return 3;
  *out = (dim == 2) ? rest_density * dx * dx : rest_density * dx * dx * dx;
  return 0;
}

int32_t pef_mass_calibrated_to_rest(double rest_density, double dx, double h, int32_t dim, double* out) {
  if (!out) return 1;
  if ((dim != 2 && dim != 3) || !(dx > 0.0) || !(h > 0.0) || !std::isfinite(rest_density)) return 3;
  const double support = 2.0 * h;
  int n = (int)std::ceil(support / dx) + 1;
  double sum_w = pef::cubic_w_f64(0.0, h, dim);
  for (int ix = -n; ix <= n; ++ix)
    for (int iy = -n; iy <= n; ++iy) {
      int zlo = (dim == 2) ? 0 : -n, zhi = (dim == 2) ? 0 : n;
      for (int iz = zlo; iz <= zhi; ++iz) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        double r = std::sqrt(((double)ix * dx) * ((double)ix * dx) + ((double)iy * dx) * ((double)iy * dx) +
                             ((double)iz * dx) * ((double)iz * dx));
        if (r < support) sum_w += pef::cubic_w_f64(r, h, dim);
      }
    }
  if (sum_w < 1e-30) return pef_mass_from_spacing(rest_density, dx, dim, out);
  *out = rest_density / sum_w;
  return 0;
}

int32_t pef_fill_box(int32_t dim, const double origin[3], uint64_t nx, uint64_t ny, uint64_t nz, double dx,
                     double* out, uint64_t cap, uint64_t* n) {
  if (!origin || !n) return 1;
  if (dim != 2 && dim != 3) return 3;
  if (dim == 2) nz = 1;
  uint64_t total = nx * ny * nz;
  *n = total;
  if (!out || cap < 3 * total) return 2;
  uint64_t k = 0;
  for (uint64_t iz = 0; iz < nz; ++iz)
    for (uint64_t iy = 0; iy < ny; ++iy)
      for (uint64_t ix = 0; ix < nx; ++ix) {
        out[3 * k] = origin[0] + (double)ix * dx;
        out[3 * k + 1] = origin[1] + (double)iy * dx;
        out[3 * k + 2] = origin[2] + (double)iz * dx;
        ++k;
      }
  return 0;
}

static uint64_t round_count(double span, double dx) {
  double r = std::round(span / dx);  // f64::round: half away from zero; `as usize` saturates
  if (!(r > 0.0)) return 0;
  return (uint64_t)r;
}

int32_t pef_fill_box_boundary(int32_t dim, const double mn[3], const double mx[3], double dx, uint64_t layers,
                              double* out, uint64_t cap, uint64_t* n) {
  if (!mn || !mx || !n) return 1;
  if (dim != 2 && dim != 3) return 3;
  if (layers < 1) layers = 1;
  std::vector<double> pts;
  auto push = [&pts](double a, double b, double c) {
    pts.push_back(a);
    pts.push_back(b);
    pts.push_back(c);
  };
  if (dim == 2) {
    uint64_t nx = round_count(mx[0] - mn[0], dx) + 1, ny = round_count(mx[1] - mn[1], dx) + 1;
    for (uint64_t layer = 0; layer < layers; ++layer) {
      double off = (double)layer * dx;
      for (uint64_t i = 0; i < nx; ++i) {
        double x = mn[0] + (double)i * dx;
        push(x, mn[1] - off, 0.0);
        push(x, mx[1] + off, 0.0);
      }
      for (uint64_t j = 1; j + 1 < ny; ++j) {
        double y = mn[1] + (double)j * dx;
        push(mn[0] - off, y, 0.0);
        push(mx[0] + off, y, 0.0);
      }
    }
  } else {
    uint64_t nx = round_count(mx[0] - mn[0], dx) + 1, ny = round_count(mx[1] - mn[1], dx) + 1,
             nz = round_count(mx[2] - mn[2], dx) + 1;
    for (uint64_t layer = 0; layer < layers; ++layer) {
      double off = (double)layer * dx;
      for (uint64_t i = 0; i < nx; ++i)
        for (uint64_t j = 0; j < ny; ++j) {
          double x = mn[0] + (double)i * dx, y = mn[1] + (double)j * dx;
          push(x, y, mn[2] - off);
          push(x, y, mx[2] + off);
        }
      for (uint64_t i = 0; i < nx; ++i)
        for (uint64_t k = 1; k + 1 < nz; ++k) {
          double x = mn[0] + (double)i * dx, z = mn[2] + (double)k * dx;
          push(x, mn[1] - off, z);
          push(x, mx[1] + off, z);
        }
      for (uint64_t j = 1; j + 1 < ny; ++j)
        for (uint64_t k = 1; k + 1 < nz; ++k) {
          double y = mn[1] + (double)j * dx, z = mn[2] + (double)k * dx;
          push(mn[0] - off, y, z);
          push(mx[0] + off, y, z);
        }
    }
  }
  *n = pts.size() / 3;
  if (!out || cap < pts.size()) return 2;
  std::memcpy(out, pts.data(), pts.size() * sizeof(double));
  return 0;
}

int32_t pef_recomputed_akinci_psi(int32_t dim, double h, uint64_t n, const double* x, const uint8_t* kind,
                                  double* psi) {
  if (!x || !kind || !psi) return 1;
  if ((dim != 2 && dim != 3) || !(h > 0.0)) return 3;
  try {
    std::vector<uint64_t> bound;
    for (uint64_t i = 0; i < n; ++i)
      if (kind[i] == 1) bound.push_back(i);
    std::vector<double> bx(3 * bound.size()), out(bound.size());
    for (size_t k = 0; k < bound.size(); ++k) {
      for (int a = 0; a < 3; ++a) {
        double v = x[3 * bound[k] + a];
        if (!std::isfinite(v)) return 3;
        bx[3 * k + a] = v;
      }
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 9;
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) return 9;
    int rc = pef::akinci_psi_gpu(st, dim, h, bound.size(), bx.data(), out.data());
    (void)hipStreamDestroy(st);
    if (rc != 0) return rc;
    for (size_t k = 0; k < bound.size(); ++k) psi[bound[k]] = out[k];
    return 0;
  } catch (...) {
    return 4;
  }
}

namespace {
// xoshiro256** with SplitMix64 seeding (Blackman & Vigna; rng.rs:14-72).
struct Xoshiro {
  uint64_t s[4];
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  explicit Xoshiro(uint64_t seed) {
    uint64_t sm = seed;
    for (auto& v : s) {
      sm += 0x9e3779b97f4a7c15ULL;
      uint64_t z = sm;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      v = z ^ (z >> 31);
    }
    if (!s[0] && !s[1] && !s[2] && !s[3]) s[0] = 0x9e3779b97f4a7c15ULL;
  }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform(double lo, double hi) {
    double u = (double)(next() >> 11) * (1.0 / (double)(1ULL << 53));
    return lo + (hi - lo) * u;
  }
};
}  // namespace

int32_t pef_jitter_positions(uint64_t seed, double amp, uint64_t n, const uint8_t* kind, double* x) {
  if (!kind || !x) return 1;
  if (!std::isfinite(amp)) return 3;
  Xoshiro r(seed);
  for (uint64_t i = 0; i < n; ++i) {
    if (kind[i] != 0) continue;
    x[3 * i] += r.uniform(-amp, amp);
    x[3 * i + 1] += r.uniform(-amp, amp);
    x[3 * i + 2] += r.uniform(-amp, amp);
  }
  return 0;
}

}  // extern "C"
