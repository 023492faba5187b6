// Elementwise passes of the DFSPH step (crates/pe-fluid/src/dfsph.rs:713-767)
// on the device; the pair passes (alpha, stiffness, pressure velocity update)
// are in pef_pair.hip, the density and XSPH passes are the WCSPH step's.
//
//   non-pressure acceleration + predicted velocity (dfsph.rs:355-379, 722-733)
//                                                         -> k_df_predict
//   warm start of a solver (dfsph.rs:531-545, 584-597)    -> k_df_warm_prep/zero
//   the solver's mean error and stopping rule              -> k_df_reduce
//   drift + free-slip planes (dfsph.rs:660-711, 737-744)  -> k_df_drift
//   free-slip after the divergence solver (dfsph.rs:753)  -> k_df_slip
//   v = v* (dfsph.rs:755-764)                              -> k_df_finish
//
// Velocities: B holds v (its w the density solver's kappa, the reported
// pressure); the step puts v + g dt in B, XSPH writes v* to B2, the solvers
// update B2 in place and the step ends with B = B2.
#include "pef_common.h"
#include "pef_launch.h"
#include "pef_pair.h"

namespace pef {

#define PEF_LAUNCH(kernel, grid, block, stream, ...) \
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)

static uint32_t blocks_for(uint64_t n, uint32_t bs) {
  uint64_t b = (n + bs - 1) / bs;
  return (uint32_t)(b ? b : 1);
}

__global__ void k_df_begin(DfCtl* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->done = 0u;
    c->stop = 0u;
    c->iter = 0u;
    c->arrive = 0u;
    c->avg = 0.0;
  }
}

__global__ void k_df_predict(uint32_t N, const SceneDev* __restrict__ sc, State cur, float dt) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t meta = cur.META[s];
  if (meta & kMetaBoundary) return;  // a boundary sample's v* is its velocity
  const SceneDev& S = sc[scene_of(meta)];
  float4 v = cur.B[s];
  v.x += S.g[0] * dt;
  v.y += S.g[1] * dt;
  v.z = (S.dim == 2) ? 0.0f : v.z + S.g[2] * dt;
  cur.B[s] = v;
}

// kappa = warm sum; K = kappa / max(rho, 1e-12) for the update pass.
__global__ void k_df_warm_prep(uint32_t N, State cur, int which) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  float* b2w = reinterpret_cast<float*>(cur.B2) + 4ull * s + 3;
  if (cur.META[s] & kMetaBoundary) {
    *b2w = 0.0f;
    return;
  }
  const float2 w = cur.DW[s];
  const float k = which ? w.y : w.x;
  *b2w = k / fmaxf(cur.A2[s].w, 1e-12f);
  if (!which) reinterpret_cast<float*>(cur.B)[4ull * s + 3] = k;
}

__global__ void k_df_warm_zero(uint32_t N, State cur, int which) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  float* w = reinterpret_cast<float*>(cur.DW) + 2ull * s + (which ? 1 : 0);
  *w = 0.0f;
}

// Mean of ERR over the fluid particles, FP64 in a fixed order: each workgroup
// sums a fixed slot range by a fixed tree, and the last one to finish sums the
// workgroup sums by the same tree. Then the stopping rule: avg <= tol after at
// least min_iter iterations, or max_iter iterations (dfsph.rs:573-577,
// 616-620).
constexpr uint32_t kDfReduceThreads = 256;
constexpr uint32_t kDfReduceSpan = 4096;  // slots per workgroup

__global__ void __launch_bounds__(kDfReduceThreads) k_df_reduce(uint32_t N, const double* __restrict__ ERR,
                                                                 double* __restrict__ part, DfCtl* c, double count,
                                                                 double tol, uint32_t min_iter, uint32_t max_iter) {
  if (c->done | c->stop) return;
  __shared__ double red[kDfReduceThreads];
  const uint64_t s0 = (uint64_t)blockIdx.x * kDfReduceSpan;
  const uint64_t s1 = s0 + kDfReduceSpan < N ? s0 + kDfReduceSpan : N;
  double acc = 0.0;
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += kDfReduceThreads) acc += ERR[s];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t o = kDfReduceThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  __shared__ uint32_t last;
  if (threadIdx.x == 0) {
    part[blockIdx.x] = red[0];
    __threadfence();
    last = atomicAdd(&c->arrive, 1u) == gridDim.x - 1u ? 1u : 0u;
    __threadfence();
  }
  __syncthreads();
  if (!last) return;
  // the last workgroup: the workgroup sums by the same fixed strided tree
  acc = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kDfReduceThreads) acc += part[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t o = kDfReduceThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double avg = count > 0.0 ? red[0] / count : 0.0;
  const uint32_t it = c->iter + 1u;
  c->iter = it;
  c->avg = avg;
  c->stop = ((avg <= tol && it >= min_iter) || it >= max_iter) ? 1u : 0u;
  c->arrive = 0u;
}

// Clamp a fluid position and its predicted velocity to the free-slip planes.
__device__ inline void free_slip(double& x, double& y, float& vx, float& vy, const DfSlip& p) {
  if (p.on[0] && x < p.at[0]) {
    x = p.at[0];
    if (vx < 0.0f) vx = 0.0f;
  }
  if (p.on[1] && y < p.at[1]) {
    y = p.at[1];
    if (vy < 0.0f) vy = 0.0f;
  }
  if (p.on[2] && x > p.at[2]) {
    x = p.at[2];
    if (vx > 0.0f) vx = 0.0f;
  }
  if (p.on[3] && y > p.at[3]) {
    y = p.at[3];
    if (vy > 0.0f) vy = 0.0f;
  }
}

// New position of fluid slot s: the position frame, the displacement bound
// and the skin test of the lists.
__device__ inline void move_to(State& cur, const SceneDev& S, uint32_t s, double x, double y, double z, Ctl* ctlw) {
  const double dx = x - cur.X[s], dy = y - cur.Y[s], dz = z - cur.Z[s];
  cur.X[s] = x;
  cur.Y[s] = y;
  cur.Z[s] = z;
  cur.A[s] = make_float4(__int_as_float(to_units(x, S.org[0], S.inv_unit)),
                         __int_as_float(to_units(y, S.org[1], S.inv_unit)),
                         __int_as_float(to_units(z, S.org[2], S.inv_unit)), S.mass);
  const float d = cur.DISP[s] + (float)sqrt(dx * dx + dy * dy + dz * dz);
  cur.DISP[s] = d;
  if (!(d <= S.skin_half)) ctlw->need_rebuild = 1u;  // also catches NaN
}

__global__ void k_df_drift(uint32_t N, const SceneDev* __restrict__ sc, State cur, Ctl* ctlw, double dt, DfSlip slip) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t meta = cur.META[s];
  if (meta & kMetaBoundary) return;
  const SceneDev& S = sc[scene_of(meta)];
  const float4 v = cur.B2[s];
  float vx = v.x, vy = v.y;
  double x = cur.X[s] + (double)v.x * dt, y = cur.Y[s] + (double)v.y * dt;
  const double z = (S.dim == 2) ? 0.0 : cur.Z[s] + (double)v.z * dt;
  free_slip(x, y, vx, vy, slip);
  float* b2 = reinterpret_cast<float*>(cur.B2) + 4ull * s;
  b2[0] = vx;
  b2[1] = vy;
  move_to(cur, S, s, x, y, z, ctlw);
}

__global__ void k_df_slip(uint32_t N, const SceneDev* __restrict__ sc, State cur, Ctl* ctlw, DfSlip slip) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t meta = cur.META[s];
  if (meta & kMetaBoundary) return;
  const SceneDev& S = sc[scene_of(meta)];
  const float4 v = cur.B2[s];
  float vx = v.x, vy = v.y;
  double x = cur.X[s], y = cur.Y[s];
  free_slip(x, y, vx, vy, slip);
  float* b2 = reinterpret_cast<float*>(cur.B2) + 4ull * s;
  b2[0] = vx;
  b2[1] = vy;
  if (x != cur.X[s] || y != cur.Y[s]) move_to(cur, S, s, x, y, cur.Z[s], ctlw);
}

__global__ void k_df_finish(uint32_t N, State cur) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  if (cur.META[s] & kMetaBoundary) return;
  const float4 v = cur.B2[s];
  float* b = reinterpret_cast<float*>(cur.B) + 4ull * s;
  b[0] = v.x;
  b[1] = v.y;
  b[2] = v.z;
}

// A per-slot FP32 array in original order.
__global__ void k_df_unsort_f32(uint32_t N, State cur, const float* __restrict__ src, double* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < N) out[cur.ID[s]] = (double)src[s];
}

void launch_df_begin(hipStream_t st, DfCtl* c) { PEF_LAUNCH(k_df_begin, 1, 64, st, c); }
void launch_df_unsort_f32(hipStream_t st, uint32_t N, const State& cur, const float* src, double* out) {
  PEF_LAUNCH(k_df_unsort_f32, blocks_for(N, 256), 256, st, N, cur, src, out);
}
void launch_df_predict(hipStream_t st, uint32_t N, const SceneDev* sc, const State& cur, float dt) {
  PEF_LAUNCH(k_df_predict, blocks_for(N, 256), 256, st, N, sc, cur, dt);
}
void launch_df_warm_prep(hipStream_t st, uint32_t N, const State& cur, int which) {
  PEF_LAUNCH(k_df_warm_prep, blocks_for(N, 256), 256, st, N, cur, which);
}
void launch_df_warm_zero(hipStream_t st, uint32_t N, const State& cur, int which) {
  PEF_LAUNCH(k_df_warm_zero, blocks_for(N, 256), 256, st, N, cur, which);
}
uint32_t df_reduce_blocks(uint32_t N) { return blocks_for(N, kDfReduceSpan); }
void launch_df_reduce(hipStream_t st, uint32_t N, const double* ERR, double* part, DfCtl* c, double count, double tol,
                      uint32_t min_iter, uint32_t max_iter) {
  PEF_LAUNCH(k_df_reduce, df_reduce_blocks(N), kDfReduceThreads, st, N, ERR, part, c, count, tol, min_iter, max_iter);
}
void launch_df_drift(hipStream_t st, uint32_t N, const SceneDev* sc, const State& cur, Ctl* ctlw, double dt,
                     const DfSlip& slip) {
  PEF_LAUNCH(k_df_drift, blocks_for(N, 256), 256, st, N, sc, cur, ctlw, dt, slip);
}
void launch_df_slip(hipStream_t st, uint32_t N, const SceneDev* sc, const State& cur, Ctl* ctlw, const DfSlip& slip) {
  PEF_LAUNCH(k_df_slip, blocks_for(N, 256), 256, st, N, sc, cur, ctlw, slip);
}
void launch_df_finish(hipStream_t st, uint32_t N, const State& cur) {
  PEF_LAUNCH(k_df_finish, blocks_for(N, 256), 256, st, N, cur);
}

}  // namespace pef
