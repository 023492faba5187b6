// CPU f64 restatement of the learned force surrogate: the 32 neighbourhood
// features, the MLP 32 -> 128 -> 128 -> 128 -> 3 (SiLU) forward pass, its
// loss gradient and the Adam update.
//
// TEST INFRASTRUCTURE ONLY, like wcsph_oracle.cpp: only tests/ and the
// checker legs may load it; the product path never does.
//
// The surrogate has no counterpart in the reference; these functions restate
// the specification the device implements (include/pef.h), in f64:
//   features ....... brute-force O(N^2) partner search, f64 sums. Pinned by
//                    closed forms: a particle alone has zero moments, a
//                    mirror-symmetric neighbourhood has zero position moments,
//                    a uniform velocity field zero velocity moments, and a
//                    single partner puts V (x_j - x_i) / h in the shell
//                    floor(2.5 q).
//   MLP gradient ... analytic backward pass, pinned by central finite
//                    differences of the f64 loss (tests/test_oracle_mlp.py).
//   bf16 mode ...... rounds to BF16 (nearest even) where the device does:
//                    standardised inputs (computed in FP32), weights, hidden
//                    activations and every layer's output gradient. Sums stay
//                    f64, so the device differs only by FP32 accumulation.
//   Adam ........... bias-corrected Adam (Kingma and Ba, 2015, Algorithm 1),
//                    pinned by its closed form after one and two steps.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 32, kHid = 128, kOut = 3;
constexpr uint32_t kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid;
constexpr uint32_t kW3 = kB2 + kHid, kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid * kOut;
constexpr uint32_t kParams = kB4 + kOut;

float bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return f;
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
double rnd(double x, int bf) { return bf ? (double)bf16_round((float)x) : x; }
double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }
double silu(double z) { return z * sigmoid(z); }
double dsilu(double z) {
  const double s = sigmoid(z);
  return s * (1.0 + z * (1.0 - s));
}

struct Net {
  const double* p;
  int bf;
  double w(uint32_t off, uint32_t i, uint32_t o, uint32_t nout) const { return rnd(p[off + i * nout + o], bf); }
};

// Standardised input of one sample; in bf16 mode computed in FP32 as the
// device does, then rounded.
void input(const float* x, const double* norm, int bf, double out[kIn]) {
  for (int k = 0; k < kIn; ++k) {
    if (bf) {
      const float a = x[k] - (float)norm[k];
      out[k] = bf16_round(a * (float)norm[kIn + k]);
    } else {
      out[k] = ((double)x[k] - norm[k]) * norm[kIn + k];
    }
  }
}

struct Fwd {
  double h0[kIn], z1[kHid], h1[kHid], z2[kHid], h2[kHid], z3[kHid], h3[kHid], y[kOut];
};

void forward(const Net& n, const float* x, const double* norm, Fwd& f) {
  input(x, norm, n.bf, f.h0);
  for (int o = 0; o < kHid; ++o) {
    double s = n.p[kB1 + o];
    for (int i = 0; i < kIn; ++i) s += f.h0[i] * n.w(kW1, i, o, kHid);
    f.z1[o] = s;
    f.h1[o] = rnd(silu(s), n.bf);
  }
  for (int o = 0; o < kHid; ++o) {
    double s = n.p[kB2 + o];
    for (int i = 0; i < kHid; ++i) s += f.h1[i] * n.w(kW2, i, o, kHid);
    f.z2[o] = s;
    f.h2[o] = rnd(silu(s), n.bf);
  }
  for (int o = 0; o < kHid; ++o) {
    double s = n.p[kB3 + o];
    for (int i = 0; i < kHid; ++i) s += f.h2[i] * n.w(kW3, i, o, kHid);
    f.z3[o] = s;
    f.h3[o] = rnd(silu(s), n.bf);
  }
  for (int o = 0; o < kOut; ++o) {
    double s = n.p[kB4 + o];
    for (int i = 0; i < kHid; ++i) s += f.h3[i] * n.w(kW4, i, o, kOut);
    f.y[o] = s;
  }
}

}  // namespace

extern "C" {

// Predicted accelerations [n][3] = network output * out_scale. norm holds
// in_mean[32], in_inv_std[32], out_scale[3].
void ora_mlp_forward(const double* params, const double* norm, const float* X, uint64_t n, int bf16, double* Y) {
  const Net net{params, bf16};
  Fwd f;
  for (uint64_t s = 0; s < n; ++s) {
    forward(net, X + s * kIn, norm, f);
    for (int o = 0; o < kOut; ++o) Y[s * kOut + o] = f.y[o] * norm[2 * kIn + o];
  }
}

// Loss = mean over n samples and 3 outputs of (y - t / out_scale)^2 and its
// gradient with respect to every parameter (grad: kParams doubles).
void ora_mlp_grad(const double* params, const double* norm, const float* X, const float* T, uint64_t n, int bf16,
                  double* grad, double* loss) {
  const Net net{params, bf16};
  for (uint32_t k = 0; k < kParams; ++k) grad[k] = 0.0;
  double L = 0.0;
  const double scale = 2.0 / (3.0 * (double)n);
  Fwd f;
  double d4[kOut], d3[kHid], d2[kHid], d1[kHid];
  for (uint64_t s = 0; s < n; ++s) {
    forward(net, X + s * kIn, norm, f);
    for (int o = 0; o < kOut; ++o) {
      double t;
      if (bf16) t = (double)(T[s * kOut + o] / (float)norm[2 * kIn + o]);
      else t = (double)T[s * kOut + o] / norm[2 * kIn + o];
      const double d = f.y[o] - t;
      L += d * d;
      d4[o] = rnd(scale * d, bf16);
    }
    for (int i = 0; i < kHid; ++i) {
      double a = 0.0;
      for (int o = 0; o < kOut; ++o) a += d4[o] * net.w(kW4, i, o, kOut);
      d3[i] = rnd(a * dsilu(f.z3[i]), bf16);
    }
    for (int i = 0; i < kHid; ++i) {
      double a = 0.0;
      for (int o = 0; o < kHid; ++o) a += d3[o] * net.w(kW3, i, o, kHid);
      d2[i] = rnd(a * dsilu(f.z2[i]), bf16);
    }
    for (int i = 0; i < kHid; ++i) {
      double a = 0.0;
      for (int o = 0; o < kHid; ++o) a += d2[o] * net.w(kW2, i, o, kHid);
      d1[i] = rnd(a * dsilu(f.z1[i]), bf16);
    }
    for (int i = 0; i < kIn; ++i)
      for (int o = 0; o < kHid; ++o) grad[kW1 + i * kHid + o] += f.h0[i] * d1[o];
    for (int i = 0; i < kHid; ++i)
      for (int o = 0; o < kHid; ++o) {
        grad[kW2 + i * kHid + o] += f.h1[i] * d2[o];
        grad[kW3 + i * kHid + o] += f.h2[i] * d3[o];
      }
    for (int i = 0; i < kHid; ++i)
      for (int o = 0; o < kOut; ++o) grad[kW4 + i * kOut + o] += f.h3[i] * d4[o];
    for (int o = 0; o < kHid; ++o) {
      grad[kB1 + o] += d1[o];
      grad[kB2 + o] += d2[o];
      grad[kB3 + o] += d3[o];
    }
    for (int o = 0; o < kOut; ++o) grad[kB4 + o] += d4[o];
  }
  *loss = L / (3.0 * (double)n);
}

// Bias-corrected Adam, step t >= 1, in place on w, m, v.
void ora_adam(double* w, double* m, double* v, const double* g, uint64_t n, double lr, double b1, double b2,
              double eps, uint32_t t) {
  const double c1 = 1.0 / (1.0 - std::pow(b1, (double)t)), c2 = 1.0 / (1.0 - std::pow(b2, (double)t));
  for (uint64_t k = 0; k < n; ++k) {
    m[k] = b1 * m[k] + (1.0 - b1) * g[k];
    v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k];
    w[k] -= lr * (m[k] * c1) / (std::sqrt(v[k] * c2) + eps);
  }
}

// Features of every fluid particle (boundary rows zero), brute force:
// x [n][3], v [n][3], rho [n], p [n], kind [n] (0 fluid), psi [n]; scalars
// rho0, mass, h, c (sound speed), B (Tait stiffness). edge[i] = 1 when some
// partner of i sits within tol of a shell edge, 2.5 q in {1, 2, 3, 4, 5},
// where a rounding difference can move it across.
void ora_features(uint64_t n, const double* x, const double* v, const double* rho, const double* p,
                  const uint8_t* kind, const double* psi, double rho0, double mass, double h, double c, double B,
                  double tol, double* out, uint8_t* edge) {
  for (uint64_t i = 0; i < n; ++i) {
    double* f = out + i * kIn;
    for (int k = 0; k < kIn; ++k) f[k] = 0.0;
    edge[i] = 0;
    if (kind[i] != 0) continue;
    f[0] = rho[i] / rho0 - 1.0;
    f[1] = p[i] / B;
    for (uint64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double rx = x[3 * j] - x[3 * i], ry = x[3 * j + 1] - x[3 * i + 1], rz = x[3 * j + 2] - x[3 * i + 2];
      const double q = std::sqrt(rx * rx + ry * ry + rz * rz) / h;
      const double sq = 2.5 * q;
      const double near = std::fabs(sq - std::round(sq));
      if (sq < 5.0 + tol && sq > 0.5 && near < tol) edge[i] = 1;
      if (!(q < 2.0)) continue;
      const int sh = (int)std::floor(sq);
      const double vj = kind[j] == 0 ? rho0 / rho[j] : psi[j] * rho0 / mass;
      const double d[6] = {rx / h, ry / h, rz / h, (v[3 * j] - v[3 * i]) / c, (v[3 * j + 1] - v[3 * i + 1]) / c,
                           (v[3 * j + 2] - v[3 * i + 2]) / c};
      for (int k = 0; k < 6; ++k) f[2 + 6 * sh + k] += vj * d[k];
    }
  }
}

}  // extern "C"
