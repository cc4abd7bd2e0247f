// guess.h -- learned Picard initial iterate: a per-column 1-D convolutional
// network evaluated and trained with bf16 MFMA (fp32 accumulation).
//
// A soil column is the NSTR+1 nodes under one surface node (CATHY node id
// = level * NNOD + surface id).  The network maps, per level, 9 input
// channels (psi_n, theta_n, log10 Ks, alpha, n, rainfall rate, log10 dt,
// depth below the surface, E; each standardised; stored as 16 channels)
// to a correction of the extrapolated increment
//   E = dt_n (psi_n - psi_{n-1}) / dt_{n-1},
//   psi_guess = psi_n + beta E + s_col * out * y,
//   s_col = rms over the column's levels of E + floor,  |y| <= 10,
// so the network's output is relative to the column's own rate of change.
// Layers (kernel 3, zero padding at the column ends):
//   conv 16 -> 64, ReLU; three conv 64 -> 64, ReLU; conv 64 -> 1.
// Parameters (fp32, 40385): per layer W[k][ci][co] then b[co].
//
// Device layout: activations bf16 [column][Lp][channels], Lp = levels
// rounded up to 16 (levels >= L hold zeros).  One wavefront computes one
// 16-level block of one column with v_mfma_f32_16x16x16bf16_1k tiles; the
// weights are pre-packed in the MFMA B-operand order.
#pragma once

#include <array>
#include <memory>
#include <mutex>

#include "common.h"

namespace cg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));

constexpr int kGuessIn = 16;   // stored input channels
constexpr int kGuessLive = 9;  // channels carrying features
constexpr int kGuessWidth = 64;
constexpr int kGuessLayers = 5;
constexpr int kGuessParams = (3 * 16 * 64 + 64) + 3 * (3 * 64 * 64 + 64) + (3 * 64 * 1 + 1);
constexpr int kGuessScale = 2 * kGuessIn + 4;

struct GuessScale {
  float mu[kGuessIn];
  float sd[kGuessIn];
  float out;    // correction per unit of network output, in units of s_col
  float floor;  // length (m) added to the column scale: s_col += floor
  float beta;   // weight of the extrapolated increment E
  float band;   // nodes with psi_n or the guess above -band keep psi_n (< 0: off)
};

struct GuessLayer {
  int cin, cout, nb;    // input / output channels, 16-wide output blocks
  int K, nchunk;        // 3*cin, forward K chunks of 16
  int nchunk_w;         // dW K chunks (3*cin + bias row)
  int off_w, off_b;     // parameter offsets
  int cin_b, nchunk_b;  // backward-data conv: input channels (padded cout), K chunks
  size_t fpack, bpack;  // offsets into the packs (bf16x4 units)
  size_t part;          // offset of this layer's dW partials (floats)
};

struct GuessNet {
  int L = 0, Lp = 0;
  hipStream_t st = nullptr;
  std::array<GuessLayer, kGuessLayers> ly{};
  GuessScale scale{};
  DBuf<float> w, m, v, grad, lossp, hist;
  DBuf<bf16x4> pack;
  DBuf<unsigned short> act[4], gb[2], ghead, xs, X;
  DBuf<float> y32, ts, T, part;
  DBuf<int> bidx;
  int cap = 0, nsplit_cap = 0;
  int64_t nsamp = 0;
  long adam_t = 0;
  // serialises users of the work buffers (flows on several threads may
  // share one network)
  std::mutex mu;

  explicit GuessNet(int levels);
  ~GuessNet();
  void set_params(const float *p, const GuessScale *s);
  void get_params(float *p, GuessScale *s);
  void repack();
  void ensure(int ncol);
  // device pointers; x: [ncol or idx-range][Lp][8] bf16, y: [ncol][Lp] fp32
  void forward(const unsigned short *x, const int *idx, int ncol, float *y);
  // one gradient of the mean squared error over levels < L into grad
  void gradient(const unsigned short *x, const int *idx, const float *t, int ncol, float *loss_slot);
  void adam(float lr);
};

std::shared_ptr<GuessNet> guess_lookup(int64_t h);

// Flow-side kernels (features, apply, target), launched by the flow driver.
void guess_features(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const double *ptimep,
                    const double *pprev, double ratio, const double *swtimep, const float *stat,
                    const double *atmpot, const double *arenod, double dt, const GuessScale &sc, float *raw,
                    unsigned short *x);
void guess_apply(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const int *ifatm,
                 const double *ptimep, const double *pprev, double ratio, const float *y, const GuessScale &sc,
                 double *pnew);
void guess_target(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const double *ptimep,
                  const double *pnew, float *target);

}  // namespace cg
