// guess.hip -- the learned initial iterate's network on gfx950: bf16 MFMA
// convolution (inference, training forward, backward-data), weight
// gradients, Adam, and the cathy_gpu_guess_* C entries.
//
// MFMA operand layout (v_mfma_f32_16x16x16bf16_1k, D = A B + C, 16x16x16):
//   lane l holds A[l%16][4(l/16)+t], B[4(l/16)+t][l%16], D[4(l/16)+t][l%16],
//   t = 0..3.
// Convolution as a GEMM per 16-level block of one column: rows = levels,
// K = (tap k, input channel ci) with kk = k*CIN + ci, A(l, kk) =
// in[l + k - 1][ci] (zero outside the column), N = output channels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>

#include "cathy_gpu.h"
#include "guess.h"

static_assert(CATHY_GPU_EINVAL == -1 && CATHY_GPU_ENOMEM == -3, "status codes of guarded()");

namespace cg {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ int wave_id() { return (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6); }

// ------------------------------------------------------------- packs --
// Forward B operand of one layer: entry (j, b, lane) holds W[kk][co],
// kk = 16j + 4(lane/16) + t, co = 16b + lane%16 (W[k][ci][co] flattened
// as kk*cout + co).
__global__ void k_gpack_fwd(const float *__restrict__ w, int off_w, int cin, int cout, int nchunk, int nb,
                            bf16x4 *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nchunk * nb * 64) return;
  const int lane = e & 63, jn = e >> 6, j = jn / nb, b = jn - j * nb;
  const int co = 16 * b + (lane & 15);
  bf16x4 r;
  for (int t = 0; t < 4; ++t) {
    const int kk = 16 * j + 4 * (lane >> 4) + t;
    const float val = (kk < 3 * cin && co < cout) ? w[off_w + kk * cout + co] : 0.f;
    r[t] = (short)f2bf(val);
  }
  out[e] = r;
}

// Backward-data operand: the input gradient is the convolution of the
// output gradient (cin_b channels, cout real) with W'[k'][c][ci] =
// W[2-k'][ci][c]; output channels = cin (4 blocks of 16).
__global__ void k_gpack_bwd(const float *__restrict__ w, int off_w, int cin, int cout, int cin_b, int nchunk_b,
                            bf16x4 *__restrict__ out) {
  const int nb = cin / 16;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nchunk_b * nb * 64) return;
  const int lane = e & 63, jn = e >> 6, j = jn / nb, b = jn - j * nb;
  const int ci = 16 * b + (lane & 15);
  bf16x4 r;
  for (int t = 0; t < 4; ++t) {
    const int kk = 16 * j + 4 * (lane >> 4) + t;
    const int kp = kk / cin_b, c = kk - kp * cin_b;
    const float val = (kp < 3 && c < cout) ? w[off_w + ((2 - kp) * cin + ci) * cout + c] : 0.f;
    r[t] = (short)f2bf(val);
  }
  out[e] = r;
}

// -------------------------------------------------------- convolution --
// One wavefront: levels [16 rb, 16 rb + 16) of one column, all 16*NB
// output channels.  out = act(conv + bias), zero at levels >= L; with a
// mask tensor (same layout as out) entries whose mask is not > 0 are zero
// (the ReLU derivative in the backward pass).  out32 takes channel 0 in
// fp32 (the network output).
template <int CIN, int NB>
__global__ __launch_bounds__(256) void k_gconv(int ncol, int Lp, int L, const int *__restrict__ idx,
                                               const unsigned short *__restrict__ in,
                                               const bf16x4 *__restrict__ wp, const float *__restrict__ bias,
                                               int cout, int relu, const unsigned short *__restrict__ mask,
                                               unsigned short *__restrict__ out, float *__restrict__ out32) {
  constexpr int K = 3 * CIN, NCH = (K + 15) / 16, OS = 16 * NB;
  const int lane = threadIdx.x & 63;
  const int wv = wave_id();
  const int rbs = Lp >> 4;
  if (wv >= ncol * rbs) return;
  const int col = wv / rbs, rb = wv - col * rbs;
  const size_t src = idx ? (size_t)idx[col] : (size_t)col;
  const unsigned short *xin = in + src * Lp * CIN;
  const int grp = lane >> 4, r16 = lane & 15;
  const int l = rb * 16 + r16;
  f32x4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < NCH; ++j) {
    const int kk = 16 * j + 4 * grp;
    bf16x4 a = {0, 0, 0, 0};
    if (kk < K) {
      const int k = kk / CIN, ci = kk - k * CIN;
      const int lv = l + k - 1;
      if (lv >= 0 && lv < L) a = *(const bf16x4 *)(xin + (size_t)lv * CIN + ci);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
      acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wp[(j * NB + b) * 64 + lane], acc[b], 0, 0, 0);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int co = 16 * b + r16;
    const float bv = (bias && co < cout) ? bias[co] : 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int lo = rb * 16 + 4 * grp + t;
      const size_t o = ((size_t)col * Lp + lo) * OS + co;
      float v = acc[b][t] + bv;
      if (relu) v = fmaxf(v, 0.f);
      if (lo >= L) v = 0.f;
      if (mask && !(bf2f(mask[o]) > 0.f)) v = 0.f;
      if (out) out[o] = f2bf(v);
      if (out32 && co == 0) out32[(size_t)col * Lp + lo] = v;
    }
  }
}

// ---------------------------------------------------- weight gradient --
// dW[kk][co] = sum over positions of A(pos, kk) g(pos, co); row kk = 3*CIN
// is the bias (A = 1 at live levels).  MFMA with M = kk, K = 16 positions
// (one level block), N = co.  A wavefront owns one 16x16 tile over a
// contiguous range of level blocks (split); partials [split][rows][GS].
template <int CIN, int GS>
__global__ __launch_bounds__(256) void k_gdw(int ncol, int Lp, int L, const int *__restrict__ idx,
                                             const unsigned short *__restrict__ in,
                                             const unsigned short *__restrict__ g, int nsplit,
                                             float *__restrict__ part) {
  constexpr int K = 3 * CIN, NCW = (K + 1 + 15) / 16, NB = GS / 16, T = NCW * NB;
  const int lane = threadIdx.x & 63;
  const int wv = wave_id();
  if (wv >= T * nsplit) return;
  const int tile = wv % T, split = wv / T;
  const int j = tile / NB, b = tile - j * NB;
  const int rbs = Lp >> 4;
  const long R = (long)ncol * rbs;
  const long r0 = R * split / nsplit, r1 = R * (split + 1) / nsplit;
  const int grp = lane >> 4, r16 = lane & 15;
  const int kk = 16 * j + r16;
  const bool isw = kk < K, isb = kk == K;
  const int kdx = isw ? kk / CIN : 0, ci = isw ? kk - kdx * CIN : 0;
  const int co = 16 * b + r16;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (long r = r0; r < r1; ++r) {
    const int col = (int)(r / rbs), rb = (int)(r - (long)col * rbs);
    const size_t src = idx ? (size_t)idx[col] : (size_t)col;
    const unsigned short *xin = in + src * Lp * CIN;
    const unsigned short *gc = g + (size_t)col * Lp * GS;
    bf16x4 a, bb;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int lv = rb * 16 + 4 * grp + t;
      unsigned short av = 0;
      if (isw) {
        const int ls = lv + kdx - 1;
        if (ls >= 0 && ls < L) av = xin[(size_t)ls * CIN + ci];
      } else if (isb && lv < L) {
        av = 0x3F80;  // 1.0
      }
      a[t] = (short)av;
      bb[t] = (short)gc[(size_t)lv * GS + co];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
    part[((size_t)split * NCW * 16 + 16 * j + 4 * grp + t) * GS + co] = acc[t];
}

// Sums the split partials in split order into the gradient vector.
__global__ void k_gdw_reduce(const float *__restrict__ part, int nsplit, int rows, int gs, int K, int cout,
                             int off_w, int off_b, float *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (K + 1) * cout) return;
  const int kk = e / cout, c = e - kk * cout;
  float s = 0.f;
  for (int p = 0; p < nsplit; ++p) s += part[((size_t)p * rows + kk) * gs + c];
  if (kk < K) grad[off_w + kk * cout + c] = s;
  else grad[off_b + c] = s;
}

// ------------------------------------------------------- loss, Adam --
// Mean squared error over live levels: gradient 2 d / (ncol L) into the
// head's 16-channel gradient tensor (channel 0), per-block sums of d^2.
__global__ __launch_bounds__(256) void k_gloss(int ncol, int Lp, int L, const float *__restrict__ y,
                                               const int *__restrict__ idx, const float *__restrict__ t,
                                               float inv, unsigned short *__restrict__ gh,
                                               float *__restrict__ lossp) {
  __shared__ float red[256];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.f;
  if (e < ncol * Lp) {
    const int col = e / Lp, l = e - col * Lp;
    float gv = 0.f;
    if (l < L) {
      const size_t src = idx ? (size_t)idx[col] : (size_t)col;
      const float d = y[e] - t[src * Lp + l];
      d2 = d * d;
      gv = 2.f * d * inv;
    }
    unsigned short *o = gh + (size_t)e * 16;
    o[0] = f2bf(gv);
    for (int c = 1; c < 16; ++c) o[c] = 0;
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_gloss_sum(const float *__restrict__ lossp, int nb, float inv,
                                                   float *__restrict__ slot) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < nb; i += 256) s += lossp[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) slot[0] = red[0] * inv;
}

__global__ void k_gadam(int np, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                        const float *__restrict__ g, float lr_t, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// ------------------------------------------------------ flow-side ops --
// Features of (surface node s, level l): raw fp32 values and the
// standardised bf16 network input; levels >= L are zero.
__global__ void k_guess_features(int nnod, int L, int Lp, const int *__restrict__ colrow,
                                 const double *__restrict__ ptimep, const double *__restrict__ pprev,
                                 double ratio, const double *__restrict__ swtimep,
                                 const float *__restrict__ stat, const double *__restrict__ atmpot,
                                 const double *__restrict__ arenod, float ldt, GuessScale sc,
                                 float *__restrict__ raw, unsigned short *__restrict__ x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnod * Lp) return;
  const int s = e / Lp, l = e - s * Lp;
  float f[kGuessIn];
  for (int c = 0; c < kGuessIn; ++c) f[c] = 0.f;
  const bool live = l < L;
  if (live) {
    const int row = colrow[s * L + l];
    const float *sn = stat + ((size_t)s * L + l) * 5;
    f[0] = (float)ptimep[row];
    f[1] = (float)(swtimep[row] * (double)sn[4]);
    f[2] = sn[0];
    f[3] = sn[1];
    f[4] = sn[2];
    f[5] = (float)(atmpot[s] / arenod[s]);
    f[6] = ldt;
    f[7] = sn[3];
    f[8] = (float)(ratio * (ptimep[row] - pprev[row]));
  }
  for (int c = 0; c < kGuessIn; ++c) {
    if (raw) raw[(size_t)e * kGuessIn + c] = f[c];
    if (x) x[(size_t)e * kGuessIn + c] = live ? f2bf((f[c] - sc.mu[c]) / sc.sd[c]) : (unsigned short)0;
  }
}

// One thread per column: E_l = ratio (PTIMEP - PPREV), s = rms_l E_l +
// floor, PNEW = PTIMEP + beta E + s out clamp(y, -10, 10) on every node
// except the surface nodes that are Dirichlet at the start of the attempt
// (IFATM != 0) and, with band >= 0, the nodes where PTIMEP or the guess is
// above -band (at or near saturation, where the Picard linearisation of
// the storage term changes branch).
__global__ void k_guess_apply(int nnod, int L, int Lp, const int *__restrict__ colrow,
                              const int *__restrict__ ifatm, const double *__restrict__ ptimep,
                              const double *__restrict__ pprev, double ratio, const float *__restrict__ y,
                              double beta, double out, double floor_abs, double band,
                              double *__restrict__ pnew) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnod) return;
  const int *cr = colrow + (size_t)s * L;
  double e2 = 0.0;
  for (int l = 0; l < L; ++l) {
    const double d = ratio * (ptimep[cr[l]] - pprev[cr[l]]);
    e2 += d * d;
  }
  const double scol = sqrt(e2 / L) + floor_abs;
  for (int l = (ifatm[s] != 0) ? 1 : 0; l < L; ++l) {
    const int row = cr[l];
    const double pt = ptimep[row];
    const double yl = fmin(fmax((double)y[(size_t)s * Lp + l], -10.0), 10.0);
    const double gv = pt + beta * (ratio * (pt - pprev[row])) + scol * out * yl;
    if (band >= 0.0 && (pt > -band || gv > -band)) continue;
    pnew[row] = gv;
  }
}

__global__ void k_guess_target(int nnod, int L, int Lp, const int *__restrict__ colrow,
                               const double *__restrict__ ptimep, const double *__restrict__ pnew,
                               float *__restrict__ target) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnod * Lp) return;
  const int s = e / Lp, l = e - s * Lp;
  float v = 0.f;
  if (l < L) {
    const int row = colrow[s * L + l];
    v = (float)(pnew[row] - ptimep[row]);
  }
  target[e] = v;
}

void guess_features(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const double *ptimep,
                    const double *pprev, double ratio, const double *swtimep, const float *stat,
                    const double *atmpot, const double *arenod, double dt, const GuessScale &sc, float *raw,
                    unsigned short *x) {
  hipLaunchKernelGGL(k_guess_features, dim3(grid_for((size_t)nnod * Lp)), dim3(kBlock), 0, st, nnod, L, Lp,
                     colrow, ptimep, pprev, ratio, swtimep, stat, atmpot, arenod, (float)std::log10(dt), sc, raw,
                     x);
}

void guess_apply(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const int *ifatm,
                 const double *ptimep, const double *pprev, double ratio, const float *y, const GuessScale &sc,
                 double *pnew) {
  hipLaunchKernelGGL(k_guess_apply, dim3(grid_for((size_t)nnod)), dim3(kBlock), 0, st, nnod, L, Lp, colrow, ifatm,
                     ptimep, pprev, ratio, y, (double)sc.beta, (double)sc.out, (double)sc.floor, (double)sc.band,
                     pnew);
}

void guess_target(hipStream_t st, int nnod, int L, int Lp, const int *colrow, const double *ptimep,
                  const double *pnew, float *target) {
  hipLaunchKernelGGL(k_guess_target, dim3(grid_for((size_t)nnod * Lp)), dim3(kBlock), 0, st, nnod, L, Lp,
                     colrow, ptimep, pnew, target);
}

// ------------------------------------------------------------ GuessNet --
static int waves_blocks(long waves) { return (int)((waves * 64 + kBlock - 1) / kBlock); }

static int split_count(int ncol, int Lp) {
  const long r = (long)ncol * (Lp / 16);
  long s = r / 8;
  if (s < 1) s = 1;
  if (s > 256) s = 256;
  return (int)s;
}

GuessNet::GuessNet(int levels) : L(levels) {
  CG_REQUIRE(levels >= 2 && levels <= 4096, "guess net: levels out of range");
  Lp = (L + 15) / 16 * 16;
  CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const int cin[kGuessLayers] = {kGuessIn, 64, 64, 64, 64}, cout[kGuessLayers] = {64, 64, 64, 64, 1};
  int off = 0;
  size_t fp = 0;
  for (int l = 0; l < kGuessLayers; ++l) {
    GuessLayer &g = ly[l];
    g.cin = cin[l];
    g.cout = cout[l];
    g.nb = (g.cout + 15) / 16;
    g.K = 3 * g.cin;
    g.nchunk = (g.K + 15) / 16;
    g.nchunk_w = (g.K + 1 + 15) / 16;
    g.off_w = off;
    off += g.K * g.cout;
    g.off_b = off;
    off += g.cout;
    g.cin_b = g.nb * 16;
    g.nchunk_b = l == 0 ? 0 : (3 * g.cin_b + 15) / 16;
    g.fpack = fp;
    fp += (size_t)g.nchunk * g.nb * 64;
  }
  for (int l = 1; l < kGuessLayers; ++l) {
    ly[l].bpack = fp;
    fp += (size_t)ly[l].nchunk_b * (ly[l].cin / 16) * 64;
  }
  CG_REQUIRE(off == kGuessParams, "guess net: parameter layout");
  pack.alloc(fp);
  for (DBuf<float> *b : {&w, &m, &v, &grad}) b->alloc(kGuessParams);
  CG_HIP(hipMemsetAsync(w.p, 0, sizeof(float) * kGuessParams, st));
  CG_HIP(hipMemsetAsync(m.p, 0, sizeof(float) * kGuessParams, st));
  CG_HIP(hipMemsetAsync(v.p, 0, sizeof(float) * kGuessParams, st));
  for (int c = 0; c < kGuessIn; ++c) { scale.mu[c] = 0.f; scale.sd[c] = 1.f; }
  scale.out = 1.f;
  scale.floor = 1e-5f;
  scale.beta = 1.f;
  scale.band = 0.1f;
  repack();
  CG_HIP(hipStreamSynchronize(st));
}

GuessNet::~GuessNet() {
  if (st) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
}

void GuessNet::repack() {
  for (int l = 0; l < kGuessLayers; ++l) {
    const GuessLayer &g = ly[l];
    const int nf = g.nchunk * g.nb * 64;
    hipLaunchKernelGGL(k_gpack_fwd, dim3(grid_for(nf)), dim3(kBlock), 0, st, w.p, g.off_w, g.cin, g.cout,
                       g.nchunk, g.nb, pack.p + g.fpack);
    if (l > 0) {
      const int nbk = g.nchunk_b * (g.cin / 16) * 64;
      hipLaunchKernelGGL(k_gpack_bwd, dim3(grid_for(nbk)), dim3(kBlock), 0, st, w.p, g.off_w, g.cin, g.cout,
                         g.cin_b, g.nchunk_b, pack.p + g.bpack);
    }
  }
}

void GuessNet::set_params(const float *p, const GuessScale *s) {
  if (p) {
    w.upload(p, kGuessParams, st);
    CG_HIP(hipMemsetAsync(m.p, 0, sizeof(float) * kGuessParams, st));
    CG_HIP(hipMemsetAsync(v.p, 0, sizeof(float) * kGuessParams, st));
    adam_t = 0;
    repack();
  }
  if (s) {
    for (int c = 0; c < kGuessIn; ++c) CG_REQUIRE(s->sd[c] > 0.f, "guess net: feature scales must be > 0");
    CG_REQUIRE(s->floor >= 0.f, "guess net: floor must be >= 0");
    scale = *s;
  }
  CG_HIP(hipStreamSynchronize(st));
}

void GuessNet::get_params(float *p, GuessScale *s) {
  if (p) {
    w.download(p, kGuessParams, st);
    CG_HIP(hipStreamSynchronize(st));
  }
  if (s) *s = scale;
}

void GuessNet::ensure(int ncol) {
  CG_REQUIRE(ncol > 0, "guess net: no columns");
  if (ncol <= cap) return;
  const size_t pos = (size_t)ncol * Lp;
  for (auto &a : act) a.alloc(pos * kGuessWidth);
  for (auto &a : gb) a.alloc(pos * kGuessWidth);
  ghead.alloc(pos * 16);
  xs.alloc(pos * kGuessIn);
  y32.alloc(pos);
  ts.alloc(pos);
  lossp.alloc(grid_for(pos));
  nsplit_cap = split_count(ncol, Lp);
  size_t po = 0;
  for (int l = 0; l < kGuessLayers; ++l) {
    ly[l].part = po;
    po += (size_t)nsplit_cap * ly[l].nchunk_w * 16 * ly[l].nb * 16;
  }
  part.alloc(po);
  if (!hist.p) hist.alloc(1);
  cap = ncol;
}

void GuessNet::forward(const unsigned short *x, const int *idx, int ncol, float *y) {
  ensure(ncol);
  const int gw = waves_blocks((long)ncol * (Lp / 16));
  const float *W = w.p;
  hipLaunchKernelGGL((k_gconv<kGuessIn, 4>), dim3(gw), dim3(kBlock), 0, st, ncol, Lp, L, idx, x, pack.p + ly[0].fpack,
                     W + ly[0].off_b, 64, 1, (const unsigned short *)nullptr, act[0].p, (float *)nullptr);
  for (int l = 1; l < 4; ++l)
    hipLaunchKernelGGL((k_gconv<64, 4>), dim3(gw), dim3(kBlock), 0, st, ncol, Lp, L, (const int *)nullptr,
                       act[l - 1].p, pack.p + ly[l].fpack, W + ly[l].off_b, 64, 1, (const unsigned short *)nullptr,
                       act[l].p, (float *)nullptr);
  hipLaunchKernelGGL((k_gconv<64, 1>), dim3(gw), dim3(kBlock), 0, st, ncol, Lp, L, (const int *)nullptr, act[3].p,
                     pack.p + ly[4].fpack, W + ly[4].off_b, 1, 0, (const unsigned short *)nullptr,
                     (unsigned short *)nullptr, y);
}

void GuessNet::gradient(const unsigned short *x, const int *idx, const float *t, int ncol, float *loss_slot) {
  forward(x, idx, ncol, y32.p);
  const size_t pos = (size_t)ncol * Lp;
  const int gl = grid_for(pos);
  const float inv = 1.0f / ((float)ncol * (float)L);
  hipLaunchKernelGGL(k_gloss, dim3(gl), dim3(kBlock), 0, st, ncol, Lp, L, y32.p, idx, t, inv, ghead.p, lossp.p);
  hipLaunchKernelGGL(k_gloss_sum, dim3(1), dim3(kBlock), 0, st, lossp.p, gl, 1.0f / ((float)ncol * (float)L),
                     loss_slot);
  // the loss partials hold sums of d^2; k_gloss_sum scales by 1/(ncol L)
  const int ns = split_count(ncol, Lp);
  const int gw = waves_blocks((long)ncol * (Lp / 16));
  auto dw_blocks = [&](const GuessLayer &g) { return waves_blocks((long)g.nchunk_w * g.nb * ns); };
  // head
  hipLaunchKernelGGL((k_gdw<64, 16>), dim3(dw_blocks(ly[4])), dim3(kBlock), 0, st, ncol, Lp, L,
                     (const int *)nullptr, act[3].p, ghead.p, ns, part.p + ly[4].part);
  hipLaunchKernelGGL((k_gconv<16, 4>), dim3(gw), dim3(kBlock), 0, st, ncol, Lp, L, (const int *)nullptr, ghead.p,
                     pack.p + ly[4].bpack, (const float *)nullptr, 64, 0, act[3].p, gb[0].p, (float *)nullptr);
  // hidden layers 3, 2, 1: gradient in gb[cur], input activation act[l-1]
  int cur = 0;
  for (int l = 3; l >= 1; --l) {
    hipLaunchKernelGGL((k_gdw<64, 64>), dim3(dw_blocks(ly[l])), dim3(kBlock), 0, st, ncol, Lp, L,
                       (const int *)nullptr, act[l - 1].p, gb[cur].p, ns, part.p + ly[l].part);
    hipLaunchKernelGGL((k_gconv<64, 4>), dim3(gw), dim3(kBlock), 0, st, ncol, Lp, L, (const int *)nullptr,
                       gb[cur].p, pack.p + ly[l].bpack, (const float *)nullptr, 64, 0, act[l - 1].p,
                       gb[1 - cur].p, (float *)nullptr);
    cur = 1 - cur;
  }
  hipLaunchKernelGGL((k_gdw<kGuessIn, 64>), dim3(dw_blocks(ly[0])), dim3(kBlock), 0, st, ncol, Lp, L, idx, x, gb[cur].p,
                     ns, part.p + ly[0].part);
  for (int l = 0; l < kGuessLayers; ++l) {
    const GuessLayer &g = ly[l];
    hipLaunchKernelGGL(k_gdw_reduce, dim3(grid_for((size_t)(g.K + 1) * g.cout)), dim3(kBlock), 0, st,
                       part.p + g.part, ns, g.nchunk_w * 16, g.nb * 16, g.K, g.cout, g.off_w, g.off_b, grad.p);
  }
}

void GuessNet::adam(float lr) {
  ++adam_t;
  const double b1 = 0.9, b2 = 0.999;
  const double lr_t = lr * std::sqrt(1.0 - std::pow(b2, (double)adam_t)) / (1.0 - std::pow(b1, (double)adam_t));
  hipLaunchKernelGGL(k_gadam, dim3(grid_for(kGuessParams)), dim3(kBlock), 0, st, kGuessParams, w.p, m.p, v.p,
                     grad.p, (float)lr_t, (float)b1, (float)b2, 1e-8f);
  repack();
}

// ------------------------------------------------------------ handles --
static std::mutex g_gmu;
static std::map<int64_t, std::shared_ptr<GuessNet>> g_nets;
static int64_t g_gnext = 1;

std::shared_ptr<GuessNet> guess_lookup(int64_t h) {
  std::lock_guard<std::mutex> lk(g_gmu);
  auto it = g_nets.find(h);
  if (it == g_nets.end()) throw Error(CATHY_GPU_EHANDLE, "unknown guess-net handle");
  return it->second;
}

}  // namespace cg

using namespace cg;

static GuessScale scale_from(const float *s) {
  GuessScale g;
  for (int c = 0; c < kGuessIn; ++c) {
    g.mu[c] = s[c];
    g.sd[c] = s[kGuessIn + c];
  }
  g.out = s[2 * kGuessIn];
  g.floor = s[2 * kGuessIn + 1];
  g.beta = s[2 * kGuessIn + 2];
  g.band = s[2 * kGuessIn + 3];
  return g;
}

extern "C" {

int cathy_gpu_guess_create(int64_t *net, const int *levels, int *nparam) {
  return guarded([&] {
    CG_REQUIRE(net && levels, "null argument");
    auto g = std::make_shared<GuessNet>(*levels);
    std::lock_guard<std::mutex> lk(g_gmu);
    *net = g_gnext++;
    g_nets[*net] = std::move(g);
    if (nparam) *nparam = kGuessParams;
    return 0;
  });
}

int cathy_gpu_guess_destroy(int64_t *net) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_gmu);
    if (!net || !g_nets.count(*net)) throw Error(CATHY_GPU_EHANDLE, "unknown guess-net handle");
    g_nets.erase(*net);
    *net = 0;
    return 0;
  });
}

int cathy_gpu_guess_set(int64_t *net, const float *params, const float *scale) {
  return guarded([&] {
    CG_REQUIRE(net, "null argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    GuessScale s;
    if (scale) s = scale_from(scale);
    g->set_params(params, scale ? &s : nullptr);
    return 0;
  });
}

int cathy_gpu_guess_get(int64_t *net, float *params, float *scale) {
  return guarded([&] {
    CG_REQUIRE(net, "null argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    GuessScale s;
    g->get_params(params, &s);
    if (scale) {
      for (int c = 0; c < kGuessIn; ++c) {
        scale[c] = s.mu[c];
        scale[kGuessIn + c] = s.sd[c];
      }
      scale[2 * kGuessIn] = s.out;
      scale[2 * kGuessIn + 1] = s.floor;
      scale[2 * kGuessIn + 2] = s.beta;
      scale[2 * kGuessIn + 3] = s.band;
    }
    return 0;
  });
}

int cathy_gpu_guess_forward(int64_t *net, const int *ncol, const uint16_t *x, float *y) {
  return guarded([&] {
    CG_REQUIRE(net && ncol && x && y && *ncol > 0, "bad argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    g->ensure(*ncol);
    const size_t pos = (size_t)*ncol * g->Lp;
    g->xs.upload(x, pos * kGuessIn, g->st);
    g->forward(g->xs.p, nullptr, *ncol, g->y32.p);
    g->y32.download(y, pos, g->st);
    CG_HIP(hipStreamSynchronize(g->st));
    return 0;
  });
}

int cathy_gpu_guess_gradient(int64_t *net, const int *ncol, const uint16_t *x, const float *t, float *loss,
                             float *grad) {
  return guarded([&] {
    CG_REQUIRE(net && ncol && x && t && *ncol > 0, "bad argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    g->ensure(*ncol);
    const size_t pos = (size_t)*ncol * g->Lp;
    g->xs.upload(x, pos * kGuessIn, g->st);
    g->ts.upload(t, pos, g->st);
    g->gradient(g->xs.p, nullptr, g->ts.p, *ncol, g->hist.p);
    if (loss) g->hist.download(loss, 1, g->st);
    if (grad) g->grad.download(grad, kGuessParams, g->st);
    CG_HIP(hipStreamSynchronize(g->st));
    return 0;
  });
}

int cathy_gpu_guess_load(int64_t *net, const int64_t *nsamp, const uint16_t *x, const float *t) {
  return guarded([&] {
    CG_REQUIRE(net && nsamp && x && t && *nsamp > 0, "bad argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    const size_t pos = (size_t)*nsamp * g->Lp;
    g->X.alloc(pos * kGuessIn);
    g->T.alloc(pos);
    g->X.upload(x, pos * kGuessIn, g->st);
    g->T.upload(t, pos, g->st);
    g->nsamp = *nsamp;
    CG_HIP(hipStreamSynchronize(g->st));
    return 0;
  });
}

int cathy_gpu_guess_train(int64_t *net, const int *nsteps, const int *batch, const int *idx, const float *lr,
                          float *loss) {
  return guarded([&] {
    CG_REQUIRE(net && nsteps && batch && idx && lr && *nsteps > 0 && *batch > 0, "bad argument");
    auto g = guess_lookup(*net);
    std::lock_guard<std::mutex> lk(g->mu);
    CG_REQUIRE(g->nsamp > 0, "no training set loaded");
    const size_t nidx = (size_t)*nsteps * *batch;
    for (size_t i = 0; i < nidx; ++i)
      CG_REQUIRE(idx[i] >= 0 && idx[i] < g->nsamp, "sample index out of range");
    g->ensure(*batch);
    g->bidx.alloc(nidx);
    g->bidx.upload(idx, nidx, g->st);
    DBuf<float> h(*nsteps);
    for (int s = 0; s < *nsteps; ++s) {
      g->gradient(g->X.p, g->bidx.p + (size_t)s * *batch, g->T.p, *batch, h.p + s);
      g->adam(*lr);
    }
    if (loss) h.download(loss, *nsteps, g->st);
    CG_HIP(hipStreamSynchronize(g->st));
    CG_HIP(hipGetLastError());
    return 0;
  });
}

}  // extern "C"
