// kernels.h -- gfx950 kernels of the CATHY subsurface Picard step.
//
// Device numbering: node ids are permuted at set-up (ordering NATURAL keeps
// CATHY's, COLUMN3 groups whole soil columns by a 3-colouring of the
// surface mesh so IC(0) sweeps run column-parallel).  Every vector below is
// in device numbering; the full CSR (rowptr/col, sorted columns) holds the
// symmetric system matrix (FP64) and, in the same pattern, the IC(0) factor
// (lower half = L_rj, upper half = transpose, diagonal = L_rr).
#pragma once

#include "common.h"

namespace cg {

// ------------------------------------------------------------------ vG --
// van Genuchten / Mualem closed forms exactly as fvgse.f, fvgkr.f,
// fvgdse.f (psi >= -1e-14 is saturated).
__device__ __forceinline__ double dev_fvgse(double psi, double psat, double vgn, double vgm) {
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / psat), vgn);
    return pow(fabs(1.0 / (beta + 1.0)), vgm);
  }
  return 1.0;
}
__device__ __forceinline__ double dev_fvgkr(double psi, double se, double vgmr, double vgm) {
  if (psi < -1.0e-14) {
    double omega = pow(fabs(se), vgmr);
    double v1 = 1.0 - pow(fabs(1.0 - omega), vgm);
    return sqrt(se) * v1 * v1;
  }
  return 1.0;
}
__device__ __forceinline__ double dev_fvgdse(double psi, double psat, double vgn, double vgn1,
                                             double vgpsn, double vgnr) {
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / psat), vgn);
    double b1 = beta + 1.0;
    double b1r = 1.0 / b1;
    return vgn1 * (pow(fabs(psi), vgn1) / vgpsn) * pow(fabs(b1), vgnr) * b1r * b1r;
  }
  return 0.0;
}

// d2 Se / d psi2 (fvgdds.f) and the kr derivative term of fvgdkr.f, the
// formulas the reference's Newton Jacobian uses (fvgdkr.f's is not the
// exact derivative of fvgkr.f; it is reproduced as written).
__device__ __forceinline__ double dev_fvgdds(double psi, double psat, double vgn, double vgn1, double vgm) {
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / psat), vgn);
    double b1 = beta + 1.0;
    double b1r = 1.0 / b1;
    return vgn1 * (beta / psi) * (1.0 / psi) * ((1.0 + vgn * (beta - 1.0)) / pow(fabs(b1), vgm)) * b1r * b1r;
  }
  return 0.0;
}
__device__ __forceinline__ double dev_fvgdkr(double psi, double psat, double vgn, double vgn1, double vgm) {
  if (psi < -1.0e-14) {
    double beta = pow(fabs(psi / psat), vgn);
    double b1 = beta + 1.0;
    double v1 = pow(fabs(b1), vgm) - pow(fabs(beta), vgm);
    double v2 = psat / psi;
    double v3 = vgn1 * beta * v2 * pow(fabs(1.0 / b1), 2.5 * vgm) / psat;
    double v4 = v2 * ((2.5 / b1) * beta - 2.0) - 0.5 * pow(fabs(b1), vgm - 1.0);
    return v3 * v1 * v4;
  }
  return 0.0;
}

// Node soil table, structure of arrays (device numbering).
struct SoilSoA {
  const double *psat, *vgn, *vgm, *vgn1, *vgnr, *vgmr, *vgpsn, *pnot, *rmcp, *snodi, *pnodi;
};

// ------------------------------------------------- block reductions --
// MI355X dispatches the work-groups of a launch round-robin over its 8
// XCDs (group g runs on XCD g % 8), each XCD with its own L2.  xcd_group
// maps the physical group id to a logical one so that XCD x owns the
// contiguous logical range [start(x), start(x) + count(x)): the rows an XCD
// processes, and the neighbours they gather, then stay in one L2.  The map
// is a bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_group() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int x = b % kXcds, k = b / kXcds;
  const int q = nb / kXcds, rem = nb % kXcds;
  return x * q + (x < rem ? x : rem) + k;
}

// Logical work-group of a launch.  A single-member kernel takes its own
// group (b = blockIdx.x or xcd_group(), nb = gridDim.x); a batched kernel
// runs every member of an ensemble in one grid and hands each member's
// share of groups to the same body as groups 0..nb-1 of that member, so a
// member's rows, partial sums and their order are those of its own launch.
struct Blk {
  int b, nb;
};
__device__ __forceinline__ Blk self_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }
__device__ __forceinline__ Blk self_xcd_blk() { return Blk{xcd_group(), (int)gridDim.x}; }

// Batched launches: member m owns the groups [start[m], start[m+1]) of the
// grid (after the XCD remap, so each XCD serves contiguous runs of every
// member); returns the member of logical group g (start[0] = 0, g <
// start[nmem]; members without groups are skipped).
__device__ __forceinline__ int batch_member(const int *__restrict__ start, int nmem, int g) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Deterministic block sum: returns the total in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < BS / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

// Sum of `cnt` partials in a fixed order by one block (result in all
// threads).  Used by consumers of a previous kernel's partial sums.
template <int BS>
__device__ __forceinline__ double reduce_partials(const double *part, int cnt, double *sh) {
  // four independent accumulators keep four loads in flight
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * BS < cnt; i += 4 * BS) {
    v0 += part[i];
    v1 += part[i + BS];
    v2 += part[i + 2 * BS];
    v3 += part[i + 3 * BS];
  }
  for (; i < cnt; i += BS) v0 += part[i];
  double v = block_sum<BS>((v0 + v1) + (v2 + v3), sh);
  if (threadIdx.x == 0) sh[BS / 64] = v;
  __syncthreads();
  double r = sh[BS / 64];
  __syncthreads();
  return r;
}

// Deterministic block sums of two values: the totals in thread 0 (sh:
// 2*BS/64 doubles).
template <int BS>
__device__ __forceinline__ double2 block_sum2(double a, double b, double *sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[2 * w] = a;
    sh[2 * w + 1] = b;
  }
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  if (threadIdx.x == 0)
    for (int k = 0; k < BS / 64; ++k) {
      t.x += sh[2 * k];
      t.y += sh[2 * k + 1];
    }
  __syncthreads();
  return t;
}

// Sums of two partial arrays of `cnt` entries in a fixed order, in every
// thread (sh: 2*BS/64 doubles).  Eight strided loads of each array are in
// flight per round; the tail round re-reads the last entry for strides
// past cnt and masks it (cnt >= 1).
template <int BS>
__device__ __forceinline__ double2 reduce_partials2(const double *pa, const double *pb, int cnt, double *sh) {
  double a[8], b[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) a[u] = b[u] = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * BS < cnt; i += 8 * BS) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] += pa[i + u * BS];
      b[u] += pb[i + u * BS];
    }
  }
  double ta[8], tb[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = min(i + u * BS, cnt - 1);
    ta[u] = pa[j];
    tb[u] = pb[j];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const bool in = i + u * BS < cnt;
    a[u] += in ? ta[u] : 0.0;
    b[u] += in ? tb[u] : 0.0;
  }
  double va = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  double vb = ((b[0] + b[1]) + (b[2] + b[3])) + ((b[4] + b[5]) + (b[6] + b[7]));
  va = wave_sum(va);
  vb = wave_sum(vb);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[2 * w] = va;
    sh[2 * w + 1] = vb;
  }
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  for (int k = 0; k < BS / 64; ++k) {
    t.x += sh[2 * k];
    t.y += sh[2 * k + 1];
  }
  __syncthreads();
  return t;
}

}  // namespace cg
