// build.hip - DetectorStatistics from the moments, on the device in FP64.
//
// Two workgroups of 8 waves turn (count, sum, sum x x^T) into what the
// scorers read, as build_statistics does (detector.cpp:47-152) - one for the
// covariance half, one for the independent autocorrelation half:
//   mean = sum / N;
//   S = M2/N - mean mean^T + loading I   (covariance, linalg.cpp:66-79)
//   R = M2/N + loading I                 (autocorrelation, linalg.cpp:52-64)
//   L_S L_S^T = S, whitener W = L_S^-1   (linalg.cpp:105-122, 150-162)
//   u_t = W (d_t - mean), |u_t|^2        (detector.cpp:111-128)
//   L_R L_R^T = R, w_t = R^-1 d_t / (d_t^T R^-1 d_t)  (detector.cpp:133-147)
// and narrows to float. A pivot that is not positive (or is below 4*256*eps
// of its diagonal, see diag_factor) sets the device status word instead.
//
// The factorisations are blocked over 16x16 tiles so that everything but
// the 16x16 diagonal tiles is a tile product on v_mfma_f64_16x16x4:
//   Cholesky, per panel K: one wave factors L_KK and inverts it in registers
//   (lane i holds row i; cross-lane values by shuffle); the panel below is
//   L_IK = A_IK L_KK^-T; the trailing tiles take A_IJ -= L_IK L_JK^T.
//   Inverse, per tile row I: X_IJ = -L_II^-1 sum_K L_IK X_KJ for all J < I.
// That is three workgroup barriers per panel and one per tile row, against
// two per column for an unblocked sweep. Tiles live in LDS (bands <= 128) or
// in global scratch, row-major with an XOR swizzle (element (r, c) at
// r*16 + (c ^ r)) so that reading a tile column across 16 lanes touches 16
// different LDS banks.
//
// MFMA tile products: lane l = 16g + m supplies A[m][k] and B[k][m] with the
// K index of K-step s taken as k = g + 4s for both operands, and
// v_mfma_f64_16x16x4 leaves C[g + 4r][m] in lane l, register r - so a result
// can be fed back as a B operand without a transpose (register s is exactly
// B[g + 4s][m]).
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / kWave;
constexpr int kTile = 16;
constexpr int kTileElems = kTile * kTile;
constexpr int kLdsTiles = 8;  // up to 128 bands in LDS
constexpr int kLdsBlocks = kLdsTiles * (kLdsTiles + 1) / 2;

__device__ __forceinline__ int sw(int r, int c) { return r * kTile + (c ^ r); }
__device__ __forceinline__ int tile_index(int I, int J) { return I * (I + 1) / 2 + J; }

// Tiles are reached through a pointer whose address space is part of the
// type: LDS-resident matrices use address_space(3) pointers so every access
// compiles to ds_read/ds_write, not to generic (flat) loads.
typedef __attribute__((address_space(3))) double lds_double;

template <typename P>
struct Mat {
  P* base;
  __device__ P* tile(int I, int J) const { return base + tile_index(I, J) * kTileElems; }
};

__device__ __forceinline__ void wave_sync() {
  // Order this wave's LDS/global writes before its later reads of other
  // lanes' data (a wavefront executes in lockstep; this waits for the
  // memory operations and keeps the compiler from reordering them).
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

// C (4 regs, layout C[g+4r][m]) <- tile t
template <typename P>
__device__ __forceinline__ d4 load_c(const P* t, int lane) {
  const int g = lane >> 4, m = lane & 15;
  return d4{t[sw(g, m)], t[sw(g + 4, m)], t[sw(g + 8, m)], t[sw(g + 12, m)]};
}
template <typename P>
__device__ __forceinline__ void store_c(P* t, int lane, const d4& c) {
  const int g = lane >> 4, m = lane & 15;
  t[sw(g, m)] = c[0];
  t[sw(g + 4, m)] = c[1];
  t[sw(g + 8, m)] = c[2];
  t[sw(g + 12, m)] = c[3];
}
// Operand value for K-step s: M[m][g + 4s] (A of M, or B of M^T).
template <typename P>
__device__ __forceinline__ double op(const P* t, int lane, int s) {
  return t[sw(lane & 15, (lane >> 4) + 4 * s)];
}
// Operand value for K-step s from a matrix used untransposed as B: M[g+4s][m].
template <typename P>
__device__ __forceinline__ double op_t(const P* t, int lane, int s) {
  return t[sw((lane >> 4) + 4 * s, lane & 15)];
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Lane `src`'s value (src uniform across the wave) through v_readlane into
// scalar registers: a few cycles, against a round trip through LDS for a
// shuffle, on the serial pivot chain below.
__device__ __forceinline__ double lane_value(double v, int src) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffu), src));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), src));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// 1/sqrt(p) to within an ulp or so: the hardware estimate and one Newton
// step (its relative error squares, from ~1e-15 to rounding level).
__device__ __forceinline__ double rsqrt_refined(double p) {
  const double r = __builtin_amdgcn_rsq(p);
  return r * __builtin_fma(-0.5 * p * r, r, 1.5);
}

// Factor the 16x16 diagonal tile and invert it, in one wave; both halves are
// serial chains, so each step is kept short.
//   Cholesky: lane i < 16 holds row i, and the columns go in panels of
//   four. Within a panel, pivot j's value and L[k][j] reach the other lanes
//   through v_readlane; sqrt and the division by the pivot come from one
//   refined reciprocal square root (d = p / sqrt(p) = p * r). A finished
//   panel reaches every lane through LDS at once (each row's four values
//   stored, every lane reading all rows as broadcasts), and the columns
//   right of it take its four updates in column order - the same operations
//   in the same order as a column-at-a-time sweep, so L is identical to it.
//   Inverse: lane c < 16 computes column c of L^-1 by forward substitution,
//   reading L's rows as LDS broadcasts, so no lane waits on another.
// Returns 0, or the first failing pivot's kind: kPivotNotPositive (not
// > 0, the reference's own rule, linalg.cpp:115-117) or kPivotBelowTolerance
// (positive, but at most 4*256*eps of its original diagonal: rounding noise
// of a rank-deficient matrix, which the blocked order would otherwise accept
// where the reference's sequential order lands on exactly zero). Writes L_KK
// into `lt` and L_KK^-1 into `xt` (both swizzled, upper triangle zero).
constexpr int kPivotNotPositive = 1;
constexpr int kPivotBelowTolerance = 2;
template <typename P>
__device__ __forceinline__ int diag_factor(P* lt, P* xt, const lds_double* diag0, int lane) {
  double row[kTile];
  double rinv[kTile];
  const int i = lane & 15;
#pragma unroll
  for (int c = 0; c < kTile; ++c) row[c] = (lane < 16 && c <= i) ? lt[sw(i, c)] : 0.0;
  bool ok = true;
  int why = 0;
  const double tol = 4.0 * 256 * 2.220446049250313e-16;
  constexpr int kPanel = 4;
  __shared__ double pan[kTile * kPanel];
  typedef __attribute__((address_space(3))) double lds_d;
  lds_d* pn = (lds_d*)pan;
  // The pivots' tolerances, read before the serial chain starts.
  double bound[kTile];
#pragma unroll
  for (int j = 0; j < kTile; ++j) bound[j] = tol * diag0[j];
#pragma unroll
  for (int p0 = 0; p0 < kTile; p0 += kPanel) {
#pragma unroll
    for (int jj = 0; jj < kPanel; ++jj) {
      const int j = p0 + jj;
      const double pivot = lane_value(row[j], j);
      // branch-free: the first failing pivot's kind is kept
      const int kind = !(pivot > 0.0) ? kPivotNotPositive : (!(pivot > bound[j]) ? kPivotBelowTolerance : 0);
      why = ok ? kind : why;
      ok = ok && kind == 0;
      const double r = ok ? rsqrt_refined(pivot) : 1.0;
      rinv[j] = r;
      if (i == j) row[j] = ok ? pivot * r : 1.0;
      if (i > j) row[j] = row[j] * r;
      const double lij = row[j];
#pragma unroll
      for (int k = j + 1; k < p0 + kPanel; ++k) {
        const double lkj = lane_value(lij, k);
        row[k] -= lij * lkj;  // entries above the diagonal are never read
      }
    }
    if (p0 + kPanel < kTile) {
      if (lane < 16) {
#pragma unroll
        for (int q = 0; q < kPanel; ++q) pn[kPanel * i + q] = row[p0 + q];
      }
      wave_sync();
#pragma unroll
      for (int k = p0 + kPanel; k < kTile; ++k) {
        double l[kPanel];
#pragma unroll
        for (int q = 0; q < kPanel; ++q) l[q] = pn[kPanel * k + q];
#pragma unroll
        for (int q = 0; q < kPanel; ++q) row[k] -= row[p0 + q] * l[q];
      }
      wave_sync();
    }
  }
  if (lane < 16) {
#pragma unroll
    for (int c = 0; c < kTile; ++c) lt[sw(i, c)] = c <= i ? row[c] : 0.0;
  }
  wave_sync();
  // Column c of X = L^-1: X[c][c] = 1/L[c][c]; X[r][c] = -(sum_{k<r} L[r][k] X[k][c]) / L[r][r],
  // with each X[k][c] subtracted from every later row's sum as soon as it
  // is known (b[r] holds minus the partial sum, so the chain from one
  // X[k][c] to the next is one FMA and one product; the sums take their
  // terms in the same order, so X is the same as summing row by row).
  const int c = lane & 15;
  double x[kTile], bsum[kTile];
#pragma unroll
  for (int r = 0; r < kTile; ++r) bsum[r] = r == c ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < kTile; ++k) {
    x[k] = k < c ? 0.0 : bsum[k] * rinv[k];
#pragma unroll
    for (int r = k + 1; r < kTile; ++r) bsum[r] = __builtin_fma(-static_cast<double>(lt[sw(r, k)]), x[k], bsum[r]);
  }
  if (lane < 16) {
#pragma unroll
    for (int r = 0; r < kTile; ++r) xt[sw(r, c)] = x[r];
  }
  return why;
}

// Blocked Cholesky of A (nt x nt tiles, lower) in place; the inverses of the
// diagonal tiles go to X's diagonal. Returns 0, or (uniformly across the
// workgroup) the kind of the first failing pivot.
// Lookahead: in panel K's trailing update wave 0 updates the next diagonal
// tile first and factors it at once, while the other waves finish the rest
// of the trailing update, so the serial diagonal factor of K + 1 is off the
// critical path.
template <typename P>
__device__ __attribute__((noinline)) int block_cholesky(Mat<P> A, Mat<P> X, int nt, const lds_double* diag0,
                                                        __attribute__((address_space(3))) int* flag) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  auto factor = [&](int K) {
    const int why = diag_factor(A.tile(K, K), X.tile(K, K), diag0 + K * kTile, lane);
    if (lane == 0 && why != 0) *flag = why;
  };
  auto update = [&](int K, int q) {
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= q) ++r;
    const int I = K + 1 + r;
    const int J = K + 1 + (q - r * (r + 1) / 2);
    P* aij = A.tile(I, J);
    const P* lik = A.tile(I, K);
    const P* ljk = A.tile(J, K);
    d4 c = load_c(aij, lane);
#pragma unroll
    for (int s = 0; s < 4; ++s) c = mfma(-op(lik, lane, s), op(ljk, lane, s), c);
    store_c(aij, lane, c);
  };
  if (wave == 0) factor(0);
  __syncthreads();
  if (*flag) return *flag;
  for (int K = 0; K < nt; ++K) {
    // Panel: L_IK = A_IK * Linv_KK^T.
    const P* xkk = X.tile(K, K);
    for (int I = K + 1 + wave; I < nt; I += kWaves) {
      P* aik = A.tile(I, K);
      d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) c = mfma(op(aik, lane, s), op(xkk, lane, s), c);
      wave_sync();
      store_c(aik, lane, c);
    }
    __syncthreads();
    if (K + 1 == nt) break;
    // Trailing: A_IJ -= L_IK L_JK^T for K < J <= I. Tile q = 0 is the next
    // diagonal tile (K + 1, K + 1).
    const int rest = nt - K - 1;
    const int count = rest * (rest + 1) / 2;
    if (wave == 0) {
      update(K, 0);
      wave_sync();
      factor(K + 1);
    } else {
      for (int q = wave; q < count; q += kWaves - 1) update(K, q);
    }
    __syncthreads();
    if (*flag) return *flag;
  }
  return 0;
}

// X = L^-1 given L (tiles) and the diagonal tiles of X already inverted.
template <typename P>
__device__ __attribute__((noinline)) void block_inverse(Mat<P> L, Mat<P> X, int nt) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int I = 1; I < nt; ++I) {
    for (int J = wave; J < I; J += kWaves) {
      d4 s = {0.0, 0.0, 0.0, 0.0};
      for (int K = J; K < I; ++K) {
        const P* lik = L.tile(I, K);
        const P* xkj = X.tile(K, J);
#pragma unroll
        for (int t = 0; t < 4; ++t) s = mfma(op(lik, lane, t), op_t(xkj, lane, t), s);
      }
      // X_IJ = -Linv_II * S, with S fed straight back as the B operand.
      const P* xii = X.tile(I, I);
      d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) c = mfma(-op(xii, lane, t), s[t], c);
      store_c(X.tile(I, J), lane, c);
    }
    __syncthreads();
  }
}

template <typename P>
__device__ __forceinline__ double tile_get(Mat<P> M, int i, int j) {
  return M.tile(i >> 4, j >> 4)[sw(i & 15, j & 15)];
}

// Fill tiles of an n-band matrix padded to nt*16 with identity on the padded
// diagonal: S (covariance) if centred, else R (autocorrelation). One wave per
// tile; each lane loads its four entries at once (C layout).
template <typename P>
__device__ __attribute__((noinline)) void fill(Mat<P> A, int n, int nt, const double* so, double count,
                                               const lds_double* mu, double loading, bool centred,
                                               lds_double* diag0) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, m = lane & 15;
  const int tiles = nt * (nt + 1) / 2;
  for (int q = wave; q < tiles; q += kWaves) {
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= q) ++I;
    const int J = q - I * (I + 1) / 2;
    const int j = 16 * J + m;
    double raw[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * I + g + 4 * r;
      raw[r] = (i < n && j <= i) ? so[static_cast<std::size_t>(i) * n + j] : 0.0;
    }
    d4 c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * I + g + 4 * r;
      double v = 0.0;
      if (j <= i) {
        if (i < n) {
          v = raw[r] / count;
          if (centred) v = v - mu[i] * mu[j];
          if (i == j) v += loading;
        } else if (i == j) {
          v = 1.0;
        }
      }
      c[r] = v;
      if (i == j) diag0[i] = v;
    }
    store_c(A.tile(I, J), lane, c);
  }
}

// out[t][i] = sum_j M[i][j] V(t, j)   (TRANS false: M lower, j <= i), or
// out[t][i] = sum_j M[j][i] V(t, j)   (TRANS true:  j >= i),
// for t < T, i < n, as 16x16 tile products on the matrix cores: A from the
// LDS/global tiles of M, B = V(t, j) loaded per lane (independent loads).
template <bool TRANS, typename P, typename VF>
__device__ __attribute__((noinline)) void tile_matvec(Mat<P> M, int nt, int n, int T, VF v_at, double* out) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, col = lane & 15;
  const int tt_count = (T + 15) / 16;
  for (int item = wave; item < nt * tt_count; item += kWaves) {
    const int I = item / tt_count;
    const int tt = item % tt_count;
    const int t = 16 * tt + col;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    const int j0 = TRANS ? I : 0;
    const int j1 = TRANS ? nt : I + 1;
    // Four column tiles at a time: their B values are loaded together, so
    // the global-memory latency is paid once per four tile products.
    for (int J0 = j0; J0 < j1; J0 += 4) {
      double bv[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int j = 16 * (J0 + q) + g + 4 * s;
          bv[q][s] = (J0 + q < j1 && t < T && j < n) ? v_at(t, j) : 0.0;
        }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (J0 + q < j1) {
          const P* tile = TRANS ? M.tile(J0 + q, I) : M.tile(I, J0 + q);
#pragma unroll
          for (int s = 0; s < 4; ++s)
            acc = mfma(TRANS ? op_t(tile, lane, s) : op(tile, lane, s), bv[q][s], acc);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * I + g + 4 * r;
      if (t < T && i < n) out[static_cast<std::size_t>(t) * n + i] = acc[r];
    }
  }
}

// Per-target dot products over n, one wave per target: result[t] = sum_i f(t, i).
template <typename F>
__device__ __forceinline__ double wave_sum(int n, F f, int t) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int i = lane; i < n; i += kWave) s += f(t, i);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, kWave);
  return __shfl(s, 0, kWave);
}

typedef __attribute__((address_space(3))) int lds_int;

// One of the two independent halves of build_statistics, each on its own
// workgroup (CU): path 0 factors the covariance (mean, whitener, whitened
// targets and their norms - RX and ACE); path 1 factors the autocorrelation
// (CEM filters). Each writes its own status slot.
template <typename P>
__device__ __forceinline__ void build_path(const BuildArgs& a, int path, Mat<P> A, Mat<P> X,
                                           lds_double* mu, lds_double* diag0, lds_int* flag,
                                           double* vec, double* solved) {
  const int n = a.bands;
  const int nt = (n + kTile - 1) / kTile;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int T = a.num_targets;
  unsigned* status = a.status + path;
  const double count = *a.count;
  if (!(count > 0.0)) {
    if (tid == 0) *status = path == 0 ? HSI_STATUS_COV_NOT_PD : HSI_STATUS_CORR_NOT_PD;
    return;
  }
  if (path == 1 && T == 0) {
    if (tid == 0) *status = 0u;
    return;
  }
  if (tid == 0) *flag = 0;
  for (int i = tid; i < n; i += kThreads) mu[i] = a.sum[i] / count;
  __syncthreads();
  fill(A, n, nt, a.sum_outer, count, mu, a.loading, path == 0, diag0);
  __syncthreads();
  const int why = block_cholesky(A, X, nt, diag0, flag);
  if (why != 0) {
    if (tid == 0)
      *status = (path == 0 ? HSI_STATUS_COV_NOT_PD : HSI_STATUS_CORR_NOT_PD) |
                (why == kPivotBelowTolerance ? HSI_STATUS_BELOW_TOLERANCE : 0u);
    return;
  }
  block_inverse(A, X, nt);
  const float* tg = a.targets;

  if (path == 0) {
    for (int i = tid; i < n; i += kThreads) a.mean[i] = static_cast<float>(mu[i]);
    for (int e = tid; e < n * n; e += kThreads) {
      const int i = e / n, k = e % n;
      a.whitener[e] = (k <= i) ? static_cast<float>(tile_get(X, i, k)) : 0.0f;
    }
    if (T > 0) {
      // u_t = W (d_t - mean) with the double whitener, and |u_t|^2.
      tile_matvec<false>(X, nt, n, T,
                         [&](int t, int j) { return static_cast<double>(tg[static_cast<std::size_t>(t) * n + j]) - mu[j]; },
                         vec);
      __syncthreads();
      for (int e = tid; e < T * n; e += kThreads) a.target_white[e] = static_cast<float>(vec[e]);
      for (int t = wave; t < T; t += kWaves) {
        const double s = wave_sum(n, [&](int tt, int i) { return vec[tt * n + i] * vec[tt * n + i]; }, t);
        if (lane == 0) a.target_norm2[t] = static_cast<float>(s);
      }
    }
  } else {
    // R^-1 d = L_R^-T (L_R^-1 d); w = R^-1 d / (d^T R^-1 d), zero if the gain is not positive.
    tile_matvec<false>(X, nt, n, T,
                       [&](int t, int j) { return static_cast<double>(tg[static_cast<std::size_t>(t) * n + j]); },
                       vec);
    __syncthreads();
    tile_matvec<true>(X, nt, n, T, [&](int t, int j) { return vec[static_cast<std::size_t>(t) * n + j]; },
                      solved);
    __syncthreads();
    for (int t = wave; t < T; t += kWaves) {
      const double gain = wave_sum(
          n, [&](int tt, int i) { return static_cast<double>(tg[static_cast<std::size_t>(tt) * n + i]) * solved[tt * n + i]; },
          t);
      for (int i = lane; i < n; i += kWave) {
        a.cem_weight[static_cast<std::size_t>(t) * n + i] =
            gain > 0.0 ? static_cast<float>(solved[t * n + i] / gain) : 0.0f;
      }
    }
  }
  if (tid == 0) *status = 0u;
}

template <bool IN_LDS>
__device__ __forceinline__ void build_body(const BuildArgs& a, int path, double* lds, double* mu, double* diag0,
                                           int* flag) {
  const int nt = (a.bands + kTile - 1) / kTile;
  const std::size_t mat = static_cast<std::size_t>(nt) * (nt + 1) / 2 * kTileElems;
  // work: [path 0: A, X][path 1: A, X][vec 0][vec 1][solved]
  double* vec = a.work + 4 * mat + static_cast<std::size_t>(path) * kMaxTargets * kMaxBands;
  double* solved = a.work + 4 * mat + 2 * static_cast<std::size_t>(kMaxTargets) * kMaxBands;
  lds_double* l_mu = (lds_double*)mu;
  lds_double* l_diag0 = (lds_double*)diag0;
  lds_int* l_flag = (lds_int*)flag;
  if constexpr (IN_LDS) {
    lds_double* base = (lds_double*)lds;
    build_path(a, path, Mat<lds_double>{base}, Mat<lds_double>{base + kLdsBlocks * kTileElems}, l_mu,
               l_diag0, l_flag, vec, solved);
  } else {
    (void)lds;
    double* base = a.work + 2 * mat * path;
    build_path(a, path, Mat<double>{base}, Mat<double>{base + mat}, l_mu, l_diag0, l_flag, vec, solved);
  }
}

// One scene: workgroup 0 takes the covariance path, workgroup 1 the
// autocorrelation path.
template <bool IN_LDS>
__global__ __launch_bounds__(kThreads) void build_kernel(BuildArgs a) {
  __shared__ double lds[IN_LDS ? 2 * kLdsBlocks * kTileElems : 1];
  __shared__ double mu[kMaxBands];
  __shared__ double diag0[kMaxBands];
  __shared__ int flag;
  build_body<IN_LDS>(a, blockIdx.x, lds, mu, diag0, &flag);
}

// Many scenes in one launch: workgroups 2i and 2i + 1 build scene i of the
// table (its two paths), with the loading given here.
template <bool IN_LDS>
__global__ __launch_bounds__(kThreads) void build_batch_kernel(const BuildArgs* __restrict__ table, double loading) {
  __shared__ double lds[IN_LDS ? 2 * kLdsBlocks * kTileElems : 1];
  __shared__ double mu[kMaxBands];
  __shared__ double diag0[kMaxBands];
  __shared__ int flag;
  BuildArgs a = table[blockIdx.x / 2];
  a.loading = loading;
  build_body<IN_LDS>(a, blockIdx.x % 2, lds, mu, diag0, &flag);
}

}  // namespace

std::size_t build_work_doubles(int bands, int num_targets) {
  (void)num_targets;
  const std::size_t nt = (static_cast<std::size_t>(bands) + kTile - 1) / kTile;
  return 4 * nt * (nt + 1) / 2 * kTileElems + 3 * static_cast<std::size_t>(kMaxTargets) * kMaxBands;
}

bool build_in_lds(int bands) { return bands <= kLdsTiles * kTile; }

hipError_t launch_build_batch(const BuildArgs* d_table, int count, bool in_lds, double loading, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  if (in_lds)
    hipLaunchKernelGGL(build_batch_kernel<true>, dim3(2 * count), dim3(kThreads), 0, stream, d_table, loading);
  else
    hipLaunchKernelGGL(build_batch_kernel<false>, dim3(2 * count), dim3(kThreads), 0, stream, d_table, loading);
  return hipGetLastError();
}

hipError_t launch_build(const BuildArgs& a, hipStream_t stream) {
  if (a.bands <= kLdsTiles * kTile) {
    hipLaunchKernelGGL(build_kernel<true>, dim3(2), dim3(kThreads), 0, stream, a);
  } else {
    hipLaunchKernelGGL(build_kernel<false>, dim3(2), dim3(kThreads), 0, stream, a);
  }
  return hipGetLastError();
}

}  // namespace hsi_amd
