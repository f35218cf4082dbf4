// bd_cv3d.hip -- force contraction (Contracting_Vector3D.cu:11-304) on gfx950.
//
// For every particle m, the unit vectors to its neighbours within `range`
// are summed per distance bin into hit vectors h_m[bin] (x, y, z); the
// normal equations A^T A (per axis, bins x bins) and A^T b (b = the
// particle's force) are accumulated over particles, and calculate_vector
// solves them with a truncated pseudo-inverse and averages the three axes.
//
// Kernels (FP64 VALU; the matrix cores are reserved for the surrogate):
//   k_cv3d_hits  one thread per particle, candidates in ascending index
//                order through LDS tiles; the particle's own hit row is
//                accumulated in that order (Contracting_Vector3D.cu:85-131).
//   k_cv3d_syrk  the normal equations as a SYRK of the augmented hit matrix
//                [H | f] per axis: 64x64 tiles, 4x4 entries per thread, over
//                blocks of 2048 particles taken in ascending order (the
//                first block continuing the previous samples' sums);
//                k_cv3d_reduce adds the blocks in ascending order.  With one
//                block (n <= 2048) every entry is the serial loop's sequence
//                of IEEE operations (Contracting_Vector3D.cu:135-156); with
//                more, the same products summed block by block with fused
//                multiply-adds (deterministic, within 1e-12 of the entry's
//                scale sqrt(C_ii C_jj)).
//   k_cv3d_hits_wave  the same rows with one wavefront per particle (bins <=
//                2474), up to four per workgroup sharing LDS tiles of candidates: 64
//                candidates tested per round, the round's hits added to an LDS
//                row in ascending j.
//   (A^T b is the force column of the augmented tiles, :152-156.)
// The solve (bins^3) runs on the host: cyclic Jacobi eigen-decomposition of
// the symmetric matrix, singular values |lambda| truncated at 1e-6 of the
// largest (:194-282).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/bd_abi.h"
#include "bd_internal.h"

#pragma clang fp contract(off)

struct bd_cv3d {
  int bins = 0;
  double range = 0.0;
  int samples = 0;
  bool solved = false;
  double* ata = nullptr;  // 3 x bins x bins
  double* atb = nullptr;  // 3 x bins
  double* hit = nullptr;  // n x 3 x bins scratch
  size_t hit_cap = 0;
  double* in = nullptr;   // 6 x n staging of host inputs
  size_t in_cap = 0;
  double* part = nullptr;  // per particle block partial sums of the normal equations
  size_t part_cap = 0;
  hipStream_t stream = nullptr;
  std::vector<double> result;
};

namespace {

constexpr int kHitThreads = 256;

#define CV_TRY(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bdi::set_error(BD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

__global__ void __launch_bounds__(kHitThreads)
    k_cv3d_hits(const double* __restrict__ px, const double* __restrict__ py,
                const double* __restrict__ pz, int n, double side, int periodic,
                double range_sq, double bin_size_inv, int bins, double* __restrict__ hit) {
  __shared__ double tx[kHitThreads], ty[kHitThreads], tz[kHitThreads];
  const int m = (int)(blockIdx.x * kHitThreads + threadIdx.x);
  const bool active = m < n;
  double xm = 0, ym = 0, zm = 0;
  if (active) {
    xm = px[m];
    ym = py[m];
    zm = pz[m];
  }
  const double side_inv = periodic ? 1.0 / side : 0.0;
  double* row = hit + (size_t)(active ? m : 0) * 3 * bins;
  for (int jt = 0; jt < n; jt += kHitThreads) {
    const int jl = jt + (int)threadIdx.x;
    if (jl < n) {
      tx[threadIdx.x] = px[jl];
      ty[threadIdx.x] = py[jl];
      tz[threadIdx.x] = pz[jl];
    }
    __syncthreads();
    const int tn = min(kHitThreads, n - jt);
    if (active) {
      for (int k = 0; k < tn; k++) {
        if (jt + k == m) continue;
        double dx = xm - tx[k];
        double dy = ym - ty[k];
        double dz = zm - tz[k];
        if (periodic) {
          dx -= side * floor(side_inv * dx + 0.5);
          dy -= side * floor(side_inv * dy + 0.5);
          dz -= side * floor(side_inv * dz + 0.5);
        }
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < range_sq) {
          const double dist = sqrt(d2);
          int nb = (int)(bin_size_inv * dist);
          if (nb >= bins) nb = bins - 1;
          row[nb] += dx / dist;
          row[bins + nb] += dy / dist;
          row[2 * bins + nb] += dz / dist;
        }
      }
    }
    __syncthreads();
  }
}

// The hit rows with one wavefront per particle m (bins <= 2474): the
// lanes test 64 candidates j per round, in ascending order, and the round's
// hits are added to the particle's row in LDS one after the other in lane
// order, so every bin receives the serial loop's additions in the serial
// order (ascending j).  Each separation and unit vector is computed with
// the operations of k_cv3d_hits.
// LDS of the wavefront kernel: a 256-candidate tile and one row per wavefront,
// within the 64 KB a workgroup gets without an opt-in
constexpr size_t kHitLds = 64 * 1024;
__host__ inline int hit_waves(int bins) {
  const size_t row = 3 * (size_t)bins * sizeof(double), tile = 768 * sizeof(double);
  const int w = row > 0 && kHitLds > tile ? (int)((kHitLds - tile) / row) : 0;
  return w < 4 ? w : 4;
}

__device__ inline double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

__global__ void __launch_bounds__(4 * 64)
    k_cv3d_hits_wave(const double* __restrict__ px, const double* __restrict__ py,
                     const double* __restrict__ pz, int n, double side, int periodic,
                     double range_sq, double bin_size_inv, int bins, double* __restrict__ hit) {
  // one particle per wavefront, the workgroup's wavefronts sharing each tile
  // of 256 candidates staged in LDS; then their rows (3 x bins each)
  extern __shared__ double sh[];
  double* tx = sh;
  double* ty = sh + 256;
  double* tz = sh + 512;
  const int waves = (int)blockDim.x / 64;
  const int w = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
  const int m = (int)blockIdx.x * waves + w;
  const bool active = m < n;
  double* row = sh + 768 + (size_t)w * 3 * bins;
  for (int k = lane; k < 3 * bins; k += 64) row[k] = 0.0;
  double xm = 0.0, ym = 0.0, zm = 0.0;
  if (active) {
    xm = px[m];
    ym = py[m];
    zm = pz[m];
  }
  const double side_inv = periodic ? 1.0 / side : 0.0;
  for (int jt = 0; jt < n; jt += 256) {
    __syncthreads();
    for (int k = (int)threadIdx.x; k < 256; k += (int)blockDim.x) {
      const int jl = jt + k;
      if (jl < n) {
        tx[k] = px[jl];
        ty[k] = py[jl];
        tz[k] = pz[jl];
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int r = 0; r < 4; r++) {
      const int t = r * 64 + lane;
      const int j = jt + t;
      bool h = false;
      int nb = 0;
      double ux = 0.0, uy = 0.0, uz = 0.0;
      if (j < n && j != m) {
        double dx = xm - tx[t];
        double dy = ym - ty[t];
        double dz = zm - tz[t];
        if (periodic) {
          dx -= side * floor(side_inv * dx + 0.5);
          dy -= side * floor(side_inv * dy + 0.5);
          dz -= side * floor(side_inv * dz + 0.5);
        }
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < range_sq) {
          const double dist = sqrt(d2);
          nb = (int)(bin_size_inv * dist);
          if (nb >= bins) nb = bins - 1;
          ux = dx / dist;
          uy = dy / dist;
          uz = dz / dist;
          h = true;
        }
      }
      unsigned long long mask = __ballot(h);
      while (mask) {
        const int l = __ffsll(mask) - 1;
        mask &= mask - 1;
        const int b = __builtin_amdgcn_readlane(nb, l);
        const double vx = lane_value(ux, l), vy = lane_value(uy, l), vz = lane_value(uz, l);
        if (lane < 3) row[lane * bins + b] += lane == 0 ? vx : (lane == 1 ? vy : vz);
      }
    }
  }
  if (!active) return;
  double* out = hit + (size_t)m * 3 * bins;
  for (int k = lane; k < 3 * bins; k += 64) out[k] = row[k];
}

constexpr int kSyrkTile = 64;        // a workgroup's tile of the normal equations
constexpr int kSyrkChunk = 16;       // particles staged in LDS at a time
constexpr int kBlockParticles = 2048;  // particles per partial sum

__device__ inline double aug_value(const double* __restrict__ hit, const double* __restrict__ f,
                                   int m, int a, int c, int bins) {
  return c < bins ? hit[((size_t)m * 3 + a) * bins + c] : (c == bins ? f[m] : 0.0);
}

// Tile pair t (0 .. nt (nt + 1) / 2 - 1) -> (ti, tj), ti >= tj.
__device__ inline void tile_pair(int t, int& ti, int& tj) {
  ti = 0;
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  tj = t - ti * (ti + 1) / 2;
}

// One 64x64 tile (tile pair blockIdx.x) of axis blockIdx.y over particle
// block blockIdx.z: entries (i, j), i >= j, of the augmented normal
// equations (rows 0..bins-1: A^T A, row bins: A^T b), each a running sum of
// aug[m][i] * aug[m][j] over the block's particles in ascending order
// (product, then sum).  Block 0 starts from the previous samples' sums.  With
// one block (FMA false) the results are final (mirrored into the upper
// triangle) and equal to the serial loop's; with several (FMA true: each
// product and sum one fused operation, the summation order differing from
// the serial one anyway) they go to the block's partial array.
template <bool FMA>
__global__ void __launch_bounds__(256)
    k_cv3d_syrk(const double* __restrict__ hit, const double* __restrict__ fx,
                const double* __restrict__ fy, const double* __restrict__ fz, int n, int bins,
                int nblocks, double* __restrict__ ata, double* __restrict__ atb,
                double* __restrict__ part) {
  __shared__ double si[kSyrkChunk][kSyrkTile], sj[kSyrkChunk][kSyrkTile];
  int ti, tj;
  tile_pair((int)blockIdx.x, ti, tj);
  const int a = (int)blockIdx.y, pb = (int)blockIdx.z;
  const double* f = a == 0 ? fx : (a == 1 ? fy : fz);
  const int naug = bins + 1;
  const int ty = (int)threadIdx.x / 16, tx = (int)threadIdx.x % 16;
  const int i0 = ti * kSyrkTile + ty * 4, j0 = tj * kSyrkTile + tx * 4;
  double* A = ata + (size_t)a * bins * bins;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int i = i0 + r, j = j0 + c;
      acc[r][c] = (pb == 0 && i < naug && j < bins && j <= i)
                      ? (i < bins ? A[(size_t)i * bins + j] : atb[(size_t)a * bins + j])
                      : 0.0;
    }
  const int m_lo = pb * kBlockParticles, m_hi = min(n, m_lo + kBlockParticles);
  // the staging of a chunk: thread t loads columns 4 (t % 16) .. + 3 of row t / 16
  const int lr = (int)threadIdx.x / 16, lc = ((int)threadIdx.x % 16) * 4;
  for (int m0 = m_lo; m0 < m_hi; m0 += kSyrkChunk) {
    const int m = m0 + lr;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      si[lr][lc + q] = m < m_hi ? aug_value(hit, f, m, a, ti * kSyrkTile + lc + q, bins) : 0.0;
      sj[lr][lc + q] = m < m_hi ? aug_value(hit, f, m, a, tj * kSyrkTile + lc + q, bins) : 0.0;
    }
    __syncthreads();
    const int kend = min(kSyrkChunk, m_hi - m0);
    for (int k = 0; k < kend; k++) {
      double hi[4], hj[4];
#pragma unroll
      for (int r = 0; r < 4; r++) hi[r] = si[k][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; c++) hj[c] = sj[k][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (FMA) {
            acc[r][c] = fma(hi[r], hj[c], acc[r][c]);
          } else {
            const double aux = hi[r] * hj[c];
            acc[r][c] += aux;
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int i = i0 + r, j = j0 + c;
      if (!(i < naug && j < bins && j <= i)) continue;
      if (nblocks > 1) {
        part[(((size_t)pb * 3 + a) * naug + i) * naug + j] = acc[r][c];
      } else if (i < bins) {
        A[(size_t)i * bins + j] = acc[r][c];
        A[(size_t)j * bins + i] = acc[r][c];
      } else {
        atb[(size_t)a * bins + j] = acc[r][c];
      }
    }
}

// The partial sums of the particle blocks added in ascending block order
// (block 0 holds the previous samples' sums), one thread per entry (a, i, j),
// i >= j, of the augmented lower triangle.
__global__ void __launch_bounds__(256)
    k_cv3d_reduce(const double* __restrict__ part, int nblocks, int bins,
                  double* __restrict__ ata, double* __restrict__ atb) {
  const int naug = bins + 1;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int a = (int)blockIdx.y;
  const int i = (int)(e / naug), j = (int)(e % naug);
  if (i >= naug || j >= bins || j > i) return;
  const size_t stride = (size_t)3 * naug * naug;
  const size_t k = ((size_t)a * naug + i) * naug + j;
  double acc = part[k];
  for (int pb = 1; pb < nblocks; pb++) acc += part[(size_t)pb * stride + k];
  if (i < bins) {
    double* A = ata + (size_t)a * bins * bins;
    A[(size_t)i * bins + j] = acc;
    A[(size_t)j * bins + i] = acc;
  } else {
    atb[(size_t)a * bins + j] = acc;
  }
}

int accumulate(bd_cv3d* cv, const double* px, const double* py, const double* pz,
               const double* fx, const double* fy, const double* fz, int n, double side,
               int periodic, hipStream_t st) {
  const size_t need = (size_t)n * 3 * cv->bins;
  if (need > cv->hit_cap) {
    if (cv->hit) (void)hipFree(cv->hit);
    cv->hit = nullptr;
    cv->hit_cap = 0;
    CV_TRY(hipMalloc((void**)&cv->hit, need * sizeof(double)));
    cv->hit_cap = need;
  }
  // the wavefront kernel writes whole rows; the fallback accumulates into zeros
  if (hit_waves(cv->bins) == 0) CV_TRY(hipMemsetAsync(cv->hit, 0, need * sizeof(double), st));
  const double bin_size = cv->range / (double)cv->bins;
  const double bin_size_inv = 1.0 / bin_size;
  const int waves = hit_waves(cv->bins);
  if (waves > 0 && n > 0)
    k_cv3d_hits_wave<<<(n + waves - 1) / waves, waves * 64,
                       (768 + 3 * (size_t)waves * cv->bins) * sizeof(double), st>>>(
        px, py, pz, n, side, periodic, cv->range * cv->range, bin_size_inv, cv->bins, cv->hit);
  else
    k_cv3d_hits<<<(n + kHitThreads - 1) / kHitThreads, kHitThreads, 0, st>>>(
        px, py, pz, n, side, periodic, cv->range * cv->range, bin_size_inv, cv->bins, cv->hit);
  CV_TRY(hipGetLastError());
  const int naug = cv->bins + 1;
  const int nt = (naug + kSyrkTile - 1) / kSyrkTile;
  const int nblocks = (n + kBlockParticles - 1) / kBlockParticles;
  if (nblocks > 1) {
    const size_t pneed = (size_t)nblocks * 3 * naug * naug;
    if (pneed > cv->part_cap) {
      if (cv->part) (void)hipFree(cv->part);
      cv->part = nullptr;
      cv->part_cap = 0;
      CV_TRY(hipMalloc((void**)&cv->part, pneed * sizeof(double)));
      cv->part_cap = pneed;
    }
  }
  if (n > 0) {
    const dim3 grid(nt * (nt + 1) / 2, 3, nblocks);
    if (nblocks > 1)
      k_cv3d_syrk<true><<<grid, 256, 0, st>>>(cv->hit, fx, fy, fz, n, cv->bins, nblocks, cv->ata,
                                              cv->atb, cv->part);
    else
      k_cv3d_syrk<false><<<grid, 256, 0, st>>>(cv->hit, fx, fy, fz, n, cv->bins, nblocks,
                                               cv->ata, cv->atb, cv->part);
    CV_TRY(hipGetLastError());
    if (nblocks > 1) {
      k_cv3d_reduce<<<dim3((unsigned)(((size_t)naug * naug + 255) / 256), 3), 256, 0, st>>>(
          cv->part, nblocks, cv->bins, cv->ata, cv->atb);
      CV_TRY(hipGetLastError());
    }
  }
  CV_TRY(hipStreamSynchronize(st));
  cv->samples++;
  cv->solved = false;
  return BD_OK;
}

// Cyclic Jacobi eigen-decomposition of a symmetric matrix (Rutishauser's
// rotation, threshold sweeps): a (destroyed) -> eigenvalues w, vectors v
// (columns).
void jacobi_eigen(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double q = A(i, j) * A(i, j);
        tot += q;
        if (i != j) off += q;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A(p, q);
        if (apq == 0.0) continue;
        const double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        const double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double* vr = &v[(size_t)k * n];
          const double vkp = vr[p], vkq = vr[q];
          vr[p] = c * vkp - s * vkq;
          vr[q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; i++) w[i] = A(i, i);
}

}  // namespace

extern "C" {

int bd_cv3d_create(int32_t bins, double range, bd_cv3d** out) {
  if (!out || bins < 1 || bins > 4096 || !(range > 0.0))
    return bdi::set_error(BD_E_ARG, "bins must be in 1..4096 and range > 0");
  bd_cv3d* cv = new bd_cv3d();
  cv->bins = bins;
  cv->range = range;
  const size_t bb = (size_t)bins * bins;
  hipError_t e = hipMalloc((void**)&cv->ata, 3 * bb * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&cv->atb, 3 * (size_t)bins * sizeof(double));
  if (e == hipSuccess) e = hipMemset(cv->ata, 0, 3 * bb * sizeof(double));
  if (e == hipSuccess) e = hipMemset(cv->atb, 0, 3 * (size_t)bins * sizeof(double));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&cv->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (cv->ata) (void)hipFree(cv->ata);
    if (cv->atb) (void)hipFree(cv->atb);
    delete cv;
    return bdi::set_error(BD_E_HIP, std::string("bd_cv3d_create: ") + hipGetErrorString(e));
  }
  *out = cv;
  return BD_OK;
}

int bd_cv3d_destroy(bd_cv3d* cv) {
  if (!cv) return BD_OK;
  if (cv->stream) (void)hipStreamSynchronize(cv->stream);
  double* p[] = {cv->ata, cv->atb, cv->hit, cv->in, cv->part};
  for (double* q : p)
    if (q) (void)hipFree(q);
  if (cv->stream) (void)hipStreamDestroy(cv->stream);
  delete cv;
  return BD_OK;
}

int bd_cv3d_add(bd_cv3d* cv, bd_sys* s, int32_t replica, int32_t periodic) {
  if (!cv) return bdi::set_error(BD_E_ARG, "null contraction handle");
  bdi::DevState d;
  int rc = bdi::device_state(s, replica, &d);
  if (rc != BD_OK) return rc;
  CV_TRY(hipStreamSynchronize(d.stream));  // positions and forces of the last step
  return accumulate(cv, d.px, d.py, d.pz, d.fx, d.fy, d.fz, d.n, d.side, periodic ? 1 : 0,
                    cv->stream);
}

int bd_cv3d_add_host(bd_cv3d* cv, const double* xyz, const double* f, int32_t n, double side) {
  if (!cv || !xyz || !f || n < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  const size_t need = (size_t)6 * n;
  if (need > cv->in_cap) {
    if (cv->in) (void)hipFree(cv->in);
    cv->in = nullptr;
    cv->in_cap = 0;
    CV_TRY(hipMalloc((void**)&cv->in, need * sizeof(double)));
    cv->in_cap = need;
  }
  std::vector<double> h(need);
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      h[(size_t)a * n + i] = xyz[3 * i + a];
      h[(size_t)(3 + a) * n + i] = f[3 * i + a];
    }
  CV_TRY(hipMemcpyAsync(cv->in, h.data(), need * sizeof(double), hipMemcpyHostToDevice,
                        cv->stream));
  const double* b = cv->in;
  return accumulate(cv, b, b + n, b + 2 * (size_t)n, b + 3 * (size_t)n, b + 4 * (size_t)n,
                    b + 5 * (size_t)n, n, side, side > 0.0 ? 1 : 0, cv->stream);
}

int bd_cv3d_get_normal(bd_cv3d* cv, double* ata, double* atb) {
  if (!cv) return bdi::set_error(BD_E_ARG, "null contraction handle");
  const size_t bb = (size_t)cv->bins * cv->bins;
  if (ata) CV_TRY(hipMemcpy(ata, cv->ata, 3 * bb * sizeof(double), hipMemcpyDeviceToHost));
  if (atb) CV_TRY(hipMemcpy(atb, cv->atb, 3 * (size_t)cv->bins * sizeof(double),
                            hipMemcpyDeviceToHost));
  return BD_OK;
}

int bd_cv3d_samples(const bd_cv3d* cv, int32_t* out) {
  if (!cv || !out) return bdi::set_error(BD_E_ARG, "null argument");
  *out = cv->samples;
  return BD_OK;
}

// calculate_vector (Contracting_Vector3D.cu:166-297).  Unlike the reference,
// a second call recomputes rather than adds to the previous result.
int bd_cv3d_solve(bd_cv3d* cv) {
  if (!cv) return bdi::set_error(BD_E_ARG, "null contraction handle");
  if (cv->samples <= 0)
    return bdi::set_error(BD_E_STATE, "Number of samples taken must be greater than zero.");
  const int n = cv->bins;
  const size_t bb = (size_t)n * n;
  std::vector<double> ata(3 * bb), atb(3 * (size_t)n);
  int rc = bd_cv3d_get_normal(cv, ata.data(), atb.data());
  if (rc != BD_OK) return rc;
  cv->result.assign(n, 0.0);
  std::vector<double> a, w, v;
  for (int ax = 0; ax < 3; ax++) {
    a.assign(ata.begin() + ax * bb, ata.begin() + (ax + 1) * bb);
    jacobi_eigen(n, a, w, v);
    double wmax = 0.0;
    for (int i = 0; i < n; i++) wmax = std::max(wmax, std::fabs(w[i]));
    const double* b = &atb[(size_t)ax * n];
    for (int k = 0; k < n; k++) {
      if (!(wmax > 0.0) || !(std::fabs(w[k] / wmax) > 1.0e-6)) continue;
      double proj = 0.0;
      for (int i = 0; i < n; i++) proj += v[(size_t)i * n + k] * b[i];
      const double c = proj / w[k];
      for (int i = 0; i < n; i++) cv->result[i] += v[(size_t)i * n + k] * c;
    }
  }
  for (int i = 0; i < n; i++) cv->result[i] /= 3.0;
  cv->solved = true;
  return BD_OK;
}

int bd_cv3d_get(const bd_cv3d* cv, double* out) {
  if (!cv || !out) return bdi::set_error(BD_E_ARG, "null argument");
  if (!cv->solved)
    return bdi::set_error(BD_E_STATE, "Retrieved contracted vector must be firstly calculated.");
  for (int i = 0; i < cv->bins; i++) out[i] = cv->result[i];
  return BD_OK;
}

}  // extern "C"
