// bd_engine.hip -- MI355X (gfx950) Brownian-dynamics engine behind the C ABI
// of include/bd_abi.h.
//
// One engine holds R independent systems ("replicas"); the plain single
// system is R = 1.  Every kernel covers all replicas in one launch through a
// block table (replica, first item, end item, grid) so that a workgroup never
// straddles two replicas.
//
// Device layout (all FP64 state structure-of-arrays, replicas concatenated):
//   pos/raw/ini/frc  x[N] y[N] z[N]          24 B per particle each
//   vir, pot         [N]
//   nocup            int per cell   atomic insertion counters (zero between steps)
//   cnt              int per cell   occupancy of the last binning
//   ins              int per slot   insertion-order particle ids (ntags slots per
//                                   cell; cells z-fastest, index z + n (x + n y),
//                                   for nocup and cnt too)
//   cpos             double4 per slot  x, y, z, replica-local index, in ROW
//                    layout: row (x, y) owns the n * ntags slots from
//                    (x + n y) n ntags; its particles are packed from the
//                    start in (z, ascending index) order, i.e. the order in
//                    which the reference's innermost k_del loop meets them
//   rs               int per (row, z) + 1  slot where cell z of the row starts
//                    (rs[row (n+1) + n] = end of the row), grid-relative
//   ring             u64 [250][lanes_total]  rannew64 lagged-Fibonacci rings
//   lcg              u64 [lanes_total]
//   draws            u64 per replica  draws taken by every lane of the replica
//   ni               int64  step index of the next update (shared)
//   flags            int per replica  bit0 cell overflow, bit1 non-finite position
//
// One BD step (main.cu:494-586) is three kernels:
//   k_update_bin  Ermak-McCammon update fused with the rannew64 Gaussian stream
//                 and the atomic cell insertion (update_ermak.cu:91-160 +
//                 cell_locate.cu:155-209)
//   k_cells       per-row packing of the cells in canonical order (ascending
//                 index within a cell, the order of cell_locate_hst) with
//                 their positions, counter reset and the step/draw counters
//   k_forces      per-particle WCA forces over the same-species then the
//                 cross-species stencil in the reference's traversal order
//                 (get_forces.cu:297-558), written (not accumulated)
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bd_abi.h"
#include "bd_device.h"
#include "bd_internal.h"

#pragma clang fp contract(off)

// Fixed histogram of bd_gder_noalloc: the four partial histograms of up to
// 512 bins (the bd_gder limit), 16 KiB per GPU, held until the process ends.
// The host-side name is only a stand-in: host code uses the per-device
// address from hipGetSymbolAddress (gder_scratch_ptr).
constexpr int kScratchWords = 4 * 512;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

namespace {

// One wavefront per workgroup for the per-particle, per-lane and per-row
// kernels: a 32768-particle system then spreads over 512 workgroups, i.e.
// every CU of the chip, instead of 128.
constexpr int kBlk = 64;
constexpr int kFlagOverflow = 1;
constexpr int kFlagNonFinite = 2;
constexpr int kFlagBrick = 4;

struct RepDev {
  int32_t n, nb;
  int32_t p_off;
  int32_t lanes, per_lane, lane_off;
  int32_t ncell[2];
  int32_t ntags[2];
  int32_t cell_off[2];
  int32_t tag_off[2];
  int32_t rs_off[2];        // row-start tables, ncell^2 * (ncell + 1) ints per grid
  int32_t nrange_same[2];   // bb, ss
  int32_t nrange_cross[2];  // big particle -> small grid (bs), small -> big (sb)
  int32_t potential, temp_protocol;
  int32_t fast;  // both grids >= 2*nrange+3 cells and nrange <= 2: scan_fast
  int32_t pad_;
  double side, side_inv;
  double cell_inv[2];
  double cell_side[2];
  double sigma[2];
  double sigma_bs;
  double rc2_same[2];
  double rc2_bs;
  double dt, kappa;
  double amp_t, inv_t, sigma_pol, phi_pol;  // BD_POT_WCA49_50(_AO)
  double T0, Tf, t0, tf, period;
  int64_t first_step;
  uint64_t seed;
};

// Error record of one replica: flag bits (kFlag*) and the step index ni of
// the first update (or initial binning) whose result raised one.
struct RepFlags {
  int bits;
  int pad_;
  long long first_step;
};
constexpr long long kNoFailStep = 0x7fffffffffffffffLL;

__device__ inline void raise_flag(RepFlags* f, int rep, int bit, int64_t ni) {
  atomicOr(&f[rep].bits, bit);
  atomicMin(&f[rep].first_step, (long long)ni);
}

// Id-ordered copy of a particle's position with its cell in its own
// species' grid, written by the binning for the neighbour-list walk (one
// 32-byte gather per listed candidate).
struct __align__(32) PosCell {
  double x, y, z;
  unsigned int cell;  // x | y << 10 | z << 20
  unsigned int pad_;
};

__device__ inline unsigned int pack_cell(int cx, int cy, int cz) {
  return (unsigned int)cx | ((unsigned int)cy << 10) | ((unsigned int)cz << 20);
}

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(BD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// Temperature used by the update at step ni (set_temp.cu:7-23,
// main.cu:155,520-524).  tf <= t0 (instant quench) gives Tf for t >= t0
// instead of the reference's 0/0; the sine uses an exact 2*pi turn.
__device__ double temperature_at(const RepDev& R, int64_t ni) {
  if (ni == R.first_step) return R.T0;
  double t = R.dt * (1.0 + (double)(ni - 1));
  if (t < R.t0) return R.T0;
  if (R.temp_protocol == BD_TEMP_SINE) {
    if (t > R.tf) return R.Tf;
    double a = 0.5 * (R.T0 + R.Tf);
    double b = 0.5 * (R.Tf - R.T0);
    double u = t / R.period - 0.25;
    u = u - floor(u);
    double s, c;
    bd::sincos_turn(u, s, c);
    return a + b * s;
  }
  if (t > R.tf || !(R.tf > R.t0)) return R.Tf;
  double m = (R.Tf - R.T0) / (R.tf - R.t0);
  return R.T0 + m * (t - R.t0);
}

// Cell coordinates of a position in grid g (cell_locate.cu:191-197);
// false when the position is not finite.
__device__ inline bool cell_coords(const RepDev& R, int g, double x, double y, double z,
                                   int& ii, int& jj, int& kk) {
  const double inv = R.cell_inv[g];
  const int n = R.ncell[g];
  double fx = inv * x, fy = inv * y, fz = inv * z;
  const double lo = -1.0, hi = (double)n + 1.0;
  if (!(fx > lo && fx < hi && fy > lo && fy < hi && fz > lo && fz < hi)) return false;
  ii = (int)fx;
  jj = (int)fy;
  kk = (int)fz;
  if (ii == n) ii--;
  if (jj == n) jj--;
  if (kk == n) kk--;
  return true;
}

__device__ inline void insert_cell(const RepDev& R, int rep, int i, double x, double y,
                                   double z, int* nocup, int* ins, RepFlags* flags,
                                   int64_t ni) {
  const int g = i < R.nb ? 0 : 1;
  const int loc = g == 0 ? i : i - R.nb;
  int ii, jj, kk, io, jo, ko;
  // a position the force walk cannot place in both grids (non-finite, or far
  // outside the box) is an error of the whole replica
  if (!cell_coords(R, g, x, y, z, ii, jj, kk) || !cell_coords(R, 1 - g, x, y, z, io, jo, ko)) {
    raise_flag(flags, rep, kFlagNonFinite, ni);
    return;
  }
  const int n = R.ncell[g];
  const int c = kk + n * (ii + n * jj);  // z fastest: a row's cells are adjacent
  const int slot = atomicAdd(&nocup[R.cell_off[g] + c], 1);
  if (slot < R.ntags[g]) ins[(size_t)R.tag_off[g] + (size_t)c * R.ntags[g] + slot] = loc;
  // the insertion that fills the cell to ntags (the main.cu:655-659 rule)
  if (slot >= R.ntags[g] - 1) raise_flag(flags, rep, kFlagOverflow, ni);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Seed every lane: rannew64 initialisation (rannew64.cu:69-97) with
// idum = lane_seed(replica seed, lane).
__global__ void k_rng_seed(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                           uint64_t* __restrict__ ring, uint64_t* __restrict__ lcg,
                           int total_lanes) {
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;  // padding entry of an XCD-grouped table
  const int l = t.y + (int)threadIdx.x;
  if (l >= t.z) return;
  const RepDev& R = reps[t.x];
  const int gl = R.lane_off + l;
  uint64_t u = bd::lane_seed(R.seed, (uint64_t)l);
  for (int i = 0; i < 1000; i++) u = u * bd::kSeedA + bd::kSeedC;
  for (int i = 0; i < bd::kRingLen; i++) {
    u = u * bd::kSeedA + bd::kSeedC;
    ring[(size_t)i * total_lanes + gl] = u;
  }
  lcg[gl] = u;
}

// Initial binning without an update (main.cu:536-543 outside the loop).
__global__ void __launch_bounds__(kBlk)
    k_bin(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
          const double* __restrict__ px, const double* __restrict__ py,
          const double* __restrict__ pz, int* __restrict__ nocup, int* __restrict__ ins,
          RepFlags* __restrict__ flags, int64_t ni) {
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;
  const int i = t.y + (int)threadIdx.x;
  if (i >= t.z) return;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + i;
  insert_cell(R, t.x, i, px[p], py[p], pz[p], nocup, ins, flags, ni);
}

// Four rannew64 draws of one lane (rannew64.cu:99-126): ring slot np and
// its lag partner, the LCG; np steps down the ring.  The four slots np..np-3
// and their partners (np+103..np+100, mod 250) never coincide, so all eight
// ring words are loaded before any is rewritten: the same values as the
// serial recurrence, with one memory latency instead of four.
static_assert(bd::kLagGap > 3 && bd::kRingLen - bd::kLagGap > 3,
              "a draw block's slots and lag partners must not overlap");
__device__ inline void draw4(uint64_t* rl, size_t stride, int& np, uint64_t& lcg,
                             uint64_t d[4]) {
  int sp[4], sq[4];
  uint64_t a[4], b[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    sp[j] = np;
    sq[j] = np >= bd::kLagGap ? np - bd::kLagGap : np + (bd::kRingLen - bd::kLagGap);
    np = np == 0 ? bd::kRingLen - 1 : np - 1;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    a[j] = rl[(size_t)sp[j] * stride];
    b[j] = rl[(size_t)sq[j] * stride];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint64_t x = a[j] - b[j];
    rl[(size_t)sp[j] * stride] = x;
    lcg = bd::kLcgA * lcg + bd::kLcgC;
    d[j] = x - lcg;
  }
}

// Neighbour-list bookkeeping of the moves (list walk only): raw positions
// at the last list build, and per replica two rebuild flags used on
// alternate steps (flag[2 r + parity]); a move that takes a particle more
// than half the skin from its build position raises the flag of the step
// whose forces come next.
struct ListMove {
  const double *rbx, *rby, *rbz;
  int* lflag;
  double half_skin2;
};

// Step parity of a replica's rebuild flags, from its draw counter (four
// draws per lane slot and step; the counter advances in the binning, before
// the forces): ahead = 0 for the forces of the current positions, 1 for the
// move that produces the next ones.
__device__ inline int list_parity(const RepDev& R, uint64_t draws, int ahead) {
  return (int)(((draws / (4ULL * (uint64_t)R.per_lane)) + (uint64_t)ahead) & 1ULL);
}

// Ermak-McCammon move of replica-local particle i (update_ermak.cu:91-160)
// from (x, y, z), raw position (ox, oy, oz), with force (fx, fy, fz) and
// normals (xi, yi, zi) at temperature T: raw and wrapped positions written,
// the wrapped one inserted
// in its cell (do_bin).  With lm, the rebuild flag lflag[2 rep + par] is
// raised for a move past half the skin.  Returns false for a non-finite
// result.
__device__ inline bool ermak_move(const RepDev& R, int rep, int i, double T, double x, double y,
                                  double z, double fx, double fy, double fz, double xi,
                                  double yi, double zi, double ox, double oy, double oz,
                                  double* __restrict__ px,
                                  double* __restrict__ py, double* __restrict__ pz,
                                  double* __restrict__ rx, double* __restrict__ ry,
                                  double* __restrict__ rz, int* __restrict__ nocup,
                                  int* __restrict__ ins, RepFlags* __restrict__ flags, int do_bin,
                                  int64_t ni, const ListMove* lm = nullptr, int par = 0) {
  const int p = R.p_off + i;
  const double side = R.side, side_inv = R.side_inv, dt = R.dt;
  const double diameter = i < R.nb ? R.sigma[0] : R.sigma[1];
  const double dinv = 1.0 / diameter;
  const double amplitude = sqrt(2.0 * T * dinv * dt);
  const double dx = fx * dinv * dt + amplitude * xi;
  const double dy = fy * dinv * dt + amplitude * yi;
  const double dz = fz * dinv * dt + amplitude * zi;
  x = x + dx;
  y = y + dy;
  z = z + dz;
  const double qx = ox + dx, qy = oy + dy, qz = oz + dz;
  rx[p] = qx;
  ry[p] = qy;
  rz[p] = qz;
  if (lm) {
    const double ex = qx - lm->rbx[p], ey = qy - lm->rby[p], ez = qz - lm->rbz[p];
    if (!(ex * ex + ey * ey + ez * ez <= lm->half_skin2)) atomicOr(&lm->lflag[2 * rep + par], 1);
  }
  x -= side * floor(side_inv * x);
  y -= side * floor(side_inv * y);
  z -= side * floor(side_inv * z);
  px[p] = x;
  py[p] = y;
  pz[p] = z;
  if (do_bin) insert_cell(R, rep, i, x, y, z, nocup, ins, flags, ni);
  return isfinite(x) && isfinite(y) && isfinite(z);
}

// Ermak-McCammon update (update_ermak.cu:91-160) with rannew64 lanes.
// One thread per lane; lane l serves particles l, l+L, l+2L ... and takes
// four draws per served slot, so every lane advances 4*ceil(N/L) draws per
// step and the ring pointers np, nq are common to the replica.
// One workgroup's share (block-table entry vb) of the update; the kernel
// below runs it for blockIdx.x, the persistent step kernel for every entry.
__device__ __forceinline__ void k_update_bin_block(
    int vb, const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                 double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                 double* __restrict__ rx, double* __restrict__ ry, double* __restrict__ rz,
                 const double* __restrict__ fx, const double* __restrict__ fy,
                 const double* __restrict__ fz, uint64_t* __restrict__ ring,
                 uint64_t* __restrict__ lcgs, const uint64_t* __restrict__ draws,
                 const int64_t* __restrict__ step, int total_lanes, int* __restrict__ nocup,
                 int* __restrict__ ins, RepFlags* __restrict__ flags, int do_bin,
                 ListMove lm = ListMove{nullptr, nullptr, nullptr, nullptr, 0.0}) {
  const int4 t = tbl[vb];
  if (t.x < 0) return;
  const int l = t.y + (int)threadIdx.x;
  if (l >= t.z) return;
  const RepDev& R = reps[t.x];
  const int gl = R.lane_off + l;
  const int64_t ni = *step;
  const double T = temperature_at(R, ni);
  const uint64_t dr = draws[t.x];
  const int par = list_parity(R, dr, 1);
  uint64_t lcg = lcgs[gl];
  int np = bd::kRingLen - 1 - (int)(dr % (uint64_t)bd::kRingLen);
  uint64_t* rl = ring + gl;
  const size_t stride = (size_t)total_lanes;
  bool bad = false;
  for (int k = 0; k < R.per_lane; k++) {
    const int i = l + k * R.lanes;
    const bool live = i < R.n;
    // the particle's state is loaded ahead of the draws, so that its memory
    // latency overlaps theirs (a slot past the end reads particle 0)
    const int p = R.p_off + (live ? i : 0);
    const double x = px[p], y = py[p], z = pz[p];
    const double gx = fx[p], gy = fy[p], gz = fz[p];
    const double ox = rx[p], oy = ry[p], oz = rz[p];
    uint64_t d[4];
    draw4(rl, stride, np, lcg, d);
    if (!live) continue;
    double xi, yi, zi;
    bd::gauss3(d[0], d[1], d[2], d[3], xi, yi, zi);
    if (!ermak_move(R, t.x, i, T, x, y, z, gx, gy, gz, xi, yi, zi, ox, oy, oz, px, py, pz, rx,
                    ry, rz, nocup, ins, flags, do_bin, ni, lm.lflag ? &lm : nullptr, par))
      bad = true;
  }
  lcgs[gl] = lcg;
  if (bad) raise_flag(flags, t.x, kFlagNonFinite, ni);
}

// The normals of particle i's next move, for a replica whose lanes are its
// particles (one draw block per lane and step): the same draws as
// k_update_bin_block, taken by the force kernel ahead of the move.
__device__ inline void lane_normals(const RepDev& R, int rep, int i, uint64_t* __restrict__ ring,
                                    uint64_t* __restrict__ lcgs,
                                    const uint64_t* __restrict__ draws, int total_lanes,
                                    double& xi, double& yi, double& zi) {
  const int gl = R.lane_off + i;
  uint64_t lcg = lcgs[gl];
  int np = bd::kRingLen - 1 - (int)(draws[rep] % (uint64_t)bd::kRingLen);
  uint64_t d[4];
  draw4(ring + gl, (size_t)total_lanes, np, lcg, d);
  lcgs[gl] = lcg;
  bd::gauss3(d[0], d[1], d[2], d[3], xi, yi, zi);
}

__global__ void __launch_bounds__(kBlk)
    k_update_bin(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                 double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                 double* __restrict__ rx, double* __restrict__ ry, double* __restrict__ rz,
                 const double* __restrict__ fx, const double* __restrict__ fy,
                 const double* __restrict__ fz, uint64_t* __restrict__ ring,
                 uint64_t* __restrict__ lcgs, const uint64_t* __restrict__ draws,
                 const int64_t* __restrict__ step, int total_lanes, int* __restrict__ nocup,
                 int* __restrict__ ins, RepFlags* __restrict__ flags, int do_bin, ListMove lm) {
  k_update_bin_block((int)blockIdx.x, reps, tbl, px, py, pz, rx, ry, rz, fx, fy, fz, ring, lcgs,
                     draws, step, total_lanes, nocup, ins, flags, do_bin, lm);
}

// Inclusive prefix sum over the 64 lanes of a wavefront: within each row of
// 16 lanes by shifts of 1, 2, 4 and 8 lanes (lanes without a source add 0),
// then the last lane of row 0 (and of row 2) added to row 1 (and row 3), and
// the last lane of row 1 to rows 2 and 3.
__device__ inline int wave_prefix_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Row packing.  A wavefront packs rpw = 64 / n rows (x, y) of a grid with n
// <= 64 cells per row (lane j n + z holds cell z of its j-th row), or one
// row in chunks of 64 cells when n > 64.  Each lane reads its cell's
// occupancy, a segmented wavefront prefix sum places the cells of a row one
// after the other from the row's first slot, and the lane writes its cell's
// particles ranked by id (ascending, the serial order of cell_locate_hst,
// cell_locate.cu:72-101) with their positions.  The insertion counters are
// reset for the next binning.  With advance != 0 the first workgroup also
// advances the step index, and the workgroup holding a replica's first
// big-grid rows that replica's draw counter.
constexpr int kRowWaves = kBlk / 64;

__host__ __device__ inline int rows_per_wave(int n) { return n <= 64 ? 64 / n : 1; }

__device__ __forceinline__ void k_cells_block(
    int vb, const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
            const double* __restrict__ px, const double* __restrict__ py,
            const double* __restrict__ pz, int* __restrict__ nocup, int* __restrict__ cnt,
            const int* __restrict__ ins, int* __restrict__ rs, double4* __restrict__ cpos,
            RepFlags* __restrict__ flags, int64_t* __restrict__ step,
            uint64_t* __restrict__ draws, int advance, int* __restrict__ lflag = nullptr,
            PosCell* __restrict__ pos4 = nullptr) {
  const int4 t = tbl[vb];
  if (advance && threadIdx.x == 0) {
    if (vb == 0) *step = *step + 1;
    // each replica's draw counter by the workgroup of its first big-grid rows,
    // which also clears the list rebuild flag that the next move will set
    if (t.x >= 0 && t.y == 0 && t.w == 0) {
      const uint64_t d = draws[t.x];
      if (lflag) lflag[2 * t.x + list_parity(reps[t.x], d, 0)] = 0;
      draws[t.x] = d + 4ULL * (uint64_t)reps[t.x].per_lane;
    }
  }
  if (t.x < 0) return;
  const int g = t.w;
  const RepDev& R = reps[t.x];
  const int n = R.ncell[g];
  const int rpw = rows_per_wave(n);
  const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
  const int row0 = t.y + wave * rpw;
  if (row0 >= t.z) return;  // whole wavefront
  const int j = n <= 64 ? lane / n : 0;
  const int seg0 = j * n;  // first lane of this lane's row (n <= 64), else 0
  const int zl = lane - seg0;
  const int row = row0 + j;  // x + n y
  const bool row_live = j < rpw && row < t.z;
  const int seg_last = min(seg0 + min(n, 64) - 1, 63);
  const int ntags = R.ntags[g];
  const int sp_off = g == 0 ? 0 : R.nb;
  const int row_base = row * n * ntags;  // grid-relative slot of the row
  int* rrow = rs + R.rs_off[g] + row * (n + 1);
  double4* cp = cpos + R.tag_off[g];
  const int* ins_g = ins + R.tag_off[g];
  int carry = 0;
  bool over = false;
  for (int z0 = 0; z0 < n; z0 += 64) {
    const int z = z0 + zl;
    const bool live = row_live && z < n;
    const int c = row * n + z;  // z + n (x + n y): coalesced across the lanes
    const bool fast8 = ntags == 8 && (R.tag_off[g] & 3) == 0;
    const int* slots = ins_g + (size_t)c * ntags;
    int occ = 0;
    int4 slo = make_int4(0, 0, 0, 0), shi = make_int4(0, 0, 0, 0);
    if (live) {
      occ = nocup[R.cell_off[g] + c];
      if (fast8) {  // the cell's eight slots, loaded beside its occupancy
        slo = reinterpret_cast<const int4*>(slots)[0];
        shi = reinterpret_cast<const int4*>(slots)[1];
      }
    }
    const int m = occ < ntags ? occ : ntags;
    const int incl = wave_prefix_sum(m);
    const int before = __builtin_amdgcn_ds_bpermute((seg0 > 0 ? seg0 - 1 : 0) << 2, incl);
    const int prior = seg0 > 0 ? before : 0;
    const int row_sum = __builtin_amdgcn_ds_bpermute(seg_last << 2, incl) - prior;
    const int start = row_base + carry + (incl - prior) - m;
    if (live) {
      rrow[z] = start;
      if (fast8) {
        // the cell's eight slots (two aligned 16-byte loads), ranked in registers
        int id[8] = {slo.x, slo.y, slo.z, slo.w, shi.x, shi.y, shi.z, shi.w};
#pragma unroll
        for (int k = 0; k < 8; k++) id[k] = k < m ? id[k] : 0x7fffffff;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          if (k < m) {
            int rank = 0;
#pragma unroll
            for (int q = 0; q < 8; q++) rank += id[q] < id[k] ? 1 : 0;
            const int p = R.p_off + sp_off + id[k];
            const double qx = px[p], qy = py[p], qz = pz[p];
            cp[start + rank] = make_double4(qx, qy, qz, (double)(sp_off + id[k]));
            if (pos4) pos4[p] = PosCell{qx, qy, qz, pack_cell(row % n, row / n, z), 0u};
          }
        }
      } else {
        for (int k = 0; k < m; k++) {
          const int idk = slots[k];
          int rank = 0;
          for (int q = 0; q < m; q++) rank += slots[q] < idk ? 1 : 0;
          const int p = R.p_off + sp_off + idk;
          const double qx = px[p], qy = py[p], qz = pz[p];
          cp[start + rank] = make_double4(qx, qy, qz, (double)(sp_off + idk));
          if (pos4) pos4[p] = PosCell{qx, qy, qz, pack_cell(row % n, row / n, z), 0u};
        }
      }
      cnt[R.cell_off[g] + c] = occ;
      nocup[R.cell_off[g] + c] = 0;
      over = over || occ >= ntags;
    }
    carry += row_sum;
  }
  if (row_live && zl == 0) rrow[n] = row_base + carry;
  if (over) atomicOr(&flags[t.x].bits, kFlagOverflow);
}

__global__ void __launch_bounds__(kBlk)
    k_cells(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
            const double* __restrict__ px, const double* __restrict__ py,
            const double* __restrict__ pz, int* __restrict__ nocup, int* __restrict__ cnt,
            const int* __restrict__ ins, int* __restrict__ rs, double4* __restrict__ cpos,
            RepFlags* __restrict__ flags, int64_t* __restrict__ step,
            uint64_t* __restrict__ draws, int advance, int* __restrict__ lflag,
            PosCell* __restrict__ pos4) {
  k_cells_block((int)blockIdx.x, reps, tbl, px, py, pz, nocup, cnt, ins, rs, cpos, flags, step,
                draws, advance, lflag, pos4);
}

// Distance along one axis from a point to the cell at offset d of its home
// cell (u = coordinate within the home cell), over the three periodic images.
__device__ inline double axis_gap(int d, double cs, double u, double L) {
  const double lo = (double)d * cs - u;
  const double hi = lo + cs;
  double g = lo > 0.0 ? lo : (hi < 0.0 ? -hi : 0.0);
  const double lo2 = lo - L, hi2 = hi - L;
  const double g2 = lo2 > 0.0 ? lo2 : (hi2 < 0.0 ? -hi2 : 0.0);
  const double lo3 = lo + L, hi3 = hi + L;
  const double g3 = lo3 > 0.0 ? lo3 : (hi3 < 0.0 ? -hi3 : 0.0);
  g = g2 < g ? g2 : g;
  return g3 < g ? g3 : g;
}

// Scan one grid around position (x, y, z) in the reference's order
// (get_forces.cu:360-416 / 495-550): i_del, j_del, k_del nested, slot order.
// A stencil cell whose nearest point is farther than the cutoff (plus a
// 1e-9 margin that covers rounding) cannot hold an interacting partner and
// is skipped; the interacting pairs and their order are unchanged.
template <int POT, bool NEIGH>
__device__ inline void scan_grid(const RepDev& R, int g, int nrange, double sigma, double rc2,
                                 int self, double x, double y, double z,
                                 const int* __restrict__ rs, const double4* __restrict__ cpos,
                                 double& fx, double& fy, double& fz, double& vir, double& pot,
                                 int* nlist, int& ncount, int maxn) {
  int ii, jj, kk;
  if (!cell_coords(R, g, x, y, z, ii, jj, kk)) return;
  const int n = R.ncell[g];
  const double side = R.side, side_inv = R.side_inv;
  const double cs = R.cell_side[g];
  const double ux = x - (double)ii * cs, uy = y - (double)jj * cs, uz = z - (double)kk * cs;
  const double lim = rc2 + 1e-9;
  const int* rs_g = rs + R.rs_off[g];
  const double4* cp_g = cpos + R.tag_off[g];
  for (int di = -nrange; di <= nrange; di++) {
    const double gx = axis_gap(di, cs, ux, side);
    const double gx2 = gx * gx;
    if (gx2 > lim) continue;
    int a = ii + di;
    a = a < 0 ? a + n : (a >= n ? a - n : a);
    for (int dj = -nrange; dj <= nrange; dj++) {
      const double gy = axis_gap(dj, cs, uy, side);
      const double gxy2 = gx2 + gy * gy;
      if (gxy2 > lim) continue;
      int b = jj + dj;
      b = b < 0 ? b + n : (b >= n ? b - n : b);
      for (int dk = -nrange; dk <= nrange; dk++) {
        const double gz = axis_gap(dk, cs, uz, side);
        if (gxy2 + gz * gz > lim) continue;
        int c = kk + dk;
        c = c < 0 ? c + n : (c >= n ? c - n : c);
        const int ri = (a + n * b) * (n + 1) + c;
        const int s0 = rs_g[ri];
        const int occ = rs_g[ri + 1] - s0;
        const double4* cell = cp_g + s0;
        for (int s = 0; s < occ; s++) {
          const double4 q = cell[s];
          if (self >= 0 && q.w == (double)self) continue;
          double dx = q.x - x;
          double dy = q.y - y;
          double dz = q.z - z;
          dx -= side * floor(side_inv * dx + 0.5);
          dy -= side * floor(side_inv * dy + 0.5);
          dz -= side * floor(side_inv * dz + 0.5);
          const double dist2 = dx * dx + dy * dy + dz * dz;
          if (dist2 < rc2) {
            const double dist_inv = sqrt(1.0 / dist2);
            double nf;
            bd::potential<POT>(R, sigma, dist_inv, pot, nf);
            const double qx = nf * dist_inv * dx;
            const double qy = nf * dist_inv * dy;
            const double qz = nf * dist_inv * dz;
            fx += qx;
            fy += qy;
            fz += qz;
            vir -= 0.5 * (dx * qx + dy * qy + dz * qz);
            if (NEIGH) {
              if (ncount < maxn) nlist[ncount] = (int)q.w;
              ncount++;
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fast stencil walk.
//
// Same pairs, same order and same IEEE operations as scan_grid, with less
// bookkeeping:
//  * the distance from the particle to a stencil slab along one axis is
//    d*cs - u (d > 0), u + (|d|-1)*cs (d < 0) or 0, u the coordinate within
//    the home cell; the innermost offsets' squared gaps are precomputed;
//  * the minimum image: for a candidate in the cell at unwrapped coordinate
//    h+d, floor(dx/L + 0.5) equals -K with K = -1 (h+d < 0), +1 (h+d >= n) or
//    0, because |dx + K L| <= (NR+1) cs <= L/2 - cs/2 when n >= 2 NR + 3; so
//    dx - L*floor(...) is exactly dx + K*L (and dx itself when K == 0).
// Replicas whose grids are narrower than 2 NR + 3 cells (or NR > 2) use
// scan_grid.  ACC maps accessor-frame cell coordinates to a slot range.
// ---------------------------------------------------------------------------
struct GlobalCells {
  const int* rs;  // the grid's row-start table
  const double4* cp;
  int n;
  __device__ int index(int a, int b, int c) const {
    a = a < 0 ? a + n : (a >= n ? a - n : a);
    b = b < 0 ? b + n : (b >= n ? b - n : b);
    c = c < 0 ? c + n : (c >= n ? c - n : c);
    return (a + n * b) * (n + 1) + c;
  }
  __device__ void cell(int idx, int& occ, const double4*& ptr) const {
    const int s0 = rs[idx];
    occ = rs[idx + 1] - s0;
    ptr = cp + s0;
  }
};

struct LdsCells {
  const unsigned int* tbl;  // (start << 6) | occupancy
  const double4* spos;
  int lo0, lo1, lo2, d0, d1;
  __device__ int index(int a, int b, int c) const {
    return (a - lo0) + d0 * ((b - lo1) + d1 * (c - lo2));
  }
  __device__ void cell(int idx, int& occ, const double4*& ptr) const {
    const unsigned int e = tbl[idx];
    occ = (int)(e & 63u);
    ptr = spos + (e >> 6);
  }
};

__device__ inline double slab_gap(int d, double cs, double u) {
  return d > 0 ? (double)d * cs - u : (d < 0 ? u + (double)(-d - 1) * cs : 0.0);
}

__device__ inline double image_shift(int hd, int n, double L) {
  return hd < 0 ? -L : (hd >= n ? L : 0.0);
}

// Walk one grid's stencil (nrange NR) around (x, y, z).  hw: home cell
// (wrapped), hu: the same cell in the accessor's frame.  The innermost (dk)
// row is unrolled: its occupancies and first slots are loaded together
// before the pairs are visited in order.  SPLIT: the lane walks only the
// cells whose running index cbase + flat is congruent to `sub` modulo G.
template <int POT, int NR, bool SPLIT, class ACC>
__device__ inline void scan_fast(const RepDev& R, int n, double cs, double sigma, double rc2,
                                 int self, double x, double y, double z, const int hw[3],
                                 const int hu[3], const ACC& acc, double& fx, double& fy,
                                 double& fz, double& vir, double& pot, int sub, int G,
                                 int cbase) {
  const double L = R.side;
  const double lim = rc2 + 1e-9;
  const double ux = x - (double)hw[0] * cs, uy = y - (double)hw[1] * cs,
               uz = z - (double)hw[2] * cs;
  double g2z[2 * NR + 1], shz[2 * NR + 1];
#pragma unroll
  for (int k = 0; k <= 2 * NR; k++) {
    const double gz = slab_gap(k - NR, cs, uz);
    g2z[k] = gz * gz;
    shz[k] = image_shift(hw[2] + k - NR, n, L);
  }
  for (int di = -NR; di <= NR; di++) {
    const double gx = slab_gap(di, cs, ux);
    const double gx2 = gx * gx;
    if (gx2 > lim) continue;
    const double shx = image_shift(hw[0] + di, n, L);
    for (int dj = -NR; dj <= NR; dj++) {
      const double gy = slab_gap(dj, cs, uy);
      const double gxy2 = gx2 + gy * gy;
      if (gxy2 > lim) continue;
      const double shy = image_shift(hw[1] + dj, n, L);
      const int rowflat = cbase + ((di + NR) * (2 * NR + 1) + (dj + NR)) * (2 * NR + 1);
      int occk[2 * NR + 1];
      const double4* cellk[2 * NR + 1];
#pragma unroll
      for (int k = 0; k <= 2 * NR; k++) {
        occk[k] = 0;
        cellk[k] = nullptr;
        const bool mine = !SPLIT || (rowflat + k) % G == sub;
        if (mine && gxy2 + g2z[k] <= lim)
          acc.cell(acc.index(hu[0] + di, hu[1] + dj, hu[2] + k - NR), occk[k], cellk[k]);
      }
      double4 first[2 * NR + 1];
#pragma unroll
      for (int k = 0; k <= 2 * NR; k++)
        if (occk[k] > 0) first[k] = cellk[k][0];
#pragma unroll
      for (int k = 0; k <= 2 * NR; k++) {
        const int occ = occk[k];
        if (occ == 0) continue;
        const int dk = k - NR;
        const double shzk = shz[k];
        const double4* cell = cellk[k];
        const bool center = self >= 0 && di == 0 && dj == 0 && dk == 0;
        for (int s = 0; s < occ; s++) {
          const double4 q = s == 0 ? first[k] : cell[s];
          if (center && q.w == (double)self) continue;
          double dx = q.x - x;
          double dy = q.y - y;
          double dz = q.z - z;
          if (shx != 0.0) dx = dx + shx;
          if (shy != 0.0) dy = dy + shy;
          if (shzk != 0.0) dz = dz + shzk;
          const double dist2 = dx * dx + dy * dy + dz * dz;
          if (dist2 < rc2) {
            const double dist_inv = sqrt(1.0 / dist2);
            double nf;
            bd::potential<POT>(R, sigma, dist_inv, pot, nf);
            const double qx = nf * dist_inv * dx;
            const double qy = nf * dist_inv * dy;
            const double qz = nf * dist_inv * dz;
            fx += qx;
            fy += qy;
            fz += qz;
            vir -= 0.5 * (dx * qx + dy * qy + dz * qz);
          }
        }
      }
    }
  }
}

template <int POT, bool SPLIT, class ACC>
__device__ inline void scan_fast_nr(int nr, const RepDev& R, int n, double cs, double sigma,
                                    double rc2, int self, double x, double y, double z,
                                    const int hw[3], const int hu[3], const ACC& acc,
                                    double& fx, double& fy, double& fz, double& vir,
                                    double& pot, int sub, int G, int cbase) {
  if (nr == 1)
    scan_fast<POT, 1, SPLIT>(R, n, cs, sigma, rc2, self, x, y, z, hw, hu, acc, fx, fy, fz, vir,
                             pot, sub, G, cbase);
  else
    scan_fast<POT, 2, SPLIT>(R, n, cs, sigma, rc2, self, x, y, z, hw, hu, acc, fx, fy, fz, vir,
                             pot, sub, G, cbase);
}

// One pair term inside the cutoff (get_forces.cu:396-411): potential, force
// along the separation, accumulations.
template <int POT>
__device__ inline void pair_add(const RepDev& R, double sigma, double dist2, double dx, double dy,
                                double dz, double& fx, double& fy, double& fz, double& vir,
                                double& pot) {
  const double dist_inv = sqrt(1.0 / dist2);
  double nf;
  bd::potential<POT>(R, sigma, dist_inv, pot, nf);
  const double qx = nf * dist_inv * dx;
  const double qy = nf * dist_inv * dy;
  const double qz = nf * dist_inv * dz;
  fx += qx;
  fy += qy;
  fz += qz;
  vir -= 0.5 * (dx * qx + dy * qy + dz * qz);
}

// Deferred pair terms: the walk records each partner inside the cutoff as
// slot | image codes (2 bits per axis, two's complement: 0 none, 1 +L,
// 3 -L) | pass bit in a
// per-thread LDS column, for both stencil passes of the particle; drain()
// evaluates the recorded terms in recording order, recomputing the
// separation with the same operations.  The potential is then evaluated by
// all lanes together on interacting pairs only, instead of by the whole
// wavefront whenever one lane meets a partner.
constexpr unsigned int kSlotBits = 25;
constexpr unsigned int kSlotMask = (1u << kSlotBits) - 1u;
constexpr unsigned int kPassBit = 1u << 31;

__device__ inline unsigned int image_code(double sh) {
  return sh == 0.0 ? 0u : (sh < 0.0 ? 3u : 1u);
}

// The shift of the 2-bit image code at bit `bit` of `word`: L times the
// code's sign-extended value (-1, 0, +1).  L * -1 = -L and L * 0 = +0 exactly,
// so the shift equals the one the code was made from (a zero shift at most
// flips the sign of a zero difference, which no squared distance, force or
// virial term can see), in three instructions (extract, convert, multiply).
__device__ inline double image_at(unsigned int word, int bit, double L) {
  return L * (double)((int)(word << (30 - bit)) >> 30);
}

struct Defer {
  const double4* cp0;  // slots of the same-species pass (pass bit 0)
  const double4* cp1;  // slots of the cross-species pass (pass bit 1)
  double sigma0, sigma1;
  unsigned int* buf;   // LDS column, stride kBlk
  int nd;
};

template <int POT>
__device__ inline void drain(const RepDev& R, double x, double y, double z, Defer& d,
                             double& fx, double& fy, double& fz, double& vir, double& pot) {
  const double L = R.side;
  const int nd = d.nd;
  d.nd = 0;
  for (int e = 0; e < nd; e++) {
    const unsigned int wc = d.buf[e * kBlk];
    const double4 qc = ((wc & kPassBit) ? d.cp1 : d.cp0)[wc & kSlotMask];
    const double dx = (qc.x - x) + image_at(wc, kSlotBits, L);
    const double dy = (qc.y - y) + image_at(wc, kSlotBits + 2, L);
    const double dz = (qc.z - z) + image_at(wc, kSlotBits + 4, L);
    const double dist2 = dx * dx + dy * dy + dz * dz;
    pair_add<POT>(R, (wc & kPassBit) ? d.sigma1 : d.sigma0, dist2, dx, dy, dz, fx, fy, fz, vir,
                  pot);
  }
}

// Row-range stencil walk, the default force path.  For every stencil row
// (di, dj) that passes the slab-gap test, the cells k_del = klo..khi that
// can hold a partner are consecutive z cells of one packed row, i.e. one
// slot range, or two when the run crosses the periodic boundary (taken in
// k_del order: the wrapped low cells first, or the wrapped high cells
// last).  The ranges are walked as one loop: same pairs, same order and the
// same IEEE operations as scan_fast, with one loop per row instead of one
// per cell.  DEPTH > 0 defers the pair terms through `buf` (DEPTH entries,
// stride kBlk), draining whenever it fills and at the end.
template <int POT, int NR, int DEPTH>
__device__ inline void scan_rows(const RepDev& R, int n, double cs, double sigma, double rc2,
                                 int self, double x, double y, double z, const int hw[3],
                                 const int* __restrict__ rs, const double4* __restrict__ cp,
                                 Defer& df, unsigned int passbit, double& fx, double& fy,
                                 double& fz, double& vir, double& pot) {
  const double L = R.side;
  const double lim = rc2 + 1e-9;
  const double ux = x - (double)hw[0] * cs, uy = y - (double)hw[1] * cs,
               uz = z - (double)hw[2] * cs;
  double g2lo[NR], g2hi[NR];  // squared z gaps of k_del = -(k+1) and +(k+1)
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double lo = slab_gap(-(k + 1), cs, uz);
    const double hi = slab_gap(k + 1, cs, uz);
    g2lo[k] = lo * lo;
    g2hi[k] = hi * hi;
  }
  for (int di = -NR; di <= NR; di++) {
    const double gx = slab_gap(di, cs, ux);
    const double gx2 = gx * gx;
    if (gx2 > lim) continue;
    const double shx = image_shift(hw[0] + di, n, L);
    int a = hw[0] + di;
    a = a < 0 ? a + n : (a >= n ? a - n : a);
    for (int dj = -NR; dj <= NR; dj++) {
      const double gy = slab_gap(dj, cs, uy);
      const double gxy2 = gx2 + gy * gy;
      if (gxy2 > lim) continue;
      const double shy = image_shift(hw[1] + dj, n, L);
      int b = hw[1] + dj;
      b = b < 0 ? b + n : (b >= n ? b - n : b);
      // the gaps grow with |k_del|, so the passing offsets are contiguous
      int klo = 0, khi = 0;
#pragma unroll
      for (int k = 0; k < NR; k++) {
        if (gxy2 + g2lo[k] <= lim) klo = -(k + 1);
        if (gxy2 + g2hi[k] <= lim) khi = k + 1;
      }
      const int* rrow = rs + (a + n * b) * (n + 1);
      const int z0 = hw[2] + klo, z1 = hw[2] + khi;
      int s0a, s1a, s0b = 0, s1b = 0;
      double sha = 0.0, shb = 0.0;
      if (z0 < 0) {  // cells z0+n .. n-1 seen through the -L image, then 0 .. z1
        s0a = rrow[z0 + n];
        s1a = rrow[n];
        sha = -L;
        s0b = rrow[0];
        s1b = rrow[z1 + 1];
      } else if (z1 >= n) {  // cells z0 .. n-1, then 0 .. z1-n through the +L image
        s0a = rrow[z0];
        s1a = rrow[n];
        s0b = rrow[0];
        s1b = rrow[z1 - n + 1];
        shb = L;
      } else {
        s0a = rrow[z0];
        s1a = rrow[z1 + 1];
      }
      const int na = s1a - s0a;
      const int tot = na + (s1b - s0b);
      const bool center = self >= 0 && di == 0 && dj == 0;
      // the next candidate is loaded while this one is evaluated
      double4 qn = make_double4(0.0, 0.0, 0.0, -1.0);
      if (tot > 0) qn = cp[na > 0 ? s0a : s0b];
      if (DEPTH == 0) {
        for (int t = 0; t < tot; t++) {
          const bool in_a = t < na;
          const double4 q = qn;
          if (t + 1 < tot) qn = cp[t + 1 < na ? s0a + t + 1 : s0b + (t + 1 - na)];
          if (center && q.w == (double)self) continue;
          const double shz = in_a ? sha : shb;
          // adding a zero shift changes at most the sign of a zero difference,
          // which no squared distance, force or virial term can see
          const double dx = (q.x - x) + shx;
          const double dy = (q.y - y) + shy;
          const double dz = (q.z - z) + shz;
          const double dist2 = dx * dx + dy * dy + dz * dz;
          if (dist2 < rc2) pair_add<POT>(R, sigma, dist2, dx, dy, dz, fx, fy, fz, vir, pot);
        }
      } else {
        // partners inside the cutoff are recorded in visiting order
        const unsigned int cxy =
            (image_code(shx) << kSlotBits) | (image_code(shy) << (kSlotBits + 2));
        const unsigned int ca = cxy | (image_code(sha) << (kSlotBits + 4));
        const unsigned int cb = cxy | (image_code(shb) << (kSlotBits + 4));
        const double selfw = center ? (double)self : -1.0;
        for (int t = 0; t < tot; t++) {
          const bool in_a = t < na;
          const int sl = in_a ? s0a + t : s0b + (t - na);
          const double4 q = qn;
          if (t + 1 < tot) qn = cp[t + 1 < na ? s0a + t + 1 : s0b + (t + 1 - na)];
          const double dx = (q.x - x) + shx;
          const double dy = (q.y - y) + shy;
          const double dz = (q.z - z) + (in_a ? sha : shb);
          const double dist2 = dx * dx + dy * dy + dz * dz;
          if (dist2 < rc2 && q.w != selfw) {
            df.buf[df.nd * kBlk] = (unsigned int)sl | (in_a ? ca : cb) | passbit;
            if (++df.nd == DEPTH) drain<POT>(R, x, y, z, df, fx, fy, fz, vir, pot);
          }
        }
      }
    }
  }
}

// Slab walk: the same rows, ranges, order and arithmetic as scan_rows, but
// the candidate loop is flattened over a whole slab (fixed di, all dj):
// the lane first lists the slab's non-empty slot ranges in LDS (`seg`,
// stride kBlk: start, then length | y image code << 16 | z image code << 18
// | centre-row flag << 20), then walks them as one stream.  A wavefront
// then loops max(sum of the slab's range lengths) times instead of the sum
// over rows of each row's longest range, which is where the row walk loses
// most of its lanes.
constexpr int kSlabSegs = 10;  // 2 ranges x (2 NR + 1) rows, NR <= 2

template <int POT, int NR, int DEPTH, bool REC = false>
__device__ inline void scan_slabs(const RepDev& R, int n, double cs, double sigma, double rc2,
                                  int self, double x, double y, double z, const int hw[3],
                                  const int* __restrict__ rs, const double4* __restrict__ cp,
                                  Defer& df, unsigned int passbit, unsigned int* seg, double& fx,
                                  double& fy, double& fz, double& vir, double& pot,
                                  int di0 = -NR, int di1 = NR) {
  const double L = R.side;
  const double lim = rc2 + 1e-9;
  const double ux = x - (double)hw[0] * cs, uy = y - (double)hw[1] * cs,
               uz = z - (double)hw[2] * cs;
  double g2lo[NR], g2hi[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double lo = slab_gap(-(k + 1), cs, uz);
    const double hi = slab_gap(k + 1, cs, uz);
    g2lo[k] = lo * lo;
    g2hi[k] = hi * hi;
  }
  const double selfw = self >= 0 ? (double)self : -1.0;
  for (int di = di0; di <= di1; di++) {
    const double gx = slab_gap(di, cs, ux);
    const double gx2 = gx * gx;
    const double shx = image_shift(hw[0] + di, n, L);
    int a = hw[0] + di;
    a = a < 0 ? a + n : (a >= n ? a - n : a);
    // 1. the slab's ranges
    int nseg = 0;
    if (gx2 <= lim) {
      for (int dj = -NR; dj <= NR; dj++) {
        const double gy = slab_gap(dj, cs, uy);
        const double gxy2 = gx2 + gy * gy;
        if (gxy2 > lim) continue;
        const unsigned int ycode = image_code(image_shift(hw[1] + dj, n, L));
        int b = hw[1] + dj;
        b = b < 0 ? b + n : (b >= n ? b - n : b);
        int klo = 0, khi = 0;
#pragma unroll
        for (int k = 0; k < NR; k++) {
          if (gxy2 + g2lo[k] <= lim) klo = -(k + 1);
          if (gxy2 + g2hi[k] <= lim) khi = k + 1;
        }
        const int* rrow = rs + (a + n * b) * (n + 1);
        const int z0 = hw[2] + klo, z1 = hw[2] + khi;
        int s0a, s1a, s0b = 0, s1b = 0;
        unsigned int za = 0u, zb = 0u;
        if (z0 < 0) {
          s0a = rrow[z0 + n];
          s1a = rrow[n];
          za = 3u;  // -L
          s0b = rrow[0];
          s1b = rrow[z1 + 1];
        } else if (z1 >= n) {
          s0a = rrow[z0];
          s1a = rrow[n];
          s0b = rrow[0];
          s1b = rrow[z1 - n + 1];
          zb = 1u;  // +L
        } else {
          s0a = rrow[z0];
          s1a = rrow[z1 + 1];
        }
        const unsigned int ctr = (self >= 0 && di == 0 && dj == 0) ? (1u << 20) : 0u;
        if (s1a > s0a) {
          seg[(2 * nseg) * kBlk] = (unsigned int)s0a;
          seg[(2 * nseg + 1) * kBlk] = (unsigned int)(s1a - s0a) | (ycode << 16) | (za << 18) | ctr;
          nseg++;
        }
        if (s1b > s0b) {
          seg[(2 * nseg) * kBlk] = (unsigned int)s0b;
          seg[(2 * nseg + 1) * kBlk] = (unsigned int)(s1b - s0b) | (ycode << 16) | (zb << 18) | ctr;
          nseg++;
        }
      }
    }
    // 2. one stream over the slab's ranges, in listing order; the next
    //    candidate is loaded before the current one is evaluated
    const unsigned int xbits = image_code(shx) << kSlotBits;
    bool more = nseg > 0;
    int k = 0, t = 0, len = 0, st = 0;
    unsigned int meta = 0u;
    double4 qn = make_double4(0.0, 0.0, 0.0, -1.0);
    if (more) {
      st = (int)seg[0];
      meta = seg[kBlk];
      len = (int)(meta & 0xffffu);
      k = 1;
      qn = cp[st];
    }
    while (more) {
      const double4 q = qn;
      const int sl = st + t;
      const unsigned int mq = meta;
      if (++t == len) {
        if (k < nseg) {
          st = (int)seg[(2 * k) * kBlk];
          meta = seg[(2 * k + 1) * kBlk];
          len = (int)(meta & 0xffffu);
          t = 0;
          k++;
        } else {
          more = false;
        }
      }
      if (more) qn = cp[st + t];
      const double dx = (q.x - x) + shx;
      const double dy = (q.y - y) + image_at(mq, 16, L);
      const double dz = (q.z - z) + image_at(mq, 18, L);
      const double dist2 = dx * dx + dy * dy + dz * dz;
      if (dist2 < rc2 && !((mq >> 20) != 0u && q.w == selfw)) {
        if (DEPTH == 0) {
          pair_add<POT>(R, sigma, dist2, dx, dy, dz, fx, fy, fz, vir, pot);
        } else {
          if (!REC || df.nd < DEPTH)
            df.buf[df.nd * kBlk] = (unsigned int)sl | xbits |
                                   (((mq >> 16) & 3u) << (kSlotBits + 2)) |
                                   (((mq >> 18) & 3u) << (kSlotBits + 4)) | passbit;
          if (++df.nd == DEPTH && !REC) drain<POT>(R, x, y, z, df, fx, fy, fz, vir, pot);
        }
      }
    }
  }
}

// Chunked slab walk: the slab walk's rows, ranges, order and arithmetic with
// the memory latency batched.  The range bounds of all rows of a slab are
// loaded before any is used (one latency per slab instead of one per row),
// and the slab's candidate stream is loaded CH slots at a time before the
// chunk is evaluated in stream order (one latency per CH candidates instead
// of one per candidate).  A candidate equal to the particle itself can only
// occur in its own cell, so the self test needs no centre-row flag.  The
// deferral buffer is drained before a chunk that might overfill it, which
// moves drains but not the order of the pair terms.  REC (both slab walks):
// record the partners only, never drain; df.nd then counts every partner and
// exceeds DEPTH when the column overflowed.
template <int POT, int NR, int DEPTH, int CH, bool REC = false>
__device__ inline void scan_chunks(const RepDev& R, int n, double cs, double sigma, double rc2,
                                   int self, double x, double y, double z, const int hw[3],
                                   const int* __restrict__ rs, const double4* __restrict__ cp,
                                   Defer& df, unsigned int passbit, unsigned int* seg, double& fx,
                                   double& fy, double& fz, double& vir, double& pot,
                                   int di0 = -NR, int di1 = NR) {
  static_assert(DEPTH == 0 || DEPTH >= CH, "a chunk must fit the deferral buffer");
  constexpr int W = 2 * NR + 1;
  const double L = R.side;
  const double lim = rc2 + 1e-9;
  const double ux = x - (double)hw[0] * cs, uy = y - (double)hw[1] * cs,
               uz = z - (double)hw[2] * cs;
  double g2lo[NR], g2hi[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double lo = slab_gap(-(k + 1), cs, uz);
    const double hi = slab_gap(k + 1, cs, uz);
    g2lo[k] = lo * lo;
    g2hi[k] = hi * hi;
  }
  const double selfw = self >= 0 ? (double)self : -1.0;
  for (int di = di0; di <= di1; di++) {
    const double gx = slab_gap(di, cs, ux);
    const double gx2 = gx * gx;
    if (gx2 > lim) continue;
    const double shx = image_shift(hw[0] + di, n, L);
    int a = hw[0] + di;
    a = a < 0 ? a + n : (a >= n ? a - n : a);
    // 1. every row's range bounds, loads issued back to back
    int lo[W], hi[W], lo2[W], hi2[W];
    unsigned int code[W];  // y image << 4 | z image of range b << 2 | z image of range a
#pragma unroll
    for (int r = 0; r < W; r++) {
      const int dj = r - NR;
      const double gy = slab_gap(dj, cs, uy);
      const double gxy2 = gx2 + gy * gy;
      int b = hw[1] + dj;
      b = b < 0 ? b + n : (b >= n ? b - n : b);
      int klo = 0, khi = 0;
#pragma unroll
      for (int k = 0; k < NR; k++) {
        if (gxy2 + g2lo[k] <= lim) klo = -(k + 1);
        if (gxy2 + g2hi[k] <= lim) khi = k + 1;
      }
      const int* rrow = rs + (a + n * b) * (n + 1);
      const int z0 = hw[2] + klo, z1 = hw[2] + khi;
      const bool wlo = z0 < 0, whi = z1 >= n;
      lo[r] = hi[r] = lo2[r] = hi2[r] = 0;
      if (gxy2 <= lim) {
        lo[r] = rrow[wlo ? z0 + n : z0];
        hi[r] = rrow[(wlo || whi) ? n : z1 + 1];
        if (wlo || whi) {
          lo2[r] = rrow[0];
          hi2[r] = rrow[wlo ? z1 + 1 : z1 - n + 1];
        }
      }
      code[r] = (image_code(image_shift(hw[1] + dj, n, L)) << 4) | (whi ? 4u : 0u) |
                (wlo ? 3u : 0u);
    }
    // 2. the slab's non-empty ranges in visiting order
    int nseg = 0;
#pragma unroll
    for (int r = 0; r < W; r++) {
      const unsigned int ycode = code[r] >> 4;
      if (hi[r] > lo[r]) {
        seg[(2 * nseg) * kBlk] = (unsigned int)lo[r];
        seg[(2 * nseg + 1) * kBlk] =
            (unsigned int)(hi[r] - lo[r]) | (ycode << 16) | ((code[r] & 3u) << 18);
        nseg++;
      }
      if (hi2[r] > lo2[r]) {
        seg[(2 * nseg) * kBlk] = (unsigned int)lo2[r];
        seg[(2 * nseg + 1) * kBlk] =
            (unsigned int)(hi2[r] - lo2[r]) | (ycode << 16) | (((code[r] >> 2) & 3u) << 18);
        nseg++;
      }
    }
    // 3. the stream, CH candidates per round; the next range's descriptor is
    //    read one range ahead
    const unsigned int xbits = image_code(shx) << kSlotBits;
    int k = 0, st = 0, rem = 0;
    unsigned int mc = 0u, nst = 0u, nmc = 0u;
    double sy = 0.0, sz = 0.0;
    if (nseg > 0) {
      nst = seg[0];
      nmc = seg[kBlk];
    }
    while (true) {
      int sl[CH];
      unsigned int mu[CH];
      double syu[CH], szu[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        if (rem == 0 && k < nseg) {
          st = (int)nst;
          mc = nmc;
          rem = (int)(mc & 0xffffu);
          sy = image_at(mc, 16, L);
          sz = image_at(mc, 18, L);
          k++;
          if (k < nseg) {
            nst = seg[(2 * k) * kBlk];
            nmc = seg[(2 * k + 1) * kBlk];
          }
        }
        sl[u] = rem > 0 ? st : -1;
        mu[u] = mc;
        syu[u] = sy;
        szu[u] = sz;
        if (rem > 0) {
          st++;
          rem--;
        }
      }
      if (sl[0] < 0) break;
      double4 q[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) q[u] = cp[sl[u] >= 0 ? sl[u] : sl[0]];
      if (DEPTH > 0 && !REC && df.nd > DEPTH - CH)
        drain<POT>(R, x, y, z, df, fx, fy, fz, vir, pot);
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const double dx = (q[u].x - x) + shx;
        const double dy = (q[u].y - y) + syu[u];
        const double dz = (q[u].z - z) + szu[u];
        const double dist2 = dx * dx + dy * dy + dz * dz;
        if (sl[u] >= 0 && dist2 < rc2 && q[u].w != selfw) {
          if (DEPTH == 0) {
            pair_add<POT>(R, sigma, dist2, dx, dy, dz, fx, fy, fz, vir, pot);
          } else {
            if (!REC || df.nd < DEPTH)
              df.buf[df.nd * kBlk] = (unsigned int)sl[u] | xbits |
                                     (((mu[u] >> 16) & 3u) << (kSlotBits + 2)) |
                                     (((mu[u] >> 18) & 3u) << (kSlotBits + 4)) | passbit;
            df.nd++;
          }
        }
      }
    }
  }
}

// Both stencils of replica-local particle i (same species, then the other
// species, main.cu:576-585) from global memory, as one rolled pass loop.
// SPLIT walks cells (scan_fast) so that G lanes can share the stencil; the
// single-lane walk uses row ranges (scan_rows), deferring pair terms through
// `buf` when DEPTH > 0.
template <int POT, bool SPLIT, int DEPTH = 0, int WALK = 0>
__device__ inline void forces_global(const RepDev& R, int i, double x, double y, double z,
                                     const int* __restrict__ rs,
                                     const double4* __restrict__ cpos, double& fx, double& fy,
                                     double& fz, double& v, double& u, int sub = 0, int G = 1,
                                     unsigned int* buf = nullptr, unsigned int* seg = nullptr) {
  const int g = i < R.nb ? 0 : 1;
  int cbase = 0;
  Defer df{cpos + (g == 0 ? R.tag_off[0] : R.tag_off[1]),
           cpos + (g == 0 ? R.tag_off[1] : R.tag_off[0]), g == 0 ? R.sigma[0] : R.sigma[1],
           R.sigma_bs, buf, 0};
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int gg = pass == 0 ? g : 1 - g;
    const int nr = pass == 0 ? (g == 0 ? R.nrange_same[0] : R.nrange_same[1])
                             : (g == 0 ? R.nrange_cross[0] : R.nrange_cross[1]);
    const double sigma = pass == 0 ? (g == 0 ? R.sigma[0] : R.sigma[1]) : R.sigma_bs;
    const double rc2 = pass == 0 ? (g == 0 ? R.rc2_same[0] : R.rc2_same[1]) : R.rc2_bs;
    const int self = pass == 0 ? i : -1;
    if (R.fast) {
      int hg[3];
      if (!cell_coords(R, gg, x, y, z, hg[0], hg[1], hg[2])) return;
      const int n = gg == 0 ? R.ncell[0] : R.ncell[1];
      const double cs = gg == 0 ? R.cell_side[0] : R.cell_side[1];
      const int* rs_g = rs + (gg == 0 ? R.rs_off[0] : R.rs_off[1]);
      const double4* cp_g = cpos + (gg == 0 ? R.tag_off[0] : R.tag_off[1]);
      if (SPLIT) {
        const GlobalCells acc{rs_g, cp_g, n};
        scan_fast_nr<POT, true>(nr, R, n, cs, sigma, rc2, self, x, y, z, hg, hg, acc, fx, fy,
                                fz, v, u, sub, G, cbase);
        cbase += (2 * nr + 1) * (2 * nr + 1) * (2 * nr + 1);
      } else if (WALK == 1) {
        if (nr == 1)
          scan_slabs<POT, 1, DEPTH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                    pass == 0 ? 0u : kPassBit, seg, fx, fy, fz, v, u);
        else
          scan_slabs<POT, 2, DEPTH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                    pass == 0 ? 0u : kPassBit, seg, fx, fy, fz, v, u);
      } else if (WALK >= 2) {
        constexpr int CH = WALK == 2 ? 2 : 4;
        if (nr == 1)
          scan_chunks<POT, 1, DEPTH, CH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                         pass == 0 ? 0u : kPassBit, seg, fx, fy, fz, v, u);
        else
          scan_chunks<POT, 2, DEPTH, CH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                         pass == 0 ? 0u : kPassBit, seg, fx, fy, fz, v, u);
      } else if (nr == 1) {
        scan_rows<POT, 1, DEPTH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                 pass == 0 ? 0u : kPassBit, fx, fy, fz, v, u);
      } else {
        scan_rows<POT, 2, DEPTH>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df,
                                 pass == 0 ? 0u : kPassBit, fx, fy, fz, v, u);
      }
    } else {
      int nc = 0;
      scan_grid<POT, false>(R, gg, nr, sigma, rc2, self, x, y, z, rs, cpos, fx, fy, fz, v, u,
                            nullptr, nc, 0);
    }
  }
  if (DEPTH > 0 && !SPLIT) drain<POT>(R, x, y, z, df, fx, fy, fz, v, u);
}

// Forces, virials and potentials of every particle: same-species stencil
// then cross-species stencil (main.cu:576-585), accumulators start at zero.
// NEIGH records the interacting partners in traversal order (test hook).
// DEPTH: per-thread LDS entries for deferred pair terms (0 = direct);
// WALK: 0 row walk, 1 slab walk, 2 / 3 chunked slab walk (2 / 4 candidates per load round).
// What a force kernel needs to also move its particles (the update of the
// next step, fused): rannew64 rings and LCGs, draw counters, the step index
// (temperature), raw positions and the binning arrays.
struct MoveArgs {
  uint64_t* ring;
  uint64_t* lcgs;
  const uint64_t* draws;
  const int64_t* step;
  int total_lanes;
  double *rx, *ry, *rz;
  int *nocup, *ins;
  RepFlags* flags;
  ListMove lm;  // lm.lflag == nullptr: no neighbour lists
};

// The fused move of particle i once its force is known (replicas with one
// particle per lane only).
__device__ inline void fused_move(const RepDev& R, int rep, int i, double x, double y, double z,
                                  double fx, double fy, double fz, double xi, double yi,
                                  double zi, double* px, double* py, double* pz,
                                  const MoveArgs& mv) {
  const int64_t ni = *mv.step;
  const double T = temperature_at(R, ni);
  const int p = R.p_off + i;
  if (!ermak_move(R, rep, i, T, x, y, z, fx, fy, fz, xi, yi, zi, mv.rx[p], mv.ry[p], mv.rz[p],
                  px, py, pz, mv.rx, mv.ry,
                  mv.rz, mv.nocup, mv.ins, mv.flags, 1, ni, mv.lm.lflag ? &mv.lm : nullptr,
                  list_parity(R, mv.draws[rep], 1)))
    raise_flag(mv.flags, rep, kFlagNonFinite, ni);
}

template <int POT, bool NEIGH, int DEPTH, int WALK, bool UPD = false>
__device__ __forceinline__ void k_forces_block(
    int vb, const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
             double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
             const int* __restrict__ rs, const double4* __restrict__ cpos,
             double* __restrict__ ffx, double* __restrict__ ffy, double* __restrict__ ffz,
             double* __restrict__ vir, double* __restrict__ pot, int* __restrict__ ncounts,
             int* __restrict__ nlist, int maxn, const MoveArgs* mv = nullptr) {
  __shared__ unsigned int dbuf[DEPTH > 0 ? DEPTH * kBlk : 1];
  __shared__ unsigned int sbuf[WALK >= 1 ? 2 * kSlabSegs * kBlk : 1];
  const int4 t = tbl[vb];
  if (t.x < 0) return;
  const int i = t.y + (int)threadIdx.x;
  if (i >= t.z) return;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + i;
  const int g = i < R.nb ? 0 : 1;
  const double x = px[p], y = py[p], z = pz[p];
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  if (NEIGH) {
    int nc = 0;
    int* nl = nlist + (size_t)i * maxn;
    scan_grid<POT, true>(R, g, R.nrange_same[g], R.sigma[g], R.rc2_same[g], i, x, y, z, rs,
                         cpos, fx, fy, fz, v, u, nl, nc, maxn);
    scan_grid<POT, true>(R, 1 - g, R.nrange_cross[g], R.sigma_bs, R.rc2_bs, -1, x, y, z, rs,
                         cpos, fx, fy, fz, v, u, nl, nc, maxn);
    ncounts[i] = nc;
  } else {
    forces_global<POT, false, DEPTH, WALK>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u, 0, 1,
                                           dbuf + threadIdx.x, sbuf + threadIdx.x);
  }
  ffx[p] = fx;
  ffy[p] = fy;
  ffz[p] = fz;
  vir[p] = v;
  pot[p] = u;
  if (UPD && !NEIGH) {
    // the walks read the packed copy (cpos), never px/py/pz, so the move can
    // overwrite this particle's position while other lanes still walk
    double xi, yi, zi;
    lane_normals(R, t.x, i, mv->ring, mv->lcgs, mv->draws, mv->total_lanes, xi, yi, zi);
    fused_move(R, t.x, i, x, y, z, fx, fy, fz, xi, yi, zi, px, py, pz, *mv);
  }
}

// Forces of the current positions; with UPD also the next step's update of
// every particle from those forces (the binning into the insertion lists
// included), which k_update_bin would otherwise do in a launch of its own.
// The fused slab walk at depth <= 16 is held to the unfused kernel's four
// wavefronts per SIMD (the move alone would take it two registers past 128);
// the 4-candidate chunked walk to three (batch 347 instead of 407 us).
template <int POT, bool NEIGH, int DEPTH, int WALK, bool UPD>
__global__ void __launch_bounds__(kBlk, (UPD && WALK == 1 && DEPTH <= 16) ? 4
                                       : (WALK == 3 && DEPTH <= 16)       ? 3
                                                                          : 1)
    k_forces(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
             double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
             const int* __restrict__ rs, const double4* __restrict__ cpos,
             double* __restrict__ ffx, double* __restrict__ ffy, double* __restrict__ ffz,
             double* __restrict__ vir, double* __restrict__ pot, int* __restrict__ ncounts,
             int* __restrict__ nlist, int maxn, MoveArgs mv) {
  k_forces_block<POT, NEIGH, DEPTH, WALK, UPD>((int)blockIdx.x, reps, tbl, px, py, pz, rs, cpos,
                                               ffx, ffy, ffz, vir, pot, ncounts, nlist, maxn,
                                               &mv);
}

// ---------------------------------------------------------------------------
// Flat slab walk (BD_WALK_FLAT): the slab walk's ranges, partners and order,
// with the candidates of a wavefront's 64 particles tested together.
//
// Per stencil slab (pass, di), every lane lists its particle's ranges (the
// slab walk's rows, gap tests and z ranges) in its LDS column, one word per
// range: slot | y image code << 22 | z image code << 24 | length << 26, the
// slot counted from the replica's first (big-grid) slot.
// A wavefront prefix sum of the lanes' candidate counts places every
// particle's candidates, in its own visiting order and after those of the
// lanes before it, in one LDS array (owner lane in the top six bits), up to
// kFlatCap at a time; the 64 lanes then test 64 candidates per trip, whoever
// owns them, so no lane idles while another still has candidates.  A partner
// inside the cutoff goes to its owner's record column (global memory, the
// deferral record of the slab walk: slot | image codes | pass bit) at the
// owner's running count plus its rank among the owner's partners in the
// trip; an owner's candidates are contiguous and in visiting order, so the
// records come out in the reference order.  A candidate's position is
// gathered from its staged entry alone, so the entries and gathers of
// kFlatBatch trips are issued together before the first is tested.  After
// the last slab every lane evaluates its particle's records in that order
// with the reference operations (the deferral drain, four records loaded at
// a time), or, when more partners were found than the column holds, walks
// its stencil again with direct evaluation.  Needs replicas of fewer than
// 2^22 slots (both grids) and ranges of at most 63 slots (ntags <= 12).
// ---------------------------------------------------------------------------
constexpr int kFlatCap = 960;               // staged candidates per round (15 trips)
constexpr int kFlatDepth = 32;              // record column per particle (maximum depth)
constexpr unsigned int kFlatSlotBits = 22;  // grid slot bits of a staged entry
constexpr unsigned int kFlatSlotMask = (1u << kFlatSlotBits) - 1u;
constexpr int kFlatMaxTags = 12;            // 5 cells x 12 slots fit the 6-bit length
constexpr int kFlatBatch = 4;               // trips whose gathers are issued together

// Element idx of an array addressed by a 32-bit byte offset from its base
// (one scalar base, one 32-bit vector offset per access); idx * sizeof(T)
// must stay below 2^32.
template <class T>
__device__ inline T& at32(T* base, unsigned int idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (idx * (unsigned int)sizeof(T)));
}
template <class T>
__device__ inline const T& at32(const T* base, unsigned int idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     (idx * (unsigned int)sizeof(T)));
}

// The ranges of one slab of the flat walk (rows dj = -NRW..NRW, those beyond
// the particle's own nr or the cutoff skipped), in the slab walk's visiting
// order, with every row's bounds loaded before any range is written.  rs is
// the engine's row-start table and rsa the first entry of row (a, 0) of this
// grid in it (32-bit offsets from one base).  A range is written at the next
// position whether or not it is empty; only a non-empty one advances nseg.
template <int NRW>
__device__ inline void flat_list(int nr, double gx2, double lim, const double* g2lo,
                                 const double* g2hi, double uy, double cs, int n, const int hw[3],
                                 int rsa, const int* __restrict__ rs, int goff,
                                 unsigned int* myseg, int& nseg, int& tot) {
  constexpr int W = 2 * NRW + 1;
  int lo[W], hi[W], lo2[W], hi2[W];
  unsigned int code[W];  // y image << 22 | z image of range b << 2 | z image of range a
#pragma unroll
  for (int r = 0; r < W; r++) {
    const int dj = r - NRW;
    const double gy = slab_gap(dj, cs, uy);
    const double gxy2 = gx2 + gy * gy;
    const int hb = hw[1] + dj;
    const int b = hb < 0 ? hb + n : (hb >= n ? hb - n : hb);
    int klo = 0, khi = 0;
#pragma unroll
    for (int k = 0; k < NRW; k++) {
      if (k < nr && gxy2 + g2lo[k] <= lim) klo = -(k + 1);
      if (k < nr && gxy2 + g2hi[k] <= lim) khi = k + 1;
    }
    const unsigned int row = (unsigned int)(rsa + b * n * (n + 1));
    const int z0 = hw[2] + klo, z1 = hw[2] + khi;
    const bool wlo = z0 < 0, whi = z1 >= n;
    lo[r] = hi[r] = lo2[r] = hi2[r] = 0;
    if (dj >= -nr && dj <= nr && gxy2 <= lim) {
      lo[r] = at32(rs, row + (unsigned int)(wlo ? z0 + n : z0));
      hi[r] = at32(rs, row + (unsigned int)((wlo || whi) ? n : z1 + 1));
      if (wlo || whi) {
        lo2[r] = at32(rs, row);
        hi2[r] = at32(rs, row + (unsigned int)(wlo ? z1 + 1 : z1 - n + 1));
      }
    }
    code[r] = ((hb < 0 ? 3u : (hb >= n ? 1u : 0u)) << 22) | (whi ? 4u : 0u) | (wlo ? 3u : 0u);
  }
#pragma unroll
  for (int r = 0; r < W; r++) {
    const int la = hi[r] - lo[r], lb = hi2[r] - lo2[r];
    const unsigned int yc = code[r] & (3u << 22);
    myseg[nseg * kBlk] = (unsigned int)(lo[r] + goff) | yc | ((code[r] & 3u) << 24) |
                         ((unsigned int)la << 26);
    nseg += la > 0 ? 1 : 0;
    myseg[nseg * kBlk] = (unsigned int)(lo2[r] + goff) | yc | (((code[r] >> 2) & 3u) << 24) |
                         ((unsigned int)lb << 26);
    nseg += lb > 0 ? 1 : 0;
    tot += la + lb;
  }
}

template <int POT, bool UPD>
__global__ void __launch_bounds__(kBlk, 4)
    k_forces_flat(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                  double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                  const int* __restrict__ rs,
                  const double4* __restrict__ cpos, double* __restrict__ ffx,
                  double* __restrict__ ffy, double* __restrict__ ffz, double* __restrict__ vir,
                  double* __restrict__ pot, unsigned int* __restrict__ recs, int rstride,
                  int depth, MoveArgs mv) {
  __shared__ unsigned int seg[kSlabSegs * kBlk];
  __shared__ unsigned int cand[kFlatCap];
  __shared__ double4 opos[kBlk];  // owner x, y, z and the slab's x shift
  __shared__ double2 orc[kBlk];   // owner cutoff^2 of the pass and self index (-1: none)
  __shared__ int2 oinf[kBlk];     // owner x code | pass bit, first flat index
  __shared__ int ond[kBlk];       // partners recorded so far
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;
  const int lane = (int)threadIdx.x;
  const int i = t.y + lane;
  const bool live = i < t.z;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + (live ? i : t.y);
  const double x = px[p], y = py[p], z = pz[p];
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  const int g = i < R.nb ? 0 : 1;
  if (!R.fast) {  // grids too narrow for the slab walks (one replica per workgroup)
    if (live) {
      forces_global<POT, false>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u);
      ffx[p] = fx;
      ffy[p] = fy;
      ffz[p] = fz;
      vir[p] = v;
      pot[p] = u;
      if (UPD) {
        double xi, yi, zi;
        lane_normals(R, t.x, i, mv.ring, mv.lcgs, mv.draws, mv.total_lanes, xi, yi, zi);
        fused_move(R, t.x, i, x, y, z, fx, fy, fz, xi, yi, zi, px, py, pz, mv);
      }
    }
    return;
  }
  const double L = R.side;
  ond[lane] = 0;
  opos[lane] = make_double4(x, y, z, 0.0);
  unsigned int* myseg = seg + lane;
  unsigned int* rbase = recs + R.p_off + t.y;  // record column of lane 0's particle
  const double4* cbase = cpos + R.tag_off[0];  // the replica's slots
  for (int pass = 0; pass < 2; pass++) {
    const int gg = pass == 0 ? g : 1 - g;
    const int nr = pass == 0 ? R.nrange_same[g] : R.nrange_cross[g];
    const double rc2 = pass == 0 ? R.rc2_same[g] : R.rc2_bs;
    const double lim = rc2 + 1e-9;
    int hw[3] = {0, 0, 0};
    const bool ok = live && cell_coords(R, gg, x, y, z, hw[0], hw[1], hw[2]);
    const int n = R.ncell[gg];
    const double cs = R.cell_side[gg];
    const int goff = R.tag_off[gg] - R.tag_off[0];  // the grid's first slot in the replica
    const double ux = x - (double)hw[0] * cs, uy = y - (double)hw[1] * cs,
                 uz = z - (double)hw[2] * cs;
    double g2lo[2], g2hi[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double lo = slab_gap(-(k + 1), cs, uz);
      const double hi = slab_gap(k + 1, cs, uz);
      g2lo[k] = lo * lo;
      g2hi[k] = hi * hi;
    }
    const int nrw = __any(ok && nr > 1) ? 2 : 1;
    const unsigned int pb = pass == 0 ? 0u : kPassBit;
    orc[lane] = make_double2(rc2, pass == 0 ? (double)i : -1.0);
    for (int di = -nrw; di <= nrw; di++) {
      // 1. this particle's ranges of the slab, in visiting order
      int nseg = 0, tot = 0;
      double shx = 0.0;
      if (ok && di >= -nr && di <= nr) {
        const double gx = slab_gap(di, cs, ux);
        const double gx2 = gx * gx;
        shx = image_shift(hw[0] + di, n, L);
        if (gx2 <= lim) {
          int a = hw[0] + di;
          a = a < 0 ? a + n : (a >= n ? a - n : a);
          if (nrw == 1)
            flat_list<1>(nr, gx2, lim, g2lo, g2hi, uy, cs, n, hw, R.rs_off[gg] + a * (n + 1), rs,
                         goff, myseg, nseg, tot);
          else
            flat_list<2>(nr, gx2, lim, g2lo, g2hi, uy, cs, n, hw, R.rs_off[gg] + a * (n + 1), rs,
                         goff, myseg, nseg, tot);
        }
      }
      // 2. the lanes' candidates one after the other: exclusive prefix sum
      const int incl = wave_prefix_sum(tot);
      const int T = __builtin_amdgcn_readlane(incl, 63);
      const int base = incl - tot;
      opos[lane].w = shx;
      oinf[lane] = make_int2((int)((image_code(shx) << kSlotBits) | pb), base);
      for (int w0 = 0; w0 < T; w0 += kFlatCap) {
        // 3. stage the candidates with flat index in [w0, w0 + kFlatCap)
        int f = base;
        for (int k = 0; k < nseg; k++) {
          const unsigned int sw = myseg[k * kBlk];
          const int len = (int)(sw >> 26);
          const int lo = max(f, w0), hi = min(f + len, w0 + kFlatCap);
          const unsigned int e0 = (sw & 0x3ffffffu) | ((unsigned int)lane << 26);
          for (int ff = lo; ff < hi; ff++) cand[ff - w0] = e0 + (unsigned int)(ff - f);
          f += len;
        }
        __syncthreads();
        // 4. 64 candidates per trip, kFlatBatch trips at a time: their entries
        //    and position gathers are all issued before the first is tested
        const int m = min(kFlatCap, T - w0);
        for (int it0 = 0; it0 < m; it0 += 64 * kFlatBatch) {
          unsigned int eb[kFlatBatch];
          double4 qb[kFlatBatch];
#pragma unroll
          for (int k = 0; k < kFlatBatch; k++) {
            const int e = it0 + 64 * k + lane;
            eb[k] = cand[e < m ? e : 0];
          }
#pragma unroll
          for (int k = 0; k < kFlatBatch; k++) qb[k] = at32(cbase, eb[k] & kFlatSlotMask);
#pragma unroll
          for (int k = 0; k < kFlatBatch; k++) {
            const int it = it0 + 64 * k;
            const unsigned int ent = eb[k];
            const double4 q = qb[k];
            const bool valid = it + lane < m;
            const int o = (int)(ent >> 26);
            const double4 op = opos[o];
            const double2 oc = orc[o];
            const int2 oi = oinf[o];
            const double dx = (q.x - op.x) + op.w;
            const double dy = (q.y - op.y) + image_at(ent, 22, L);
            const double dz = (q.z - op.z) + image_at(ent, 24, L);
            const double dist2 = dx * dx + dy * dy + dz * dz;
            const bool hit = valid && dist2 < oc.x && q.w != oc.y;
            const unsigned long long hm = __ballot(hit);
            if (hm != 0ull) {
              // partners below this lane in the trip, and below the owner's first lane
              const int mb = (int)__builtin_amdgcn_mbcnt_hi(
                  (unsigned int)(hm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)hm, 0u));
              const int fo = max(0, oi.y - w0 - it);  // <= this lane for every valid lane
              const int before = __builtin_amdgcn_ds_bpermute(fo << 2, mb);
              if (hit) {
                const int pos = ond[o] + mb - before;
                if (pos < depth)
                  at32(rbase, (unsigned int)pos * (unsigned int)rstride + (unsigned int)o) =
                      (ent & kFlatSlotMask) | ((ent << 5) & 0x78000000u) | (unsigned int)oi.x;
                atomicAdd(&ond[o], 1);
              }
            }
          }
        }
        __syncthreads();  // the next round restages cand
      }
    }
  }
  __syncthreads();  // the records of the other lanes are visible
  if (!live) return;
  const int nd = ond[lane];
  if (nd > depth) {
    // more partners than the column holds: the direct walk, same order
    forces_global<POT, false>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u);
  } else {
    const double s0 = R.sigma[g], s1 = R.sigma_bs;
    for (int e0 = 0; e0 < nd; e0 += 4) {
      unsigned int wc[4];
      double4 qc[4];
#pragma unroll
      for (int k = 0; k < 4; k++)
        wc[k] = at32(rbase, (unsigned int)(e0 + k < nd ? e0 + k : e0) * (unsigned int)rstride +
                                (unsigned int)lane);
#pragma unroll
      for (int k = 0; k < 4; k++) qc[k] = at32(cbase, wc[k] & kSlotMask);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (e0 + k < nd) {
          const double dx = (qc[k].x - x) + image_at(wc[k], kSlotBits, L);
          const double dy = (qc[k].y - y) + image_at(wc[k], kSlotBits + 2, L);
          const double dz = (qc[k].z - z) + image_at(wc[k], kSlotBits + 4, L);
          const double dist2 = dx * dx + dy * dy + dz * dz;
          pair_add<POT>(R, (wc[k] & kPassBit) ? s1 : s0, dist2, dx, dy, dz, fx, fy, fz, v, u);
        }
      }
    }
  }
  ffx[p] = fx;
  ffy[p] = fy;
  ffz[p] = fz;
  vir[p] = v;
  pot[p] = u;
  if (UPD) {
    // the next step's move: the walks read the packed copy (cpos), never
    // px/py/pz, and every lane of this workgroup read its position at the start
    double xi, yi, zi;
    lane_normals(R, t.x, i, mv.ring, mv.lcgs, mv.draws, mv.total_lanes, xi, yi, zi);
    fused_move(R, t.x, i, x, y, z, fx, fy, fz, xi, yi, zi, px, py, pz, mv);
  }
}

// One stencil pass (0: same species, 1: the other species, main.cu:576-585)
// of replica-local particle i with a slab walk (WALK 1..3), on grids of at
// least 2 nrange + 3 cells.  part selects the slabs di: -1 all, 0 = [-nrange,
// 0], 1 = [1, nrange], 2 = [lo, hi].  Returns false when the particle has no
// cell in that grid.
// REC: record the partners in df only (see scan_chunks).
template <int POT, int DEPTH, int WALK, bool REC>
__device__ inline bool walk_pass(const RepDev& R, int i, int pass, double x, double y, double z,
                                 const int* __restrict__ rs, const double4* __restrict__ cpos,
                                 Defer& df, unsigned int* seg, double& fx, double& fy,
                                 double& fz, double& v, double& u, int part = -1, int lo = 0,
                                 int hi = 0) {
  const int g = i < R.nb ? 0 : 1;
  const int gg = pass == 0 ? g : 1 - g;
  const int nr = pass == 0 ? R.nrange_same[g] : R.nrange_cross[g];
  const double sigma = pass == 0 ? R.sigma[g] : R.sigma_bs;
  const double rc2 = pass == 0 ? R.rc2_same[g] : R.rc2_bs;
  const int self = pass == 0 ? i : -1;
  int hg[3];
  if (!cell_coords(R, gg, x, y, z, hg[0], hg[1], hg[2])) return false;
  const int n = R.ncell[gg];
  const double cs = R.cell_side[gg];
  const int* rs_g = rs + R.rs_off[gg];
  const double4* cp_g = cpos + R.tag_off[gg];
  const unsigned int pb = pass == 0 ? 0u : kPassBit;
  const int di0 = part == 2 ? lo : (part == 1 ? 1 : -nr);
  const int di1 = part == 2 ? hi : (part == 0 ? 0 : nr);
  if (WALK == 1) {
    if (nr == 1)
      scan_slabs<POT, 1, DEPTH, REC>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df, pb,
                                     seg, fx, fy, fz, v, u, di0, di1);
    else
      scan_slabs<POT, 2, DEPTH, REC>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g, df, pb,
                                     seg, fx, fy, fz, v, u, di0, di1);
  } else {
    constexpr int CH = WALK == 2 ? 2 : 4;
    if (nr == 1)
      scan_chunks<POT, 1, DEPTH, CH, REC>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g,
                                          df, pb, seg, fx, fy, fz, v, u, di0, di1);
    else
      scan_chunks<POT, 2, DEPTH, CH, REC>(R, n, cs, sigma, rc2, self, x, y, z, hg, rs_g, cp_g,
                                          df, pb, seg, fx, fy, fz, v, u, di0, di1);
  }
  return true;
}

// Slabs [b, e) of a particle's visiting order (the same-species slabs
// -nr0..nr0, then the cross-species slabs -nr1..nr1) with walk_pass.  False
// when the particle has no cell in a grid it needs.
template <int POT, int DEPTH, int WALK, bool REC>
__device__ inline bool walk_slab_run(const RepDev& R, int i, int b, int e, double x, double y,
                                     double z, const int* __restrict__ rs,
                                     const double4* __restrict__ cpos, Defer& df,
                                     unsigned int* seg, double& fx, double& fy, double& fz,
                                     double& v, double& u) {
  const int g = i < R.nb ? 0 : 1;
  const int nr0 = R.nrange_same[g], nr1 = R.nrange_cross[g];
  const int s0 = 2 * nr0 + 1;
  bool ok = true;
  if (b < s0 && e > b)
    ok = walk_pass<POT, DEPTH, WALK, REC>(R, i, 0, x, y, z, rs, cpos, df, seg, fx, fy, fz, v, u,
                                          2, b - nr0, min(e, s0) - 1 - nr0);
  if (ok && e > s0)
    ok = walk_pass<POT, DEPTH, WALK, REC>(R, i, 1, x, y, z, rs, cpos, df, seg, fx, fy, fz, v, u,
                                          2, max(b, s0) - s0 - nr1, e - 1 - s0 - nr1);
  return ok;
}

// Slab counts of particle i's visiting order: same-species, all.
__device__ inline void slab_counts(const RepDev& R, int i, int& s0, int& s) {
  const int g = i < R.nb ? 0 : 1;
  s0 = 2 * R.nrange_same[g] + 1;
  s = s0 + 2 * R.nrange_cross[g] + 1;
}

// The five terms one recorded pair adds (pair_add's operations, split from
// the accumulation): force components, virial term, energy term; qc is the
// partner's slot, already loaded.
template <int POT>
__device__ inline void pair_terms_q(const RepDev& R, double x, double y, double z, const Defer& d,
                                    unsigned int wc, const double4& qc, double& qx, double& qy,
                                    double& qz, double& wv, double& e) {
  const double L = R.side;
  const double dx = (qc.x - x) + image_at(wc, kSlotBits, L);
  const double dy = (qc.y - y) + image_at(wc, kSlotBits + 2, L);
  const double dz = (qc.z - z) + image_at(wc, kSlotBits + 4, L);
  const double dist2 = dx * dx + dy * dy + dz * dz;
  const double dist_inv = sqrt(1.0 / dist2);
  double nf;
  bd::pair_energy<POT>(R, (wc & kPassBit) ? d.sigma1 : d.sigma0, dist_inv, e, nf);
  qx = nf * dist_inv * dx;
  qy = nf * dist_inv * dy;
  qz = nf * dist_inv * dz;
  wv = 0.5 * (dx * qx + dy * qy + dz * qz);
}

__device__ inline double4 record_slot(const Defer& d, unsigned int wc) {
  return ((wc & kPassBit) ? d.cp1 : d.cp0)[wc & kSlotMask];
}

// Pair terms of the records e0 .. e1-1 of d's column, added in order; the
// partner slots are loaded U at a time before they are used (one memory
// latency per U records).
template <int POT, int U = 1>
__device__ inline void drain_range(const RepDev& R, double x, double y, double z, const Defer& d,
                                   int e0, int e1, double& fx, double& fy, double& fz,
                                   double& vir, double& pot) {
  for (int e = e0; e < e1; e += U) {
    unsigned int wc[U];
    double4 qc[U];
#pragma unroll
    for (int k = 0; k < U; k++) wc[k] = d.buf[(e + k < e1 ? e + k : e) * kBlk];
#pragma unroll
    for (int k = 0; k < U; k++) qc[k] = record_slot(d, wc[k]);
#pragma unroll
    for (int k = 0; k < U; k++) {
      if (e + k < e1) {
        double qx, qy, qz, wv, en;
        pair_terms_q<POT>(R, x, y, z, d, wc[k], qc[k], qx, qy, qz, wv, en);
        pot = pot + en;
        fx += qx;
        fy += qy;
        fz += qz;
        vir -= wv;
      }
    }
  }
}

// Records whose partner slots the split kernels load together before
// evaluating them (their occupancy is set by LDS, so the registers are free).
constexpr int kDrainBatch = 4;

// Stencil passes split over two wavefronts (no reference counterpart): for
// the same 64 particles, wavefront 0 walks the first half of the slabs in
// visiting order (the same-species stencil, and for a big particle of the
// standard grids the first cross-species slab too) and evaluates its pair
// terms, while wavefront 1 walks the rest and only records its partners in
// LDS; after the barrier wavefront 0 evaluates those records after its own,
// so every pair term is added in the single-wavefront kernel's order
// (bit-identical).  A particle whose second-half partners overflow the
// DEPTH-entry column has that half walked again by wavefront 0.  Twice the wavefronts of k_forces for the same work,
// for systems too small to fill the GPU with one lane per particle.
template <int POT, int DEPTH, int WALK, bool UPD>
__global__ void __launch_bounds__(2 * kBlk, (WALK == 1 && DEPTH <= 16) ? 4 : 1)
    k_forces_pair(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                  double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                  const int* __restrict__ rs, const double4* __restrict__ cpos,
                  double* __restrict__ ffx, double* __restrict__ ffy, double* __restrict__ ffz,
                  double* __restrict__ vir, double* __restrict__ pot, MoveArgs mv) {
  static_assert(DEPTH > 0 && WALK >= 1, "the split needs a slab walk with deferral");
  __shared__ unsigned int dbuf[2 * DEPTH * kBlk];
  __shared__ unsigned int sbuf[2 * 2 * kSlabSegs * kBlk];
  __shared__ int n1s[kBlk];
  __shared__ double nrm[UPD ? 3 * kBlk : 1];
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;
  const int lane = (int)threadIdx.x % kBlk;
  const int w = (int)threadIdx.x / kBlk;
  const int i = t.y + lane;
  const bool live = i < t.z;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + i;
  const int g = i < R.nb ? 0 : 1;
  double x = 0.0, y = 0.0, z = 0.0;
  if (live) {
    x = px[p];
    y = py[p];
    z = pz[p];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  Defer df{cpos + R.tag_off[g], cpos + R.tag_off[1 - g], R.sigma[g], R.sigma_bs,
           dbuf + w * DEPTH * kBlk + lane, 0};
  unsigned int* seg = sbuf + w * 2 * kSlabSegs * kBlk + lane;
  bool ok = true;
  int s0, ns;
  slab_counts(R, i, s0, ns);
  const int half = ns / 2;  // wavefront 0: slabs [0, half) of the visiting order
  if (w == 1) {
    int n1 = -1;  // -1: wavefront 0 walks the run itself; -2: no cell in a grid
    if (live && R.fast) {
      if (walk_slab_run<POT, DEPTH, WALK, true>(R, i, half, ns, x, y, z, rs, cpos, df, seg, fx,
                                                fy, fz, v, u))
        n1 = df.nd <= DEPTH ? df.nd : -1;
      else
        n1 = -2;
    }
    n1s[lane] = n1;
    if (UPD && live) {  // the next move's normals, while wavefront 0 still walks
      lane_normals(R, t.x, i, mv.ring, mv.lcgs, mv.draws, mv.total_lanes, nrm[lane],
                   nrm[kBlk + lane], nrm[2 * kBlk + lane]);
    }
  } else if (live) {
    if (R.fast)
      ok = walk_slab_run<POT, DEPTH, WALK, false>(R, i, 0, half, x, y, z, rs, cpos, df, seg, fx,
                                                  fy, fz, v, u);
    else
      forces_global<POT, false>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u);
  }
  __syncthreads();
  if (w != 0 || !live) return;
  if (R.fast && ok) {
    // as the single-wavefront walk: a particle without a cell in the second
    // grid ends its walk there, before the final drain
    const int n1 = n1s[lane];
    if (n1 != -2) {
      drain_range<POT, kDrainBatch>(R, x, y, z, df, 0, df.nd, fx, fy, fz, v, u);
      df.nd = 0;
      if (n1 >= 0) {
        Defer d1{df.cp0, df.cp1, df.sigma0, df.sigma1, dbuf + DEPTH * kBlk + lane, n1};
        drain_range<POT, kDrainBatch>(R, x, y, z, d1, 0, n1, fx, fy, fz, v, u);
      } else if (walk_slab_run<POT, DEPTH, WALK, false>(R, i, half, ns, x, y, z, rs, cpos, df,
                                                        seg, fx, fy, fz, v, u)) {
        drain<POT>(R, x, y, z, df, fx, fy, fz, v, u);
      }
    }
  }
  ffx[p] = fx;
  ffy[p] = fy;
  ffz[p] = fz;
  vir[p] = v;
  pot[p] = u;
  if (UPD)
    fused_move(R, t.x, i, x, y, z, fx, fy, fz, nrm[lane], nrm[kBlk + lane], nrm[2 * kBlk + lane],
               px, py, pz, mv);
}

// Wavefront w (1..3) of the four-wavefront split: its share of the slabs
// after wavefront 0's same-species slabs [-nr0, 0], in visiting order (the
// same-species slabs [1, nr0], then the cross-species slabs [-nr1, nr1]),
// cut into three contiguous runs of nearly equal length (the first runs take
// the remainder).  False when the particle has no cell in a grid it needs.
template <int POT, int DEPTH, int WALK, bool REC>
__device__ inline bool walk_quad_part(const RepDev& R, int i, int w, double x, double y,
                                      double z, const int* __restrict__ rs,
                                      const double4* __restrict__ cpos, Defer& df,
                                      unsigned int* seg, double& fx, double& fy, double& fz,
                                      double& v, double& u) {
  int s0, s;
  slab_counts(R, i, s0, s);
  const int first = (s0 + 1) / 2;  // wavefront 0's slabs [-nr0, 0]
  const int rest = s - first;
  const int q = rest / 3, rem = rest % 3;
  const int b = first + (w - 1) * q + min(w - 1, rem);
  const int e = b + q + (w - 1 < rem ? 1 : 0);
  return walk_slab_run<POT, DEPTH, WALK, REC>(R, i, b, e, x, y, z, rs, cpos, df, seg, fx, fy, fz,
                                              v, u);
}

// Stencil passes split over four wavefronts (no reference counterpart): for
// the same 64 particles, wavefront 0 walks the same-species slabs [-nr0, 0]
// and wavefronts 1..3 share the remaining slabs (walk_quad_part: two each
// for a big particle of the standard grids, whose cross-species pass has five
// slabs), so the walk of one particle runs on four lanes at once.
// Wavefronts 1..3 record their partners and evaluate the terms of the first
// kTermCap of them into LDS; after the barrier wavefront 0 adds, in walking
// order, its own terms, then each other wavefront's stored terms and the
// terms of its records beyond kTermCap.  A part whose partners overflow the
// DEPTH-entry column is walked again by wavefront 0.  Bit-identical to the
// single-wavefront kernel.  Wavefront 0 takes the part with the centre slab
// of the same-species pass, whose terms it adds as it goes.
constexpr int kTermCap = 5;

template <int POT, int DEPTH, int WALK, bool UPD>
__global__ void __launch_bounds__(4 * kBlk)
    k_forces_quad(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                  double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                  const int* __restrict__ rs, const double4* __restrict__ cpos,
                  double* __restrict__ ffx, double* __restrict__ ffy, double* __restrict__ ffz,
                  double* __restrict__ vir, double* __restrict__ pot, MoveArgs mv) {
  static_assert(DEPTH >= kTermCap && WALK >= 1, "the split needs a slab walk with deferral");
  __shared__ unsigned int dbuf[4 * DEPTH * kBlk];
  __shared__ unsigned int sbuf[4 * 2 * kSlabSegs * kBlk];
  __shared__ double terms[3 * kTermCap * 5 * kBlk];
  __shared__ int nrec[3 * kBlk];
  __shared__ double nrm[UPD ? 3 * kBlk : 1];
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;
  const int lane = (int)threadIdx.x % kBlk;
  const int w = (int)threadIdx.x / kBlk;
  const int i = t.y + lane;
  const bool live = i < t.z;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + i;
  const int g = i < R.nb ? 0 : 1;
  double x = 0.0, y = 0.0, z = 0.0;
  if (live) {
    x = px[p];
    y = py[p];
    z = pz[p];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  Defer df{cpos + R.tag_off[g], cpos + R.tag_off[1 - g], R.sigma[g], R.sigma_bs,
           dbuf + w * DEPTH * kBlk + lane, 0};
  unsigned int* seg = sbuf + w * 2 * kSlabSegs * kBlk + lane;
  bool ok = true;
  if (w > 0) {
    int nw = -1;  // -1: wavefront 0 walks the part itself; -2: no cell in that grid
    if (live && R.fast) {
      if (walk_quad_part<POT, DEPTH, WALK, true>(R, i, w, x, y, z, rs, cpos, df, seg, fx, fy,
                                                 fz, v, u)) {
        nw = df.nd <= DEPTH ? df.nd : -1;
        double* tq = terms + (w - 1) * kTermCap * 5 * kBlk + lane;
        const int m = nw < kTermCap ? nw : kTermCap;
        // all partner slots loaded before the first is used
        unsigned int wc[kTermCap];
        double4 qc[kTermCap];
#pragma unroll
        for (int e = 0; e < kTermCap; e++)
          if (e < m) wc[e] = df.buf[e * kBlk];
#pragma unroll
        for (int e = 0; e < kTermCap; e++)
          if (e < m) qc[e] = record_slot(df, wc[e]);
#pragma unroll
        for (int e = 0; e < kTermCap; e++) {
          if (e < m) {
            double qx, qy, qz, wv, en;
            pair_terms_q<POT>(R, x, y, z, df, wc[e], qc[e], qx, qy, qz, wv, en);
            tq[(5 * e + 0) * kBlk] = qx;
            tq[(5 * e + 1) * kBlk] = qy;
            tq[(5 * e + 2) * kBlk] = qz;
            tq[(5 * e + 3) * kBlk] = wv;
            tq[(5 * e + 4) * kBlk] = en;
          }
        }
      } else {
        nw = -2;
      }
    }
    nrec[(w - 1) * kBlk + lane] = nw;
    if (UPD && w == 3 && live) {  // the next move's normals (a part of at most two slabs)
      lane_normals(R, t.x, i, mv.ring, mv.lcgs, mv.draws, mv.total_lanes, nrm[lane],
                   nrm[kBlk + lane], nrm[2 * kBlk + lane]);
    }
  } else if (live) {
    if (R.fast) {
      ok = walk_pass<POT, DEPTH, WALK, false>(R, i, 0, x, y, z, rs, cpos, df, seg, fx, fy, fz, v,
                                              u, 0);
      if (ok) {
        drain_range<POT, kDrainBatch>(R, x, y, z, df, 0, df.nd, fx, fy, fz, v, u);
        df.nd = 0;
      }
    } else {
      forces_global<POT, false>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u);
    }
  }
  __syncthreads();
  if (w != 0 || !live) return;
  const int n1 = nrec[lane], n2 = nrec[kBlk + lane], n3 = nrec[2 * kBlk + lane];
  // a particle without a cell in a grid (non-finite position) skips the rest
  if (R.fast && ok && n1 != -2 && n2 != -2 && n3 != -2) {
    const int nws[3] = {n1, n2, n3};
#pragma unroll
    for (int k = 1; k < 4; k++) {
      const int nw = nws[k - 1];
      if (nw >= 0) {
        const double* tq = terms + (k - 1) * kTermCap * 5 * kBlk + lane;
        const int m = nw < kTermCap ? nw : kTermCap;
        for (int e = 0; e < m; e++) {
          u = u + tq[(5 * e + 4) * kBlk];
          fx += tq[(5 * e + 0) * kBlk];
          fy += tq[(5 * e + 1) * kBlk];
          fz += tq[(5 * e + 2) * kBlk];
          v -= tq[(5 * e + 3) * kBlk];
        }
        if (nw > kTermCap) {
          const Defer dk{df.cp0, df.cp1, df.sigma0, df.sigma1, dbuf + k * DEPTH * kBlk + lane,
                         nw};
          drain_range<POT, kDrainBatch>(R, x, y, z, dk, kTermCap, nw, fx, fy, fz, v, u);
        }
      } else if (walk_quad_part<POT, DEPTH, WALK, false>(R, i, k, x, y, z, rs, cpos, df, seg,
                                                         fx, fy, fz, v, u)) {
        drain<POT>(R, x, y, z, df, fx, fy, fz, v, u);
      }
    }
  }
  ffx[p] = fx;
  ffy[p] = fy;
  ffz[p] = fz;
  vir[p] = v;
  pot[p] = u;
  if (UPD)
    fused_move(R, t.x, i, x, y, z, fx, fy, fz, nrm[lane], nrm[kBlk + lane], nrm[2 * kBlk + lane],
               px, py, pz, mv);
}

// ---------------------------------------------------------------------------
// Neighbour-list walk (BD_WALK_LIST, no reference counterpart).  Each
// particle keeps the ids of every particle of either species within its
// pass cutoff plus a skin.  A move that takes some particle of a replica
// more than half the skin from its position at the replica's last build
// raises the replica's rebuild flag (Verlet's criterion: while no particle
// has moved that far, every pair inside the cutoff is listed).  In the step
// whose binning follows that move, the replica's forces are computed by the
// cell walk while k_list_build rebuilds its lists from the same cells, on a
// second stream beside the force kernel; every other replica tests only its
// listed candidates, gathered by id from pos4, the copy of the positions
// that k_cells writes with each particle's cell in its own species' grid.
// The reference order of the pair terms
// is recovered exactly: a partner inside the cutoff gets the key (pass,
// offset of its cell from the particle's cell in the stencil's (di, dj, dk)
// order, id), its place in the cell walk (cells in (di, dj, dk) order, ids
// ascending within a cell, get_forces.cu:360-416,495-550); the partners are
// insertion-sorted by key in LDS and their terms added in that order.  The
// pass's grid is the partner's own species' grid in both passes, so pos4
// carries the partner's cell there.  For a pair
// inside the cutoff the compare-based minimum image equals the reference's
// dx - L floor(dx/L + 0.5) (the cutoff lies below L/2), so separations and
// terms are the cell walk's to the bit.  A particle whose list or sorted
// partners overflow walks its cells instead.
// ---------------------------------------------------------------------------
constexpr int kListCap = 48;   // listed candidates per particle
constexpr int kHitCap = 32;    // partners inside the cutoff sorted per particle
constexpr int kListChunk = 4;  // list entries whose positions are gathered together

struct ListArgs {
  unsigned int* nlist;      // [kListCap][n_total]: replica-local ids
  int* ncount;              // listed candidates per particle, -1: walk the cells
  double *rbx, *rby, *rbz;  // raw positions at the build
  const double *rx, *ry, *rz;
  const int* lflag;         // rebuild flags, [2 R]
  const int* list_ok;       // [R]: the grids hold the list stencil
  const PosCell* pos4;      // by particle, written by k_cells
  int n_total;
  double skin;
};

__device__ inline double min_image(double e, double L, double halfL) {
  return e > halfL ? -L : (e < -halfL ? L : 0.0);
}

// List build, kBuildLanes lanes per particle: lane `sub` walks the rows
// (di, dj) of both passes' stencils whose running index is congruent to sub,
// appending each particle of the pass's grid closer than the pass cutoff
// plus the skin to the particle's list through an LDS counter (the list
// order is free: the force kernel sorts the partners it finds).  The stencil
// reaches up to 2 cells each way on grids of at least 2 reach + 3 cells (so
// cell images are minimum images); rows and cells whose nearest point lies
// beyond the list radius are skipped, and a row's candidates are loaded
// kListChunk at a time.
constexpr int kBuildLanes = 8;
constexpr int kBuildThreads = kBlk * kBuildLanes;

struct BuildPass {
  int h[3];        // home cell
  int n, nr;       // cells per side, stencil reach
  double cs, rl2;  // cell side, squared list radius
  double ux, uy, uz;
  double g2z[5];   // squared z gaps of dk = -2 .. 2
  double selfw;    // replica-local index skipped (-1: none)
  const int* rs;
  const double4* cp;
};

__device__ inline bool build_pass(const RepDev& R, int g, int pass, int i, double x, double y,
                                  double z, double skin, const int* __restrict__ rs,
                                  const double4* __restrict__ cpos, BuildPass& b) {
  const int gg = pass == 0 ? g : 1 - g;
  const double rl = sqrt(pass == 0 ? R.rc2_same[g] : R.rc2_bs) + skin;
  // a margin over (rc + skin)^2: a pair exactly that far at the build stays
  // listed whatever the rounding
  b.rl2 = rl * rl * (1.0 + 1e-12);
  b.cs = R.cell_side[gg];
  b.n = R.ncell[gg];
  b.nr = (int)ceil(rl / b.cs);
  if (!cell_coords(R, gg, x, y, z, b.h[0], b.h[1], b.h[2])) return false;
  b.ux = x - (double)b.h[0] * b.cs;
  b.uy = y - (double)b.h[1] * b.cs;
  b.uz = z - (double)b.h[2] * b.cs;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double gz = slab_gap(k - 2, b.cs, b.uz);
    b.g2z[k] = gz * gz;
  }
  b.selfw = pass == 0 ? (double)i : -1.0;
  b.rs = rs + R.rs_off[gg];
  b.cp = cpos + R.tag_off[gg];
  return true;
}

// Row r (0 .. (2 nr + 1)^2 - 1, di major) of a pass: its candidates within
// the list radius appended to the list column `out` (stride) through *ctr.
__device__ inline void build_row(const RepDev& R, const BuildPass& b, int r, double x, double y,
                                 double z, unsigned int* __restrict__ out, size_t stride,
                                 int* ctr) {
  const double L = R.side;
  const double lim = b.rl2 + 1e-9;
  const int w = 2 * b.nr + 1;
  const int di = r / w - b.nr, dj = r % w - b.nr;
  const double gx = slab_gap(di, b.cs, b.ux), gy = slab_gap(dj, b.cs, b.uy);
  const double gxy2 = gx * gx + gy * gy;
  if (gxy2 > lim) return;
  const int n = b.n;
  const double shx = image_shift(b.h[0] + di, n, L), shy = image_shift(b.h[1] + dj, n, L);
  int a = b.h[0] + di, c = b.h[1] + dj;
  a = a < 0 ? a + n : (a >= n ? a - n : a);
  c = c < 0 ? c + n : (c >= n ? c - n : c);
  int klo = 0, khi = 0;
  for (int k = 1; k <= b.nr; k++) {
    if (gxy2 + b.g2z[2 - k] <= lim) klo = -k;
    if (gxy2 + b.g2z[2 + k] <= lim) khi = k;
  }
  const int* rrow = b.rs + (a + n * c) * (n + 1);
  const int z0 = b.h[2] + klo, z1 = b.h[2] + khi;
  const bool wlo = z0 < 0, whi = z1 >= n;
  int lo[2], hi[2];
  double shz[2];
  lo[0] = rrow[wlo ? z0 + n : z0];
  hi[0] = rrow[(wlo || whi) ? n : z1 + 1];
  shz[0] = wlo ? -L : 0.0;
  lo[1] = hi[1] = 0;
  shz[1] = whi ? L : 0.0;
  if (wlo || whi) {
    lo[1] = rrow[0];
    hi[1] = rrow[wlo ? z1 + 1 : z1 - n + 1];
  }
  for (int part = 0; part < 2; part++) {
    for (int s0 = lo[part]; s0 < hi[part]; s0 += kListChunk) {
      double4 q[kListChunk];
#pragma unroll
      for (int u = 0; u < kListChunk; u++) q[u] = b.cp[s0 + u < hi[part] ? s0 + u : s0];
#pragma unroll
      for (int u = 0; u < kListChunk; u++) {
        const double dx = (q[u].x - x) + shx;
        const double dy = (q[u].y - y) + shy;
        const double dz = (q[u].z - z) + shz[part];
        if (s0 + u < hi[part] && dx * dx + dy * dy + dz * dz < b.rl2 && q[u].w != b.selfw) {
          const int k = atomicAdd(ctr, 1);
          if (k < kListCap) out[(size_t)k * stride] = (unsigned int)q[u].w;
        }
      }
    }
  }
}

// Rebuilds the lists of the replicas whose rebuild flag is raised.  The
// grid strides over the particle block table: each thread checks one entry,
// the flagged entries are compacted in LDS, then the workgroup builds them
// one after the other (kBuildLanes lanes per particle) and records the build
// positions.  A step without rebuilds costs the checks only.
__global__ void __launch_bounds__(kBuildThreads)
    k_list_build(const RepDev* __restrict__ reps, const int4* __restrict__ tbl, int ntbl,
                 const double* __restrict__ px, const double* __restrict__ py,
                 const double* __restrict__ pz, const int* __restrict__ rs,
                 const double4* __restrict__ cpos, const uint64_t* __restrict__ draws,
                 ListArgs la) {
  __shared__ int todo[kBuildThreads];
  __shared__ int ntodo;
  __shared__ int lcnt[kBlk];
  if (threadIdx.x == 0) ntodo = 0;
  __syncthreads();
  {
    const int e = (int)blockIdx.x + (int)threadIdx.x * (int)gridDim.x;
    if (e < ntbl) {
      const int4 t = tbl[e];
      if (t.x >= 0 && la.list_ok[t.x] &&
          la.lflag[2 * t.x + list_parity(reps[t.x], draws[t.x], 0)])
        todo[atomicAdd(&ntodo, 1)] = e;
    }
  }
  __syncthreads();
  const int nt = ntodo;
  const int pl = (int)threadIdx.x / kBuildLanes, sub = (int)threadIdx.x % kBuildLanes;
  for (int k = 0; k < nt; k++) {
    const int4 t = tbl[todo[k]];
    const RepDev& R = reps[t.x];
    if (threadIdx.x < kBlk) lcnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = t.y + pl;
    const bool live = i < t.z;
    const int p = R.p_off + i;
    if (live) {
      const int g = i < R.nb ? 0 : 1;
      const double x = px[p], y = py[p], z = pz[p];
      BuildPass b0, b1;
      if (build_pass(R, g, 0, i, x, y, z, la.skin, rs, cpos, b0) &&
          build_pass(R, g, 1, i, x, y, z, la.skin, rs, cpos, b1)) {
        const int w0 = (2 * b0.nr + 1) * (2 * b0.nr + 1);
        const int w1 = (2 * b1.nr + 1) * (2 * b1.nr + 1);
        for (int r = sub; r < w0 + w1; r += kBuildLanes) {
          if (r < w0)
            build_row(R, b0, r, x, y, z, la.nlist + p, (size_t)la.n_total, &lcnt[pl]);
          else
            build_row(R, b1, r - w0, x, y, z, la.nlist + p, (size_t)la.n_total, &lcnt[pl]);
        }
      } else if (sub == 0) {
        atomicAdd(&lcnt[pl], kListCap + 1);  // no cell: the particle walks its cells
      }
    }
    __syncthreads();
    if (live && sub == 0) {
      const int c = lcnt[pl];
      la.ncount[p] = c <= kListCap ? c : -1;
      la.rbx[p] = la.rx[p];
      la.rby[p] = la.ry[p];
      la.rbz[p] = la.rz[p];
    }
    __syncthreads();
  }
}

// Forces of every particle from its list; replicas whose lists are being
// rebuilt in this step walk their cells (the row walk, no LDS).  The listed
// candidates' positions are gathered kListChunk at a time, the partners
// inside the cutoff are insertion-sorted by their cell-walk key, then their
// terms are added in key order (pair_add's operations on the same
// separations), the partners' positions again gathered kListChunk at a time.
template <int POT>
__global__ void __launch_bounds__(kBlk)
    k_forces_list(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                  const double* __restrict__ px, const double* __restrict__ py,
                  const double* __restrict__ pz, const int* __restrict__ rs,
                  const double4* __restrict__ cpos, double* __restrict__ ffx,
                  double* __restrict__ ffy, double* __restrict__ ffz, double* __restrict__ vir,
                  double* __restrict__ pot, const uint64_t* __restrict__ draws, ListArgs la) {
  __shared__ unsigned int keys[kHitCap * kBlk];
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;
  const int i = t.y + (int)threadIdx.x;
  if (i >= t.z) return;
  const RepDev& R = reps[t.x];
  const int p = R.p_off + i;
  const int g = i < R.nb ? 0 : 1;
  const double x = px[p], y = py[p], z = pz[p];
  const double L = R.side, halfL = 0.5 * R.side;
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  bool walk = !R.fast || !la.list_ok[t.x] ||
              la.lflag[2 * t.x + list_parity(R, draws[t.x], 0)] != 0;
  if (!walk) {
    const int cnt = la.ncount[p];
    int ho[2][3];  // the particle's cell in its own grid (pass 0) and the other (pass 1)
    walk = cnt < 0 || !cell_coords(R, g, x, y, z, ho[0][0], ho[0][1], ho[0][2]) ||
           !cell_coords(R, 1 - g, x, y, z, ho[1][0], ho[1][1], ho[1][2]);
    const int nr0 = R.nrange_same[g], nr1 = R.nrange_cross[g];
    const double rc2s = R.rc2_same[g], rc2x = R.rc2_bs;
    const PosCell* pc = la.pos4 + R.p_off;
    unsigned int* kc = keys + threadIdx.x;
    int nh = 0;
    for (int c0 = 0; c0 < cnt && !walk; c0 += kListChunk) {
      unsigned int id[kListChunk];
#pragma unroll
      for (int k = 0; k < kListChunk; k++)
        id[k] = la.nlist[(size_t)(c0 + k < cnt ? c0 + k : c0) * la.n_total + p];
      PosCell q[kListChunk];
#pragma unroll
      for (int k = 0; k < kListChunk; k++) q[k] = pc[id[k]];
#pragma unroll
      for (int k = 0; k < kListChunk; k++) {
        if (c0 + k >= cnt) break;
        const bool cross = ((int)id[k] < R.nb) != (g == 0);
        const double ex = q[k].x - x, ey = q[k].y - y, ez = q[k].z - z;
        const double dx = ex + min_image(ex, L, halfL);
        const double dy = ey + min_image(ey, L, halfL);
        const double dz = ez + min_image(ez, L, halfL);
        if (dx * dx + dy * dy + dz * dz < (cross ? rc2x : rc2s)) {
          // the partner's cell in its own grid, the pass's grid
          const int gj = cross ? 1 - g : g;
          const unsigned int cw = q[k].cell;
          const int cj[3] = {(int)(cw & 1023u), (int)((cw >> 10) & 1023u),
                             (int)((cw >> 20) & 1023u)};
          const int nr = cross ? nr1 : nr0;
          const int n = R.ncell[gj];
          const int* h = ho[cross ? 1 : 0];
          int d[3];
#pragma unroll
          for (int a = 0; a < 3; a++) {
            const int o = cj[a] - h[a];
            d[a] = (o > nr ? o - n : (o < -nr ? o + n : o)) + nr;
          }
          const int w = 2 * nr + 1;
          const unsigned int key = (cross ? kPassBit : 0u) |
                                   ((unsigned int)((d[0] * w + d[1]) * w + d[2]) << 24) | id[k];
          if (nh == kHitCap) {
            walk = true;
            break;
          }
          int pos = nh;
          while (pos > 0 && kc[(pos - 1) * kBlk] > key) {
            kc[pos * kBlk] = kc[(pos - 1) * kBlk];
            pos--;
          }
          kc[pos * kBlk] = key;
          nh++;
        }
      }
    }
    if (!walk) {
      for (int e0 = 0; e0 < nh; e0 += kListChunk) {
        unsigned int key[kListChunk];
        PosCell q[kListChunk];
#pragma unroll
        for (int k = 0; k < kListChunk; k++) key[k] = kc[(e0 + k < nh ? e0 + k : e0) * kBlk];
#pragma unroll
        for (int k = 0; k < kListChunk; k++) q[k] = pc[key[k] & 0xffffffu];
#pragma unroll
        for (int k = 0; k < kListChunk; k++) {
          if (e0 + k >= nh) break;
          const double ex = q[k].x - x, ey = q[k].y - y, ez = q[k].z - z;
          const double dx = ex + min_image(ex, L, halfL);
          const double dy = ey + min_image(ey, L, halfL);
          const double dz = ez + min_image(ez, L, halfL);
          const double dist2 = dx * dx + dy * dy + dz * dz;
          pair_add<POT>(R, (key[k] & kPassBit) ? R.sigma_bs : R.sigma[g], dist2, dx, dy, dz, fx,
                        fy, fz, v, u);
        }
      }
    }
  }
  if (walk) {
    fx = fy = fz = v = u = 0.0;
    forces_global<POT, false>(R, i, x, y, z, rs, cpos, fx, fy, fz, v, u);
  }
  ffx[p] = fx;
  ffy[p] = fy;
  ffz[p] = fz;
  vir[p] = v;
  pot[p] = u;
}

// ---------------------------------------------------------------------------
// Persistent step kernel: nsteps whole steps (update -> cells -> forces) in
// one cooperative launch.  Every workgroup walks the three block tables with
// stride gridDim.x, and grid-wide barriers separate the phases, so the state
// after each barrier is exactly the state after the corresponding separate
// kernel.  The forces use the slab walk with deferral depth 16 (every
// force-kernel shape gives bit-identical results).
// ---------------------------------------------------------------------------
struct StepArgs {
  const RepDev* reps;
  const int4 *tbl_lane, *tbl_cell, *tbl_part;
  int nb_lane, nb_cell, nb_part, total_lanes;
  double *px, *py, *pz, *rx, *ry, *rz, *fx, *fy, *fz, *vir, *pot;
  uint64_t *ring, *lcg, *draws;
  int64_t* step;
  int *nocup, *cnt, *ins, *rs;
  RepFlags* flags;
  double4* cpos;
};

template <int POT>
__global__ void __launch_bounds__(kBlk) k_persistent_steps(StepArgs a, int nsteps) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (int st = 0; st < nsteps; st++) {
    for (int vb = (int)blockIdx.x; vb < a.nb_lane; vb += (int)gridDim.x)
      k_update_bin_block(vb, a.reps, a.tbl_lane, a.px, a.py, a.pz, a.rx, a.ry, a.rz, a.fx, a.fy,
                         a.fz, a.ring, a.lcg, a.draws, a.step, a.total_lanes, a.nocup, a.ins,
                         a.flags, 1);
    grid.sync();
    for (int vb = (int)blockIdx.x; vb < a.nb_cell; vb += (int)gridDim.x)
      k_cells_block(vb, a.reps, a.tbl_cell, a.px, a.py, a.pz, a.nocup, a.cnt, a.ins, a.rs,
                    a.cpos, a.flags, a.step, a.draws, 1);
    grid.sync();
    for (int vb = (int)blockIdx.x; vb < a.nb_part; vb += (int)gridDim.x)
      k_forces_block<POT, false, 16, 1>(vb, a.reps, a.tbl_part, a.px, a.py, a.pz, a.rs, a.cpos,
                                        a.fx, a.fy, a.fz, a.vir, a.pot, nullptr, nullptr, 0);
    grid.sync();
  }
}

// FAST force order: G consecutive lanes share one particle, each walking
// every G-th stencil cell (same-species cells first, then cross-species, in
// the reference's cell order), and the G partial sums are combined by a
// fixed xor butterfly.  Deterministic and identical for every replica and
// launch, but the additions are grouped differently from the reference's
// single running sum, so results agree with it to rounding (<= ~1e-15
// relative), not bit for bit.  Needs R.fast for every replica.
template <int POT, int G>
__global__ void __launch_bounds__(kBlk)
    k_forces_split(const RepDev* __restrict__ reps, const int4* __restrict__ tbl,
                   const double* __restrict__ px, const double* __restrict__ py,
                   const double* __restrict__ pz, const int* __restrict__ rs,
                   const double4* __restrict__ cpos, double* __restrict__ ffx,
                   double* __restrict__ ffy, double* __restrict__ ffz, double* __restrict__ vir,
                   double* __restrict__ pot) {
  const int4 t = tbl[blockIdx.x];
  if (t.x < 0) return;  // whole-block padding entry (uniform)
  const int i = t.y + (int)threadIdx.x / G;
  const int sub = (int)threadIdx.x % G;
  const bool live = i < t.z;
  const RepDev& R = reps[t.x];
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  if (live) {
    const int p = R.p_off + i;
    forces_global<POT, true>(R, i, px[p], py[p], pz[p], rs, cpos, fx, fy, fz, v, u, sub, G);
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, G);
    fy += __shfl_xor(fy, off, G);
    fz += __shfl_xor(fz, off, G);
    v += __shfl_xor(v, off, G);
    u += __shfl_xor(u, off, G);
  }
  if (live && sub == 0) {
    const int p = R.p_off + i;
    ffx[p] = fx;
    ffy[p] = fy;
    ffz[p] = fz;
    vir[p] = v;
    pot[p] = u;
  }
}

// ---------------------------------------------------------------------------
// LDS-staged brick force kernel.
//
// A workgroup owns a brick of B x B x B cells of the big grid.  Its home
// particles are the big particles binned in those cells and the small
// particles whose big-grid cell (cell_coords on grid 0) lies in the brick.
// The workgroup stages into LDS every cell either species can reach from
// the brick: the big cells of the brick widened by max(nrange_bb, nrange_sb)
// and the small cells a home particle can occupy widened by
// max(nrange_ss, nrange_bs).  Cells are addressed by unwrapped coordinates
// relative to the box and staged as z runs (each a contiguous slot range of
// a packed row, or two where it wraps).  Home
// particles are listed in cell order, then each thread walks one particle's
// two stencils from LDS (scan_fast), so the forces are bit-identical to the
// global-memory kernels.  A brick that would stage more than CAP positions
// walks its particles from global memory instead.
// ---------------------------------------------------------------------------
constexpr int kBrickThreads = 128;
constexpr int kBrickMaxCells = 2048;
constexpr int kBrickMaxRows = 512;

struct BrickBox {
  int lo[3];
  int dim[3];
};

__device__ inline int wrap_cell(int u, int n) {
  int w = u % n;
  return w < 0 ? w + n : w;
}

// exclusive scan of vals[0..n) in place (n <= 4 * kBrickThreads); returns the total
__device__ inline int block_scan(int* vals, int n, int* part) {
  const int tid = (int)threadIdx.x;
  const int chunk = (n + kBrickThreads - 1) / kBrickThreads;
  const int c0 = min(n, tid * chunk), c1 = min(n, c0 + chunk);
  int mysum = 0;
  for (int c = c0; c < c1; c++) mysum += vals[c];
  part[tid] = mysum;
  __syncthreads();
  for (int off = 1; off < kBrickThreads; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  const int total = part[kBrickThreads - 1];
  int run = part[tid] - mysum;
  for (int c = c0; c < c1; c++) {
    const int v = vals[c];
    vals[c] = run;
    run += v;
  }
  __syncthreads();
  return total;
}

template <int POT, int CAP>
__global__ void __launch_bounds__(kBrickThreads)
    k_forces_brick(const RepDev* __restrict__ reps, const int4* __restrict__ btbl, int B,
                   const double* __restrict__ px, const double* __restrict__ py,
                   const double* __restrict__ pz, const int* __restrict__ rs,
                   const double4* __restrict__ cpos,
                   double* __restrict__ ffx, double* __restrict__ ffy, double* __restrict__ ffz,
                   double* __restrict__ vir, double* __restrict__ pot, RepFlags* __restrict__ flags) {
  __shared__ unsigned int tbl[kBrickMaxCells];
  __shared__ double4 spos[CAP];
  __shared__ int home[CAP];
  __shared__ int rows[kBrickMaxRows];
  __shared__ int part[kBrickThreads];
  const int tid = (int)threadIdx.x;
  const int4 t = btbl[blockIdx.x];
  if (t.x < 0) return;  // padding entry of the XCD-grouped table
  const RepDev& R = reps[t.x];
  const int b0[3] = {t.y, t.z, t.w};
  const int rb = max(R.nrange_same[0], R.nrange_cross[1]);
  const int rsm = max(R.nrange_same[1], R.nrange_cross[0]);
  const int nbg = R.ncell[0], nsg = R.ncell[1];
  BrickBox bb, sb;
  int bext[3], shl[3], shh[3];
  for (int a = 0; a < 3; a++) {
    bext[a] = min(B, nbg - b0[a]);
    bb.lo[a] = b0[a] - rb;
    bb.dim[a] = bext[a] + 2 * rb;
    // small-grid coordinates a home particle can have (1e-9 covers rounding)
    const int l = (int)floor((double)b0[a] * R.cell_side[0] * R.cell_inv[1] - 1e-9);
    const int h = (int)floor((double)(b0[a] + bext[a]) * R.cell_side[0] * R.cell_inv[1] + 1e-9);
    shl[a] = l;
    shh[a] = min(h, l + nsg - 1);
    sb.lo[a] = l - rsm;
    sb.dim[a] = h - l + 2 * rsm + 1;
  }
  const int nbig = bb.dim[0] * bb.dim[1] * bb.dim[2];
  const int ncells = nbig + sb.dim[0] * sb.dim[1] * sb.dim[2];
  const int nbr = bb.dim[0] * bb.dim[1];
  const int nrows = nbr + sb.dim[0] * sb.dim[1];
  if (ncells > kBrickMaxCells || nrows > kBrickMaxRows) {  // excluded on the host
    if (tid == 0) atomicOr(&flags[t.x].bits, kFlagBrick);
    return;
  }
  // 1. occupancy of every staged cell, one z-run (packed row) per thread
  for (int r = tid; r < nrows; r += kBrickThreads) {
    const int g = r < nbr ? 0 : 1;
    const BrickBox& bx = g == 0 ? bb : sb;
    const int n = g == 0 ? nbg : nsg;
    const int rr = g == 0 ? r : r - nbr;
    const int rx = rr % bx.dim[0], ry = rr / bx.dim[0];
    const int* rrow = rs + (g == 0 ? R.rs_off[0] : R.rs_off[1]) +
                      (wrap_cell(bx.lo[0] + rx, n) + n * wrap_cell(bx.lo[1] + ry, n)) * (n + 1);
    const int cbase = (g == 0 ? 0 : nbig) + rx + bx.dim[0] * ry;
    const int cstride = bx.dim[0] * bx.dim[1];
    int wz = wrap_cell(bx.lo[2], n), sum = 0;
    for (int rz = 0; rz < bx.dim[2]; rz++) {
      const int occ = rrow[wz + 1] - rrow[wz];
      tbl[cbase + cstride * rz] = (unsigned int)occ;
      sum += occ;
      wz = wz + 1 == n ? 0 : wz + 1;
    }
    rows[r] = sum;
  }
  __syncthreads();
  // 2. row offsets into the staged position array
  const int total = block_scan(rows, nrows, part);
  const bool staged = total <= CAP;
  // 3. cell offsets and the gather of positions (run by run)
  for (int r = tid; r < nrows; r += kBrickThreads) {
    const int g = r < nbr ? 0 : 1;
    const BrickBox& bx = g == 0 ? bb : sb;
    const int n = g == 0 ? nbg : nsg;
    const int rr = g == 0 ? r : r - nbr;
    const int rx = rr % bx.dim[0], ry = rr / bx.dim[0];
    const int* rrow = rs + (g == 0 ? R.rs_off[0] : R.rs_off[1]) +
                      (wrap_cell(bx.lo[0] + rx, n) + n * wrap_cell(bx.lo[1] + ry, n)) * (n + 1);
    const double4* src = cpos + (g == 0 ? R.tag_off[0] : R.tag_off[1]);
    const int cbase = (g == 0 ? 0 : nbig) + rx + bx.dim[0] * ry;
    const int cstride = bx.dim[0] * bx.dim[1];
    int wz = wrap_cell(bx.lo[2], n), start = rows[r];
    for (int rz = 0; rz < bx.dim[2]; rz++) {
      const int idx = cbase + cstride * rz;
      const int occ = (int)tbl[idx];
      tbl[idx] = ((unsigned int)start << 6) | (unsigned int)occ;
      if (staged) {
        const double4* cs = src + rrow[wz];
        for (int s = 0; s < occ; s++) spos[start + s] = cs[s];
      }
      start += occ;
      wz = wz + 1 == n ? 0 : wz + 1;
    }
  }
  __syncthreads();
  const LdsCells lb{tbl, spos, bb.lo[0], bb.lo[1], bb.lo[2], bb.dim[0], bb.dim[1]};
  const LdsCells ls{tbl + nbig, spos, sb.lo[0], sb.lo[1], sb.lo[2], sb.dim[0], sb.dim[1]};
  // 4. home particles in cell order: big rows of the brick, then small rows
  //    holding particles whose big cell is in the brick (filtered).  The
  //    list holds LDS slots (staged) or replica-local indices (not staged).
  const GlobalCells gb{rs + R.rs_off[0], cpos + R.tag_off[0], nbg};
  const GlobalCells gs{rs + R.rs_off[1], cpos + R.tag_off[1], nsg};
  const int hb = bext[1] * bext[2];
  const int sh1 = shh[1] - shl[1] + 1, sh2 = shh[2] - shl[2] + 1;
  const int nhr = hb + sh1 * sh2;
#pragma unroll 1
  for (int pass = 0; pass < 2; pass++) {  // 0: count, 1: write
#pragma unroll 1
    for (int r = tid; r < nhr; r += kBrickThreads) {
      int at = pass == 1 ? rows[r] : 0;
      const bool big = r < hb;
      int ry, rz, x0, x1;
      if (big) {
        ry = b0[1] + r % bext[1];
        rz = b0[2] + r / bext[1];
        x0 = b0[0];
        x1 = b0[0] + bext[0] - 1;
      } else {
        ry = shl[1] + (r - hb) % sh1;
        rz = shl[2] + (r - hb) / sh1;
        x0 = shl[0];
        x1 = shh[0];
      }
#pragma unroll 1
      for (int cx = x0; cx <= x1; cx++) {
        int occ, st = 0;
        const double4* cell;
        if (staged) {
          const LdsCells& lc = big ? lb : ls;
          const unsigned int e = lc.tbl[lc.index(cx, ry, rz)];
          occ = (int)(e & 63u);
          st = (int)(e >> 6);
          cell = spos + st;
        } else {
          const GlobalCells& gc = big ? gb : gs;
          gc.cell(gc.index(cx, ry, rz), occ, cell);
        }
#pragma unroll 1
        for (int s = 0; s < occ; s++) {
          const double4 q = cell[s];
          if (!big) {
            int ii, jj, kk;
            if (!cell_coords(R, 0, q.x, q.y, q.z, ii, jj, kk)) continue;
            if (ii < b0[0] || ii >= b0[0] + bext[0] || jj < b0[1] || jj >= b0[1] + bext[1] ||
                kk < b0[2] || kk >= b0[2] + bext[2])
              continue;
          }
          if (pass == 1 && at < CAP) home[at] = staged ? st + s : (int)q.w;
          at++;
        }
      }
      if (pass == 0) rows[r] = at;
    }
    __syncthreads();
    if (pass == 0) {
      const int nh = block_scan(rows, nhr, part);
      if (nh > CAP) {  // cannot happen below 32 particles per cell
        if (tid == 0) atomicOr(&flags[t.x].bits, kFlagBrick);
        return;
      }
      if (tid == 0) part[0] = nh;  // keep the count (part is free after the scan)
      __syncthreads();
    }
  }
  const int nhome = part[0];
  // 5. forces of the home particles in the reference order
#pragma unroll 1
  for (int h = tid; h < nhome; h += kBrickThreads) {
    double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
    int i;
    if (staged) {
      const double4 me = spos[home[h]];
      i = (int)me.w;
      const int g = i < R.nb ? 0 : 1;
      int hwb[3], hws[3], hus[3];
      cell_coords(R, 0, me.x, me.y, me.z, hwb[0], hwb[1], hwb[2]);
      cell_coords(R, 1, me.x, me.y, me.z, hws[0], hws[1], hws[2]);
      for (int a = 0; a < 3; a++) {
        int uu = hws[a];
        if (uu < shl[a]) uu += nsg;
        if (uu >= shl[a] + nsg) uu -= nsg;
        hus[a] = uu;
      }
#pragma unroll 1
      for (int pass = 0; pass < 2; pass++) {
        const int gg = pass == 0 ? g : 1 - g;
        const int nr = pass == 0 ? (g == 0 ? R.nrange_same[0] : R.nrange_same[1])
                                 : (g == 0 ? R.nrange_cross[0] : R.nrange_cross[1]);
        const double sigma = pass == 0 ? (g == 0 ? R.sigma[0] : R.sigma[1]) : R.sigma_bs;
        const double rc2 = pass == 0 ? (g == 0 ? R.rc2_same[0] : R.rc2_same[1]) : R.rc2_bs;
        const LdsCells acc = gg == 0 ? lb : ls;
        const int hw[3] = {gg == 0 ? hwb[0] : hws[0], gg == 0 ? hwb[1] : hws[1],
                           gg == 0 ? hwb[2] : hws[2]};
        const int hu[3] = {gg == 0 ? hwb[0] : hus[0], gg == 0 ? hwb[1] : hus[1],
                           gg == 0 ? hwb[2] : hus[2]};
        scan_fast_nr<POT, false>(nr, R, gg == 0 ? nbg : nsg,
                                 gg == 0 ? R.cell_side[0] : R.cell_side[1], sigma, rc2,
                                 pass == 0 ? i : -1, me.x, me.y, me.z, hw, hu, acc, fx, fy, fz, v,
                                 u, 0, 1, 0);
      }
    } else {
      i = home[h];
      const int p = R.p_off + i;
      forces_global<POT, false>(R, i, px[p], py[p], pz[p], rs, cpos, fx, fy, fz, v, u);
    }
    const int p = R.p_off + i;
    ffx[p] = fx;
    ffy[p] = fy;
    ffz[p] = fz;
    vir[p] = v;
    pot[p] = u;
  }
}

// Per-replica reductions for fself_isotropic (fself.cu:3-21), the MSD
// (main.cu:776-797), the energies (main.cu:627-638) and the virial sum
// (main.cu:801-806).  One workgroup per replica, fixed-order tree.
constexpr int kObsThreads = 1024;
__global__ void __launch_bounds__(kObsThreads)
    k_observables(const RepDev* __restrict__ reps, const int* __restrict__ which,
                  const double* __restrict__ rx, const double* __restrict__ ry,
                  const double* __restrict__ rz, const double* __restrict__ ix,
                  const double* __restrict__ iy, const double* __restrict__ iz,
                  const double* __restrict__ vir, const double* __restrict__ pot, double q,
                  double* __restrict__ out) {
  const int rep = which[blockIdx.x];
  const RepDev& R = reps[rep];
  double acc[6] = {0, 0, 0, 0, 0, 0};  // fs, msd_b, msd_s, e_b, e_s, vir
  for (int i = (int)threadIdx.x; i < R.n; i += kObsThreads) {
    const int p = R.p_off + i;
    const double dx = rx[p] - ix[p], dy = ry[p] - iy[p], dz = rz[p] - iz[p];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (i < R.nb) {
      const double a = q * sqrt(d2);
      acc[0] += a == 0.0 ? 1.0 : sin(a) / a;
      acc[1] += d2;
      acc[3] += 0.5 * pot[p];
    } else {
      acc[2] += d2;
      acc[4] += 0.5 * pot[p];
    }
    acc[5] += vir[p];
  }
  __shared__ double red[6][kObsThreads];
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = kObsThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) out[blockIdx.x * 6 + k] = red[k][0];
}

// F_s(q) and MSD of every replica against every stored waiting-time origin
// (raw positions saved by bd_set_origin): workgroup (replica, slot),
// fixed-order tree reduction.
constexpr int kOriginThreads = 256;
__global__ void __launch_bounds__(kOriginThreads)
    k_fself_origins(const RepDev* __restrict__ reps, const double* __restrict__ rx,
                    const double* __restrict__ ry, const double* __restrict__ rz,
                    const double* __restrict__ org, size_t slot_stride, int n_total, double q,
                    double* __restrict__ out) {
  const int rep = (int)blockIdx.x, slot = (int)blockIdx.y;
  const RepDev& R = reps[rep];
  const double* ox = org + slot * slot_stride;
  const double* oy = ox + n_total;
  const double* oz = oy + n_total;
  double fs = 0.0, mb = 0.0, ms = 0.0;
  for (int i = (int)threadIdx.x; i < R.n; i += kOriginThreads) {
    const int p = R.p_off + i;
    const double dx = rx[p] - ox[p], dy = ry[p] - oy[p], dz = rz[p] - oz[p];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (i < R.nb) {
      const double a = q * sqrt(d2);
      fs += a == 0.0 ? 1.0 : sin(a) / a;
      mb += d2;
    } else {
      ms += d2;
    }
  }
  __shared__ double red[3][kOriginThreads];
  red[0][threadIdx.x] = fs;
  red[1][threadIdx.x] = mb;
  red[2][threadIdx.x] = ms;
  __syncthreads();
  for (int w = kOriginThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const size_t o = ((size_t)rep * gridDim.y + slot) * 3;
    out[o] = red[0][0] / (double)R.nb;
    out[o + 1] = red[1][0] / (double)R.nb;
    out[o + 2] = red[2][0] / (double)(R.n - R.nb);
  }
}

// Ordered-pair distance histograms of one replica (get_gder.cu:69-127):
// pair (i, j), i != j, histogram 2*species(i) + species(j).  j tiles are
// staged in LDS, counts go to wavefront-private LDS histograms (integer
// atomics), then one global 64-bit atomic per bin per workgroup.
constexpr int kGderThreads = 256;
constexpr int kGderMaxBins = 512;

// Minimum image of one axis (get_gder.cu:97-99): d - L floor(d / L + 0.5).
// With U the image index K is the same for every pair of the wavefront and
// tile, and sk = L K (the same product) is subtracted directly.
template <bool U>
__device__ inline double gder_image(double d, double side, double side_inv, double sk) {
  if (U) return d - sk;
  return d - side * floor(side_inv * d + 0.5);
}

template <int MASK>
__device__ inline void gder_tile(const double* tx, const double* ty, const double* tz, int tn,
                                 int jt, int i, double xi, double yi, double zi, double side,
                                 double side_inv, double sx, double sy, double sz, double guard,
                                 double range, double bin_size, int nbins, int si, int nb,
                                 unsigned int* wh) {
  for (int k = 0; k < tn; k++) {
    const int j = jt + k;
    if (j == i) continue;
    const double dx = gder_image<(MASK & 1) != 0>(xi - tx[k], side, side_inv, sx);
    const double dy = gder_image<(MASK & 2) != 0>(yi - ty[k], side, side_inv, sy);
    const double dz = gder_image<(MASK & 4) != 0>(zi - tz[k], side, side_inv, sz);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < guard)) continue;
    const double dist = sqrt(d2);
    if (dist < range) {
      int bin = (int)(dist / bin_size);
      if (bin == nbins) bin--;
      const int pr = 2 * si + (j < nb ? 0 : 1);
      atomicAdd(&wh[pr * nbins + bin], 1u);
    }
  }
}

// Wavefront-wide min and max of v.
__device__ inline void wave_minmax(double v_lo, double v_hi, double& lo, double& hi) {
  lo = v_lo;
  hi = v_hi;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off, 64));
    hi = fmax(hi, __shfl_xor(hi, off, 64));
  }
}
__global__ void __launch_bounds__(kGderThreads)
    k_gder(const RepDev* __restrict__ reps, int rep, const double* __restrict__ px,
           const double* __restrict__ py, const double* __restrict__ pz, int nbins,
           double range, double bin_size, int j_chunk,
           unsigned long long* __restrict__ counts) {
  // rep < 0: one replica per blockIdx.z, counts offset by 4*nbins per replica
  if (rep < 0) {
    rep = (int)blockIdx.z;
    counts += (size_t)rep * 4 * nbins;
  }
  const RepDev& R = reps[rep];
  if ((int)blockIdx.x * kGderThreads >= R.n || (int)blockIdx.y * j_chunk >= R.n) return;
  extern __shared__ unsigned int hist[];  // [waves][4][nbins]
  __shared__ double tx[kGderThreads], ty[kGderThreads], tz[kGderThreads];
  const int waves = kGderThreads / 64;
  const int wave = (int)threadIdx.x / 64;
  for (int k = (int)threadIdx.x; k < waves * 4 * nbins; k += kGderThreads) hist[k] = 0u;
  const int i = (int)blockIdx.x * kGderThreads + (int)threadIdx.x;
  const bool active = i < R.n;
  double xi = 0, yi = 0, zi = 0;
  int si = 0;
  if (active) {
    xi = px[R.p_off + i];
    yi = py[R.p_off + i];
    zi = pz[R.p_off + i];
    si = i < R.nb ? 0 : 1;
  }
  const double side = R.side, side_inv = R.side_inv;
  // sqrt(d2) < range implies d2 < range^2 <= guard: pairs at or beyond the
  // guard are never counted, and their square root is not taken
  const double guard = range * range * (1.0 + 1e-12);
  unsigned int* wh = hist + (size_t)wave * 4 * nbins;
  const int j0 = (int)blockIdx.y * j_chunk;
  const int j1 = min(R.n, j0 + j_chunk);
  // the wavefront's bounding box (its active particles)
  const double inf = __builtin_inf();
  double ilo[3], ihi[3];
  wave_minmax(active ? xi : inf, active ? xi : -inf, ilo[0], ihi[0]);
  wave_minmax(active ? yi : inf, active ? yi : -inf, ilo[1], ihi[1]);
  wave_minmax(active ? zi : inf, active ? zi : -inf, ilo[2], ihi[2]);
  const int lane = (int)threadIdx.x % 64;
  __syncthreads();
  for (int jt = j0; jt < j1; jt += kGderThreads) {
    const int jl = jt + (int)threadIdx.x;
    if (jl < j1) {
      tx[threadIdx.x] = px[R.p_off + jl];
      ty[threadIdx.x] = py[R.p_off + jl];
      tz[threadIdx.x] = pz[R.p_off + jl];
    }
    __syncthreads();
    const int tn = min(kGderThreads, j1 - jt);
    // Per axis: the image index K(d) = floor(RN(RN(d / L) + 0.5)) is
    // monotonic in d, and every pair's d = RN(x_i - x_j) lies between
    // RN(min x_i - max x_j) and RN(max x_i - min x_j); where both ends give
    // the same K, so does every pair of this wavefront and tile.
    int mask = 0;
    double sh[3] = {0.0, 0.0, 0.0};
    {
      const double* tc[3] = {tx, ty, tz};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        double lo = inf, hi = -inf;
        for (int k = lane; k < tn; k += 64) {
          lo = fmin(lo, tc[a][k]);
          hi = fmax(hi, tc[a][k]);
        }
        double jlo, jhi;
        wave_minmax(lo, hi, jlo, jhi);
        const double kmin = floor(side_inv * (ilo[a] - jhi) + 0.5);
        const double kmax = floor(side_inv * (ihi[a] - jlo) + 0.5);
        if (kmin == kmax) {
          mask |= 1 << a;
          sh[a] = side * kmin;
        }
      }
    }
    if (active) {
      switch (mask) {
#define BD_GDER_CASE(M)                                                                        \
  case M:                                                                                      \
    gder_tile<M>(tx, ty, tz, tn, jt, i, xi, yi, zi, side, side_inv, sh[0], sh[1], sh[2], guard, \
                 range, bin_size, nbins, si, R.nb, wh);                                        \
    break;
        BD_GDER_CASE(0)
        BD_GDER_CASE(1)
        BD_GDER_CASE(2)
        BD_GDER_CASE(3)
        BD_GDER_CASE(4)
        BD_GDER_CASE(5)
        BD_GDER_CASE(6)
        BD_GDER_CASE(7)
#undef BD_GDER_CASE
      }
    }
    __syncthreads();
  }
  for (int k = (int)threadIdx.x; k < 4 * nbins; k += kGderThreads) {
    unsigned long long s = 0;
    for (int w = 0; w < waves; w++) s += hist[(size_t)w * 4 * nbins + k];
    if (s) atomicAdd(&counts[k], s);
  }
}

// Stream copy for the bandwidth probe: b = a over n double4, grid-stride.
__global__ void __launch_bounds__(256)
    k_stream_copy(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) b[k] = a[k];
}

// Grid barriers only, for the barrier-cost probe: every workgroup of a
// cooperative launch passes `barriers` grid-wide barriers.
__global__ void __launch_bounds__(kBlk) k_grid_barrier(int barriers, int* __restrict__ done) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (int k = 0; k < barriers; k++) grid.sync();
  if (blockIdx.x == 0 && threadIdx.x == 0) done[0] = barriers;
}

// A kernel that does next to nothing, for the launch-floor probe.
__global__ void __launch_bounds__(1024) k_launch_floor(int* __restrict__ done) {
  if (blockIdx.x == 0 && threadIdx.x == 0) done[0] = 1;
}

// Stand-alone rannew64 lanes for the stream test.
__global__ void k_rng_stream(uint64_t seed, int lanes, int ndraws, uint64_t* __restrict__ ring,
                             uint64_t* __restrict__ out) {
  const int l = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (l >= lanes) return;
  uint64_t u = bd::lane_seed(seed, (uint64_t)l);
  for (int i = 0; i < 1000; i++) u = u * bd::kSeedA + bd::kSeedC;
  for (int i = 0; i < bd::kRingLen; i++) {
    u = u * bd::kSeedA + bd::kSeedC;
    ring[(size_t)i * lanes + l] = u;
  }
  uint64_t lcg = u;
  int np = bd::kRingLen - 1;
  for (int k = 0; k < ndraws; k++) {
    const int nq = np >= bd::kLagGap ? np - bd::kLagGap : np + (bd::kRingLen - bd::kLagGap);
    const uint64_t x = ring[(size_t)np * lanes + l] - ring[(size_t)nq * lanes + l];
    ring[(size_t)np * lanes + l] = x;
    lcg = bd::kLcgA * lcg + bd::kLcgC;
    out[(size_t)l * ndraws + k] = x - lcg;
    np = np == 0 ? bd::kRingLen - 1 : np - 1;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return fail(BD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return BD_OK;
}

}  // namespace

struct bd_sys {
  int R = 0;
  int device = 0;  // the GPU current at bd_create
  std::vector<bd_params> params;
  std::vector<bd_geometry> geom;
  std::vector<RepDev> reps;
  std::vector<bool> have_pos;
  int n_total = 0, lanes_total = 0, cells_total = 0, rs_total = 0;
  size_t tags_total = 0;
  int64_t ni = 0;
  // launch tables
  std::vector<int4> tbl_part_h, tbl_lane_h, tbl_cell_h;
  int4 *tbl_part = nullptr, *tbl_lane = nullptr, *tbl_cell = nullptr;
  RepDev* reps_d = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;
  double *rx = nullptr, *ry = nullptr, *rz = nullptr;
  double *ix = nullptr, *iy = nullptr, *iz = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *vir = nullptr, *pot = nullptr;
  int *nocup = nullptr, *cnt = nullptr, *ins = nullptr, *rs = nullptr;
  double4* cpos = nullptr;
  uint64_t *ring = nullptr, *lcg = nullptr, *draws = nullptr;
  int64_t* step = nullptr;
  RepFlags* flags = nullptr;
  int pot_kind = BD_POT_WCA;
  bool located = false;
  hipStream_t stream = nullptr;
  // cached step graphs: whole steps, and the fused path's (forces + next
  // move, binning) pairs; each kept until a setting or the length changes
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    int steps = 0;
  } graph_plain, graph_fused;
  int graph_mode = 16;       // steps per graph launch; 0 = eager
  int persist = 0;           // steps per persistent launch; 0 = off
  int persist_grid = 0;      // co-resident workgroups of the persistent kernel
  // brick force kernel (0 = thread-per-particle global kernel)
  int brick = 0;
  std::vector<int4> tbl_brick_h;
  int4* tbl_brick = nullptr;
  // force summation order: 1 = reference (bit-exact), G = 4 or 8 = split
  int order = 1;
  // deferred pair terms of the walk: 0 (direct), 16 or 32 LDS entries
  int defer = 16;
  // walk of the default force kernel: 0 rows, 1 slabs, 2 / 3 chunked slabs
  int walk = 1;
  // stencil passes of the thread-per-particle kernel: 0 one wavefront, 1 the
  // two passes on two wavefronts, 2 the two halves of each pass on four
  // wavefronts (slab walks with deferral only)
  int pass_split = 0;
  // update inside the force kernel: 0 never, 1 wherever the shape allows it,
  // 2 (auto) with the multi-wavefront pass splits only
  int fuse = 2;
  // kernel timing: when set, launch() has the runtime stamp these at the
  // kernel's start and end
  hipEvent_t tev[2] = {nullptr, nullptr};
  // neighbour lists of the list walk (allocated on first use)
  double skin = 0.3;
  unsigned int* nlist = nullptr;
  int* ncount = nullptr;
  double *rbx = nullptr, *rby = nullptr, *rbz = nullptr;
  int* lflag = nullptr;    // [2 R] rebuild flags
  // the list rebuild runs on a second stream beside the force kernel
  hipStream_t stream2 = nullptr;
  PosCell* pos4 = nullptr;  // positions with own-grid cells, by particle
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int* list_ok = nullptr;  // [R]
  bool defer_ok = true;  // every grid's slot index fits the 26-bit record
  // flat walk: every grid below 2^22 slots with at most kFlatMaxTags per
  // cell; its partner records (kFlatDepth per particle, allocated on first use)
  bool flat_ok = true;
  unsigned int* frec = nullptr;
  // the flat walk's block tables of 32 and 16 particles (pass split 1, 2)
  std::vector<int4> tbl_flat_h[2];
  int4* tbl_flat[2] = {nullptr, nullptr};
  int flat_depth = 32;  // records used per particle (kFlatDepth at most)
  std::vector<int4> tbl_split_h;
  int4* tbl_split = nullptr;
  // waiting-time origins: BD_MAX_ORIGINS slots of raw positions (x, y, z)
  double* origins = nullptr;
  // device buffer of the measurement entry points' per-call results (grown
  // on demand, then reused: a sampling loop allocates nothing after its
  // first call)
  void* callbuf = nullptr;
  size_t callbuf_bytes = 0;
};

namespace {

// Every kernel of a step is launched through launch().  With timing events
// set on the engine (bd_kernel_times, bd_kernel_times_fused) the runtime
// stamps them at the kernel's own start and end (hipExtLaunchKernel), so the
// figures are kernel durations, without the dispatch gaps that an event pair
// recorded around the launch would add.
template <typename... P, typename... A>
void launch(bd_sys* s, void (*k)(P...), int nblk, int nthr, A... a) {
  std::tuple<P...> args{a...};
  std::apply(
      [&](auto&... e) {
        void* ptrs[] = {static_cast<void*>(&e)...};
        if (s->tev[0] || s->tev[1])
          (void)hipExtLaunchKernel(reinterpret_cast<const void*>(k), dim3(nblk), dim3(nthr), ptrs,
                                   0, s->stream, s->tev[0], s->tev[1], 0);
        else
          (void)hipLaunchKernel(reinterpret_cast<const void*>(k), dim3(nblk), dim3(nthr), ptrs, 0,
                                s->stream);
      },
      args);
}

int launch_bin(bd_sys* s) {
  const int nb = (int)s->tbl_part_h.size();
  launch(s, k_bin, nb, kBlk, s->reps_d, s->tbl_part, s->px, s->py, s->pz, s->nocup,
                                    s->ins, s->flags, (int64_t)s->ni);
  return BD_OK;
}

bool list_on(const bd_sys* s) {
  return s->walk == BD_WALK_LIST && s->brick == 0 && s->order == BD_ORDER_REFERENCE;
}

// The moves' list bookkeeping: live only while the list walk is the force path.
ListMove list_move(const bd_sys* s) {
  if (!list_on(s)) return ListMove{nullptr, nullptr, nullptr, nullptr, 0.0};
  return ListMove{s->rbx, s->rby, s->rbz, s->lflag, 0.25 * s->skin * s->skin};
}

void launch_update(bd_sys* s, int do_bin) {
  const int nb = (int)s->tbl_lane_h.size();
  launch(s, k_update_bin, nb, kBlk, s->reps_d, s->tbl_lane, s->px, s->py, s->pz, s->rx, s->ry,
         s->rz, s->fx, s->fy, s->fz, s->ring, s->lcg, s->draws, s->step, s->lanes_total,
         s->nocup, s->ins, s->flags, do_bin, list_move(s));
}

void launch_cells(bd_sys* s, int advance) {
  const int nb = (int)s->tbl_cell_h.size();
  launch(s, k_cells, nb, kBlk, s->reps_d, s->tbl_cell, s->px, s->py, s->pz, s->nocup, s->cnt,
         s->ins, s->rs, s->cpos, s->flags, s->step, s->draws, advance,
         list_on(s) ? s->lflag : nullptr, list_on(s) ? s->pos4 : nullptr);
}

MoveArgs move_args(bd_sys* s) {
  return MoveArgs{s->ring, s->lcg,   s->draws, s->step,  s->lanes_total, s->rx,
                  s->ry,   s->rz,    s->nocup, s->ins,   s->flags,       list_move(s)};
}

// The list walk's forces: k_list_build over the flagged replicas on the
// second stream, beside the force kernel on the engine's stream (forked
// from and joined back into it, also under graph capture).  Under kernel
// timing the two run one after the other and the event pair spans both.
void launch_list(bd_sys* s, const int4* tbl, int nblk) {
  const ListArgs la{s->nlist, s->ncount, s->rbx,     s->rby,      s->rbz,
                    s->rx,    s->ry,     s->rz,      s->lflag,    s->list_ok,
                    s->pos4,  s->n_total, s->skin};
  const int bgrid = std::max(std::min(nblk, 1024), (nblk + kBuildThreads - 1) / kBuildThreads);
  const hipEvent_t e0 = s->tev[0], e1 = s->tev[1];
  const bool timed = e0 || e1;
  if (timed) {
    s->tev[1] = nullptr;  // the build stamps the start
    launch(s, k_list_build, bgrid, kBuildThreads, s->reps_d, tbl, nblk, s->px, s->py, s->pz,
           s->rs, s->cpos, s->draws, la);
    s->tev[0] = nullptr;  // the force kernel the end
    s->tev[1] = e1;
  } else {
    (void)hipEventRecord(s->ev_fork, s->stream);
    (void)hipStreamWaitEvent(s->stream2, s->ev_fork, 0);
    hipStream_t own = s->stream;
    s->stream = s->stream2;
    launch(s, k_list_build, bgrid, kBuildThreads, s->reps_d, tbl, nblk, s->px, s->py, s->pz,
           s->rs, s->cpos, s->draws, la);
    s->stream = own;
  }
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_list<BD_POT_WCA>, nblk, kBlk, s->reps_d, tbl, s->px, s->py, s->pz, s->rs,
             s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->draws, la);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_list<BD_POT_WCA_MODIFIED>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
             s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->draws, la);
      break;
    default:
      launch(s, k_forces_list<BD_POT_HARMONIC>, nblk, kBlk, s->reps_d, tbl, s->px, s->py, s->pz,
             s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->draws, la);
      break;
  }
  if (!timed) {
    (void)hipEventRecord(s->ev_join, s->stream2);
    (void)hipStreamWaitEvent(s->stream, s->ev_join, 0);
  }
  s->tev[0] = e0;
  s->tev[1] = e1;
}

// Every list is stale: the next force evaluation of each replica rebuilds.
int list_invalidate(bd_sys* s) {
  if (!s->lflag) return BD_OK;
  std::vector<int> ones(2 * (size_t)s->R, 1);
  HIP_TRY(hipMemcpyAsync(s->lflag, ones.data(), sizeof(int) * ones.size(), hipMemcpyHostToDevice,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// The flat walk's partner records: kFlatDepth per particle, allocated on first use.
int flat_setup(bd_sys* s) {
  if (!s->flat_ok)
    return fail(BD_E_ARG, "the flat walk needs replicas below 2^22 slots and ntags <= 12");
  if (!s->frec) return dev_alloc(&s->frec, (size_t)kFlatDepth * s->n_total);
  return BD_OK;
}

// Allocates the lists and decides per replica whether its grids hold the
// list stencil: ceil((rc + skin) / cell side) cells each way, without a cell
// met twice (2 nr + 1 <= ncell), rc + skin below half the box, ids below 2^24.
int list_setup(bd_sys* s) {
  if (!s->nlist) {
    int rc = dev_alloc(&s->nlist, (size_t)kListCap * s->n_total);
    if (rc == BD_OK) rc = dev_alloc(&s->ncount, (size_t)s->n_total);
    if (rc == BD_OK) rc = dev_alloc(&s->rbx, (size_t)s->n_total);
    if (rc == BD_OK) rc = dev_alloc(&s->rby, (size_t)s->n_total);
    if (rc == BD_OK) rc = dev_alloc(&s->rbz, (size_t)s->n_total);
    if (rc == BD_OK) rc = dev_alloc(&s->lflag, 2 * (size_t)s->R);
    if (rc == BD_OK) rc = dev_alloc(&s->list_ok, (size_t)s->R);
    if (rc == BD_OK) rc = dev_alloc(&s->pos4, (size_t)s->n_total);
    if (rc != BD_OK) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
  }
  std::vector<int> ok(s->R, 0);
  for (int r = 0; r < s->R; r++) {
    const RepDev& R = s->reps[r];
    bool good = R.fast && R.n < (1 << 24);
    for (int g = 0; g < 2 && good; g++)
      for (int pass = 0; pass < 2 && good; pass++) {
        const int gg = pass == 0 ? g : 1 - g;
        const double rl = std::sqrt(pass == 0 ? R.rc2_same[g] : R.rc2_bs) + s->skin;
        const int nr = (int)std::ceil(rl / R.cell_side[gg]);
        // the build's stencil reaches nr <= 2 cells and sees minimum images
        good = nr <= 2 && 2 * nr + 3 <= R.ncell[gg] && rl < 0.5 * R.side;
      }
    ok[r] = good ? 1 : 0;
  }
  HIP_TRY(hipMemcpy(s->list_ok, ok.data(), sizeof(int) * ok.size(), hipMemcpyHostToDevice));
  return list_invalidate(s);
}

// With the list walk on and cells located, bins the current positions again
// so that k_cells rewrites pos4 (same cells: the positions did not move).
int list_refresh(bd_sys* s) {
  if (!list_on(s) || !s->located) return BD_OK;
  launch_bin(s);
  launch_cells(s, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// UPD: the force kernel also performs the next step's update (see k_forces).
template <bool NEIGH, int DEPTH, int WALK, bool UPD>
void launch_forces_k(bd_sys* s, int* ncounts, int* nlist, int maxn, const int4* tbl, int nblk) {
  const MoveArgs mv = move_args(s);
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces<BD_POT_WCA, NEIGH, DEPTH, WALK, UPD>, nblk, kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, ncounts, nlist, maxn, mv);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces<BD_POT_WCA_MODIFIED, NEIGH, DEPTH, WALK, UPD>, nblk, kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, ncounts, nlist, maxn, mv);
      break;
    default:
      launch(s, k_forces<BD_POT_HARMONIC, NEIGH, DEPTH, WALK, UPD>, nblk, kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, ncounts, nlist, maxn, mv);
      break;
  }
}

template <int WALK, bool UPD>
void launch_forces_w(bd_sys* s, const int4* tbl, int nblk) {
  if (s->defer == 0)
    launch_forces_k<false, 0, WALK, UPD>(s, nullptr, nullptr, 0, tbl, nblk);
  else if (s->defer == 16)
    launch_forces_k<false, 16, WALK, UPD>(s, nullptr, nullptr, 0, tbl, nblk);
  else
    launch_forces_k<false, 32, WALK, UPD>(s, nullptr, nullptr, 0, tbl, nblk);
}

template <int DEPTH, int WALK, bool UPD>
void launch_pair_k(bd_sys* s, const int4* tbl, int nblk) {
  const MoveArgs mv = move_args(s);
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_pair<BD_POT_WCA, DEPTH, WALK, UPD>, nblk, 2 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_pair<BD_POT_WCA_MODIFIED, DEPTH, WALK, UPD>, nblk, 2 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
    default:
      launch(s, k_forces_pair<BD_POT_HARMONIC, DEPTH, WALK, UPD>, nblk, 2 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
  }
}

template <int DEPTH, int WALK, bool UPD>
void launch_quad_k(bd_sys* s, const int4* tbl, int nblk) {
  const MoveArgs mv = move_args(s);
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_quad<BD_POT_WCA, DEPTH, WALK, UPD>, nblk, 4 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_quad<BD_POT_WCA_MODIFIED, DEPTH, WALK, UPD>, nblk, 4 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
    default:
      launch(s, k_forces_quad<BD_POT_HARMONIC, DEPTH, WALK, UPD>, nblk, 4 * kBlk,
          s->reps_d, tbl, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir,
          s->pot, mv);
      break;
  }
}

template <int WALK, bool UPD>
void launch_pair_w(bd_sys* s, const int4* tbl, int nblk) {
  if (s->pass_split == 2) {
    if (s->defer == 16)
      launch_quad_k<16, WALK, UPD>(s, tbl, nblk);
    else
      launch_quad_k<32, WALK, UPD>(s, tbl, nblk);
  } else if (s->defer == 16) {
    launch_pair_k<16, WALK, UPD>(s, tbl, nblk);
  } else {
    launch_pair_k<32, WALK, UPD>(s, tbl, nblk);
  }
}

template <bool UPD>
void launch_slab_forces(bd_sys* s, const int4* tbl, int nblk) {
  if (s->pass_split && s->defer > 0) {
    if (s->walk == 1)
      launch_pair_w<1, UPD>(s, tbl, nblk);
    else if (s->walk == 2)
      launch_pair_w<2, UPD>(s, tbl, nblk);
    else
      launch_pair_w<3, UPD>(s, tbl, nblk);
  } else if (s->walk == 1) {
    launch_forces_w<1, UPD>(s, tbl, nblk);
  } else if (s->walk == 2) {
    launch_forces_w<2, UPD>(s, tbl, nblk);
  } else {
    launch_forces_w<3, UPD>(s, tbl, nblk);
  }
}

// The 49-50 potentials run in three kernels: the single-wavefront slab walk
// with deferral depth 16 (the row walk without deferral where slot indexes
// outgrow the deferral records), and the row walk that records the
// neighbours.  Every walk, split and brick gives the same bits, so the shape
// settings need no further instances.
template <int POT>
void launch_forces_phs_t(bd_sys* s, int* ncounts, int* nlist, int maxn, const int4* tbl,
                         int nblk) {
  const MoveArgs mv = move_args(s);
  if (ncounts)
    launch(s, k_forces<POT, true, 0, 0, false>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
           s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, ncounts, nlist, maxn, mv);
  else if (s->defer_ok)
    launch(s, k_forces<POT, false, 16, 1, false>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
           s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, ncounts, nlist, maxn, mv);
  else
    launch(s, k_forces<POT, false, 0, 0, false>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
           s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, ncounts, nlist, maxn, mv);
}

bool phs_kind(int pot) { return pot == BD_POT_WCA49_50 || pot == BD_POT_WCA49_50_AO; }

// The flat walk (bd_set_walk has allocated its records); UPD: with the next
// step's move.  Pass split 1 / 2 gives each wavefront 32 / 16 particles (all
// 64 lanes still test the staged candidates): more wavefronts for systems too
// small to fill the GPU.
template <bool UPD>
void launch_flat(bd_sys* s, const int4* tbl, int nblk) {
  const MoveArgs mv = move_args(s);
  if (s->pass_split > 0 && tbl == s->tbl_part) {  // 32 or 16 particles per wavefront
    tbl = s->tbl_flat[s->pass_split - 1];
    nblk = (int)s->tbl_flat_h[s->pass_split - 1].size();
  }
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_flat<BD_POT_WCA, UPD>, nblk, kBlk, s->reps_d, tbl, s->px, s->py, s->pz,
             s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->frec, s->n_total,
             s->flat_depth, mv);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_flat<BD_POT_WCA_MODIFIED, UPD>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
             s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->frec, s->n_total,
             s->flat_depth, mv);
      break;
    default:
      launch(s, k_forces_flat<BD_POT_HARMONIC, UPD>, nblk, kBlk, s->reps_d, tbl, s->px, s->py,
             s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz, s->vir, s->pot, s->frec, s->n_total,
             s->flat_depth, mv);
      break;
  }
}

template <bool NEIGH>
void launch_forces_t(bd_sys* s, int* ncounts, int* nlist, int maxn, const int4* tbl, int nblk) {
  if (s->pot_kind == BD_POT_WCA49_50)
    launch_forces_phs_t<BD_POT_WCA49_50>(s, NEIGH ? ncounts : nullptr, nlist, maxn, tbl, nblk);
  else if (s->pot_kind == BD_POT_WCA49_50_AO)
    launch_forces_phs_t<BD_POT_WCA49_50_AO>(s, NEIGH ? ncounts : nullptr, nlist, maxn, tbl,
                                            nblk);
  else if (NEIGH)
    launch_forces_k<true, 0, 0, false>(s, ncounts, nlist, maxn, tbl, nblk);
  else if (s->walk == BD_WALK_LIST)
    launch_list(s, tbl, nblk);
  else if (s->walk == BD_WALK_FLAT)
    launch_flat<false>(s, tbl, nblk);
  else if (s->walk >= 1)
    launch_slab_forces<false>(s, tbl, nblk);
  else
    launch_forces_w<0, false>(s, tbl, nblk);
}

template <int CAP>
void launch_brick_t(bd_sys* s) {
  const int nblk = (int)s->tbl_brick_h.size();
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_brick<BD_POT_WCA, CAP>, nblk, kBrickThreads,
          s->reps_d, s->tbl_brick, s->brick, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy,
          s->fz, s->vir, s->pot, s->flags);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_brick<BD_POT_WCA_MODIFIED, CAP>, nblk, kBrickThreads,
          s->reps_d, s->tbl_brick, s->brick, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy,
          s->fz, s->vir, s->pot, s->flags);
      break;
    default:
      launch(s, k_forces_brick<BD_POT_HARMONIC, CAP>, nblk, kBrickThreads,
          s->reps_d, s->tbl_brick, s->brick, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy,
          s->fz, s->vir, s->pot, s->flags);
      break;
  }
}

template <int G>
void launch_split_t(bd_sys* s) {
  const int nblk = (int)s->tbl_split_h.size();
  switch (s->pot_kind) {
    case BD_POT_WCA:
      launch(s, k_forces_split<BD_POT_WCA, G>, nblk, kBlk,
          s->reps_d, s->tbl_split, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz,
          s->vir, s->pot);
      break;
    case BD_POT_WCA_MODIFIED:
      launch(s, k_forces_split<BD_POT_WCA_MODIFIED, G>, nblk, kBlk,
          s->reps_d, s->tbl_split, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz,
          s->vir, s->pot);
      break;
    default:
      launch(s, k_forces_split<BD_POT_HARMONIC, G>, nblk, kBlk,
          s->reps_d, s->tbl_split, s->px, s->py, s->pz, s->rs, s->cpos, s->fx, s->fy, s->fz,
          s->vir, s->pot);
      break;
  }
}

// Whether the next step's update runs inside the force kernel: a
// thread-per-particle slab walk in reference order, and one particle per
// rannew64 lane in every replica.  In auto mode only with a pass split: the
// split kernels leave the GPU room for the move (plain N=32768: 1.13e9 vs
// 9.8e8 particle-steps/s), while the single-wavefront walk on a full batch
// loses more to the added registers than the launch it saves (2.88e9 vs
// 3.02e9 on 256 replicas).
bool fusable(const bd_sys* s) {
  if (s->fuse == 0 || s->order != BD_ORDER_REFERENCE || s->brick != 0 || s->walk < 1 ||
      s->walk == BD_WALK_LIST || phs_kind(s->pot_kind))
    return false;
  if (s->fuse == 2 && s->walk != BD_WALK_FLAT && (s->pass_split == 0 || s->defer == 0))
    return false;
  for (int r = 0; r < s->R; r++)
    if (s->reps[r].per_lane != 1) return false;
  return true;
}

// Forces of the current positions and the next step's update in one launch.
void launch_forces_move(bd_sys* s) {
  if (s->walk == BD_WALK_FLAT)
    launch_flat<true>(s, s->tbl_part, (int)s->tbl_part_h.size());
  else
    launch_slab_forces<true>(s, s->tbl_part, (int)s->tbl_part_h.size());
}

void launch_forces(bd_sys* s) {
  if (phs_kind(s->pot_kind)) {
    launch_forces_t<false>(s, nullptr, nullptr, 0, s->tbl_part, (int)s->tbl_part_h.size());
  } else if (s->order == 4) {
    launch_split_t<4>(s);
  } else if (s->order == 8) {
    launch_split_t<8>(s);
  } else if (s->brick == 0) {
    launch_forces_t<false>(s, nullptr, nullptr, 0, s->tbl_part, (int)s->tbl_part_h.size());
  } else if (s->brick <= 3) {
    launch_brick_t<1024>(s);
  } else {
    launch_brick_t<2048>(s);
  }
}

template <int POT>
int launch_persistent_t(bd_sys* s, int nsteps) {
  const void* fn = reinterpret_cast<const void*>(&k_persistent_steps<POT>);
  if (s->persist_grid == 0) {
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return fail(BD_E_STATE, "the device does not support cooperative launches");
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlk, 0));
    if (per_cu < 1) return fail(BD_E_STATE, "the persistent kernel does not fit on a CU");
    s->persist_grid = per_cu * ncu;
  }
  StepArgs a{s->reps_d, s->tbl_lane, s->tbl_cell, s->tbl_part,
             (int)s->tbl_lane_h.size(), (int)s->tbl_cell_h.size(), (int)s->tbl_part_h.size(),
             s->lanes_total, s->px, s->py, s->pz, s->rx, s->ry, s->rz, s->fx, s->fy,
             s->fz, s->vir, s->pot, s->ring, s->lcg, s->draws, s->step, s->nocup, s->cnt,
             s->ins, s->rs, s->flags, s->cpos};
  void* args[] = {&a, &nsteps};
  HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(s->persist_grid), dim3(kBlk), args, 0, s->stream));
  return BD_OK;
}

int launch_persistent(bd_sys* s, int nsteps) {
  switch (s->pot_kind) {
    case BD_POT_WCA:
      return launch_persistent_t<BD_POT_WCA>(s, nsteps);
    case BD_POT_WCA_MODIFIED:
      return launch_persistent_t<BD_POT_WCA_MODIFIED>(s, nsteps);
    case BD_POT_WCA49_50:
      return launch_persistent_t<BD_POT_WCA49_50>(s, nsteps);
    case BD_POT_WCA49_50_AO:
      return launch_persistent_t<BD_POT_WCA49_50_AO>(s, nsteps);
    default:
      return launch_persistent_t<BD_POT_HARMONIC>(s, nsteps);
  }
}


// Largest number of cells a brick of B big cells stages (both boxes).
int brick_cells_bound(const RepDev& R, int B) {
  const int rb = std::max(R.nrange_same[0], R.nrange_cross[1]);
  const int rs = std::max(R.nrange_same[1], R.nrange_cross[0]);
  const int bd = B + 2 * rb;
  const int sd = (int)((double)B * R.cell_side[0] * R.cell_inv[1]) + 3 + 2 * rs;
  return bd * bd * bd + sd * sd * sd;
}

// Deal consecutive table entries (same replica, neighbouring particles or
// cells) to the same XCD: workgroups are dispatched round-robin over the 8
// XCDs, so entry k goes to slot (k % C) * 8 + k / C; padding entries have
// replica -1.
std::vector<int4> xcd_group(const std::vector<int4>& natural) {
  const int M = (int)natural.size();
  const int C = (M + 7) / 8;
  std::vector<int4> tbl((size_t)C * 8, make_int4(-1, 0, 0, 0));
  for (int k = 0; k < M; k++) tbl[(size_t)(k % C) * 8 + k / C] = natural[k];
  return tbl;
}

// Brick table: every replica tiled by B^3 big-cell bricks, then grouped so
// that consecutive bricks (spatial neighbours sharing halo cells) land on
// the same XCD (workgroups are dealt round-robin over the 8 XCDs).
int setup_bricks(bd_sys* s, int B) {
  if (B == 0) {
    s->brick = 0;
    return BD_OK;
  }
  if (B < 1 || B > 4) return fail(BD_E_ARG, "brick must be 0..4");
  std::vector<int4> natural;
  for (int r = 0; r < s->R; r++) {
    const RepDev& R = s->reps[r];
    if (!R.fast || B > R.ncell[0] || brick_cells_bound(R, B) > kBrickMaxCells)
      return fail(BD_E_ARG, "brick too large for replica " + std::to_string(r));
    for (int bz = 0; bz < R.ncell[0]; bz += B)
      for (int by = 0; by < R.ncell[0]; by += B)
        for (int bx = 0; bx < R.ncell[0]; bx += B) natural.push_back(make_int4(r, bx, by, bz));
  }
  const std::vector<int4> tbl = xcd_group(natural);
  if (s->tbl_brick) (void)hipFree(s->tbl_brick);
  s->tbl_brick = nullptr;
  int rc = dev_alloc(&s->tbl_brick, tbl.size());
  if (rc != BD_OK) return rc;
  HIP_TRY(hipMemcpy(s->tbl_brick, tbl.data(), sizeof(int4) * tbl.size(), hipMemcpyHostToDevice));
  s->tbl_brick_h = tbl;
  s->brick = B;
  return BD_OK;
}

void enqueue_step(bd_sys* s) {
  launch_update(s, 1);
  launch_cells(s, 1);
  launch_forces(s);
}

// k whole steps.  When fusable, the update of steps 2..k runs inside the
// force kernel of the step before (same operations on the same values): one
// launch fewer per step, and the state after the k steps is the same.
void enqueue_steps(bd_sys* s, int k) {
  if (k <= 0) return;
  if (!fusable(s)) {
    for (int j = 0; j < k; j++) enqueue_step(s);
    return;
  }
  launch_update(s, 1);
  launch_cells(s, 1);
  for (int j = 1; j < k; j++) {
    launch_forces_move(s);
    launch_cells(s, 1);
  }
  launch_forces(s);
}

// Error records after the enqueued work: the first replica with a flag
// raised fails the call, naming the step whose update (or binning) raised it.
// A record stays set until that replica's state is replaced
// (bd_set_positions, bd_checkpoint_load).
int check_flags(bd_sys* s) {
  std::vector<RepFlags> f(s->R);
  HIP_TRY(hipMemcpyAsync(f.data(), s->flags, sizeof(RepFlags) * s->R, hipMemcpyDeviceToHost,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int r = 0; r < s->R; r++) {
    const std::string where =
        "replica " + std::to_string(r) +
        (f[r].first_step != kNoFailStep ? " at step " + std::to_string(f[r].first_step) : "");
    if (f[r].bits & kFlagNonFinite)
      return fail(BD_E_NAN, where + ": non-finite position, or one outside the cell grids");
    if (f[r].bits & kFlagBrick)
      return fail(BD_E_STATE, where + ": brick exceeds its LDS table");
    if (f[r].bits & kFlagOverflow)
      return fail(BD_E_CELL_OVERFLOW, where + ": a cell reached ntags particles");
  }
  return BD_OK;
}

// Clears the error records of replicas [r0, r1) (queued on the engine's
// stream, before any later launch).
int clear_flags(bd_sys* s, int r0, int r1) {
  if (r1 <= r0) return BD_OK;
  std::vector<RepFlags> f((size_t)(r1 - r0), RepFlags{0, 0, kNoFailStep});
  HIP_TRY(hipMemcpyAsync(s->flags + r0, f.data(), sizeof(RepFlags) * f.size(),
                         hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// Drops both cached step graphs once the stream has drained, so that no
// queued replay of either can outlive its executable graph.
int drop_graphs(bd_sys* s) {
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (bd_sys::StepGraph* g : {&s->graph_plain, &s->graph_fused}) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
    g->steps = 0;
  }
  return BD_OK;
}

// The engine's per-call result buffer, at least `count` elements of T.
template <class T>
int call_buffer(bd_sys* s, size_t count, T** out) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (bytes > s->callbuf_bytes) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->callbuf) (void)hipFree(s->callbuf);
    s->callbuf = nullptr;
    s->callbuf_bytes = 0;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(BD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->callbuf = p;
    s->callbuf_bytes = bytes;
  }
  *out = static_cast<T*>(s->callbuf);
  return BD_OK;
}

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (bd_create).  Never the array's host-side name.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int gder_scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(BD_E_ARG, "device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return BD_OK;
  void* ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = ptr;
  return BD_OK;
}

void* gder_scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
}

int check_rep(const bd_sys* s, int r) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (r < 0 || r >= s->R) return fail(BD_E_ARG, "replica index out of range");
  return BD_OK;
}

void free_all(bd_sys* s) {
  if (s->graph_plain.exec) (void)hipGraphExecDestroy(s->graph_plain.exec);
  if (s->graph_fused.exec) (void)hipGraphExecDestroy(s->graph_fused.exec);
  if (s->tbl_brick) (void)hipFree(s->tbl_brick);
  if (s->tbl_split) (void)hipFree(s->tbl_split);
  if (s->origins) (void)hipFree(s->origins);
  if (s->callbuf) (void)hipFree(s->callbuf);
  if (s->frec) (void)hipFree(s->frec);
  for (int4* t : s->tbl_flat)
    if (t) (void)hipFree(t);
  void* lists[] = {s->nlist, s->ncount, s->rbx, s->rby, s->rbz, s->lflag, s->list_ok, s->pos4};
  for (void* p : lists)
    if (p) (void)hipFree(p);
  void* ptrs[] = {s->tbl_part, s->tbl_lane, s->tbl_cell, s->reps_d, s->px,   s->py,
                  s->pz,       s->rx,       s->ry,       s->rz,     s->ix,   s->iy,
                  s->iz,       s->fx,       s->fy,       s->fz,     s->vir,  s->pot,
                  s->nocup,    s->cnt,      s->ins,      s->rs,     s->cpos, s->ring,
                  s->lcg,      s->draws,    s->step,     s->flags};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  if (s->ev_join) (void)hipEventDestroy(s->ev_join);
  if (s->stream2) (void)hipStreamDestroy(s->stream2);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

void add_blocks(std::vector<int4>& tbl, int rep, int count, int grid, int per_block = kBlk) {
  for (int first = 0; first < count; first += per_block)
    tbl.push_back(make_int4(rep, first, std::min(count, first + per_block), grid));
}

}  // namespace

namespace bdi {
int set_error(int code, const std::string& msg) { return fail(code, msg); }

int device_state(bd_sys* s, int replica, DevState* out) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (replica < 0 || replica >= s->R) return fail(BD_E_ARG, "replica index out of range");
  const RepDev& R = s->reps[replica];
  out->px = s->px + R.p_off;
  out->py = s->py + R.p_off;
  out->pz = s->pz + R.p_off;
  out->fx = s->fx + R.p_off;
  out->fy = s->fy + R.p_off;
  out->fz = s->fz + R.p_off;
  out->n = R.n;
  out->nb = R.nb;
  out->side = R.side;
  out->stream = s->stream;
  return BD_OK;
}
}  // namespace bdi

// ===========================================================================
// C ABI
// ===========================================================================
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version;
  int32_t replicas;
  int64_t step;
  uint64_t particles, lanes;
  uint32_t has_origins, pad;
};

template <class T>
int ckpt_dev(FILE* f, T* dev, size_t count, bool save, hipStream_t st) {
  std::vector<T> h(count);
  if (save) {
    HIP_TRY(hipMemcpyAsync(h.data(), dev, sizeof(T) * count, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (std::fwrite(h.data(), sizeof(T), count, f) != count)
      return fail(BD_E_ARG, "checkpoint write failed");
  } else {
    if (std::fread(h.data(), sizeof(T), count, f) != count)
      return fail(BD_E_ARG, "checkpoint file is truncated");
    HIP_TRY(hipMemcpyAsync(dev, h.data(), sizeof(T) * count, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return BD_OK;
}

int ckpt_arrays(bd_sys* s, FILE* f, bool save) {
  const size_t N = (size_t)s->n_total;
  double* arrs[] = {s->px, s->py, s->pz, s->rx, s->ry, s->rz, s->ix,
                    s->iy, s->iz, s->fx, s->fy, s->fz, s->vir, s->pot};
  int rc = BD_OK;
  for (double* a : arrs)
    if (rc == BD_OK) rc = ckpt_dev(f, a, N, save, s->stream);
  if (rc == BD_OK)
    rc = ckpt_dev(f, s->ring, (size_t)bd::kRingLen * s->lanes_total, save, s->stream);
  if (rc == BD_OK) rc = ckpt_dev(f, s->lcg, (size_t)s->lanes_total, save, s->stream);
  if (rc == BD_OK) rc = ckpt_dev(f, s->draws, (size_t)s->R, save, s->stream);
  if (rc == BD_OK && s->origins)
    rc = ckpt_dev(f, s->origins, (size_t)BD_MAX_ORIGINS * 3 * N, save, s->stream);
  return rc;
}
}  // namespace

extern "C" {

const char* bd_last_error(void) { return g_err.c_str(); }

const char* bd_strerror(int code) {
  switch (code) {
    case BD_OK: return "ok";
    case BD_E_ARG: return "invalid argument";
    case BD_E_HIP: return "HIP runtime error";
    case BD_E_CELL_OVERFLOW: return "cell capacity (ntags) reached";
    case BD_E_NAN: return "non-finite position";
    case BD_E_STATE: return "call out of order";
    case BD_E_NOMEM: return "out of device memory";
    default: return "unknown error";
  }
}

uint64_t bd_lane_seed(uint64_t seed, uint64_t lane) { return bd::lane_seed(seed, lane); }

// One host rannew64 stream (rannew64.cu:55-126) placing particles
// uniformly: x = side * (53 high bits of a draw) * 2^-53.
int bd_random_positions(int32_t n, double side, uint64_t seed, double* xyz) {
  if (n < 1 || !(side > 0.0) || !xyz) return fail(BD_E_ARG, "bad arguments");
  uint64_t ring[bd::kRingLen];
  uint64_t u = bd::lane_seed(seed, 0);
  for (int i = 0; i < 1000; i++) u = u * bd::kSeedA + bd::kSeedC;
  for (int i = 0; i < bd::kRingLen; i++) {
    u = u * bd::kSeedA + bd::kSeedC;
    ring[i] = u;
  }
  uint64_t lcg = u;
  int np = bd::kRingLen - 1;
  for (int64_t k = 0; k < 3 * (int64_t)n; k++) {
    const int nq = np >= bd::kLagGap ? np - bd::kLagGap : np + (bd::kRingLen - bd::kLagGap);
    const uint64_t x = ring[np] - ring[nq];
    ring[np] = x;
    lcg = bd::kLcgA * lcg + bd::kLcgC;
    np = np == 0 ? bd::kRingLen - 1 : np - 1;
    xyz[k] = side * ((double)((x - lcg) >> 11) * 0x1.0p-53);
  }
  return BD_OK;
}

// Cutoff of a pair with diameter sigma: 2^(1/6) sigma (WCA), (50/49) sigma
// (49-50, get_forces.cu:47-49), sigma (1 + sigma / sigma_pol) where the
// depletion term and its force vanish (49-50 + AO), else sigma.
static double pair_cutoff(const bd_params* p, double sigma) {
  switch (p->potential) {
    case BD_POT_WCA:
      return bd::kWcaCut * sigma;
    case BD_POT_WCA49_50:
      return (50.0 / 49.0) * sigma;
    case BD_POT_WCA49_50_AO:
      return sigma * (1.0 + sigma / p->sigma_pol);
    default:
      return 1.0 * sigma;
  }
}

// main.cu:259-298 with the cutoff rc of the chosen potential in place of
// sigma: ncell = (int)(L / (1.05 rc)), cell side L/ncell, and
// nrange = 1 + (int)(rc / cell side of the scanned grid).
int bd_compute_geometry(const bd_params* p, bd_geometry* g) {
  if (!p || !g) return fail(BD_E_ARG, "null argument");
  if (!(p->side > 0.0) || !(p->sigma_big > 0.0) || !(p->sigma_sml > 0.0))
    return fail(BD_E_ARG, "side and sigmas must be positive");
  if (p->n_big < 1 || p->n_sml < 1) return fail(BD_E_ARG, "both species need particles");
  if (p->ntags_big < 1 || p->ntags_big > 32 || p->ntags_sml < 1 || p->ntags_sml > 32)
    return fail(BD_E_ARG, "ntags must be in 1..32");
  if (p->potential < BD_POT_WCA || p->potential > BD_POT_WCA49_50_AO)
    return fail(BD_E_ARG, "unknown potential");
  if (phs_kind(p->potential) &&
      !(p->pot_amp > 0.0 && p->pot_inv_t > 0.0 && std::isfinite(p->pot_amp * p->pot_inv_t)))
    return fail(BD_E_ARG, "the 49-50 potentials need finite pot_amp > 0 and pot_inv_t > 0");
  if (p->potential == BD_POT_WCA49_50_AO &&
      !(p->sigma_pol > 0.0 && p->phi_pol >= 0.0 && std::isfinite(p->phi_pol)))
    return fail(BD_E_ARG, "the depletion branch needs sigma_pol > 0 and finite phi_pol >= 0");
  const double sbs = 0.5 * (p->sigma_big + p->sigma_sml);
  g->rc_bb = pair_cutoff(p, p->sigma_big);
  g->rc_ss = pair_cutoff(p, p->sigma_sml);
  g->rc_bs = pair_cutoff(p, sbs);
  if (!std::isfinite(g->rc_bb) || !std::isfinite(g->rc_ss) || !std::isfinite(g->rc_bs))
    return fail(BD_E_ARG, "cutoff not finite");
  g->ncell_big = (int32_t)(p->side / (1.05 * g->rc_bb));
  g->ncell_sml = (int32_t)(p->side / (1.05 * g->rc_ss));
  if (g->ncell_big < 1 || g->ncell_sml < 1) return fail(BD_E_ARG, "box smaller than a cell");
  g->cell_side_big = p->side / (double)g->ncell_big;
  g->cell_side_sml = p->side / (double)g->ncell_sml;
  g->nrange_bb = 1 + (int32_t)(g->rc_bb / g->cell_side_big);
  g->nrange_ss = 1 + (int32_t)(g->rc_ss / g->cell_side_sml);
  g->nrange_bs = 1 + (int32_t)(g->rc_bs / g->cell_side_sml);
  g->nrange_sb = 1 + (int32_t)(g->rc_bs / g->cell_side_big);
  const int rb = std::max(g->nrange_bb, g->nrange_sb);
  const int rs = std::max(g->nrange_ss, g->nrange_bs);
  if (g->ncell_big < 2 * rb + 1 || g->ncell_sml < 2 * rs + 1)
    return fail(BD_E_ARG, "box too small: a stencil would visit a cell twice");
  return BD_OK;
}

int bd_create(const bd_params* params, int32_t n_replicas, bd_sys** out) {
  if (!params || !out || n_replicas < 1) return fail(BD_E_ARG, "bad arguments to bd_create");
  *out = nullptr;
  bd_sys* s = new bd_sys();
  s->R = n_replicas;
  s->params.assign(params, params + n_replicas);
  s->geom.resize(n_replicas);
  s->reps.resize(n_replicas);
  s->have_pos.assign(n_replicas, false);
  s->pot_kind = params[0].potential;
  s->ni = params[0].first_step;
  int64_t p_off = 0, lane_off = 0, cell_off = 0, rs_off = 0;
  int64_t tag_off = 0;
  for (int r = 0; r < n_replicas; r++) {
    const bd_params& p = params[r];
    int rc = bd_compute_geometry(&p, &s->geom[r]);
    if (rc != BD_OK) {
      delete s;
      return fail(rc, "replica " + std::to_string(r) + ": " + g_err);
    }
    if (p.potential != s->pot_kind || p.first_step != s->ni) {
      delete s;
      return fail(BD_E_ARG, "all replicas must share the potential and first_step");
    }
    if (!(p.dt > 0.0) || (p.temp_protocol == BD_TEMP_SINE && !(p.period > 0.0))) {
      delete s;
      return fail(BD_E_ARG, "dt (and the sine period) must be positive");
    }
    const bd_geometry& g = s->geom[r];
    RepDev& R = s->reps[r];
    std::memset(&R, 0, sizeof(R));
    R.n = p.n_big + p.n_sml;
    R.nb = p.n_big;
    R.p_off = (int32_t)p_off;
    R.lanes = p.rng_lanes > 0 ? std::min(p.rng_lanes, R.n) : R.n;
    R.per_lane = (R.n + R.lanes - 1) / R.lanes;
    R.lane_off = (int32_t)lane_off;
    R.ncell[0] = g.ncell_big;
    R.ncell[1] = g.ncell_sml;
    R.ntags[0] = p.ntags_big;
    R.ntags[1] = p.ntags_sml;
    for (int k = 0; k < 2; k++) {
      const int64_t nc3 = (int64_t)R.ncell[k] * R.ncell[k] * R.ncell[k];
      R.cell_off[k] = (int32_t)cell_off;
      R.tag_off[k] = (int32_t)tag_off;
      R.rs_off[k] = (int32_t)rs_off;
      cell_off += nc3;
      tag_off += nc3 * R.ntags[k];
      rs_off += (int64_t)R.ncell[k] * R.ncell[k] * (R.ncell[k] + 1);
    }
    R.nrange_same[0] = g.nrange_bb;
    R.nrange_same[1] = g.nrange_ss;
    R.nrange_cross[0] = g.nrange_bs;
    R.nrange_cross[1] = g.nrange_sb;
    R.potential = p.potential;
    R.temp_protocol = p.temp_protocol;
    {
      const int rb = std::max(g.nrange_bb, g.nrange_sb), rs = std::max(g.nrange_ss, g.nrange_bs);
      R.fast = rb <= 2 && rs <= 2 && g.ncell_big >= 2 * rb + 3 && g.ncell_sml >= 2 * rs + 3;
    }
    R.side = p.side;
    R.side_inv = 1.0 / p.side;
    R.cell_inv[0] = 1.0 / g.cell_side_big;
    R.cell_inv[1] = 1.0 / g.cell_side_sml;
    R.cell_side[0] = g.cell_side_big;
    R.cell_side[1] = g.cell_side_sml;
    R.sigma[0] = p.sigma_big;
    R.sigma[1] = p.sigma_sml;
    R.sigma_bs = 0.5 * (p.sigma_big + p.sigma_sml);
    R.rc2_same[0] = g.rc_bb * g.rc_bb;
    R.rc2_same[1] = g.rc_ss * g.rc_ss;
    R.rc2_bs = g.rc_bs * g.rc_bs;
    R.dt = p.dt;
    R.kappa = p.kappa;
    R.amp_t = p.pot_amp * p.pot_inv_t;
    R.inv_t = p.pot_inv_t;
    R.sigma_pol = p.sigma_pol;
    R.phi_pol = p.phi_pol;
    R.T0 = p.T0;
    R.Tf = p.Tf;
    R.t0 = p.t0;
    R.tf = p.tf;
    R.period = p.period;
    R.first_step = p.first_step;
    R.seed = p.seed;
    p_off += R.n;
    lane_off += R.lanes;
    add_blocks(s->tbl_part_h, r, R.n, 0);
    add_blocks(s->tbl_flat_h[0], r, R.n, 0, 32);
    add_blocks(s->tbl_flat_h[1], r, R.n, 0, 16);
    add_blocks(s->tbl_lane_h, r, R.lanes, 0);
    for (int k = 0; k < 2; k++)
      add_blocks(s->tbl_cell_h, r, R.ncell[k] * R.ncell[k], k,
                 kRowWaves * rows_per_wave(R.ncell[k]));
  }
  s->tbl_part_h = xcd_group(s->tbl_part_h);
  for (auto& t : s->tbl_flat_h) t = xcd_group(t);
  s->tbl_lane_h = xcd_group(s->tbl_lane_h);
  s->tbl_cell_h = xcd_group(s->tbl_cell_h);
  if (p_off > INT32_MAX / 2 || tag_off > INT32_MAX || lane_off > INT32_MAX / 2 ||
      rs_off > INT32_MAX || cell_off > INT32_MAX) {
    delete s;
    return fail(BD_E_ARG, "engine too large for 32-bit offsets");
  }
  s->n_total = (int)p_off;
  s->lanes_total = (int)lane_off;
  s->cells_total = (int)cell_off;
  s->rs_total = (int)rs_off;
  for (const RepDev& R : s->reps)
    for (int k = 0; k < 2; k++)
      if ((int64_t)R.ncell[k] * R.ncell[k] * R.ncell[k] * R.ntags[k] > (int64_t)kSlotMask)
        s->defer_ok = false;
  if (!s->defer_ok) s->defer = 0;
  for (const RepDev& R : s->reps) {
    int64_t slots = 0;
    for (int k = 0; k < 2; k++) {
      slots += (int64_t)R.ncell[k] * R.ncell[k] * R.ncell[k] * R.ntags[k];
      if (R.ntags[k] > kFlatMaxTags) s->flat_ok = false;
    }
    if (slots > (int64_t)kFlatSlotMask) s->flat_ok = false;
  }
  // 32-bit byte offsets of the row starts and of the partner records
  if ((int64_t)s->rs_total * 4 >= (1LL << 32) ||
      (int64_t)s->n_total * kFlatDepth * 4 >= (1LL << 32))
    s->flat_ok = false;
  s->tags_total = (size_t)tag_off;
  int rc = BD_OK;
#define ALLOC(ptr, n)                 \
  if (rc == BD_OK) rc = dev_alloc(&(ptr), (n));
  ALLOC(s->tbl_part, s->tbl_part_h.size());
  ALLOC(s->tbl_flat[0], s->tbl_flat_h[0].size());
  ALLOC(s->tbl_flat[1], s->tbl_flat_h[1].size());
  ALLOC(s->tbl_lane, s->tbl_lane_h.size());
  ALLOC(s->tbl_cell, s->tbl_cell_h.size());
  ALLOC(s->reps_d, (size_t)n_replicas);
  const size_t N = (size_t)s->n_total;
  ALLOC(s->px, N) ALLOC(s->py, N) ALLOC(s->pz, N);
  ALLOC(s->rx, N) ALLOC(s->ry, N) ALLOC(s->rz, N);
  ALLOC(s->ix, N) ALLOC(s->iy, N) ALLOC(s->iz, N);
  ALLOC(s->fx, N) ALLOC(s->fy, N) ALLOC(s->fz, N);
  ALLOC(s->vir, N) ALLOC(s->pot, N);
  ALLOC(s->nocup, (size_t)s->cells_total) ALLOC(s->cnt, (size_t)s->cells_total);
  ALLOC(s->ins, s->tags_total) ALLOC(s->cpos, s->tags_total);
  ALLOC(s->rs, (size_t)s->rs_total);
  ALLOC(s->ring, (size_t)bd::kRingLen * s->lanes_total) ALLOC(s->lcg, (size_t)s->lanes_total);
  ALLOC(s->draws, (size_t)n_replicas) ALLOC(s->step, 1) ALLOC(s->flags, (size_t)n_replicas);
#undef ALLOC
  if (rc != BD_OK) {
    free_all(s);
    delete s;
    return rc;
  }
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess && gder_scratch_setup(s->device) != BD_OK) {
    const std::string m = g_err;
    free_all(s);
    delete s;
    return fail(BD_E_HIP, "bd_create: " + m);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipMemcpy(s->tbl_part, s->tbl_part_h.data(), sizeof(int4) * s->tbl_part_h.size(),
                  hipMemcpyHostToDevice);
  for (int k = 0; k < 2 && e == hipSuccess; k++)
    e = hipMemcpy(s->tbl_flat[k], s->tbl_flat_h[k].data(),
                  sizeof(int4) * s->tbl_flat_h[k].size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->tbl_lane, s->tbl_lane_h.data(), sizeof(int4) * s->tbl_lane_h.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->tbl_cell, s->tbl_cell_h.data(), sizeof(int4) * s->tbl_cell_h.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->reps_d, s->reps.data(), sizeof(RepDev) * n_replicas, hipMemcpyHostToDevice);
  const size_t n8 = N * sizeof(double);
  double* zero_targets[] = {s->px, s->py, s->pz, s->rx, s->ry, s->rz, s->ix,
                            s->iy, s->iz, s->fx, s->fy, s->fz, s->vir, s->pot};
  for (double* zt : zero_targets)
    if (e == hipSuccess) e = hipMemset(zt, 0, n8);
  if (e == hipSuccess) e = hipMemset(s->nocup, 0, sizeof(int) * s->cells_total);
  if (e == hipSuccess) e = hipMemset(s->cnt, 0, sizeof(int) * s->cells_total);
  if (e == hipSuccess) e = hipMemset(s->rs, 0, sizeof(int) * s->rs_total);
  if (e == hipSuccess) e = hipMemset(s->draws, 0, sizeof(uint64_t) * n_replicas);
  if (e == hipSuccess) {
    const std::vector<RepFlags> f0((size_t)n_replicas, RepFlags{0, 0, kNoFailStep});
    e = hipMemcpy(s->flags, f0.data(), sizeof(RepFlags) * f0.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMemcpy(s->step, &s->ni, sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_rng_seed<<<(int)s->tbl_lane_h.size(), kBlk, 0, s->stream>>>(s->reps_d, s->tbl_lane,
                                                                  s->ring, s->lcg,
                                                                  s->lanes_total);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    std::string m = hipGetErrorString(e);
    free_all(s);
    delete s;
    return fail(BD_E_HIP, "bd_create: " + m);
  }
  // default force kernel: thread per particle from global memory (the
  // autotuner may pick an LDS brick where it measures faster)
  s->brick = 0;
  *out = s;
  return BD_OK;
}

int bd_destroy(bd_sys* s) {
  if (!s) return BD_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
  return BD_OK;
}

int bd_num_replicas(const bd_sys* s, int32_t* out) {
  if (!s || !out) return fail(BD_E_ARG, "null argument");
  *out = s->R;
  return BD_OK;
}

int bd_get_geometry(const bd_sys* s, int32_t replica, bd_geometry* g) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  *g = s->geom[replica];
  return BD_OK;
}

static int copy_in3(bd_sys* s, int r, const double* xyz, double* dx, double* dy, double* dz) {
  const RepDev& R = s->reps[r];
  std::vector<double> a(R.n), b(R.n), c(R.n);
  for (int i = 0; i < R.n; i++) {
    a[i] = xyz[3 * i];
    b[i] = xyz[3 * i + 1];
    c[i] = xyz[3 * i + 2];
  }
  const size_t bytes = sizeof(double) * R.n;
  HIP_TRY(hipMemcpyAsync(dx + R.p_off, a.data(), bytes, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(dy + R.p_off, b.data(), bytes, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(dz + R.p_off, c.data(), bytes, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

static int copy_out3(bd_sys* s, int r, double* xyz, const double* dx, const double* dy,
                     const double* dz) {
  const RepDev& R = s->reps[r];
  std::vector<double> a(R.n), b(R.n), c(R.n);
  const size_t bytes = sizeof(double) * R.n;
  HIP_TRY(hipMemcpyAsync(a.data(), dx + R.p_off, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(b.data(), dy + R.p_off, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(c.data(), dz + R.p_off, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int i = 0; i < R.n; i++) {
    xyz[3 * i] = a[i];
    xyz[3 * i + 1] = b[i];
    xyz[3 * i + 2] = c[i];
  }
  return BD_OK;
}

int bd_set_positions(bd_sys* s, int32_t replica, const double* xyz) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (!xyz) return fail(BD_E_ARG, "null positions");
  rc = copy_in3(s, replica, xyz, s->px, s->py, s->pz);
  if (rc == BD_OK) rc = copy_in3(s, replica, xyz, s->rx, s->ry, s->rz);
  if (rc == BD_OK) rc = copy_in3(s, replica, xyz, s->ix, s->iy, s->iz);
  if (rc == BD_OK) rc = clear_flags(s, replica, replica + 1);
  if (rc == BD_OK) {
    s->have_pos[replica] = true;
    s->located = false;
    rc = list_invalidate(s);
  }
  return rc;
}

int bd_get_positions(bd_sys* s, int32_t replica, double* xyz) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  return copy_out3(s, replica, xyz, s->px, s->py, s->pz);
}

int bd_get_raw_positions(bd_sys* s, int32_t replica, double* xyz) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  return copy_out3(s, replica, xyz, s->rx, s->ry, s->rz);
}

int bd_get_forces(bd_sys* s, int32_t replica, double* f, double* vir, double* pot) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  const RepDev& R = s->reps[replica];
  if (f) rc = copy_out3(s, replica, f, s->fx, s->fy, s->fz);
  if (rc != BD_OK) return rc;
  const size_t bytes = sizeof(double) * R.n;
  if (vir) HIP_TRY(hipMemcpyAsync(vir, s->vir + R.p_off, bytes, hipMemcpyDeviceToHost, s->stream));
  if (pot) HIP_TRY(hipMemcpyAsync(pot, s->pot + R.p_off, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

int bd_set_forces(bd_sys* s, int32_t replica, const double* f) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (!f) return fail(BD_E_ARG, "null forces");
  return copy_in3(s, replica, f, s->fx, s->fy, s->fz);
}

int bd_mark_origin(bd_sys* s) {
  if (!s) return fail(BD_E_ARG, "null handle");
  const size_t bytes = sizeof(double) * s->n_total;
  double* src[3] = {s->px, s->py, s->pz};
  double* raw[3] = {s->rx, s->ry, s->rz};
  double* ini[3] = {s->ix, s->iy, s->iz};
  for (int k = 0; k < 3; k++) {
    HIP_TRY(hipMemcpyAsync(raw[k], src[k], bytes, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(ini[k], src[k], bytes, hipMemcpyDeviceToDevice, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

int bd_get_step(const bd_sys* s, int64_t* ni) {
  if (!s || !ni) return fail(BD_E_ARG, "null argument");
  *ni = s->ni;
  return BD_OK;
}

static int require_positions(const bd_sys* s) {
  for (int r = 0; r < s->R; r++)
    if (!s->have_pos[r]) return fail(BD_E_STATE, "positions of replica " + std::to_string(r) +
                                                     " were never set");
  return BD_OK;
}

int bd_cell_locate(bd_sys* s) {
  if (!s) return fail(BD_E_ARG, "null handle");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  launch_bin(s);
  launch_cells(s, 0);
  HIP_TRY(hipGetLastError());
  rc = check_flags(s);
  if (rc == BD_OK) s->located = true;
  return rc;
}

int bd_get_cells(bd_sys* s, int32_t replica, int32_t species, int32_t* nocup,
                 int32_t* cell_vec) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (species != 0 && species != 1) return fail(BD_E_ARG, "species must be 0 or 1");
  if (!s->located) return fail(BD_E_STATE, "no binning yet");
  const RepDev& R = s->reps[replica];
  const int n = R.ncell[species];
  const size_t nc3 = (size_t)n * n * n;
  const int ntags = R.ntags[species];
  const int sp_off = species == 0 ? 0 : R.nb;
  std::vector<int32_t> rs((size_t)n * n * (n + 1)), occ_int(nc3);
  std::vector<double4> cp(nc3 * ntags);
  HIP_TRY(hipMemcpyAsync(occ_int.data(), s->cnt + R.cell_off[species], sizeof(int32_t) * nc3,
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(rs.data(), s->rs + R.rs_off[species], sizeof(int32_t) * rs.size(),
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(cp.data(), s->cpos + R.tag_off[species], sizeof(double4) * cp.size(),
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  // device cells are z-fastest with packed rows; the reference's are x-fastest
  // with ntags slots per cell
  for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n; x++) {
        const size_t c = (size_t)x + (size_t)n * (y + (size_t)n * z);
        const size_t row = (size_t)x + (size_t)n * y;
        nocup[c] = occ_int[row * n + z];
        const size_t ri = row * (n + 1) + z;
        const int s0 = rs[ri], occ = rs[ri + 1] - s0;
        for (int t = 0; t < ntags; t++)
          cell_vec[c * ntags + t] = t < occ ? (int32_t)cp[(size_t)s0 + t].w - sp_off : -1;
      }
  return BD_OK;
}

int bd_max_occupancy(bd_sys* s, int32_t replica, int32_t* big, int32_t* sml) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (!big || !sml) return fail(BD_E_ARG, "null output");
  if (!s->located) return fail(BD_E_STATE, "no binning yet");
  const RepDev& R = s->reps[replica];
  int32_t* out[2] = {big, sml};
  for (int g = 0; g < 2; g++) {
    const size_t nc3 = (size_t)R.ncell[g] * R.ncell[g] * R.ncell[g];
    std::vector<int32_t> occ(nc3);
    HIP_TRY(hipMemcpyAsync(occ.data(), s->cnt + R.cell_off[g], sizeof(int32_t) * nc3,
                           hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *out[g] = nc3 ? *std::max_element(occ.begin(), occ.end()) : 0;
  }
  return BD_OK;
}

int bd_compute_forces(bd_sys* s) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (!s->located) return fail(BD_E_STATE, "bd_cell_locate must run first");
  launch_forces(s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

int bd_neighbours(bd_sys* s, int32_t replica, int32_t maxn, int32_t* counts, int32_t* list) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (!s->located) return fail(BD_E_STATE, "bd_cell_locate must run first");
  if (maxn < 1 || !counts || !list) return fail(BD_E_ARG, "bad neighbour buffers");
  const RepDev& R = s->reps[replica];
  std::vector<int4> tbl;
  add_blocks(tbl, replica, R.n, 0);
  int4* tbl_d = nullptr;
  int *c_d = nullptr, *l_d = nullptr;
  rc = dev_alloc(&tbl_d, tbl.size());
  if (rc == BD_OK) rc = dev_alloc(&c_d, (size_t)R.n);
  if (rc == BD_OK) rc = dev_alloc(&l_d, (size_t)R.n * maxn);
  hipError_t e = hipSuccess;
  if (rc == BD_OK) {
    e = hipMemcpy(tbl_d, tbl.data(), sizeof(int4) * tbl.size(), hipMemcpyHostToDevice);
    // forces are recomputed as a side effect (same values)
    if (e == hipSuccess) {
      launch_forces_t<true>(s, c_d, l_d, maxn, tbl_d, (int)tbl.size());
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(counts, c_d, sizeof(int) * R.n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(list, l_d, sizeof(int) * (size_t)R.n * maxn, hipMemcpyDeviceToHost,
                         s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  if (tbl_d) (void)hipFree(tbl_d);
  if (c_d) (void)hipFree(c_d);
  if (l_d) (void)hipFree(l_d);
  if (rc != BD_OK) return rc;
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_neighbours: ") + hipGetErrorString(e));
  for (int i = 0; i < R.n; i++)
    if (counts[i] > maxn) return fail(BD_E_ARG, "maxn too small");
  return BD_OK;
}

int bd_update(bd_sys* s) {
  if (!s) return fail(BD_E_ARG, "null handle");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  launch_update(s, 1);
  launch_cells(s, 1);
  HIP_TRY(hipGetLastError());
  s->ni++;
  rc = check_flags(s);
  if (rc == BD_OK) s->located = true;
  return rc;
}

// Capture `steps` steps into an executable graph: whole steps, or with
// `fused` the steady state of the fused update, `steps` pairs of (forces of
// the current positions + the next move, binning).  The graph of the other
// kind stays cached.  A graph is replaced only after the stream has drained,
// so no queued replay outlives its executable graph.
static int build_graph(bd_sys* s, int steps, bool fused) {
  bd_sys::StepGraph& sg = fused ? s->graph_fused : s->graph_plain;
  if (sg.exec && sg.steps == steps) return BD_OK;
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (sg.exec) {
    HIP_TRY(hipGraphExecDestroy(sg.exec));
    sg.exec = nullptr;
    sg.steps = 0;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  if (fused) {
    for (int j = 0; j < steps; j++) {
      launch_forces_move(s);
      launch_cells(s, 1);
    }
  } else {
    enqueue_steps(s, steps);
  }
  HIP_TRY(hipStreamEndCapture(s->stream, &g));
  hipError_t e = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    sg.exec = nullptr;
    return fail(BD_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  sg.steps = steps;
  return BD_OK;
}

// How bd_step runs nsteps with the current settings: the fused path replays
// (forces + move, binning) pairs G at a time when nsteps > G (the call's
// first update and last forces run outside the graph); otherwise whole
// steps are replayed G at a time and the remainder runs eagerly.
static bool step_uses_fused(const bd_sys* s, int nsteps) {
  return s->graph_mode > 0 && nsteps > s->graph_mode && fusable(s);
}

static int prepare_graphs(bd_sys* s, int nsteps) {
  const int G = s->graph_mode;
  if (s->persist > 0 || G <= 0) return BD_OK;
  if (step_uses_fused(s, nsteps)) return build_graph(s, G, true);
  if (nsteps >= G) return build_graph(s, G, false);
  return BD_OK;
}

int bd_prepare_step(bd_sys* s, int32_t nsteps) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (nsteps < 0) return fail(BD_E_ARG, "nsteps must be >= 0");
  return prepare_graphs(s, nsteps);
}

int bd_step(bd_sys* s, int32_t nsteps) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (nsteps < 0) return fail(BD_E_ARG, "nsteps must be >= 0");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  // the graphs are built first, so that a failure leaves no step half taken
  rc = prepare_graphs(s, nsteps);
  if (rc != BD_OK) return rc;
  if (!s->located) {
    launch_bin(s);
    launch_cells(s, 0);
    launch_forces(s);
    s->located = true;
  }
  if (s->persist > 0) {
    if (s->order != BD_ORDER_REFERENCE)
      return fail(BD_E_STATE, "the persistent step kernel computes forces in reference order");
    for (int left = nsteps; left > 0;) {
      const int k = left < s->persist ? left : s->persist;
      rc = launch_persistent(s, k);
      if (rc != BD_OK) return rc;
      left -= k;
    }
    HIP_TRY(hipGetLastError());
    s->ni += nsteps;
    rc = check_flags(s);
    // the persistent kernel keeps no neighbour lists and writes no pos4
    if (rc == BD_OK) rc = list_refresh(s);
    return rc == BD_OK ? list_invalidate(s) : rc;
  }
  const int G = s->graph_mode;
  if (step_uses_fused(s, nsteps)) {
    // the kernels of enqueue_steps(s, nsteps) in one run: the first step's
    // update, nsteps - 1 fused pairs (graph-replayed G at a time), the forces
    // of the final positions
    launch_update(s, 1);
    launch_cells(s, 1);
    int pairs = nsteps - 1;
    for (; pairs >= G; pairs -= G) HIP_TRY(hipGraphLaunch(s->graph_fused.exec, s->stream));
    for (; pairs > 0; pairs--) {
      launch_forces_move(s);
      launch_cells(s, 1);
    }
    launch_forces(s);
    HIP_TRY(hipGetLastError());
    s->ni += nsteps;
    return check_flags(s);
  }
  int left = nsteps;
  if (G > 0)
    for (; left >= G; left -= G) HIP_TRY(hipGraphLaunch(s->graph_plain.exec, s->stream));
  enqueue_steps(s, left);
  HIP_TRY(hipGetLastError());
  s->ni += nsteps;
  return check_flags(s);
}

int bd_set_graph_steps(bd_sys* s, int32_t steps) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (steps < 0 || steps > 256) return fail(BD_E_ARG, "graph steps must be 0..256");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->graph_mode = steps;
  return BD_OK;
}

int bd_set_persistent(bd_sys* s, int32_t steps) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (steps < 0 || steps > 100000) return fail(BD_E_ARG, "steps must be 0..100000");
  if (steps > 0 && !s->defer_ok)
    return fail(BD_E_ARG, "a grid has too many slots for the persistent kernel's force walk");
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->persist = steps;
  return BD_OK;
}

int bd_get_persistent(const bd_sys* s, int32_t* steps) {
  if (!s || !steps) return fail(BD_E_ARG, "null argument");
  *steps = s->persist;
  return BD_OK;
}

int bd_get_graph_steps(const bd_sys* s, int32_t* steps) {
  if (!s || !steps) return fail(BD_E_ARG, "null argument");
  *steps = s->graph_mode;
  return BD_OK;
}

int bd_set_brick(bd_sys* s, int32_t brick) {
  if (!s) return fail(BD_E_ARG, "null handle");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  int rc = setup_bricks(s, brick);
  if (rc == BD_OK) rc = list_refresh(s);
  return rc == BD_OK ? list_invalidate(s) : rc;
}

int bd_set_force_order(bd_sys* s, int32_t order) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (order != BD_ORDER_REFERENCE && order != BD_ORDER_SPLIT4 && order != BD_ORDER_SPLIT8)
    return fail(BD_E_ARG, "order must be BD_ORDER_REFERENCE, _SPLIT4 or _SPLIT8");
  if (order != BD_ORDER_REFERENCE && phs_kind(s->pot_kind))
    return fail(BD_E_ARG, "the 49-50 potentials run in the reference order only");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  if (order != BD_ORDER_REFERENCE) {
    for (int r = 0; r < s->R; r++)
      if (!s->reps[r].fast)
        return fail(BD_E_ARG, "split order needs grids of at least 2*nrange+3 cells");
    std::vector<int4> tbl;
    for (int r = 0; r < s->R; r++) add_blocks(tbl, r, s->reps[r].n, 0, kBlk / order);
    tbl = xcd_group(tbl);
    if (s->tbl_split) (void)hipFree(s->tbl_split);
    s->tbl_split = nullptr;
    int rc = dev_alloc(&s->tbl_split, tbl.size());
    if (rc != BD_OK) return rc;
    HIP_TRY(hipMemcpy(s->tbl_split, tbl.data(), sizeof(int4) * tbl.size(), hipMemcpyHostToDevice));
    s->tbl_split_h = tbl;
  }
  s->order = order;
  const int rc = list_refresh(s);
  return rc == BD_OK ? list_invalidate(s) : rc;
}

// Autotuner: average duration of the force kernel for every brick shape the
// engine accepts (all give bit-identical results), forces-only launches on
// the current cells so the simulation state is unchanged.  Choices are cached
// per (replicas, N bucket, phi bucket, potential, order) in a text file.
static std::string tune_key(const bd_sys* s) {
  double vol_p = 0.0, vol_box = 0.0;
  for (const bd_params& p : s->params) {
    vol_p += M_PI / 6.0 *
             (p.n_big * std::pow(p.sigma_big, 3.0) + p.n_sml * std::pow(p.sigma_sml, 3.0));
    vol_box += p.side * p.side * p.side;
  }
  const int nb = (int)std::lround(4.0 * std::log2((double)std::max(1, s->n_total)));
  char buf[160];
  std::snprintf(buf, sizeof buf, "R%d_N%d_phi%.2f_pot%d_order%d", s->R, nb, vol_p / vol_box,
                s->pot_kind, s->order);
  return buf;
}

int bd_autotune(bd_sys* s, int32_t trial_launches, const char* cache_path, int32_t* chosen) {
  if (!s || trial_launches < 1) return fail(BD_E_ARG, "bad arguments");
  const std::string key = tune_key(s);
  if (cache_path) {
    if (FILE* f = std::fopen(cache_path, "r")) {
      // lines: key brick depth walk split (a line without walk: brick -1 =
      // slab walk; without split: one wavefront)
      char line[320], k[200];
      int found_b = -2, found_d = 0, found_w = 0, found_p = 0;
      while (std::fgets(line, sizeof line, f)) {
        int b, d, w, sp = 0;
        const int nf = std::sscanf(line, "%199s %d %d %d %d", k, &b, &d, &w, &sp);
        if (nf < 3 || key != k) continue;
        if (nf == 3) w = b < 0 ? BD_WALK_SLABS : BD_WALK_ROWS;
        found_b = b < 0 ? 0 : b, found_d = d, found_w = w, found_p = nf >= 5 ? sp : 0;
      }
      std::fclose(f);
      if (found_b >= 0 && bd_set_defer_depth(s, found_d) == BD_OK &&
          bd_set_walk(s, found_w) == BD_OK && bd_set_pass_split(s, found_p) == BD_OK &&
          bd_set_brick(s, found_b) == BD_OK) {
        if (chosen) *chosen = found_b;
        return BD_OK;
      }
    }
  }
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  if (!s->located) {
    launch_bin(s);
    launch_cells(s, 0);
    s->located = true;
  }
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  // kernel shapes: the row, slab and chunked slab walks at three deferral
  // depths, the LDS bricks, the flat walk.  The list walk is not among them: forces-only
  // launches between rebuilds would not see its rebuild cost, which decides
  // it (measured slower on the workloads here).
  struct Shape {
    int brick, defer, walk, split;
  };
  const Shape candidates[] = {
      {0, 0, BD_WALK_ROWS, 0},    {0, 16, BD_WALK_ROWS, 0},   {0, 32, BD_WALK_ROWS, 0},
      {0, 0, BD_WALK_SLABS, 0},   {0, 16, BD_WALK_SLABS, 0},  {0, 32, BD_WALK_SLABS, 0},
      {0, 0, BD_WALK_CHUNK2, 0},  {0, 16, BD_WALK_CHUNK2, 0}, {0, 32, BD_WALK_CHUNK2, 0},
      {0, 0, BD_WALK_CHUNK4, 0},  {0, 16, BD_WALK_CHUNK4, 0}, {0, 32, BD_WALK_CHUNK4, 0},
      {0, 16, BD_WALK_SLABS, 1},  {0, 32, BD_WALK_SLABS, 1},  {0, 16, BD_WALK_CHUNK2, 1},
      {0, 32, BD_WALK_CHUNK2, 1}, {0, 16, BD_WALK_CHUNK4, 1}, {0, 32, BD_WALK_CHUNK4, 1},
      {0, 16, BD_WALK_SLABS, 2},  {0, 32, BD_WALK_SLABS, 2},  {0, 16, BD_WALK_CHUNK2, 2},
      {0, 32, BD_WALK_CHUNK2, 2}, {0, 16, BD_WALK_CHUNK4, 2}, {0, 32, BD_WALK_CHUNK4, 2},
      {1, 0, BD_WALK_ROWS, 0},    {2, 0, BD_WALK_ROWS, 0},    {3, 0, BD_WALK_ROWS, 0},
      {4, 0, BD_WALK_ROWS, 0},    {0, 0, BD_WALK_FLAT, 0},    {0, 0, BD_WALK_FLAT, 1},
      {0, 0, BD_WALK_FLAT, 2}};
  Shape best{0, 0, BD_WALK_ROWS, 0};
  double best_ms = 1e300;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const Shape keep{s->brick, s->defer, s->walk, s->pass_split};
  for (const Shape& c : candidates) {
    if (s->order != BD_ORDER_REFERENCE && (c.brick != keep.brick || c.defer != keep.defer ||
                                           c.walk != keep.walk || c.split != keep.split))
      continue;
    if (c.defer != 0 && !s->defer_ok) continue;
    if (c.walk == BD_WALK_FLAT && (phs_kind(s->pot_kind) || flat_setup(s) != BD_OK)) continue;
    // the 49-50 potentials run one kernel whatever the shape: time it once
    if (phs_kind(s->pot_kind) && (c.brick != 0 || c.defer != 16 || c.walk != BD_WALK_SLABS ||
                                  c.split != 0))
      continue;
    if (setup_bricks(s, c.brick) != BD_OK) continue;
    s->walk = c.walk;
    s->defer = c.defer;
    s->pass_split = c.split;
    launch_forces(s);  // warm-up
    HIP_TRY(hipEventRecord(e0, s->stream));
    for (int k = 0; k < trial_launches; k++) launch_forces(s);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_err.clear();
  s->defer = best.defer;
  s->walk = best.walk;
  s->pass_split = best.split;
  rc = list_invalidate(s);
  if (rc == BD_OK) rc = bd_set_brick(s, best.brick);
  if (rc != BD_OK) {
    s->defer = keep.defer;
    s->walk = keep.walk;
    s->pass_split = keep.split;
    bd_set_brick(s, keep.brick);
    return rc;
  }
  if (cache_path)
    if (FILE* f = std::fopen(cache_path, "a")) {
      std::fprintf(f, "%s %d %d %d %d\n", key.c_str(), best.brick, best.defer, best.walk,
                   best.split);
      std::fclose(f);
    }
  if (chosen) *chosen = best.brick;
  return BD_OK;
}

int bd_set_flat_depth(bd_sys* s, int32_t depth) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (depth < 1 || depth > kFlatDepth) return fail(BD_E_ARG, "depth must be 1..32");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->flat_depth = depth;
  return BD_OK;
}

int bd_set_walk(bd_sys* s, int32_t walk) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (walk != BD_WALK_ROWS && walk != BD_WALK_SLABS && walk != BD_WALK_CHUNK2 &&
      walk != BD_WALK_CHUNK4 && walk != BD_WALK_LIST && walk != BD_WALK_FLAT)
    return fail(BD_E_ARG, "walk must be BD_WALK_ROWS, _SLABS, _CHUNK2, _CHUNK4, _LIST or _FLAT");
  if (walk == BD_WALK_FLAT) {
    const int rc = flat_setup(s);
    if (rc != BD_OK) return rc;
  }
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  if (walk == BD_WALK_LIST) {
    const int rc = list_setup(s);
    if (rc != BD_OK) return rc;
  }
  s->walk = walk;
  const int rc = list_refresh(s);
  return rc == BD_OK ? list_invalidate(s) : rc;
}

int bd_set_list_skin(bd_sys* s, double skin) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (!(skin > 0.0 && skin <= 2.0)) return fail(BD_E_ARG, "skin must lie in (0, 2]");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->skin = skin;
  return s->nlist ? list_setup(s) : BD_OK;
}

int bd_get_list_skin(const bd_sys* s, double* skin) {
  if (!s || !skin) return fail(BD_E_ARG, "null argument");
  *skin = s->skin;
  return BD_OK;
}

int bd_get_walk(const bd_sys* s, int32_t* walk) {
  if (!s || !walk) return fail(BD_E_ARG, "null argument");
  *walk = s->walk;
  return BD_OK;
}

int bd_set_pass_split(bd_sys* s, int32_t split) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (split < 0 || split > 2) return fail(BD_E_ARG, "pass split must be 0, 1 or 2");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->pass_split = split;
  return BD_OK;
}

int bd_get_pass_split(const bd_sys* s, int32_t* split) {
  if (!s || !split) return fail(BD_E_ARG, "null argument");
  *split = s->pass_split;
  return BD_OK;
}

int bd_set_fused_update(bd_sys* s, int32_t on) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (on < 0 || on > 2) return fail(BD_E_ARG, "fused update must be 0, 1 or 2");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->fuse = on;
  return BD_OK;
}

int bd_get_fused_update(const bd_sys* s, int32_t* on) {
  if (!s || !on) return fail(BD_E_ARG, "null argument");
  *on = fusable(s) ? 1 : 0;
  return BD_OK;
}

int bd_set_defer_depth(bd_sys* s, int32_t depth) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (depth != 0 && depth != 16 && depth != 32) return fail(BD_E_ARG, "depth must be 0, 16 or 32");
  if (depth != 0 && !s->defer_ok)
    return fail(BD_E_ARG, "a grid has too many slots for deferred pair records");
  {
    const int rc_ = drop_graphs(s);
    if (rc_ != BD_OK) return rc_;
  }
  s->defer = depth;
  return BD_OK;
}

int bd_get_defer_depth(const bd_sys* s, int32_t* depth) {
  if (!s || !depth) return fail(BD_E_ARG, "null argument");
  *depth = s->defer;
  return BD_OK;
}

int bd_get_force_order(const bd_sys* s, int32_t* order) {
  if (!s || !order) return fail(BD_E_ARG, "null argument");
  *order = s->order;
  return BD_OK;
}

int bd_get_brick(const bd_sys* s, int32_t* brick) {
  if (!s || !brick) return fail(BD_E_ARG, "null argument");
  *brick = s->brick;
  return BD_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint / resume.  File layout (native endianness):
//   "BDCKPT01", u32 version = 1, i32 replicas, i64 step, u64 particles,
//   u64 lanes, u32 has_origins, u32 0, then bd_params[replicas], then the
//   FP64 arrays pos x,y,z, raw x,y,z, ini x,y,z, F x,y,z, vir, pot
//   (particles each), the rannew64 rings (250 x lanes u64), lcg (lanes),
//   draws (replicas) and, if stored, the waiting-time origins
//   (BD_MAX_ORIGINS x 3 x particles FP64).
// ---------------------------------------------------------------------------

int bd_checkpoint_save(bd_sys* s, const char* path) {
  if (!s || !path) return fail(BD_E_ARG, "null argument");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(BD_E_ARG, std::string("cannot open ") + path);
  CkptHeader h{};
  std::memcpy(h.magic, "BDCKPT01", 8);
  h.version = 1;
  h.replicas = s->R;
  h.step = s->ni;
  h.particles = (uint64_t)s->n_total;
  h.lanes = (uint64_t)s->lanes_total;
  h.has_origins = s->origins ? 1u : 0u;
  if (std::fwrite(&h, sizeof(h), 1, f) != 1 ||
      std::fwrite(s->params.data(), sizeof(bd_params), s->params.size(), f) != s->params.size())
    rc = fail(BD_E_ARG, "checkpoint write failed");
  if (rc == BD_OK) rc = ckpt_arrays(s, f, true);
  if (std::fclose(f) != 0 && rc == BD_OK) rc = fail(BD_E_ARG, "checkpoint close failed");
  return rc;
}

int bd_checkpoint_load(bd_sys* s, const char* path) {
  if (!s || !path) return fail(BD_E_ARG, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(BD_E_ARG, std::string("cannot open ") + path);
  CkptHeader h{};
  int rc = BD_OK;
  std::vector<bd_params> ps(s->R);
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, "BDCKPT01", 8) != 0 ||
      h.version != 1)
    rc = fail(BD_E_ARG, "not a checkpoint of this engine version");
  else if (h.replicas != s->R || h.particles != (uint64_t)s->n_total ||
           h.lanes != (uint64_t)s->lanes_total)
    rc = fail(BD_E_ARG, "checkpoint holds a different number of replicas, particles or lanes");
  else if (std::fread(ps.data(), sizeof(bd_params), ps.size(), f) != ps.size() ||
           std::memcmp(ps.data(), s->params.data(), sizeof(bd_params) * ps.size()) != 0)
    rc = fail(BD_E_ARG, "checkpoint parameters differ from the engine's");
  if (rc == BD_OK) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (h.has_origins && !s->origins)
      rc = dev_alloc(&s->origins, (size_t)BD_MAX_ORIGINS * 3 * s->n_total);
    if (rc == BD_OK && !h.has_origins && s->origins) {
      (void)hipFree(s->origins);
      s->origins = nullptr;
    }
  }
  if (rc == BD_OK) rc = ckpt_arrays(s, f, false);
  std::fclose(f);
  if (rc != BD_OK) return rc;
  s->ni = h.step;
  HIP_TRY(hipMemcpy(s->step, &s->ni, sizeof(int64_t), hipMemcpyHostToDevice));
  rc = clear_flags(s, 0, s->R);
  if (rc != BD_OK) return rc;
  for (int r = 0; r < s->R; r++) s->have_pos[r] = true;
  s->located = false;  // cells are rebuilt (and the stored forces recomputed) on the next step
  return list_invalidate(s);
}

// Kernel durations for the timing entry points: each launch between
// timed_begin(k) and timed_end() is stamped into event pair k.
struct KernelTimer {
  bd_sys* s;
  std::vector<hipEvent_t> ev;
  hipError_t e = hipSuccess;
  KernelTimer(bd_sys* sys, int pairs) : s(sys), ev(2 * pairs, nullptr) {
    for (auto& x : ev)
      if (e == hipSuccess) e = hipEventCreate(&x);
  }
  ~KernelTimer() {
    s->tev[0] = s->tev[1] = nullptr;
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
  }
  void begin(int k) {
    s->tev[0] = ev[2 * k];
    s->tev[1] = ev[2 * k + 1];
  }
  void end() { s->tev[0] = s->tev[1] = nullptr; }
  // adds pair k's milliseconds to *acc (after a synchronisation)
  void add(int k, double* acc) {
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]);
    *acc += (double)t;
  }
};

int bd_stream_bandwidth(int64_t bytes, int32_t reps, double* gbs) {
  if (bytes < (int64_t)(1 << 20) || reps < 1 || !gbs) return fail(BD_E_ARG, "bad arguments");
  const size_t n = (size_t)bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&a, n * sizeof(double4));
  if (e == hipSuccess) e = hipMalloc((void**)&b, n * sizeof(double4));
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * sizeof(double4), st);
  int dev = 0, ncu = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int nblk = (int)std::min<size_t>((size_t)ncu * 16, (n + 255) / 256);
  double total_ms = 0.0;
  for (int k = 0; k <= reps && e == hipSuccess; k++) {  // launch 0 warms up
    hipExtLaunchKernelGGL(k_stream_copy, dim3(nblk), dim3(256), 0, st, e0, e1, 0,
                          (const double4*)a, b, n);
    e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (k > 0) total_ms += (double)t;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_stream_bandwidth: ") + hipGetErrorString(e));
  *gbs = 2.0 * (double)(n * sizeof(double4)) * reps / (total_ms * 1e-3) / 1e9;
  return BD_OK;
}

int bd_launch_floor(int32_t workgroups, int32_t threads, int32_t reps, double* us) {
  if (workgroups < 1 || workgroups > (1 << 20) || threads < 64 || threads > 1024 ||
      threads % 64 != 0 || reps < 1 || reps > 100000 || !us)
    return fail(BD_E_ARG, "bad arguments");
  int* done = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&done, sizeof(int));
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double total_ms = 0.0;
  for (int k = 0; k <= reps && e == hipSuccess; k++) {  // launch 0 warms up
    hipExtLaunchKernelGGL(k_launch_floor, dim3(workgroups), dim3(threads), 0, st, e0, e1, 0,
                          done);
    e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (k > 0) total_ms += (double)t;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (done) (void)hipFree(done);
  if (e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_launch_floor: ") + hipGetErrorString(e));
  *us = total_ms * 1e3 / reps;
  return BD_OK;
}

int bd_grid_barrier_cost(int32_t workgroups, int32_t barriers, double* us,
                         int32_t* max_workgroups) {
  if (workgroups < 1 || barriers < 1 || barriers > 1000000 || !us)
    return fail(BD_E_ARG, "bad arguments");
  const void* fn = reinterpret_cast<const void*>(&k_grid_barrier);
  int dev = 0, coop = 0, ncu = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return fail(BD_E_STATE, "the device does not support cooperative launches");
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlk, 0));
  if (max_workgroups) *max_workgroups = per_cu * ncu;
  // a cooperative grid must be co-resident, or its barriers would never end
  if (workgroups > per_cu * ncu)
    return fail(BD_E_ARG, "more workgroups than are co-resident on the device");
  int* done = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&done, sizeof(int));
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int nb = barriers;
  void* args[] = {&nb, &done};
  double ms = 0.0;
  for (int k = 0; k < 2 && e == hipSuccess; k++) {  // launch 0 warms up
    e = hipEventRecord(e0, st);
    if (e == hipSuccess)
      e = hipLaunchCooperativeKernel(fn, dim3(workgroups), dim3(kBlk), args, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    ms = (double)t;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (done) (void)hipFree(done);
  if (e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_grid_barrier_cost: ") + hipGetErrorString(e));
  *us = ms * 1e3 / barriers;
  return BD_OK;
}

int bd_kernel_times(bd_sys* s, int32_t nsteps, double* ms) {
  if (!s || !ms || nsteps < 0) return fail(BD_E_ARG, "bad arguments");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  if (!s->located) {
    launch_bin(s);
    launch_cells(s, 0);
    launch_forces(s);
    s->located = true;
  }
  KernelTimer kt(s, 3);
  for (int i = 0; i < nsteps && kt.e == hipSuccess; i++) {
    kt.begin(0);
    launch_update(s, 1);
    kt.begin(1);
    launch_cells(s, 1);
    kt.begin(2);
    launch_forces(s);
    kt.end();
    if (kt.e == hipSuccess) kt.e = hipStreamSynchronize(s->stream);
    for (int k = 0; k < 3; k++) kt.add(k, &ms[k]);
    s->ni++;
  }
  if (kt.e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_kernel_times: ") + hipGetErrorString(kt.e));
  return check_flags(s);
}

int bd_kernel_times_fused(bd_sys* s, int32_t nsteps, double* ms) {
  if (!s || !ms || nsteps < 2) return fail(BD_E_ARG, "bad arguments (nsteps >= 2)");
  if (!fusable(s)) return fail(BD_E_STATE, "the fused update does not apply to these settings");
  int rc = require_positions(s);
  if (rc != BD_OK) return rc;
  if (!s->located) {
    launch_bin(s);
    launch_cells(s, 0);
    launch_forces(s);
    s->located = true;
  }
  KernelTimer kt(s, 2);
  launch_update(s, 1);
  for (int i = 0; i < nsteps && kt.e == hipSuccess; i++) {
    const bool last = i + 1 == nsteps;
    kt.begin(0);
    launch_cells(s, 1);
    kt.begin(1);
    if (last)
      launch_forces(s);
    else
      launch_forces_move(s);
    kt.end();
    if (kt.e == hipSuccess) kt.e = hipStreamSynchronize(s->stream);
    kt.add(0, &ms[0]);
    if (!last) kt.add(1, &ms[1]);
    s->ni++;
  }
  if (kt.e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_kernel_times_fused: ") + hipGetErrorString(kt.e));
  return check_flags(s);
}

int bd_get_rng_lane(bd_sys* s, int32_t replica, int32_t lane, uint64_t* ring250, uint64_t* lcg,
                    uint64_t* draws) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  const RepDev& R = s->reps[replica];
  if (lane < 0 || lane >= R.lanes) return fail(BD_E_ARG, "lane out of range");
  const int gl = R.lane_off + lane;
  for (int i = 0; i < bd::kRingLen; i++)
    HIP_TRY(hipMemcpyAsync(&ring250[i], s->ring + (size_t)i * s->lanes_total + gl,
                           sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(lcg, s->lcg + gl, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(draws, s->draws + replica, sizeof(uint64_t), hipMemcpyDeviceToHost,
                         s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

int bd_rng_stream(uint64_t seed, int32_t lanes, int32_t ndraws, uint64_t* out) {
  if (lanes < 1 || ndraws < 0 || !out) return fail(BD_E_ARG, "bad arguments");
  uint64_t *ring = nullptr, *o = nullptr;
  int rc = dev_alloc(&ring, (size_t)bd::kRingLen * lanes);
  if (rc == BD_OK) rc = dev_alloc(&o, (size_t)lanes * ndraws);
  hipError_t e = hipSuccess;
  if (rc == BD_OK) {
    k_rng_stream<<<(lanes + 255) / 256, 256>>>(seed, lanes, ndraws, ring, o);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpy(out, o, sizeof(uint64_t) * (size_t)lanes * ndraws, hipMemcpyDeviceToHost);
  }
  if (ring) (void)hipFree(ring);
  if (o) (void)hipFree(o);
  if (rc != BD_OK) return rc;
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_rng_stream: ") + hipGetErrorString(e));
  return BD_OK;
}

int bd_gder(bd_sys* s, int32_t replica, int32_t nbins, double range, uint64_t* counts) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (nbins < 1 || nbins > kGderMaxBins || !(range > 0.0) || !counts)
    return fail(BD_E_ARG, "nbins must be in 1..512 and range > 0");
  const RepDev& R = s->reps[replica];
  unsigned long long* c_d = nullptr;
  rc = call_buffer(s, (size_t)4 * nbins, &c_d);
  if (rc != BD_OK) return rc;
  hipError_t e = hipMemsetAsync(c_d, 0, sizeof(unsigned long long) * 4 * nbins, s->stream);
  const int iblocks = (R.n + kGderThreads - 1) / kGderThreads;
  int jsplit = std::max(1, std::min(64, 4096 / std::max(1, iblocks)));
  const int j_chunk = ((R.n + jsplit - 1) / jsplit + kGderThreads - 1) / kGderThreads * kGderThreads;
  jsplit = (R.n + j_chunk - 1) / j_chunk;
  const size_t shmem = sizeof(unsigned int) * (kGderThreads / 64) * 4 * nbins;
  const double bin_size = range / (double)nbins;
  if (e == hipSuccess) {
    k_gder<<<dim3(iblocks, jsplit), kGderThreads, shmem, s->stream>>>(
        s->reps_d, replica, s->px, s->py, s->pz, nbins, range, bin_size, j_chunk, c_d);
    e = hipGetLastError();
  }
  std::vector<unsigned long long> h((size_t)4 * nbins);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), c_d, sizeof(unsigned long long) * 4 * nbins,
                       hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_gder: ") + hipGetErrorString(e));
  for (size_t k = 0; k < h.size(); k++) counts[k] += h[k];
  return BD_OK;
}

// bd_gder without any allocation: the histogram is the fixed device array
// `scratch` of the engine's GPU (through its looked-up address), loaded with
// the caller's counts, added to by the same kernel and copied back.  Calls
// that overlap on one GPU share that array and race.
int bd_gder_noalloc(bd_sys* s, int32_t replica, int32_t nbins, double range, uint64_t* counts) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (nbins < 1 || nbins > kGderMaxBins || 4 * nbins > kScratchWords || !(range > 0.0) ||
      !counts)
    return fail(BD_E_ARG, "nbins must be in 1..512 (the fixed histogram holds 4 x 512 bins) "
                          "and range > 0");
  unsigned long long* c_d = static_cast<unsigned long long*>(gder_scratch_ptr(s->device));
  if (!c_d) return fail(BD_E_STATE, "no fixed histogram on this device");
  const RepDev& R = s->reps[replica];
  const size_t bytes = sizeof(unsigned long long) * 4 * (size_t)nbins;
  hipError_t e = hipMemcpyAsync(c_d, counts, bytes, hipMemcpyHostToDevice, s->stream);
  const int iblocks = (R.n + kGderThreads - 1) / kGderThreads;
  int jsplit = std::max(1, std::min(64, 4096 / std::max(1, iblocks)));
  const int j_chunk = ((R.n + jsplit - 1) / jsplit + kGderThreads - 1) / kGderThreads * kGderThreads;
  jsplit = (R.n + j_chunk - 1) / j_chunk;
  const size_t shmem = sizeof(unsigned int) * (kGderThreads / 64) * 4 * nbins;
  if (e == hipSuccess) {
    k_gder<<<dim3(iblocks, jsplit), kGderThreads, shmem, s->stream>>>(
        s->reps_d, replica, s->px, s->py, s->pz, nbins, range, range / (double)nbins, j_chunk,
        c_d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(counts, c_d, bytes, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess)
    return fail(BD_E_HIP, std::string("bd_gder_noalloc: ") + hipGetErrorString(e));
  return BD_OK;
}

int bd_set_origin(bd_sys* s, int32_t slot) {
  if (!s) return fail(BD_E_ARG, "null handle");
  if (slot < 0 || slot >= BD_MAX_ORIGINS) return fail(BD_E_ARG, "origin slot out of range");
  const size_t n = (size_t)s->n_total;
  if (!s->origins) {
    int rc = dev_alloc(&s->origins, (size_t)BD_MAX_ORIGINS * 3 * n);
    if (rc != BD_OK) return rc;
  }
  double* o = s->origins + (size_t)slot * 3 * n;
  HIP_TRY(hipMemcpyAsync(o, s->rx, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(o + n, s->ry, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(o + 2 * n, s->rz, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

int bd_fself_origins(bd_sys* s, double q, int32_t nslots, double* out) {
  if (!s || !out) return fail(BD_E_ARG, "null argument");
  if (nslots < 1 || nslots > BD_MAX_ORIGINS) return fail(BD_E_ARG, "nslots out of range");
  if (!s->origins) return fail(BD_E_STATE, "no origin stored");
  double* o_d = nullptr;
  const size_t cnt = (size_t)s->R * nslots * 3;
  int rc = call_buffer(s, cnt, &o_d);
  if (rc != BD_OK) return rc;
  k_fself_origins<<<dim3(s->R, nslots), kOriginThreads, 0, s->stream>>>(
      s->reps_d, s->rx, s->ry, s->rz, s->origins, (size_t)3 * s->n_total, s->n_total, q, o_d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, o_d, cnt * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_fself_origins: ") + hipGetErrorString(e));
  return BD_OK;
}

int bd_gder_all(bd_sys* s, int32_t nbins, double range, uint64_t* counts) {
  if (!s || !counts) return fail(BD_E_ARG, "null argument");
  if (nbins < 1 || nbins > kGderMaxBins || !(range > 0.0))
    return fail(BD_E_ARG, "nbins must be in 1..512 and range > 0");
  int nmax = 0;
  for (const RepDev& R : s->reps) nmax = std::max(nmax, R.n);
  const size_t cnt = (size_t)s->R * 4 * nbins;
  unsigned long long* c_d = nullptr;
  int rc = call_buffer(s, cnt, &c_d);
  if (rc != BD_OK) return rc;
  hipError_t e = hipMemsetAsync(c_d, 0, sizeof(unsigned long long) * cnt, s->stream);
  const int iblocks = (nmax + kGderThreads - 1) / kGderThreads;
  const int j_chunk = 2048;
  const int jsplit = (nmax + j_chunk - 1) / j_chunk;
  const size_t shmem = sizeof(unsigned int) * (kGderThreads / 64) * 4 * nbins;
  if (e == hipSuccess) {
    k_gder<<<dim3(iblocks, jsplit, s->R), kGderThreads, shmem, s->stream>>>(
        s->reps_d, -1, s->px, s->py, s->pz, nbins, range, range / (double)nbins, j_chunk, c_d);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(counts, c_d, sizeof(unsigned long long) * cnt, hipMemcpyDeviceToHost,
                       s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_gder_all: ") + hipGetErrorString(e));
  return BD_OK;
}

int bd_observables(bd_sys* s, int32_t replica, double q, bd_obs* out) {
  int rc = check_rep(s, replica);
  if (rc != BD_OK) return rc;
  if (!out) return fail(BD_E_ARG, "null output");
  const RepDev& R = s->reps[replica];
  // six sums, then the replica index in the eighth word
  double* o_d = nullptr;
  rc = call_buffer(s, 8, &o_d);
  int* w_d = rc == BD_OK ? reinterpret_cast<int*>(o_d + 7) : nullptr;
  hipError_t e = hipSuccess;
  double h[6] = {0, 0, 0, 0, 0, 0};
  if (rc == BD_OK) {
    int which = replica;
    e = hipMemcpyAsync(w_d, &which, sizeof(int), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
      k_observables<<<1, kObsThreads, 0, s->stream>>>(s->reps_d, w_d, s->rx, s->ry, s->rz, s->ix,
                                                      s->iy, s->iz, s->vir, s->pot, q, o_d);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(h, o_d, sizeof(double) * 6, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  if (rc != BD_OK) return rc;
  if (e != hipSuccess) return fail(BD_E_HIP, std::string("bd_observables: ") + hipGetErrorString(e));
  const bd_params& p = s->params[replica];
  out->fself_big = h[0] / (double)p.n_big;
  out->msd_big = h[1] / (double)p.n_big;
  out->msd_sml = h[2] / (double)p.n_sml;
  out->epot_big = h[3] / (double)p.n_big;
  out->epot_sml = h[4] / (double)p.n_sml;
  out->virial = h[5];
  // pars.temp_set after the last update: the temperature of update ni
  const int64_t ni = s->ni;
  double T;
  {
    // same rule as temperature_at(), evaluated on the host for the report
    const RepDev& Rr = R;
    if (ni == Rr.first_step) {
      T = Rr.T0;
    } else {
      double t = Rr.dt * (1.0 + (double)(ni - 1));
      if (t < Rr.t0) {
        T = Rr.T0;
      } else if (Rr.temp_protocol == BD_TEMP_SINE) {
        if (t > Rr.tf) {
          T = Rr.Tf;
        } else {
          double a = 0.5 * (Rr.T0 + Rr.Tf), b = 0.5 * (Rr.Tf - Rr.T0);
          double u = t / Rr.period - 0.25;
          u = u - std::floor(u);
          T = a + b * std::sin(2.0 * M_PI * u);
        }
      } else if (t > Rr.tf || !(Rr.tf > Rr.t0)) {
        T = Rr.Tf;
      } else {
        T = Rr.T0 + (Rr.Tf - Rr.T0) / (Rr.tf - Rr.t0) * (t - Rr.t0);
      }
    }
  }
  out->temperature = T;
  out->z = 1.0 + out->virial / (3.0 * (double)(p.n_big + p.n_sml) * T);
  return BD_OK;
}

}  // extern "C"

