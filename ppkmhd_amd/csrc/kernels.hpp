// Interface between the host driver (capi.cpp) and the kernel translation units.  Each kernel
// file is compiled once per arithmetic variant into its own namespace (ppk::fast,
// ppk::strict); the driver picks one table of entry points per solver handle.
#pragma once

#include <hip/hip_runtime.h>

#include "mhd_device.hpp"

namespace ppk {

// slots of the device-resident time-loop control block (doubles)
enum {
  CTRL_T = 0,        // current time
  CTRL_DT = 1,       // dt of the step in flight
  CTRL_INVDT = 2,    // CFL maximum max_cells sum_d (c_f + |v_d|)/dx_d of the state (set_dt)
  CTRL_ITER = 3,     // iteration counter (exact in a double up to 2^53)
  CTRL_TEND = 4,
  CTRL_NSTEPMAX = 5,
  CTRL_ACTIVE = 6,   // 0 once SolverBase::finished() held before a checked step
  CTRL_SPARE = 7,
  CTRL_SLOTS = 8
};

// kernel stages bracketed by the optional per-stage GPU timers (ppk_mhd_timers_*)
enum Stage {
  ST_BOUNDARIES = 0,
  ST_PRIM_DT,
  ST_ELEC,
  ST_MAGSLOPES,
  ST_TRACE,
  ST_FLUXES,
  ST_EMF,
  ST_UPDATE,
  ST_CONTROL,
  ST_FUSED_HALF,     // fused: primitives + E + slopes + half step (compact records)
  ST_FUSED_RIEMANN,  // fused: Riemann problems at faces + EMFs at edges
  ST_FUSED_UPDATE,   // fused: update + CT + ghost copy (+ CFL of the new state)
  ST_FUSED_CFL,      // fused: standalone CFL reduction
  ST_FUSED_MARCH,    // single-pass 3D step (half step, solves, update, CT, CFL in one kernel)
  ST_SUR_FACES,      // surrogate form: the network's face solves (inside ST_FUSED_RIEMANN)
  ST_SUR_FALLBACK,   // surrogate form: exact HLLD of the refused faces (same)
  ST_COUNT
};
// host-side hook the launchers call around each stage (implemented by the driver)
struct StageTimer {
  virtual void mark(int stage, bool start) = 0;
  virtual ~StageTimer() {}
};
#define PPK_TIMED(tm, stage, launch)   \
  do {                                 \
    if (tm) (tm)->mark((stage), true); \
    launch;                            \
    if (tm) (tm)->mark((stage), false); \
  } while (0)

// scratch of the stage-per-kernel form (all over the padded box, SoA)
struct StagedScratch {
  double *faces = nullptr;  // 2*dim face states x 8 vars
  double *edges = nullptr;  // 4*(3 in 3D, 1 in 2D) edge states x 8 vars
  double *flux = nullptr;   // dim fluxes x 8 vars
  double *emf = nullptr;    // 3 (3D) or 1 (2D) EMF components
  double *elec = nullptr;   // 3D: 3 E components
  double *dbf = nullptr;    // 3D: 6 transverse face-field slopes
};

// scratch of the fused form (padded-box indexing, SoA)
constexpr int kRec3 = 38;   // half-step record per cell, 3D
constexpr int kRec2 = 26;   // half-step record per cell, 2D
constexpr int kFluxComp3 = 5;  // ID IP IU IV IW per face direction, 3D
constexpr int kFluxComp2 = 6;  // ID IP IU IV IW IC per face direction, 2D
// the surrogate form of the face solves: network, gate, counters (device memory)
// blocks of a surrogate face launch (at most): three per CU of an MI355X, as many as its LDS
// holds (measured: 2D step 0.344 -> 0.324 ms, 3D 57.8 -> 56.6 ms against two per CU)
constexpr int kSurrogateBlocks = 768;
struct SurrogateCfg {
  const void *net = nullptr;            // packed network (MlpPacked, device)
  double gate_rho = 0.0, gate_p = 0.0;  // max relative jumps |a_R - a_L| / min(a_L, a_R)
  // per launch block b: [2b] faces by the network, [2b + 1] by HLLD (2 x kSurrogateBlocks
  // counters, accumulated over launches), or null
  unsigned long long *counters = nullptr;
  unsigned char *flags = nullptr;  // dim x cells: faces left to the HLLD pass (device scratch)
};
// element type of rec and flux: double, or float in the FP32-bulk form (DevParams::bulk32)
struct FusedScratch {
  void *rec = nullptr;   // kRec{2,3} x cells
  void *flux = nullptr;  // dim x kFluxComp{2,3} x cells
  void *emf = nullptr;   // 3 (3D) or 1 (2D) x cells, double in every form
  const SurrogateCfg *sur = nullptr;  // surrogate face solves when set
};

// dt modes of the set_dt kernel
enum { SETDT_FIXED = 0, SETDT_CFL = 1, SETDT_REDUCE_ONLY = 2 };

// The CFL maximum is produced as per-block partial maxima (no same-address atomics): each
// block of a producer kernel writes its slot, set_dt reduces as many slots as the producer had
// blocks.  Staged primitives: cfl_blocks(p) blocks (grid-stride); fused update and fused CFL:
// one block per brick of the padded box, fused_cfl_parts(p).
constexpr int kCflBlock = 256;
constexpr int kCflMaxBlocks = 4096;
inline int cfl_blocks(const DevParams &p) {
  long long cells = (long long)p.isize * p.jsize * p.ksize;
  long long g = (cells + kCflBlock - 1) / kCflBlock;
  if (g > kCflMaxBlocks) g = kCflMaxBlocks;
  return g < 1 ? 1 : (int)g;
}

// Bricks: the fused per-cell kernels give each block of 256 threads a brick of cells.
// Shape 0 (default) is a run of 256 consecutive cells of the flattened range (one wave = 64
// cells of one x row); its rows are numbered in patches of 4 x 4 (y, z) rows in 3D, so the
// rows a cell's y and z neighbours live in were visited a few rows earlier (L2 hits).
// Shapes 1-3 are 3D bricks 64 x 2 x 2, 16 x 4 x 4, 32 x 4 x 2 (2D: 64 x 4), among which the
// batch's autotuner picks; single grids use shape 0.
constexpr int kPatch = 4;
struct Brick {
  int bx, by, bz;  // brick extents; bx = 0: flattened runs, by x bz = row patch
  int nx, ny, nz;  // bricks per direction; flattened: patches per direction in ny, nz
  int mx, my, mz;  // cells covered per direction
  long long nblocks;
  long long count() const { return nblocks; }
};
inline Brick make_brick(int dim, int mx, int my, int mz, int shape) {
  Brick b;
  b.mx = mx;
  b.my = my;
  b.mz = mz;
  if (shape <= 0 || shape > 3) {
    b.bx = 0;
    b.by = dim == 3 ? kPatch : 1;
    b.bz = dim == 3 ? kPatch : 1;
    b.ny = (my + b.by - 1) / b.by;
    b.nz = (mz + b.bz - 1) / b.bz;
    const long long rows = (long long)b.ny * b.by * b.nz * b.bz;
    b.nx = 0;
    b.nblocks = (rows * mx + 255) / 256;
    return b;
  }
  static const int shp[4][3] = {{0, 0, 0}, {64, 2, 2}, {16, 4, 4}, {32, 4, 2}};
  b.bx = dim == 3 ? shp[shape][0] : 64;
  b.by = dim == 3 ? shp[shape][1] : 4;
  b.bz = dim == 3 ? shp[shape][2] : 1;
  b.nx = (mx + b.bx - 1) / b.bx;
  b.ny = (my + b.by - 1) / b.by;
  b.nz = (mz + b.bz - 1) / b.bz;
  b.nblocks = (long long)b.nx * b.ny * b.nz;
  return b;
}
// bricks of the padded box (fused update, fused CFL)
inline Brick padded_bricks(const DevParams &p) {
  return make_brick(p.dim, p.isize, p.jsize, p.ksize, p.brick);
}
inline int fused_cfl_parts(const DevParams &p) { return (int)padded_bricks(p).count(); }
// the sub-box [gw, size-gw] per direction of the face and edge kernels, in bricks
inline Brick sub_bricks(const DevParams &p) {
  return make_brick(p.dim, p.nx + 1, p.ny + 1, p.dim == 3 ? p.nz + 1 : 1, p.brick);
}

// The single-pass 3D step (march3d.hip): blocks of 16 x 12 output columns marching along z
// through equal shares of the grid's tile-planes; one CFL partial per block
constexpr int kMarchOutX = 16, kMarchOutY = 12;
// Blocks of the launch: one per CU (each block needs a whole CU's LDS), fewer when a block
// would get under 16 tile-planes (small grids)
inline int march_nblk(const DevParams &p, int cus) {
  const long long total =
      (long long)((p.nx + kMarchOutX - 1) / kMarchOutX) * ((p.ny + kMarchOutY - 1) / kMarchOutY) *
      p.nz;
  long long b = cus > 0 ? cus : 256;
  if (b > total / 16) b = total / 16;
  return (int)(b < 1 ? 1 : b);
}
inline long long march_blocks(const DevParams &p) { return p.march > 0 ? p.march : 0; }

// half-step tile shapes of a batch (batch_block_counts / batch_step `tile`): 2D 32x8, 64x4,
// 16x16; 3D 16x4x4, 8x8x4, 32x4x2
constexpr int kBatchTiles = 3;

// ---- batched ensembles: many independent grids of one dimension advanced by the same
// launches.  Each batched launch covers the blocks of every member back to back; a member
// owns blocks [begin[stage], begin[stage] + its count) of that stage.
enum BatchStage {
  BS_BND = 0,                      // ghost fill of every face, edge and corner (one pass)
  BS_CFL,                          // standalone CFL of the current state (first step)
  BS_HALF, BS_RIEMANN, BS_UPDATE,  // the three fused stages
  BS_COUNT
};
struct BatchMember {
  DevParams p;
  double *U[2];        // ping-pong states; the current one is U[iteration & 1]
  void *rec, *flux, *emf;
  double *ctrl;        // the member's control block (CTRL_SLOTS)
  double *parts;       // the member's CFL partial maxima
  Brick pad, sub;      // bricks of the padded box and of the face/edge sub-box
  long long begin[BS_COUNT];
  int nparts;
  int unused;
};

#define PPK_DECLARE_KERNELS(NS)                                                                  \
  namespace NS {                                                                                 \
  /* ghost fill of U; copy (may be null) receives the same ghost values */                     \
  hipError_t staged_boundaries(const DevParams &p, double *U, double *copy, const double *ctrl,  \
                               hipStream_t s, StageTimer *tm);                                   \
  hipError_t staged_prim_dt(const DevParams &p, const double *U, double *Q, double *parts,       \
                            const double *ctrl, int do_dt, hipStream_t s, StageTimer *tm);       \
  hipError_t staged_begin(double *ctrl, int check_finished, hipStream_t s);                      \
  hipError_t staged_set_dt(double *ctrl, const double *parts, int nparts, double fixed_dt,       \
                           int mode, double cfl, hipStream_t s);                                 \
  hipError_t staged_end(double *ctrl, hipStream_t s);                                            \
  hipError_t staged_godunov(const DevParams &p, const double *Uin, double *Uout,                 \
                            const double *Q, StagedScratch &w, const double *ctrl,               \
                            hipStream_t s, StageTimer *tm);                                      \
  hipError_t fused_godunov(const DevParams &p, const double *Uin, double *Uout,                  \
                           FusedScratch &w, double *parts, const double *ctrl, int fold_cfl,     \
                           hipStream_t s, StageTimer *tm);                                       \
  hipError_t fused_cfl(const DevParams &p, const double *U, double *parts, const double *ctrl,   \
                       hipStream_t s, StageTimer *tm);                                           \
  hipError_t fused_half(const DevParams &p, const double *U, void *rec, const double *ctrl,      \
                        hipStream_t s);                                                          \
  /* surrogate form: face solves, and the training-sample dump of the current records */        \
  hipError_t surrogate_faces(const DevParams &p, const void *rec, void *flux,                    \
                             const double *ctrl, const SurrogateCfg &cfg, hipStream_t s,         \
                             StageTimer *tm);                                                    \
  hipError_t surrogate_dump(const DevParams &p, const void *rec, const double *ctrl, float *X,   \
                            float *T, long long stride, long long phase, hipStream_t s);         \
  long long surrogate_dump_slots(const DevParams &p, long long stride, long long phase);         \
  /* single-pass 3D step (p.march > 0): the interior of Uout (its ghosts come from the ghost   \
     fill's copy target); CFL partials, one per block, when fold_cfl */                         \
  hipError_t march_godunov(const DevParams &p, const double *Uin, double *Uout, double *parts,   \
                           const double *ctrl, int fold_cfl, hipStream_t s, StageTimer *tm);    \
  /* batched ensemble: counts of blocks per stage for one member, then the batched step */      \
  void batch_block_counts(const DevParams &p, int tile, long long counts[BS_COUNT]);            \
  hipError_t batch_step(const BatchMember *members, int nmembers, int dim, int bulk32, int tile, \
                        const long long totals[BS_COUNT], int with_cfl, hipStream_t s);         \
  }

PPK_DECLARE_KERNELS(fast)
PPK_DECLARE_KERNELS(strict)

// entry-point table of one arithmetic variant
struct KernelTable {
  decltype(&fast::staged_boundaries) boundaries;
  decltype(&fast::staged_prim_dt) prim_dt;
  decltype(&fast::staged_begin) begin;
  decltype(&fast::staged_set_dt) set_dt;
  decltype(&fast::staged_end) end;
  decltype(&fast::staged_godunov) godunov;
  decltype(&fast::fused_godunov) fused_godunov;
  decltype(&fast::fused_cfl) fused_cfl;
  decltype(&fast::fused_half) fused_half;
  decltype(&fast::surrogate_dump) surrogate_dump;
  decltype(&fast::surrogate_dump_slots) surrogate_dump_slots;
  decltype(&fast::march_godunov) march_godunov;
  decltype(&fast::batch_block_counts) batch_block_counts;
  decltype(&fast::batch_step) batch_step;
};

// entry-point tables of the two arithmetic variants (capi.cpp)
extern const KernelTable kFast;
extern const KernelTable kStrict;

// The ghost fill in one pass.  The reference fills x, then y, then z ghosts over the whole
// transverse range (src/shared/BoundariesFunctors.h:534-1057, SolverBase.cpp:339-537;
// Dirichlet flips the normal velocity and normal field), so a corner or edge ghost ends with
// the value its last pass reads from a ghost the earlier passes filled: the interior cell
// reached by mapping each of its ghost coordinates, times the product of the directions'
// Dirichlet signs (exact, each sign is +-1).  ghost_shell_element computes that composition
// directly: the same values, one write per ghost cell (copy, optional: a second state that
// receives the same ghost values, the single-pass step's output whose ghosts are the
// reference's U2 = U copy of the filled input).
// source coordinate of c in [0, n + 2 gw) along one direction; flip: a Dirichlet face mapped it
PPK_HD int ghost_src(const DevParams &p, int d, int c, bool &flip) {
  const int gw = p.gw, n = d == 0 ? p.nx : (d == 1 ? p.ny : p.nz);
  flip = false;
  if (c >= gw && c < n + gw) return c;
  const bool upper = c >= n + gw;
  const int bc = p.bc[2 * d + (upper ? 1 : 0)];
  if (bc == PPK_BC_DIRICHLET) {
    flip = true;
    return upper ? 2 * n + 2 * gw - 1 - c : 2 * gw - 1 - c;
  }
  if (bc == PPK_BC_NEUMANN) return upper ? n + gw - 1 : gw;
  return upper ? c - n : n + c;
}
// ghost cells of the padded box: the 2 gw z planes (3D), then per interior plane the 2 gw y
// rows, then per interior row the 2 gw x cells
PPK_HD long long ghost_shell_cells(const DevParams &p) {
  const long long gw2 = 2LL * p.gw, kin = p.dim == 3 ? p.nz : 1;
  return (p.dim == 3 ? gw2 * p.jsize * p.isize : 0) + kin * gw2 * p.isize + kin * p.ny * gw2;
}
PPK_HD void ghost_shell_element(double *__restrict__ U, const DevParams &p, long long t,
                                double *__restrict__ copy) {
  const int gw = p.gw;
  const long long gw2 = 2LL * gw;
  const long long nA = p.dim == 3 ? gw2 * p.jsize * p.isize : 0;
  const long long nB = (p.dim == 3 ? p.nz : 1) * gw2 * p.isize;
  const int koff = p.dim == 3 ? gw : 0;
  int i, j, k;
  if (t < nA) {
    const int kk = (int)(t / ((long long)p.jsize * p.isize));
    k = kk < gw ? kk : p.nz + kk;
    j = (int)((t / p.isize) % p.jsize);
    i = (int)(t % p.isize);
  } else if (t < nA + nB) {
    const long long u = t - nA;
    k = (int)(u / (gw2 * p.isize)) + koff;
    const int jj = (int)((u / p.isize) % gw2);
    j = jj < gw ? jj : p.ny + jj;
    i = (int)(u % p.isize);
  } else {
    const long long u = t - nA - nB;
    k = (int)(u / ((long long)p.ny * gw2)) + koff;
    const long long r = u % ((long long)p.ny * gw2);
    j = gw + (int)(r / gw2);
    const int ii = (int)(r % gw2);
    i = ii < gw ? ii : p.nx + ii;
  }
  bool fx, fy, fz = false;
  const int i0 = ghost_src(p, 0, i, fx), j0 = ghost_src(p, 1, j, fy);
  const int k0 = p.dim == 3 ? ghost_src(p, 2, k, fz) : k;
  const long long dst = ((long long)k * p.jsize + j) * p.isize + i;
  const long long src = ((long long)k0 * p.jsize + j0) * p.isize + i0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const bool neg = ((v == IU || v == IA) && fx) != (((v == IV || v == IB) && fy) !=
                                                      ((v == IW || v == IC) && fz));
    const double x = U[src + v * p.vstride] * (neg ? -1.0 : 1.0);
    U[dst + v * p.vstride] = x;
    if (copy) copy[dst + v * p.vstride] = x;
  }
}

// block-wide maximum of one double per thread (any block size that is a multiple of 64, up
// to 1024); every thread gets the result
__device__ __forceinline__ double block_max(double v, double *smem32) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) smem32[wave] = v;
  __syncthreads();
  double m = smem32[0];
  for (int w = 1; w < nw; ++w) m = fmax(m, smem32[w]);
  __syncthreads();
  return m;
}

}  // namespace ppk
