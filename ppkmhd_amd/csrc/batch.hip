// Batched ensemble of independent grids (same dimension, any sizes and parameters): every
// stage of the fused step is one launch over the blocks of all members, each member with its
// own time, dt (own CFL condition), iteration count and finished() test.  A block finds its
// member by binary search on the members' first-block table; the member's current state is
// U[iteration & 1], read from its own control block, so the same launch sequence serves every
// step and is captured once in a hipGraph by the host.
//
// The per-thread bodies are those of the single-grid fused step (fused_body.hpp), so every
// member's result equals the single-grid result bit for bit in the same arithmetic variant.
#include <hip/hip_runtime.h>

#include "fused_body.hpp"

namespace ppk {
namespace PPK_NS {
namespace batchk {

using namespace fusedk;

constexpr int kBlock = 256;

// member owning block b of stage st: the last member whose first block is <= b
PPK_HD int find_member(const BatchMember *__restrict__ m, int nm, int st, long long b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (m[mid].begin[st] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ bool member_active(const BatchMember &m) {
  return m.ctrl[CTRL_ACTIVE] != 0.0;
}
__device__ __forceinline__ int parity(const BatchMember &m) {
  return ((long long)m.ctrl[CTRL_ITER]) & 1;
}

// ---- per-member control: finished() test (SolverBase.cpp:195-201), dt, end of step ---------
__global__ void k_bbegin(const BatchMember *__restrict__ mem, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  double *ctrl = mem[i].ctrl;
  const bool done = ctrl[CTRL_T] >= (ctrl[CTRL_TEND] - 1e-14) || ctrl[CTRL_ITER] >= ctrl[CTRL_NSTEPMAX];
  ctrl[CTRL_ACTIVE] = done ? 0.0 : 1.0;
}

// one block per member: reduce its CFL partials, dt = cfl / max, clamped to tEnd - t
__global__ void __launch_bounds__(kBlock) k_bset_dt(const BatchMember *__restrict__ mem) {
  __shared__ double red[kBlock / 64];
  const BatchMember &m = mem[blockIdx.x];
  double *ctrl = m.ctrl;
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  double mx = 0.0;
  for (int b = threadIdx.x; b < m.nparts; b += kBlock) mx = fmax(mx, m.parts[b]);
  mx = block_max(mx, red);
  if (threadIdx.x != 0) return;
  ctrl[CTRL_INVDT] = mx;
  double dt = m.p.cfl / mx;
  if (ctrl[CTRL_T] + dt > ctrl[CTRL_TEND]) dt = ctrl[CTRL_TEND] - ctrl[CTRL_T];
  ctrl[CTRL_DT] = dt;
}

__global__ void k_bend(const BatchMember *__restrict__ mem, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  double *ctrl = mem[i].ctrl;
  if (ctrl[CTRL_ACTIVE] == 0.0) return;
  ctrl[CTRL_ITER] += 1.0;
  ctrl[CTRL_T] += ctrl[CTRL_DT];
}

// ---- ghost fill of the current state, every face, edge and corner in one pass -------------
__global__ void __launch_bounds__(kBlock) k_bbnd(const BatchMember *__restrict__ mem, int nm) {
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_BND, gb)];
  if (!member_active(m)) return;
  const long long t = (gb - m.begin[BS_BND]) * kBlock + threadIdx.x;
  if (t < ghost_shell_cells(m.p)) ghost_shell_element(m.U[parity(m)], m.p, t, nullptr);
}

// ---- standalone CFL partials of the current state ------------------------------------------
__global__ void __launch_bounds__(kCflBlock) k_bcfl(const BatchMember *__restrict__ mem, int nm) {
  __shared__ double red[kCflBlock / 64];
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_CFL, gb)];
  if (!member_active(m)) return;
  const int L = (int)(gb - m.begin[BS_CFL]);
  int i, j, k;
  double best = 0.0;
  if (brick_cell(m.pad, L, threadIdx.x, i, j, k))
    best = cfl_cell(m.U[parity(m)], m.p, cidx(m.p, i, j, k));
  best = block_max(best, red);
  if (threadIdx.x == 0) m.parts[L] = best;
}

// ---- half step -----------------------------------------------------------------------------
template <class T, class R>
__global__ void __launch_bounds__(T::NT) k_bhalf2d(const BatchMember *__restrict__ mem, int nm) {
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_HALF, gb)];
  if (!member_active(m)) return;
  __shared__ R lds[T::LDS];
  R *sF = lds, *sQ = lds + 2 * T::NF, *sE = sQ + 8 * T::NQ;
  const DevParams &p = m.p;
  const int L = (int)(gb - m.begin[BS_HALF]);
  const int ntx = (p.nx + 2 + T::TX - 1) / T::TX;
  const int b0x = p.gw - 1 + (L % ntx) * T::TX;
  const int b0y = p.gw - 1 + (L / ntx) * T::TY;
  const double dt = m.ctrl[CTRL_DT];
  const double *U = m.U[parity(m)];
  R *rec = static_cast<R *>(m.rec);
  const int tid = threadIdx.x;
  half2_load<T, R>(U, p, b0x, b0y, tid, sF, sQ);
  __syncthreads();
  half2_prims<T, R>(p, b0x, b0y, tid, sF, sQ);
  __syncthreads();
  half2_elec<T, R>(p, b0x, b0y, tid, sF, sQ, sE);
  __syncthreads();
  half2_record<T, R>(p, b0x, b0y, tid, sF, sQ, sE, dt, rec);
}

template <class T, class R>
__global__ void __launch_bounds__(T::NT) k_bhalf3d(const BatchMember *__restrict__ mem, int nm) {
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_HALF, gb)];
  if (!member_active(m)) return;
  __shared__ R lds[T::LDS];
  R *sF = lds, *sQ = lds + 3 * T::NF, *sE = sQ + 8 * T::NQ;
  const DevParams &p = m.p;
  const int L = (int)(gb - m.begin[BS_HALF]);
  const int ntx = (p.nx + 2 + T::TX - 1) / T::TX;
  const int nty = (p.ny + 2 + T::TY - 1) / T::TY;
  const int b0x = p.gw - 1 + (L % ntx) * T::TX;
  const int b0y = p.gw - 1 + ((L / ntx) % nty) * T::TY;
  const int b0z = p.gw - 1 + (L / (ntx * nty)) * T::TZ;
  const double dt = m.ctrl[CTRL_DT];
  const double *U = m.U[parity(m)];
  R *rec = static_cast<R *>(m.rec);
  const int tid = threadIdx.x;
  half3_load<T, R>(U, p, b0x, b0y, b0z, tid, sF, sQ);
  __syncthreads();
  half3_prims<T, R>(p, b0x, b0y, b0z, tid, sF, sQ);
  __syncthreads();
  half3_elec<T, R>(p, b0x, b0y, b0z, tid, sF, sQ, sE);
  __syncthreads();
  half3_record<T, R>(p, b0x, b0y, b0z, tid, sF, sQ, sE, dt, rec);
}

// ---- face and edge solves (interleaved types as in the single-grid step) ---------------------
template <int DIM, int W, class R>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W)))
k_briemann(const BatchMember *__restrict__ mem, int nm) {
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_RIEMANN, gb)];
  if (!member_active(m)) return;
  const DevParams &p = m.p;
  constexpr int nt = DIM == 3 ? 6 : 3;
  const long long B = gb - m.begin[BS_RIEMANN];
  const long long c = sub_brick_cell(p, m.sub, (int)(B / nt), threadIdx.x);
  if (c < 0) return;
  const R *rec = static_cast<const R *>(m.rec);
  R *flux = static_cast<R *>(m.flux);
  double *emf = static_cast<double *>(m.emf);
  if constexpr (DIM == 3) {
    switch ((int)(B % 6)) {
      case 0: face3<0>(rec, flux, p, c); break;
      case 1: face3<1>(rec, flux, p, c); break;
      case 2: face3<2>(rec, flux, p, c); break;
      case 3: edge3<2>(rec, emf, p, c); break;
      case 4: edge3<1>(rec, emf, p, c); break;
      default: edge3<0>(rec, emf, p, c); break;
    }
  } else {
    switch ((int)(B % 3)) {
      case 0: face2<0>(rec, flux, p, c); break;
      case 1: face2<1>(rec, flux, p, c); break;
      default: edge2(rec, emf, p, c); break;
    }
  }
}

// ---- update + CT + ghost copy + CFL partials of the new state -----------------------------
template <int DIM, class R>
__global__ void __launch_bounds__(kCflBlock) k_bupdate(const BatchMember *__restrict__ mem,
                                                       int nm) {
  __shared__ double red[kCflBlock / 64];
  const long long gb = blockIdx.x;
  const BatchMember &m = mem[find_member(mem, nm, BS_UPDATE, gb)];
  if (!member_active(m)) return;
  const DevParams &p = m.p;
  const int L = (int)(gb - m.begin[BS_UPDATE]);
  const double dt = m.ctrl[CTRL_DT];
  const double dtdx = dt / p.dx, dtdy = dt / p.dy, dtdz = DIM == 3 ? dt / p.dz : 0.0;
  const int par = parity(m);
  int i, j, k;
  double best = 0.0;
  if (brick_cell(m.pad, L, threadIdx.x, i, j, k))
    best = update_cell(m.U[par], m.U[par ^ 1], static_cast<const R *>(m.flux),
                       static_cast<const double *>(m.emf), p, cidx(p, i, j, k), dtdx, dtdy, dtdz, true);
  best = block_max(best, red);
  if (threadIdx.x == 0) m.parts[L] = best;
}

}  // namespace batchk

// half-step tiles of one member
template <int TX_, int TY_>
long long half_tiles2(const DevParams &p) {
  return (long long)((p.nx + 2 + TX_ - 1) / TX_) * ((p.ny + 2 + TY_ - 1) / TY_);
}
template <class T>
long long half_tiles3(const DevParams &p) {
  return (long long)((p.nx + 2 + T::TX - 1) / T::TX) * ((p.ny + 2 + T::TY - 1) / T::TY) *
         ((p.nz + 2 + T::TZ - 1) / T::TZ);
}

// blocks per stage of one member (the host builds the first-block tables from these)
void batch_block_counts(const DevParams &p, int tile, long long counts[BS_COUNT]) {
  using namespace fusedk;
  counts[BS_BND] = (ghost_shell_cells(p) + batchk::kBlock - 1) / batchk::kBlock;
  counts[BS_CFL] = fused_cfl_parts(p);
  if (p.dim == 3)
    counts[BS_HALF] = tile == 1 ? half_tiles3<Tile3d>(p)
                                : (tile == 2 ? half_tiles3<Tile3e>(p) : half_tiles3<Tile3>(p));
  else
    counts[BS_HALF] = tile == 1 ? half_tiles2<Tile2b::TX, Tile2b::TY>(p)
                                : (tile == 2 ? half_tiles2<Tile2c::TX, Tile2c::TY>(p)
                                             : half_tiles2<Tile2::TX, Tile2::TY>(p));
  counts[BS_RIEMANN] = (p.dim == 3 ? 6 : 3) * sub_bricks(p).count();
  counts[BS_UPDATE] = fused_cfl_parts(p);
}

// one step of every member: finished test, ghost fill, (standalone CFL on request), dt, the
// three fused stages, end of step.  Stream-ordered, no host synchronization, no memset: safe
// to capture in a graph and replay.
template <class R>
void launch_bhalf(const BatchMember *members, int nm, int dim, int tile, long long blocks,
                  hipStream_t s) {
  using namespace batchk;
  const dim3 g((unsigned)blocks);
  if (dim == 3) {
    if (tile == 1)
      hipLaunchKernelGGL((k_bhalf3d<Tile3d, R>), g, dim3(Tile3d::NT), 0, s, members, nm);
    else if (tile == 2)
      hipLaunchKernelGGL((k_bhalf3d<Tile3e, R>), g, dim3(Tile3e::NT), 0, s, members, nm);
    else
      hipLaunchKernelGGL((k_bhalf3d<Tile3, R>), g, dim3(Tile3::NT), 0, s, members, nm);
  } else {
    if (tile == 1)
      hipLaunchKernelGGL((k_bhalf2d<Tile2b, R>), g, dim3(Tile2b::NT), 0, s, members, nm);
    else if (tile == 2)
      hipLaunchKernelGGL((k_bhalf2d<Tile2c, R>), g, dim3(Tile2c::NT), 0, s, members, nm);
    else
      hipLaunchKernelGGL((k_bhalf2d<Tile2, R>), g, dim3(Tile2::NT), 0, s, members, nm);
  }
}

// half step, face/edge solves, update of every member (kernels specialised per dimension)
template <class R>
void launch_rest(const BatchMember *members, int nm, int dim, int tile,
                 const long long totals[BS_COUNT], hipStream_t s) {
  using namespace batchk;
  const dim3 b(kBlock), gr((unsigned)totals[BS_RIEMANN]), gu((unsigned)totals[BS_UPDATE]);
  launch_bhalf<R>(members, nm, dim, tile, totals[BS_HALF], s);
  if (dim == 3) {
    // 3 waves per SIMD keep the 3D double solves free of spills (the single-grid default)
    if constexpr (sizeof(R) == sizeof(double))
      hipLaunchKernelGGL((k_briemann<3, 3, R>), gr, b, 0, s, members, nm);
    else
      hipLaunchKernelGGL((k_briemann<3, 4, R>), gr, b, 0, s, members, nm);
    hipLaunchKernelGGL((k_bupdate<3, R>), gu, dim3(kCflBlock), 0, s, members, nm);
  } else {
    hipLaunchKernelGGL((k_briemann<2, 4, R>), gr, b, 0, s, members, nm);
    hipLaunchKernelGGL((k_bupdate<2, R>), gu, dim3(kCflBlock), 0, s, members, nm);
  }
}

hipError_t batch_step(const BatchMember *members, int nm, int dim, int bulk32, int tile,
                      const long long totals[BS_COUNT], int with_cfl, hipStream_t s) {
  using namespace batchk;
  const dim3 b(kBlock), bm((nm + 63) / 64 * 64 > 1024 ? 1024 : (nm + 63) / 64 * 64);
  const dim3 gm((nm + (int)bm.x - 1) / (int)bm.x);
  hipLaunchKernelGGL(k_bbegin, gm, bm, 0, s, members, nm);
  if (totals[BS_BND] > 0)
    hipLaunchKernelGGL(k_bbnd, dim3((unsigned)totals[BS_BND]), b, 0, s, members, nm);
  if (with_cfl)
    hipLaunchKernelGGL(k_bcfl, dim3((unsigned)totals[BS_CFL]), dim3(kCflBlock), 0, s, members,
                       nm);
  hipLaunchKernelGGL(k_bset_dt, dim3(nm), b, 0, s, members);
  if (bulk32)
    launch_rest<float>(members, nm, dim, tile, totals, s);
  else
    launch_rest<double>(members, nm, dim, tile, totals, s);
  hipLaunchKernelGGL(k_bend, gm, bm, 0, s, members, nm);
  return hipGetLastError();
}

}  // namespace PPK_NS
}  // namespace ppk
