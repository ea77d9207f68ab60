// Uploaded one-field hex8 / tet4 diffusion by node patches.
//
// Replaces, for meshes that arrive as connectivity (a ConnManager's elements), the same
// chain as the lattice kernels: Gather -> DOF/DOFGradient -> Integrator_GradBasisDotVector
// (k grad u), Integrator_BasisTimesScalar (source, transient rho_c udot) ->
// ScatterResidual_Tpetra (disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-650,
// Panzer_Integrator_GradBasisDotVector_impl.hpp:225-294, Panzer_Integrator_BasisTimesScalar_impl.hpp:215-270,
// Panzer_ScatterResidual_Tpetra_impl.hpp:375-520), with the element loop of
// Panzer_AssemblyEngine_impl.hpp:160-172 over worksets of any element order.
//
// Layout, built once per mesh with its graph:
// - the rows (LIDs; one field, so a row is a vertex) sorted along a Morton curve of their
//   vertex coordinates and cut into patches of P consecutive rows;
// - for each patch, its instances: every element with a vertex in the patch, listed by
//   colour, then sweep order;
// - a colouring of each patch's instances, greedy in sweep order (element centroids by z,
//   then y, then x): instances that share one of the patch's own vertices differ in colour;
// - instance records in instance order: the element's LIDs and its slot map (for row a and
//   column b, the position of LID b in LID a's CRS row), so that a patch streams its
//   instances contiguously whatever order the elements were uploaded in.
// One workgroup assembles one patch.  Each lane integrates one instance in registers
// (the sum-factorised hex8 routine of the lattice kernels, or the constant-gradient tet),
// then adds the element rows of the patch's own vertices into the patch's rows in LDS.
// The adds go colour by colour with a barrier between colours, so a row receives at
// most one add per step: the sums run in a fixed order (colour, sweep order) and the
// results do not depend on scheduling.  Finally each row is stored once, whole, with
// its residual entry.  No workspace, no global atomics, no dependence on the order in
// which the elements were uploaded beyond the colouring's priorities.  Elements on a
// patch border are integrated once per patch they touch.
#include "txa_internal.hpp"
#include "txa_q1_elem.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

namespace txa {
namespace {
using namespace q1elem;

constexpr int kPatchThreads = 256;
constexpr int kMaxColours = 64;         // colour masks are 64-bit
constexpr int kPatchLdsBytes = 65536;   // row accumulators of one patch
constexpr int kSetupBlock = 256;

inline unsigned setup_grid(int64_t n) { return static_cast<unsigned>((n + kSetupBlock - 1) / kSetupBlock); }

// LDS of one patch: [P][RS] row entries and [P] residual entries (double), then each row's
// first CRS position (int64), LID and length (int32)
inline size_t patch_lds_bytes(int P, int RS) { return static_cast<size_t>(P) * ((RS + 1) * sizeof(double) + 16); }

// ---------------------------------------------------------------------------------------
// setup

// Vertex coordinates by row (one field: the row of an element position is its vertex's LID).
__global__ void k_node_xyz(int64_t npos, const int64_t *__restrict__ conn, const int32_t *__restrict__ elem_lid,
                           const double *__restrict__ xyz, int64_t nverts, double *__restrict__ rec) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= npos) return;
  const int64_t v = conn[p];
  const int64_t r = elem_lid[p];
  rec[4 * r + 0] = xyz[v];   // every element position of a vertex writes the same values
  rec[4 * r + 1] = xyz[nverts + v];
  rec[4 * r + 2] = xyz[2 * nverts + v];
}

__device__ __forceinline__ uint64_t spread3(uint64_t x) {   // 21 bits to every third bit
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

struct Box {
  double lo[3], scale[3];
};

__global__ void k_morton(int64_t nrows, const double *__restrict__ rec, Box b, uint64_t *__restrict__ key,
                         int32_t *__restrict__ row) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  uint64_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double t = (rec[4 * r + d] - b.lo[d]) * b.scale[d] + 0.5;
    t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
    q[d] = static_cast<uint64_t>(t);
  }
  key[r] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
  row[r] = static_cast<int32_t>(r);
}

// patch_rows[pos] = row (padded with -1), the row's position in the w slot of its record, and
// in curve order: the vertex coordinates (xyzp) and the row's CRS range (r0p, lenp)
__global__ void k_place_rows(int64_t nrows, int64_t npad, const int32_t *__restrict__ sorted, int32_t *__restrict__ prow,
                             double *__restrict__ rec, double *__restrict__ xyzp, const int64_t *__restrict__ rowptr,
                             int64_t *__restrict__ r0p, int32_t *__restrict__ lenp) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const int32_t r = i < nrows ? sorted[i] : -1;
  prow[i] = r;
  for (int d = 0; d < 3; ++d) xyzp[4 * i + d] = r >= 0 ? rec[4 * static_cast<int64_t>(r) + d] : 0.0;
  xyzp[4 * i + 3] = 0.0;
  r0p[i] = r >= 0 ? rowptr[r] : 0;
  lenp[i] = r >= 0 ? static_cast<int32_t>(rowptr[r + 1] - rowptr[r]) : 0;
  if (r >= 0) rec[4 * static_cast<int64_t>(r) + 3] = __longlong_as_double(i);
}

// Keys (patch, 16-entry line of a by-LID FP64 array) of every row: the distinct keys count the
// 128-byte lines a patch touches when it reads its state by LID; padding rows get ~0.
__global__ void k_line_keys(int64_t npad, int P, const int32_t *__restrict__ prow, uint64_t *__restrict__ key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const int32_t r = prow[i];
  key[i] = r >= 0 ? (static_cast<uint64_t>(i / P) << 32) | static_cast<uint32_t>(r >> 4) : ~0ull;
}

__global__ void k_count_distinct(int64_t n, const uint64_t *__restrict__ key, unsigned long long *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || key[i] == ~0ull) return;
  if (i == 0 || key[i - 1] != key[i]) atomicAdd(out, 1ull);
}

// The solution (and its time derivative) in curve order, before each assembly: xp[pos] = x[row].
__global__ void k_state_by_pos(int64_t npad, const int32_t *__restrict__ prow, const double *__restrict__ x,
                               const double *__restrict__ xdot, double *__restrict__ xp, double *__restrict__ xdp) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const int32_t r = prow[i];
  xp[i] = r >= 0 ? x[r] : 0.0;
  if (xdp) xdp[i] = r >= 0 ? xdot[r] : 0.0;
}

// Sweep key of each element: its centroid on the Morton lattice, z then y then x (lexicographic),
// so that a greedy colouring in this order colours like a structured sweep.
__global__ void k_sweep_keys(int64_t nelem, int nb, const int32_t *__restrict__ elem_lid,
                             const double *__restrict__ rec, Box b, uint64_t *__restrict__ key,
                             int32_t *__restrict__ elem) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  uint64_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double c = 0.0;
    for (int v = 0; v < nb; ++v) c += rec[4 * static_cast<int64_t>(elem_lid[e * nb + v]) + d];
    double t = (c / nb - b.lo[d]) * b.scale[d];
    t = t < 0.0 ? 0.0 : (t > 2097151.0 ? 2097151.0 : t);
    q[d] = static_cast<uint64_t>(t);
  }
  key[e] = (q[2] << 42) | (q[1] << 21) | q[0];
  elem[e] = static_cast<int32_t>(e);
}

__global__ void k_rank_of(int64_t nelem, const int32_t *__restrict__ by_rank, int32_t *__restrict__ rank) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nelem) return;
  rank[by_rank[r]] = static_cast<int32_t>(r);
}

// The distinct patches of each element, as keys patch << 32 | sweep rank of the element; the
// other slots hold ~0 (sorted last).  Counts the instances of each patch.
__global__ void k_instance_keys(int64_t nelem, int nb, int P, const int32_t *__restrict__ elem_lid,
                                const double *__restrict__ rec, const int32_t *__restrict__ rank,
                                uint64_t *__restrict__ key, unsigned long long *__restrict__ count,
                                unsigned long long *__restrict__ per_patch) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  int64_t pt[8];
  int n = 0;
  for (int v = 0; v < nb; ++v) {
    const int64_t pos = __double_as_longlong(rec[4 * static_cast<int64_t>(elem_lid[e * nb + v]) + 3]);
    const int64_t p = pos / P;
    bool seen = false;
    for (int j = 0; j < n; ++j) seen = seen || pt[j] == p;
    if (!seen) pt[n++] = p;
  }
  for (int v = 0; v < nb; ++v)
    key[e * nb + v] = v < n ? (static_cast<uint64_t>(pt[v]) << 32) | static_cast<uint32_t>(rank[e]) : ~0ull;
  for (int j = 0; j < n; ++j) atomicAdd(per_patch + pt[j], 1ull);
  atomicAdd(count, static_cast<unsigned long long>(n));
}

// Greedy colouring of each patch's instances in sweep order (one thread per patch): an instance
// takes the smallest colour that no earlier instance of the patch sharing one of the patch's
// own vertices has.  Only the patch's rows are summed in its workgroup, so only those vertices
// conflict.  rowmask: [npatch * P] colours taken at each row, zero on entry.
__global__ void k_patch_colour(int64_t npatch, int nb, int P, const int64_t *__restrict__ pstart,
                               const uint64_t *__restrict__ key, const int32_t *__restrict__ by_rank,
                               const int32_t *__restrict__ elem_lid, const double *__restrict__ rec,
                               unsigned long long *__restrict__ rowmask, uint8_t *__restrict__ colour,
                               int *__restrict__ maxc) {
  const int64_t pb = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pb >= npatch) return;
  int cmax = 0;
  for (int64_t i = pstart[pb]; i < pstart[pb + 1]; ++i) {
    const int64_t e = by_rank[static_cast<uint32_t>(key[i])];
    int64_t own[8];
    unsigned long long used = 0;
    for (int v = 0; v < nb; ++v) {
      const int64_t pos = __double_as_longlong(rec[4 * static_cast<int64_t>(elem_lid[e * nb + v]) + 3]);
      own[v] = pos / P == pb ? pos : -1;
      if (own[v] >= 0) used |= rowmask[own[v]];
    }
    const int c = used == ~0ull ? kMaxColours : __builtin_ctzll(~used);
    colour[i] = static_cast<uint8_t>(c < kMaxColours ? c : kMaxColours - 1);
    cmax = c > cmax ? c : cmax;
    if (c < kMaxColours)
      for (int v = 0; v < nb; ++v)
        if (own[v] >= 0) rowmask[own[v]] |= 1ull << c;
  }
  atomicMax(maxc, cmax);
}

// Final keys: patch << 38 | colour << 32 | sweep rank.
__global__ void k_final_keys(int64_t ninst, const uint64_t *__restrict__ key, const uint8_t *__restrict__ colour,
                             uint64_t *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  out[i] = ((key[i] >> 32) << 38) | (static_cast<uint64_t>(colour[i]) << 32) | (key[i] & 0xffffffffull);
}

// Instances from the sorted final keys, and the count of each (patch, colour) pair.
__global__ void k_instances(int64_t ninst, int nc, const uint64_t *__restrict__ key,
                            const int32_t *__restrict__ by_rank, int32_t *__restrict__ inst,
                            uint8_t *__restrict__ icol, unsigned long long *__restrict__ cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const uint64_t k = key[i];
  const int64_t p = static_cast<int64_t>(k >> 38);
  const int c = static_cast<int>((k >> 32) & 63);
  inst[i] = by_rank[static_cast<uint32_t>(k)];
  icol[i] = static_cast<uint8_t>(c);
  atomicAdd(cnt + p * nc + c, 1ull);
}

// Instance records: the vertices of each instance's element (curve positions or LIDs) and its
// slot map,
// the position of LID b in the CRS row of LID a (uint8; 255: absent), stored in instance order
// so that a patch reads its instances as one contiguous stream whatever the element order.
__global__ void k_inst_records(int64_t ninst, int nb, const int32_t *__restrict__ inst,
                               const int32_t *__restrict__ elem_lid, const int64_t *__restrict__ rowptr,
                               const int32_t *__restrict__ col, const double *__restrict__ rec, int bypos,
                               int32_t *__restrict__ ivtx, uint8_t *__restrict__ slot) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ninst * nb) return;
  const int64_t i = t / nb, e = inst[i];
  const int a = static_cast<int>(t - i * nb);
  const int32_t ra = elem_lid[e * nb + a];
  ivtx[t] = bypos ? static_cast<int32_t>(__double_as_longlong(rec[4 * static_cast<int64_t>(ra) + 3])) : ra;
  const int64_t r0 = rowptr[ra], n = rowptr[ra + 1] - r0;
  for (int b = 0; b < nb; ++b) {
    const int32_t cb = elem_lid[e * nb + b];
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (col[r0 + mid] < cb) lo = mid + 1; else hi = mid;
    }
    slot[t * nb + b] = lo < n && col[r0 + lo] == cb ? static_cast<uint8_t>(lo) : uint8_t(255);
  }
}

// ---------------------------------------------------------------------------------------
// assembly

struct PatchParams {
  int P, RS, NC;
  const int64_t *cstart;       // [npatch*NC + 1] first instance of each (patch, colour)
  const int32_t *inst;         // element of each instance
  const uint8_t *icol;         // colour of each instance
  const int32_t *prow;         // [npatch*P] row at each curve position, -1 past the last
  const int64_t *r0p;          // [npatch*P] first CRS position of each row, in curve order
  const int32_t *lenp;         // [npatch*P] its length
  // vertices of each instance ([instances][NB]): curve positions when the state is read in curve
  // order (bypos: coordinates xyzp, state xp / xdp by position), else LIDs (coordinates and
  // position from rec, state x / xdot by LID)
  int bypos;
  const int32_t *ivtx;
  const double4 *xyzp;         // [npatch*P] vertex x, y, z in curve order (bypos)
  const double4 *rec;          // [rows] vertex x, y, z, curve position (int64 bits) by LID (otherwise)
  const uint8_t *islot;        // [instances][NB][NB] slot of LID b in the row of LID a
  const double *coef;
  const double *x, *xdot;      // by LID
  const double *xp, *xdp;      // [npatch*P] in curve order (k_state_by_pos; bypos)
  double alpha, beta, rhoc, src;
  double *f, *vals;
  int add;
  int64_t npatch;
};

// shards hex order -> binary order v = bx + 2by + 4bz (an involution)
__device__ __forceinline__ constexpr int hex_bin(int s) { return s ^ ((s >> 1) & 1); }

// One linear tet, 4-point degree-2 rule in closed form: K_ab = k |T| grad N_a . grad N_b,
// M_ab = rho_c |T| (1 + delta_ab) / 20, F_a = src |T| / 4 (the rule integrates these
// exactly; IntegrationValues2 weights carry |det J|, computeCellMeasure's sign).  R is FP64
// in both precisions: the flux k |T| grad u . grad N_a from the FP64 cofactors, then
// M udot - F; the matrix (Real) is beta K + alpha M.
template <typename Real, bool TRANS, bool UNIT>
__device__ __forceinline__ void tet4_element(const double (&X)[3][4], const double (&u)[4], const double (&ud)[4],
                                             double k, double rhoc, double src, double alpha, double beta,
                                             Real (&K)[10], double (&R)[4]) {
  double J[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int j = 0; j < 3; ++j) J[d][j] = X[d][j + 1] - X[d][0];
  double c[3][3];
  c[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  c[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  c[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  c[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  c[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  c[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  c[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  c[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  c[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = J[0][0] * c[0][0] + J[0][1] * c[0][1] + J[0][2] * c[0][2];
  const double sg = det < 0.0 ? -1.0 : 1.0;
  const double vol = det * sg / 6.0;
  const double cw = k * sg * recip(det) / 6.0;   // k |T| / det^2
  // grad N_{j+1} = c[:, j] / det, grad N_0 = -(sum); flux in FP64
  const double du0 = u[1] - u[0], du1 = u[2] - u[0], du2 = u[3] - u[0];
  double tt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tt[i] = cw * (c[i][0] * du0 + c[i][1] * du1 + c[i][2] * du2);
  double rs = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    R[j + 1] = c[0][j] * tt[0] + c[1][j] * tt[1] + c[2][j] * tt[2];
    rs += R[j + 1];
  }
  R[0] = -rs;
  const double fs = src * vol * 0.25;
  const double mw = rhoc * vol * 0.05;
  double sud = 0.0;
  if constexpr (TRANS) sud = ud[0] + ud[1] + ud[2] + ud[3];
#pragma unroll
  for (int a = 0; a < 4; ++a) R[a] = (TRANS ? R[a] + mw * (sud + ud[a]) : R[a]) - fs;
  double g[4][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    g[1][i] = c[i][0]; g[2][i] = c[i][1]; g[3][i] = c[i][2];
    g[0][i] = -(c[i][0] + c[i][1] + c[i][2]);
  }
  int n = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b, ++n) {
      double v = cw * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
      if (!TRANS && UNIT) { K[n] = Real(v); continue; }
      v *= beta;
      if (TRANS) v += alpha * mw * (a == b ? 2.0 : 1.0);
      K[n] = Real(v);
    }
}

// LDS row of a patch row: rows 32 apart along the curve (the same bank for an odd row stride)
// are spread over the banks by folding bits 5..9 of the row into bits 0..4
__device__ __forceinline__ int lds_row(int r) { return r ^ ((r >> 5) & 31); }

__device__ __forceinline__ constexpr int sym4(int a, int b) {
  return a <= b ? a * 4 - (a * (a - 1)) / 2 + (b - a) : b * 4 - (b * (b - 1)) / 2 + (a - b);
}

// One lane's instance of a chunk: geometry, state and coefficient, the rows' curve positions
// and the slot map, gathered through the instance's LIDs (which, with its element, are loaded a
// chunk ahead).
template <int NB, bool TRANS>
struct InstIn {
  double X[3][NB], u[NB], ud[TRANS ? NB : 1], k;
  int32_t pos[NB];
  uint32_t w[NB * NB / 4];   // slot map, four slots a word
};

template <int NB>
__device__ __forceinline__ void load_lids(const PatchParams &p, int64_t i, int64_t i_end, int32_t (&L)[NB],
                                          int32_t &e) {
#pragma unroll
  for (int v = 0; v < NB; ++v) L[v] = 0;
  e = 0;
  if (i >= i_end) return;
  e = p.inst[i];
  const int4 *l4 = reinterpret_cast<const int4 *>(p.ivtx + i * NB);
#pragma unroll
  for (int q = 0; q < NB / 4; ++q) {
    const int4 t = l4[q];
    L[4 * q] = t.x; L[4 * q + 1] = t.y; L[4 * q + 2] = t.z; L[4 * q + 3] = t.w;
  }
}

template <int NB, bool TRANS>
__device__ __forceinline__ void load_inst(const PatchParams &p, int64_t i, const int32_t (&L)[NB], int32_t e,
                                          InstIn<NB, TRANS> &in) {
#pragma unroll
  for (int v = 0; v < NB; ++v) {
    if (p.bypos) {   // L: curve positions
      const double4 r = p.xyzp[L[v]];
      in.X[0][v] = r.x; in.X[1][v] = r.y; in.X[2][v] = r.z;
      in.pos[v] = L[v];
      in.u[v] = p.xp[L[v]];
      if constexpr (TRANS) in.ud[v] = p.xdp[L[v]];
    } else {         // L: LIDs
      const double4 r = p.rec[L[v]];
      in.X[0][v] = r.x; in.X[1][v] = r.y; in.X[2][v] = r.z;
      in.pos[v] = static_cast<int32_t>(__double_as_longlong(r.w));
      in.u[v] = p.x[L[v]];
      if constexpr (TRANS) in.ud[v] = p.xdot[L[v]];
    }
  }
  in.k = p.coef[e];
  if (p.vals) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(p.islot + i * NB * NB);
#pragma unroll
    for (int q = 0; q < NB * NB / 16; ++q) {
      const uint4 t = s4[q];
      in.w[4 * q] = t.x; in.w[4 * q + 1] = t.y; in.w[4 * q + 2] = t.z; in.w[4 * q + 3] = t.w;
    }
  }
}

constexpr int kMaxRowsPerThread = 4;   // patches of at most 4 x 256 rows

// One workgroup per patch, its instances in chunks of one per lane.  A chunk's row positions
// load while the previous chunk adds its rows; everything is read in curve order.
// Whether every element of an uploaded hex8 mesh is a parallelepiped: the test of the lattice
// path's cell check (bilinear and trilinear Walsh coefficients of the vertex coordinates, in
// binary vertex order, below 2^-47 of the coordinate magnitude).  Any element that is not
// clears *flag.
__global__ void k_elem_affine(int64_t nelem, const int64_t *__restrict__ conn, const double *__restrict__ xyz,
                              int64_t nverts, int *flag) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  double nonaff = 0.0, mag = 0.0;
  for (int d = 0; d < 3; ++d) {
    double h[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) h[hex_bin(v)] = xyz[d * nverts + conn[8 * e + v]];
    walsh8(h);
    nonaff += 0.125 * (fabs(h[3]) + fabs(h[5]) + fabs(h[6]) + fabs(h[7]));
    mag += 0.125 * (fabs(h[0]) + fabs(h[1]) + fabs(h[2]) + fabs(h[4]));
  }
  if (!(nonaff <= 0x1p-47 * mag)) __hip_atomic_store(flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// AFF: every element is a parallelepiped (one metric per element, as the lattice kernels'
// kind for such cells).
template <int TOPO, typename Real, bool TRANS, bool UNIT, bool AFF>
__global__ __launch_bounds__(kPatchThreads) void k_patch(PatchParams p) {
  constexpr int NB = TOPO == TXA_HEX8 ? 8 : 4;
  extern __shared__ double sacc[];   // [P][RS] row entries, then [P] residual entries
  const int P = p.P, RS = p.RS, NC = p.NC;
  double *sres = sacc + P * RS;
  int64_t *srow0 = reinterpret_cast<int64_t *>(sres + P);   // first CRS position of each row
  int32_t *slid = reinterpret_cast<int32_t *>(srow0 + P);   // LID of each row (-1: none)
  int32_t *slen = slid + P;                                  // length of each row
  const int tid = threadIdx.x;
  const int64_t pb = blockIdx.x;
  const int64_t i_begin = p.cstart[pb * NC], i_end = p.cstart[(pb + 1) * NC];
  const int64_t pos0 = pb * P;
  int32_t Lc[NB];
  int32_t Ec;
  load_lids<NB>(p, i_begin + tid, i_end, Lc, Ec);
  for (int r = tid; r < P; r += kPatchThreads) {
    slid[r] = p.prow[pos0 + r];
    srow0[r] = p.r0p[pos0 + r];
    slen[r] = p.lenp[pos0 + r];
  }
  for (int i = tid; i < P * (RS + 1); i += kPatchThreads) sacc[i] = 0.0;
  __syncthreads();
  for (int64_t c0 = i_begin; c0 < i_end; c0 += kPatchThreads) {
    const int64_t ilast = (c0 + kPatchThreads < i_end ? c0 + kPatchThreads : i_end) - 1;
    const int cfirst = p.icol[c0], clast = p.icol[ilast];
    const int64_t i = c0 + tid;
    const int mycol = i < i_end ? p.icol[i] : -1;
    int lrow[NB];
    uint32_t w[NB * NB / 4];
    Real Ke[NB * (NB + 1) / 2];
    double Re[NB];
    if (mycol >= 0) {
      InstIn<NB, TRANS> nx;
      load_inst(p, i, Lc, Ec, nx);
#pragma unroll
      for (int v = 0; v < NB; ++v) {
        const int lp = nx.pos[v] - static_cast<int>(pos0);
        lrow[v] = lp >= 0 && lp < P ? lp : -1;
      }
#pragma unroll
      for (int q = 0; q < NB * NB / 4; ++q) w[q] = nx.w[q];
      if constexpr (TOPO == TXA_HEX8) {
        double Xb[3][8], ub[8], udb[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          Xb[0][hex_bin(s)] = nx.X[0][s]; Xb[1][hex_bin(s)] = nx.X[1][s]; Xb[2][hex_bin(s)] = nx.X[2][s];
          ub[hex_bin(s)] = nx.u[s];
          udb[hex_bin(s)] = TRANS ? nx.ud[s] : 0.0;
        }
        Hex8Out<Real, TRANS> o;
        hex8_element<Real, TRANS, UNIT, AFF>(Xb, ub, udb, nx.k, p.rhoc, p.src, p.alpha, p.beta, o, [] {});
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          Re[a] = o.R[hex_bin(a)];
#pragma unroll
          for (int b = a; b < 8; ++b) Ke[sym8(a, b)] = o.K[sym8(hex_bin(a), hex_bin(b))];
        }
      } else {
        double ud[NB];
#pragma unroll
        for (int v = 0; v < NB; ++v) ud[v] = TRANS ? nx.ud[v] : 0.0;
        tet4_element<Real, TRANS, UNIT>(nx.X, nx.u, ud, nx.k, p.rhoc, p.src, p.alpha, p.beta, Ke, Re);
      }
    }
    // the next chunk's LIDs
    load_lids<NB>(p, i + kPatchThreads, i_end, Lc, Ec);
    // colour by colour: within one colour no two instances share a vertex
    for (int c = cfirst; c <= clast; ++c) {
      __syncthreads();
      if (mycol == c) {
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          if (lrow[a] < 0) continue;
          if (p.f) __hip_atomic_fetch_add(sres + lrow[a], Re[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (p.vals) {
            double *row = sacc + lds_row(lrow[a]) * RS;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              const double kv = static_cast<double>(NB == 8 ? Ke[sym8(a, b)] : Ke[sym4(a, b)]);
              // one add per address and colour: the atomic is a single LDS instruction, not a race
              const uint32_t sab = (w[(a * NB + b) >> 2] >> (8 * ((a * NB + b) & 3))) & 0xffu;
              __hip_atomic_fetch_add(row + sab, kv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  // store the patch's rows whole: 32 lanes per row (values written once, not re-read here:
  // non-temporal stores)
  const int lane = tid & 31;
  if (p.add) {
    for (int r = tid >> 5; r < P; r += kPatchThreads / 32) {
      const int32_t L = slid[r];
      if (L < 0) continue;
      if (p.vals)
        for (int s = lane; s < slen[r]; s += 32) p.vals[srow0[r] + s] += sacc[lds_row(r) * RS + s];
      if (p.f && lane == 0) p.f[L] += sres[r];
    }
    return;
  }
  for (int r = tid >> 5; r < P; r += kPatchThreads / 32) {
    const int32_t L = slid[r];
    if (L < 0) continue;
    if (p.vals)
      for (int s = lane; s < slen[r]; s += 32)
        __builtin_nontemporal_store(sacc[lds_row(r) * RS + s], p.vals + srow0[r] + s);
    if (p.f && lane == 0) __builtin_nontemporal_store(sres[r], p.f + L);
  }
}

template <int TOPO, typename Real, bool AFF>
void launch_topo(const Mesh &m, const PatchParams &pp, bool tr, bool unit, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(m.npatch)), block(kPatchThreads);
  const size_t lds = patch_lds_bytes(pp.P, pp.RS);
  if (tr) hipLaunchKernelGGL((k_patch<TOPO, Real, true, false, AFF>), grid, block, lds, s, pp);
  else if (unit) hipLaunchKernelGGL((k_patch<TOPO, Real, false, true, AFF>), grid, block, lds, s, pp);
  else hipLaunchKernelGGL((k_patch<TOPO, Real, false, false, AFF>), grid, block, lds, s, pp);
  TXA_HIP(hipGetLastError());
}

}  // namespace

void patch_build(Mesh &m, hipStream_t s) {
  m.npatch = 0;
  m.patch_P = 0;
  m.patch_affine = false;
  const bool fits = !m.structured && m.nfields == 1 && m.blocks.size() == 1 &&
                    (m.topo == TXA_HEX8 || m.topo == TXA_TET4) && m.occ.p &&
                    m.nelem * (m.topo == TXA_HEX8 ? 8 : 4) < (int64_t(1) << 31);   // sort lengths are int
  if (!fits) return;
  const int nb = m.topo == TXA_HEX8 ? 8 : 4;
  const int64_t nrows = m.ndofs, nelem = m.nelem, npos = nelem * nb;
  // row stride in LDS: the longest row; patch size: the largest power of two whose rows fit
  const int rs = m.max_row;
  int P = kMaxRowsPerThread * kPatchThreads;
  while (P >= 32 && patch_lds_bytes(P, rs) > static_cast<size_t>(kPatchLdsBytes)) P /= 2;
  if (P < 32 || rs > 255) return;
  if (m.topo == TXA_HEX8) {
    DevBuf<int> flag;
    flag.alloc(1);
    const int one = 1;
    TXA_HIP(hipMemcpyAsync(flag.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_elem_affine, dim3(setup_grid(nelem)), dim3(kSetupBlock), 0, s, nelem, m.conn.p, m.xyz.p,
                       m.nverts, flag.p);
    TXA_HIP(hipGetLastError());
    int h = 0;
    TXA_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TXA_HIP(hipStreamSynchronize(s));
    m.patch_affine = h != 0;
  }
  // rows along the Morton curve of their vertices
  m.patch_rec.alloc(4 * nrows);
  hipLaunchKernelGGL(k_node_xyz, dim3(setup_grid(npos)), dim3(kSetupBlock), 0, s, npos, m.conn.p, m.elem_lid.p,
                     m.xyz.p, m.nverts, m.patch_rec.p);
  TXA_HIP(hipGetLastError());
  // integer lattice of about one vertex per unit cell (the spacing h of a lattice of as many
  // points spanning the bounding box), so that aligned Morton blocks of 2^k points are compact
  // blocks of vertices (on a brick, exactly its lattice)
  Box box;
  {
    std::vector<double> h(static_cast<size_t>(3 * m.nverts));
    TXA_HIP(hipMemcpyAsync(h.data(), m.xyz.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    TXA_HIP(hipStreamSynchronize(s));
    double ext[3], vol = 1.0;
    int dims = 0;
    for (int d = 0; d < 3; ++d) {
      const double *b = h.data() + d * m.nverts;
      const auto mm = std::minmax_element(b, b + m.nverts);
      box.lo[d] = *mm.first;
      ext[d] = *mm.second - *mm.first;
      if (ext[d] > 0.0) { vol *= ext[d]; ++dims; }
    }
    // the spacing of a lattice of nrows points spanning the box (n^(1/dims) points per axis)
    const double per_axis = dims ? std::pow(static_cast<double>(nrows), 1.0 / dims) : 1.0;
    const double hs = dims && per_axis > 1.5 ? std::pow(vol, 1.0 / dims) / (per_axis - 1.0) : 1.0;
    for (int d = 0; d < 3; ++d) box.scale[d] = ext[d] > 0.0 ? std::min(1.0 / hs, 2097150.0 / ext[d]) : 0.0;
  }
  DevBuf<uint64_t> key, key_s;
  DevBuf<int32_t> row, row_s;
  key.alloc(nrows); key_s.alloc(nrows); row.alloc(nrows); row_s.alloc(nrows);
  hipLaunchKernelGGL(k_morton, dim3(setup_grid(nrows)), dim3(kSetupBlock), 0, s, nrows, m.patch_rec.p, box, key.p,
                     row.p);
  TXA_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  TXA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key_s.p, row.p, row_s.p,
                                             static_cast<int>(nrows), 0, 63, s));
  DevBuf<char> tmp;
  tmp.alloc(tmp_bytes);
  TXA_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key_s.p, row.p, row_s.p,
                                             static_cast<int>(nrows), 0, 63, s));
  const int64_t npatch = (nrows + P - 1) / P;
  m.patch_rows.alloc(npatch * P);
  m.patch_xyz.alloc(4 * npatch * P);
  m.patch_r0.alloc(npatch * P);
  m.patch_len.alloc(npatch * P);
  hipLaunchKernelGGL(k_place_rows, dim3(setup_grid(npatch * P)), dim3(kSetupBlock), 0, s, nrows, npatch * P,
                     row_s.p, m.patch_rows.p, m.patch_rec.p, m.patch_xyz.p, m.rowptr.p, m.patch_r0.p,
                     m.patch_len.p);
  TXA_HIP(hipGetLastError());
  key.release(); key_s.release(); row.release(); row_s.release(); tmp.release();
  {   // read the state in curve order when the patches' rows scatter over the by-LID arrays (more
      // than 0.3 128-byte lines per row; a brick in its own order touches about 0.19, the same brick
      // with its elements shuffled about 0.4)
    const int64_t npad = npatch * P;
    DevBuf<uint64_t> lk, lk_s;
    lk.alloc(npad);
    lk_s.alloc(npad);
    hipLaunchKernelGGL(k_line_keys, dim3(setup_grid(npad)), dim3(kSetupBlock), 0, s, npad, P, m.patch_rows.p, lk.p);
    TXA_HIP(hipGetLastError());
    size_t lb = 0;
    TXA_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, lb, lk.p, lk_s.p, static_cast<int>(npad), 0, 64, s));
    DevBuf<char> lt;
    lt.alloc(lb);
    TXA_HIP(hipcub::DeviceRadixSort::SortKeys(lt.p, lb, lk.p, lk_s.p, static_cast<int>(npad), 0, 64, s));
    DevBuf<unsigned long long> nl;
    nl.alloc(1);
    TXA_HIP(hipMemsetAsync(nl.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_count_distinct, dim3(setup_grid(npad)), dim3(kSetupBlock), 0, s, npad, lk_s.p, nl.p);
    TXA_HIP(hipGetLastError());
    unsigned long long h = 0;
    TXA_HIP(hipMemcpyAsync(&h, nl.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TXA_HIP(hipStreamSynchronize(s));
    m.patch_bypos = 10 * static_cast<int64_t>(h) > 3 * nrows;
  }
  if (!m.patch_bypos) m.patch_xyz.release();
  // elements in sweep order (their centroids on the lattice, z, y, x)
  DevBuf<int32_t> rank, by_rank;
  {
    DevBuf<uint64_t> ek, ek_s;
    DevBuf<int32_t> el;
    ek.alloc(nelem); ek_s.alloc(nelem); el.alloc(nelem); by_rank.alloc(nelem); rank.alloc(nelem);
    hipLaunchKernelGGL(k_sweep_keys, dim3(setup_grid(nelem)), dim3(kSetupBlock), 0, s, nelem, nb, m.elem_lid.p,
                       m.patch_rec.p, box, ek.p, el.p);
    TXA_HIP(hipGetLastError());
    tmp_bytes = 0;
    TXA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ek.p, ek_s.p, el.p, by_rank.p,
                                               static_cast<int>(nelem), 0, 63, s));
    tmp.alloc(tmp_bytes);
    TXA_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, ek.p, ek_s.p, el.p, by_rank.p,
                                               static_cast<int>(nelem), 0, 63, s));
    hipLaunchKernelGGL(k_rank_of, dim3(setup_grid(nelem)), dim3(kSetupBlock), 0, s, nelem, by_rank.p, rank.p);
    TXA_HIP(hipGetLastError());
    TXA_HIP(hipStreamSynchronize(s));
    tmp.release();
  }
  // instances: (patch, sweep rank), sorted; then coloured patch by patch
  DevBuf<uint64_t> ik, ik_s;
  ik.alloc(npos);
  ik_s.alloc(npos);
  DevBuf<unsigned long long> nin, per;
  nin.alloc(1);
  per.alloc(npatch + 1);
  TXA_HIP(hipMemsetAsync(nin.p, 0, sizeof(unsigned long long), s));
  TXA_HIP(hipMemsetAsync(per.p, 0, sizeof(unsigned long long) * (npatch + 1), s));
  hipLaunchKernelGGL(k_instance_keys, dim3(setup_grid(nelem)), dim3(kSetupBlock), 0, s, nelem, nb, P, m.elem_lid.p,
                     m.patch_rec.p, rank.p, ik.p, nin.p, per.p);
  TXA_HIP(hipGetLastError());
  tmp_bytes = 0;
  TXA_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, ik.p, ik_s.p, static_cast<int>(npos), 0, 64, s));
  tmp.alloc(tmp_bytes);
  TXA_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, ik.p, ik_s.p, static_cast<int>(npos), 0, 64, s));
  unsigned long long ninst = 0;
  TXA_HIP(hipMemcpyAsync(&ninst, nin.p, sizeof(ninst), hipMemcpyDeviceToHost, s));
  TXA_HIP(hipStreamSynchronize(s));
  DevBuf<int64_t> pstart;
  pstart.alloc(npatch + 1);
  exclusive_scan_i64(reinterpret_cast<const int64_t *>(per.p), pstart.p, npatch + 1, s);
  DevBuf<unsigned long long> rowmask;
  rowmask.alloc(npatch * P);
  TXA_HIP(hipMemsetAsync(rowmask.p, 0, rowmask.bytes(), s));
  DevBuf<uint8_t> icol;
  icol.alloc(ninst);
  DevBuf<int> maxc;
  maxc.alloc(1);
  TXA_HIP(hipMemsetAsync(maxc.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_patch_colour, dim3(setup_grid(npatch)), dim3(kSetupBlock), 0, s, npatch, nb, P, pstart.p,
                     ik_s.p, by_rank.p, m.elem_lid.p, m.patch_rec.p, rowmask.p, icol.p, maxc.p);
  TXA_HIP(hipGetLastError());
  int nc = 0;
  TXA_HIP(hipMemcpyAsync(&nc, maxc.p, sizeof(int), hipMemcpyDeviceToHost, s));
  TXA_HIP(hipStreamSynchronize(s));
  if (nc >= kMaxColours) {   // more colours than the masks hold: no patch path
    m.patch_rows.release();
    m.patch_rec.release();
    return;
  }
  nc += 1;
  rowmask.release();
  // final order: (patch, colour, sweep rank)
  hipLaunchKernelGGL(k_final_keys, dim3(setup_grid(static_cast<int64_t>(ninst))), dim3(kSetupBlock), 0, s,
                     static_cast<int64_t>(ninst), ik_s.p, icol.p, ik.p);
  TXA_HIP(hipGetLastError());
  TXA_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, ik.p, ik_s.p, static_cast<int>(ninst), 0, 64, s));
  ik.release();
  tmp.release();
  m.patch_inst.alloc(ninst);
  m.patch_icol.alloc(ninst);
  DevBuf<unsigned long long> pc;
  pc.alloc(npatch * nc + 1);
  TXA_HIP(hipMemsetAsync(pc.p, 0, sizeof(unsigned long long) * (npatch * nc + 1), s));
  hipLaunchKernelGGL(k_instances, dim3(setup_grid(static_cast<int64_t>(ninst))), dim3(kSetupBlock), 0, s,
                     static_cast<int64_t>(ninst), nc, ik_s.p, by_rank.p, m.patch_inst.p, m.patch_icol.p, pc.p);
  TXA_HIP(hipGetLastError());
  m.patch_cstart.alloc(npatch * nc + 1);
  exclusive_scan_i64(reinterpret_cast<const int64_t *>(pc.p), m.patch_cstart.p, npatch * nc + 1, s);
  // instance records
  m.patch_ivtx.alloc(ninst * nb);
  m.patch_islot.alloc(ninst * nb * nb);
  hipLaunchKernelGGL(k_inst_records, dim3(setup_grid(static_cast<int64_t>(ninst) * nb)), dim3(kSetupBlock), 0, s,
                     static_cast<int64_t>(ninst), nb, m.patch_inst.p, m.elem_lid.p, m.rowptr.p, m.col.p,
                     m.patch_rec.p, int(m.patch_bypos), m.patch_ivtx.p, m.patch_islot.p);
  TXA_HIP(hipGetLastError());
  TXA_HIP(hipStreamSynchronize(s));
  if (m.patch_bypos) m.patch_rec.release();   // read by LID only when the state is
  m.patch_P = P;
  m.patch_RS = rs;
  m.patch_NC = nc;
  m.npatch = npatch;
  m.patch_ninst = static_cast<int64_t>(ninst);
}

bool patch_supported(const Mesh &m, int physics, int precision) {
  return m.npatch > 0 && physics == TXA_DIFFUSION && (precision == TXA_FP64 || precision == TXA_FP32);
}

size_t patch_workspace_bytes(const Mesh &m) {
  return m.npatch > 0 && m.patch_bypos ? 2 * sizeof(double) * static_cast<size_t>(m.npatch) * m.patch_P : 0;
}

void launch_patch(const Mesh &m, const txa_assembly_args &a, hipStream_t s, void *ws) {
  PatchParams pp;
  pp.P = m.patch_P; pp.RS = m.patch_RS; pp.NC = m.patch_NC;
  pp.cstart = m.patch_cstart.p; pp.inst = m.patch_inst.p; pp.icol = m.patch_icol.p;
  pp.prow = m.patch_rows.p;
  pp.r0p = m.patch_r0.p; pp.lenp = m.patch_len.p;
  pp.bypos = m.patch_bypos;
  pp.ivtx = m.patch_ivtx.p; pp.islot = m.patch_islot.p;
  pp.xyzp = reinterpret_cast<const double4 *>(m.patch_xyz.p);
  pp.rec = reinterpret_cast<const double4 *>(m.patch_rec.p);
  pp.coef = a.coef; pp.x = a.x; pp.xdot = a.xdot;
  pp.xp = pp.xdp = nullptr;
  if (m.patch_bypos) {   // the state in curve order, in the workspace
    const int64_t npad = m.npatch * m.patch_P;
    TXA_REQUIRE(ws != nullptr, TXA_ERR_LOGIC, "the node-patch assembly needs its workspace");
    double *xp = static_cast<double *>(ws), *xdp = a.transient ? xp + npad : nullptr;
    hipLaunchKernelGGL(k_state_by_pos, dim3(setup_grid(npad)), dim3(kSetupBlock), 0, s, npad, m.patch_rows.p, a.x,
                       a.xdot, xp, xdp);
    TXA_HIP(hipGetLastError());
    pp.xp = xp; pp.xdp = xdp;
  }
  pp.alpha = a.alpha; pp.beta = a.beta; pp.rhoc = a.rho; pp.src = a.src[0];
  pp.f = a.f; pp.vals = a.vals; pp.add = a.accumulate != 0;
  pp.npatch = m.npatch;
  const bool tr = a.transient != 0, unit = !tr && a.beta == 1.0;
  if (m.topo == TXA_HEX8) {
    if (m.patch_affine) {
      if (a.precision == TXA_FP32) launch_topo<TXA_HEX8, float, true>(m, pp, tr, unit, s);
      else launch_topo<TXA_HEX8, double, true>(m, pp, tr, unit, s);
    } else {
      if (a.precision == TXA_FP32) launch_topo<TXA_HEX8, float, false>(m, pp, tr, unit, s);
      else launch_topo<TXA_HEX8, double, false>(m, pp, tr, unit, s);
    }
  } else {
    if (a.precision == TXA_FP32) launch_topo<TXA_TET4, float, false>(m, pp, tr, unit, s);
    else launch_topo<TXA_TET4, double, false>(m, pp, tr, unit, s);
  }
}

}  // namespace txa
