// Q1 scalar diffusion on the structured brick (hex8 cells, or the 6-tet split of
// each cell): the row-owner, z-marching assembly kernels.
//
// Replaces Gather -> DOF/DOFGradient -> Integrator_GradBasisDotVector (k grad u),
// Integrator_BasisTimesScalar (source, transient rho_c udot) -> ScatterResidual_Tpetra
// (disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-650,
//  Panzer_DOFGradient_impl.hpp:60-160, Panzer_Integrator_GradBasisDotVector_impl.hpp:225-294,
//  Panzer_Integrator_BasisTimesScalar_impl.hpp:215-270, Panzer_ScatterResidual_Tpetra_impl.hpp:375-520)
// for HGRAD_HEX_C1 with 2x2x2 Gauss, with IntegrationValues2 / BasisValues2
// (disc-fe/src/Panzer_IntegrationValues2.cpp:949-1221, Panzer_BasisValues2_impl.hpp:1379-1460)
// recomputed per element in registers.
//
// The element routine (sum-factorised stiffness, FP64 flux residual of the FP32 form) is
// in txa_q1_elem.hpp.
#include "txa_dirichlet.hpp"
#include "txa_internal.hpp"
#include "txa_q1_elem.hpp"

#include <algorithm>
#include <cstring>

namespace txa {
namespace {

// Element tiles: TX vertex columns along x by TY along y, one element lane and one owner
// lane per column.  Single meshes take 32 x 8 (a 16 x 16 tile recomputes fewer halo
// elements but ran 1.5% slower at 256^3); batches take 32 x 8 or 16 x 16, whichever the
// autotuner times faster (16 x 16 covers meshes 16-64 cells wide with fewer idle lanes).
constexpr int NE = 256;                        // element lanes = owner lanes
constexpr int NT = 2 * NE;                     // block: 4 element waves + 4 owner waves
constexpr int kTX = 32;                        // tile width of the single-mesh kernels
template <int TXV>
struct Tile {
  static_assert(TXV == 16 || TXV == 32, "tile width");
  static constexpr int TX = TXV, TY = NE / TXV;
  // owner groups (the TX lanes of one x line of owners, TX-1 of them own rows): doubles per
  // group's row buffer, TX-1 rows of at most 27 entries, then a spare slot for the entries a
  // row does not have
  static constexpr int ROWBUF = 27 * TX, ROWBUF_SPARE = 27 * (TX - 1);
  static constexpr int NOWN = TX * (TY - 1);   // owner lanes with rows
};
using namespace q1elem;
// element kinds of the lattice kernels: TXA_HEX8 (general trilinear cells), TXA_TET4 (six
// tets per cell) and hex8 cells that are all parallelepipeds (constant metric per cell)
constexpr int kHex8Affine = 16 | TXA_HEX8;
constexpr int NSTAGE = 44;   // per element: 36 packed K, then 8 FP64 R (FP64) or 4 FP64 x-pair sums of R (FP32)
// Stage slots: FP32 stages are double-buffered (element waves publish layer z in one
// slot while the owner waves consume layer z-1 from the other, one barrier per
// layer); an FP64 stage takes one slot (two barriers per layer) to stay within LDS.
template <typename Real>
constexpr int stage_slots() { return sizeof(Real) == 4 ? 2 : 1; }

constexpr int kTargetBlocks = 2560;   // z slabs are sized to give about this many workgroups

struct Q1DiffParams {
  int nx, ny, nz;
  int zc;                      // vertex layers owned per workgroup (blockIdx.z selects the slab)
  const double *X, *Y, *Z;     // SoA vertex coordinates
  const int32_t *lid;          // LID of each vertex (one field)
  const double *coef;          // k per element
  const double *x, *xdot;      // by LID
  double alpha, beta, rhoc, src;
  double *f, *vals;
  const int64_t *rowptr;
  const uint8_t *perm;         // [rows][32] stencil slot of the 27 neighbours, 255 = absent
};



// The six tets of a cell (CartesianConnManager's 6-tet split,
// dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:228-235), vertices
// renumbered from shards hex order to the binary order v = bx + 2by + 4bz.
__device__ __forceinline__ constexpr int tet_v(int t, int j) {
  constexpr int T[6][4] = {{0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 5, 7, 4}, {1, 5, 7, 0}, {1, 7, 3, 0}};
  return T[t][j];
}

// The six linear tets of one cell, assembled into the cell's 8x8 matrix (pairs
// not joined by a tet edge stay zero).  Constant gradients: grad N_j = cof[:,j]/det
// (j = 1..3), grad N_0 = -(sum); measure w |det| with w = 1/6; the degree-2
// rule integrates the source (|T|/4 per vertex) and mass (|T|/20 (1 + delta)) exactly.
// k: the six tet coefficients, in element order 6*cell + t.
template <typename Real, bool TRANS, bool UNIT, class Mid>
__device__ __forceinline__ void tet6_cell(const double (&X)[3][8], const double (&u)[8], const double (&ud)[8],
                                          const double (&k)[6], double rhoc, double src, double alpha, double beta,
                                          Hex8Out<Real, TRANS> &o, Mid &&mid) {
  constexpr bool kFlux = flux_residual<Real>();
  mid();
  Real K[36], M[TRANS ? 36 : 1];
  double F[8], Rf[8];   // source load; FP64 residual K u + M udot of the FP32 form
#pragma unroll
  for (int i = 0; i < 36; ++i) { K[i] = Real(0); if (TRANS) M[i] = Real(0); }
#pragma unroll
  for (int v = 0; v < 8; ++v) { F[v] = 0.0; Rf[v] = 0.0; }
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    double J[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int j = 0; j < 3; ++j) J[d][j] = X[d][tet_v(t, j + 1)] - X[d][tet_v(t, 0)];
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
    const double vol = det * sg / 6.0;                 // sum of the weighted measure
    const double cw64 = k[t] * sg * recip(det) / 6.0;
    const Real cw = Real(cw64);
    if constexpr (kFlux) {
      // FP64: grad u (times det) = cof (u_j - u_0), R_j = cw g_j . (cof du), R_0 = -(R_1 + R_2 + R_3);
      // mass rho_c |T| / 20 (udot_a + sum_b udot_b)
      const double du0 = u[tet_v(t, 1)] - u[tet_v(t, 0)], du1 = u[tet_v(t, 2)] - u[tet_v(t, 0)],
                   du2 = u[tet_v(t, 3)] - u[tet_v(t, 0)];
      double tt[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) tt[i] = cw64 * (c[i][0] * du0 + c[i][1] * du1 + c[i][2] * du2);
      double rs = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double r = c[0][j] * tt[0] + c[1][j] * tt[1] + c[2][j] * tt[2];
        Rf[tet_v(t, j + 1)] += r;
        rs += r;
      }
      Rf[tet_v(t, 0)] -= rs;
      if constexpr (TRANS) {
        const double sud = ud[tet_v(t, 0)] + ud[tet_v(t, 1)] + ud[tet_v(t, 2)] + ud[tet_v(t, 3)];
        const double mw = rhoc * vol * 0.05;
#pragma unroll
        for (int a = 0; a < 4; ++a) Rf[tet_v(t, a)] += mw * (sud + ud[tet_v(t, a)]);
      }
    }
    Real g[4][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      g[1][i] = Real(c[i][0]); g[2][i] = Real(c[i][1]); g[3][i] = Real(c[i][2]);
      g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      F[tet_v(t, a)] += src * vol * 0.25;
#pragma unroll
      for (int b = a; b < 4; ++b) {
        const Real gg = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
        K[sym8(tet_v(t, a), tet_v(t, b))] += cw * gg;
        if (TRANS) M[sym8(tet_v(t, a), tet_v(t, b))] += Real(rhoc * vol * (a == b ? 0.1 : 0.05));
      }
    }
  }
  if constexpr (kFlux) {
#pragma unroll
    for (int a = 0; a < 8; ++a) o.R[a] = Rf[a] - F[a];
  } else {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      Real r = Real(0);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        r += K[sym8(a, b)] * Real(u[b]);
        if (TRANS) r += M[sym8(a, b)] * Real(ud[b]);
      }
      o.R[a] = double(r) - F[a];
    }
  }
#pragma unroll
  for (int i = 0; i < 36; ++i)
    o.K[i] = TRANS ? Real(beta) * K[i] + Real(alpha) * M[i] : (UNIT ? K[i] : Real(beta) * K[i]);
}

// Coefficients of one cell: k of the hex8 element, or of the six tets (element 6*cell + t).
template <int KIND>
struct CellCoef {
  double k[KIND == TXA_TET4 ? 6 : 1];
};

template <int KIND>
__device__ __forceinline__ void load_coef(const Q1DiffParams &p, int64_t cell, CellCoef<KIND> &c) {
  constexpr int n = KIND == TXA_TET4 ? 6 : 1;
#pragma unroll
  for (int t = 0; t < n; ++t) c.k[t] = p.coef[n * cell + t];
}

// Cell routine of the element kind: one hex8 element, or the six tets of the cell.
template <int KIND, typename Real, bool TRANS, bool UNIT, class Mid>
__device__ __forceinline__ void cell_element(const Q1DiffParams &p, const CellCoef<KIND> &c, const double (&X)[3][8],
                                             const double (&u)[8], const double (&ud)[8], Hex8Out<Real, TRANS> &o,
                                             Mid &&mid) {
  if constexpr (KIND == TXA_TET4) tet6_cell<Real, TRANS, UNIT>(X, u, ud, c.k, p.rhoc, p.src, p.alpha, p.beta, o, mid);
  else hex8_element<Real, TRANS, UNIT, KIND == kHex8Affine>(X, u, ud, c.k[0], p.rhoc, p.src, p.alpha, p.beta, o, mid);
}

// Global-memory view of an output pointer.  Stores through a generic pointer become
// flat stores, which the LDS counter also tracks, so an LDS wait would wait for them.
using gdouble = __attribute__((address_space(1))) double;
__device__ __forceinline__ gdouble *global_ptr(double *q) { return (gdouble *)q; }

__device__ __forceinline__ void scatter_row(const Q1DiffParams &p, int32_t L, int64_t base, const uint32_t (&w)[8],
                                            const double (&acc)[27]) {
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const uint32_t o = (w[s >> 2] >> (8 * (s & 3))) & 0xffu;
    if (o != 0xffu) global_ptr(p.vals)[base + o] = acc[s];
  }
}

// Orders LDS writes and reads between the lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes the finished rows of one vertex layer; called by the owner waves.
// Rows are written group by group (the 32 lanes of one x line of owners): when
// the owners of a group hold consecutive LIDs their rows are one contiguous CSR
// range, assembled in LDS in column order and stored with coalesced writes;
// otherwise each owner stores its own row entry by entry.
struct RowMeta {
  int32_t L = -1;           // LID of the row (-1: no row on this lane)
  int64_t base = 0, end = 0;
  uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // stencil slot map, 27 bytes used
};

// Row pointers and slot map of row L, issued ahead of use so the loads overlap the element work.
__device__ __forceinline__ RowMeta load_meta(const Q1DiffParams &p, int32_t L) {
  RowMeta m;
  m.L = L;
  if (L >= 0 && p.vals) {
    m.base = p.rowptr[L];
    m.end = p.rowptr[L + 1];
    const uint4 *pp = reinterpret_cast<const uint4 *>(p.perm + static_cast<int64_t>(L) * 32);
    const uint4 w0 = pp[0], w1 = pp[1];
    m.w[0] = w0.x; m.w[1] = w0.y; m.w[2] = w0.z; m.w[3] = w0.w;
    m.w[4] = w1.x; m.w[5] = w1.y; m.w[6] = w1.z; m.w[7] = w1.w;
  }
  return m;
}

// The range is stored in chunks of TX entries aligned to TX entries.  It removes the partial lines
// inside a range (WRITE_SIZE 4.32 -> 3.81 GB at 256^3, the bytes of f and the values):
// about 1% faster in both the FP32 and the FP64 kernels at 256^3.
template <int TXV>
__device__ __forceinline__ void emit_rows(const Q1DiffParams &p, const RowMeta &m, const double (&acc)[27], double r,
                                          double *rowbuf, int grp) {
  constexpr int GW = TXV;   // lanes per group
  constexpr int ROWBUF = Tile<TXV>::ROWBUF, ROWBUF_SPARE = Tile<TXV>::ROWBUF_SPARE;
  constexpr uint64_t GMASK = (1ull << GW) - 1;
  const int lane = threadIdx.x & 63, gl = lane & (GW - 1), gi = lane / GW;
  const bool act = m.L >= 0;
  const int32_t L = m.L;
  const int64_t base = m.base, end = m.end;
  if (act && p.f) global_ptr(p.f)[L] = r;
  if (!p.vals) return;   // uniform over the block
  // owners of a group are a prefix of its lanes; contiguous when successive LIDs differ by one
  const int32_t Lnext = __shfl_down(L, 1, GW);
  const bool next_act = __shfl_down(static_cast<int>(act), 1, GW) != 0 && gl < GW - 1;
  const bool ok = !act || !next_act || Lnext == L + 1;
  const uint64_t okm = __ballot(ok), actm = __ballot(act);
  const uint64_t gok = (okm >> (GW * gi)) & GMASK, gact = (actm >> (GW * gi)) & GMASK;
  const int nact = __popcll(gact);
  const bool contiguous = gok == GMASK && nact > 0;
  const int64_t base0 = __shfl(base, 0, GW);
  const int64_t end_last = __shfl(end, nact > 0 ? nact - 1 : 0, GW);
  double *buf = rowbuf + grp * ROWBUF;   // dereferenced only by groups with a row: grp < TY-1
  // interior rows: the 27 neighbours sit in the row in lattice order (slot s at s), so the
  // per-entry decode of the slot map is skipped
  const bool ident = m.w[0] == 0x03020100u && m.w[1] == 0x07060504u && m.w[2] == 0x0b0a0908u &&
                     m.w[3] == 0x0f0e0d0cu && m.w[4] == 0x13121110u && m.w[5] == 0x17161514u &&
                     (m.w[6] & 0x00ffffffu) == 0x001a1918u;
  const bool fast = __all(!act || (ident && contiguous));   // uniform over the wave
  if (fast) {
    if (act) {
      const int64_t off = base - base0;
#pragma unroll
      for (int s = 0; s < 27; ++s) buf[off + s] = acc[s];
    }
  } else if (contiguous && act) {
    const int64_t off = base - base0;
    // branch-free: absent stencil positions go to a spare slot past the group's rows
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      const uint32_t o = (m.w[s >> 2] >> (8 * (s & 3))) & 0xffu;
      buf[o == 0xffu ? ROWBUF_SPARE : off + o] = acc[s];
    }
  }
  wave_lds_sync();   // the group's lanes live in one wavefront
  // aligned: each group's part of a store instruction covers one aligned chunk of GW
  // entries of the range; lanes before the range's first entry idle in the first chunk,
  // so only the two end chunks of the range are partial
  const int total = static_cast<int>(end_last - base0);
  const int lead = static_cast<int>((reinterpret_cast<uintptr_t>(p.vals + base0) >> 3) & (GW - 1));
  int nchw = contiguous ? (total + lead + GW - 1) / GW : 0;   // chunks of the group's range
#pragma unroll
  for (int o = GW; o < 64; o <<= 1) nchw = max(nchw, __shfl_xor(nchw, o));   // uniform over the wave
  if (contiguous) {
    // chunk c: lane gl stores range entry j = GW c + gl - lead when 0 <= j < total.  Four
    // LDS reads are issued before their four stores (one LDS round trip per four chunks);
    // the reads of idle lanes may fall outside the group's buffer and are not used.
    const int j0 = gl - lead;
    const double *src = buf + j0;
    gdouble *dst = global_ptr(p.vals + base0 + j0);
    for (int c = 0; c < nchw; c += 4) {
      double v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = src[GW * (c + i)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (static_cast<unsigned>(j0 + GW * (c + i)) < static_cast<unsigned>(total)) {
          __builtin_nontemporal_store(v[i], dst + GW * (c + i));   // written once, not re-read
        }
    }
  } else if (act) {
    scatter_row(p, L, base, m.w, acc);
  }
}

// Row metadata loads for the element waves: no branch, so the loaded registers
// reach their first use unchanged (lanes without a row read row 0 and publish
// L = -1; their values are never used).
__device__ __forceinline__ RowMeta load_meta_flat(const Q1DiffParams &p, int32_t L) {
  RowMeta m;
  m.L = L;
  const int32_t r = L >= 0 ? L : 0;
  m.base = p.rowptr[r];
  m.end = p.rowptr[r + 1];
  const uint4 *pp = reinterpret_cast<const uint4 *>(p.perm + static_cast<int64_t>(r) * 32);
  const uint4 w0 = pp[0], w1 = pp[1];
  m.w[0] = w0.x; m.w[1] = w0.y; m.w[2] = w0.z; m.w[3] = w0.w;
  m.w[4] = w1.x; m.w[5] = w1.y; m.w[6] = w1.z; m.w[7] = w1.w;
  return m;
}

// Row metadata handed from the element waves to the owner waves through LDS, one
// slot per stage slot, structure of arrays over the owner lanes (ty < TY-1):
// [0] LID (-1: no row), [1] [2] low and high words of the row's first CRS
// position, [3..9] the slot map words 0-6 with the row length in the top byte of
// word 6 (byte 27 of the map, which carries no slot).
constexpr int MWORDS = 10;
template <int TXV>
constexpr int mslot() { return MWORDS * Tile<TXV>::NOWN; }

template <int TXV>
__device__ __forceinline__ void put_meta(uint32_t *ms, int t, const RowMeta &m) {
  constexpr int NOWN = Tile<TXV>::NOWN;
  ms[t] = static_cast<uint32_t>(m.L);
  ms[NOWN + t] = static_cast<uint32_t>(static_cast<uint64_t>(m.base));
  ms[2 * NOWN + t] = static_cast<uint32_t>(static_cast<uint64_t>(m.base) >> 32);
#pragma unroll
  for (int i = 0; i < 6; ++i) ms[(3 + i) * NOWN + t] = m.w[i];
  ms[9 * NOWN + t] = (m.w[6] & 0x00ffffffu) | (static_cast<uint32_t>(m.end - m.base) << 24);
}

template <int TXV>
__device__ __forceinline__ RowMeta get_meta(const uint32_t *ms, int t) {
  constexpr int NOWN = Tile<TXV>::NOWN;
  RowMeta m;
  m.L = static_cast<int32_t>(ms[t]);
  m.base = static_cast<int64_t>(ms[NOWN + t] | static_cast<uint64_t>(ms[2 * NOWN + t]) << 32);
#pragma unroll
  for (int i = 0; i < 7; ++i) m.w[i] = ms[(3 + i) * NOWN + t];
  m.end = m.base + (m.w[6] >> 24);
  return m;
}

// Whether corner (CX, CY) is the first of the order (1,1), (0,1), (1,0), (0,0) to reach
// stencil column (x, y): corner (CX, CY) reaches x in {1-CX, 2-CX}, y in {1-CY, 2-CY}.
__device__ __forceinline__ constexpr bool first_touch(int CX, int CY, int x, int y) {
  return (CX == 1 && CY == 1) || (CX == 0 && CY == 1 && x == 2) || (CX == 1 && CY == 0 && y == 2) ||
         (CX == 0 && CY == 0 && x == 2 && y == 2);
}

// Add rows a_lo = CX + 2CY (bottom) and a_lo + 4 (top) of the element at corner
// (CX, CY) of the owner's vertex column, read from the stage.  Called for the corners
// in the order (1,1), (0,1), (1,0), (0,0): the first contribution to a slot of the
// next layer's row (accn, rn) or to the upper stencil layer of this row (acc[18..26])
// is stored rather than added, so those slots need no clearing between layers.
template <int CX, int CY, typename Real>
__device__ __forceinline__ void take_corner(const Real *s, int col, double (&acc)[27], double (&accn)[18], double &rc,
                                            double &rn) {
  constexpr int alo = CX + 2 * CY, ahi = alo + 4;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
    const int x = bx - CX + 1, y = by - CY + 1, st = x + 3 * y;
    const bool first = first_touch(CX, CY, x, y);
    const double lo = double(s[sym8(alo, b) * NE + col]), hi = double(s[sym8(ahi, b) * NE + col]);
    if (bz == 1 && first) acc[st + 18] = lo;
    else acc[st + 9 * (bz + 1)] += lo;
    if (first) accn[st + 9 * bz] = hi;
    else accn[st + 9 * bz] += hi;
  }
  if constexpr (!flux_residual<Real>()) {   // residual rows: FP64 values, 8 per element
    const double *rs = reinterpret_cast<const double *>(s + 36 * NE);
    rc += rs[alo * NE + col];
    if (CX == 1 && CY == 1) rn = rs[ahi * NE + col];
    else rn += rs[ahi * NE + col];
  }
}

// Residual of the FP32 form: the stage holds, per element column c, the x-pair sums
// P_c[by + 2 bz] = R_c[1 + 2by + 4bz] + R_{c+1}[2by + 4bz] (FP64) of the element and its
// +x neighbour, so the owner at corner (1,1) of column c adds two values per vertex layer.
template <int TXV, typename Real>
__device__ __forceinline__ void take_pairs(const Real *s, int col, double &rc, double &rn) {
  constexpr int TX = TXV;
  if constexpr (flux_residual<Real>()) {
    const double *ps = reinterpret_cast<const double *>(s + 36 * NE);
    rc += ps[1 * NE + col] + ps[0 * NE + col + TX];
    rn = ps[3 * NE + col] + ps[2 * NE + col + TX];
  }
}

// Vertex data of one face (4 vertices) of an element column.
template <bool TRANS>
struct Face {
  double x[3][4], u[4], ud[4];
};

template <bool TRANS>
__device__ __forceinline__ void load_face(const Q1DiffParams &p, int64_t n0, int64_t sx, const int32_t (&L)[4],
                                          Face<TRANS> &f) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t n = n0 + (v & 1) + sx * (v >> 1);
    f.x[0][v] = p.X[n]; f.x[1][v] = p.Y[n]; f.x[2][v] = p.Z[n];
    f.u[v] = p.x[L[v]];
    f.ud[v] = TRANS ? p.xdot[L[v]] : 0.0;
  }
}

__device__ __forceinline__ void load_face_lids(const Q1DiffParams &p, int64_t n0, int64_t sx, int32_t (&L)[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) L[v] = p.lid[n0 + (v & 1) + sx * (v >> 1)];
}

// Block = 4 element waves + 4 owner waves.  Per element layer ez the element
// waves compute layer ez into registers while the owner waves sum layer ez-1
// from the stage and store the finished rows; then the element waves publish
// layer ez in the stage (in the other slot when the stage is double-buffered, so
// the publication needs no wait for the owners).  Loads (element waves) and
// stores (owner waves) sit in different waves, so neither waits for the other's
// memory counter.  The owner waves issue no global loads at all: one counter
// tracks both loads and stores in a wave, so a wait for a load issued after a
// layer's row stores also waits for those stores to complete, which held each
// owner wave to one layer of stores in flight.  The element waves therefore load
// the owners' LIDs, row pointers and slot maps and publish them with the layer.
// UNIT: the steady tangent with beta = 1 (the element matrix is stored unscaled).
template <int TXV, int KIND, typename Real, bool TRANS, bool UNIT>
__device__ __forceinline__ void spec_block(const Q1DiffParams &p, int bx, int by, int bz, Real *stage,
                                           double *rowbuf, uint32_t *meta) {
  constexpr int TX = Tile<TXV>::TX, TY = Tile<TXV>::TY, NOWN = Tile<TXV>::NOWN, MSLOT = mslot<TXV>();
  constexpr int NSLOT = stage_slots<Real>();
  const bool is_elem = threadIdx.x < NE;
  const int t = is_elem ? threadIdx.x : threadIdx.x - NE;
  const int tx = t % TX, ty = t / TX;
  const int x0 = -1 + bx * (TX - 1);
  const int y0 = -1 + by * (TY - 1);
  const int64_t sx = p.nx + 1, sxy = static_cast<int64_t>(p.nx + 1) * (p.ny + 1);
  // this workgroup owns vertex layers [k0, k1); element layer k0-1 is recomputed
  // (it also belongs to the slab below) because vertex layer k0 needs it
  const int k0 = bz * p.zc;
  const int k1 = min(k0 + p.zc, p.nz + 1);
  const int ez0 = max(k0 - 1, 0), ez1 = min(k1 - 1, p.nz - 1);

  if (is_elem) {
    // ---------------- element waves
    constexpr bool kPrefetch = sizeof(Real) == 4;   // FP32 integration leaves registers for it
    const int ex = x0 + tx, ey = y0 + ty;
    const bool evalid = ex >= 0 && ex < p.nx && ey >= 0 && ey < p.ny;
    const int64_t col0 = evalid ? ex + sx * ey : 0;   // vertex (ex, ey, 0)
    const int64_t cell0 = ex + static_cast<int64_t>(p.nx) * ey;   // element (ex, ey, 0)
    const int64_t sz = static_cast<int64_t>(p.nx) * p.ny;
    Face<TRANS> fb, ft, fn;        // bottom, top and prefetched next-top faces
    int32_t Ln[4] = {0, 0, 0, 0};  // LIDs of the face after the prefetched one
    CellCoef<KIND> kc, kn;         // coefficients of this layer's element and (prefetched) the next one's
    if (evalid) {
      int32_t L[4];
      load_face_lids(p, col0 + sxy * ez0, sx, L);
      load_face(p, col0 + sxy * ez0, sx, L, fb);
      load_face_lids(p, col0 + sxy * (ez0 + 1), sx, L);
      load_face(p, col0 + sxy * (ez0 + 1), sx, L, ft);
      if (kPrefetch) {
        load_face_lids(p, col0 + sxy * min(ez0 + 2, p.nz), sx, Ln);
        load_coef<KIND>(p, cell0 + sz * ez0, kc);
      }
    }
    // the vertex column of owner lane t: (ex+1, ey+1), rows of layers [k0, k1)
    const int oi = ex + 1, oj = ey + 1;
    const bool owns = tx < TX - 1 && ty < TY - 1 && oi <= p.nx && oj <= p.ny;
    const int64_t ocol = owns ? oi + sx * oj : 0;
    auto own_lid = [&](int z) -> int32_t { return owns && z >= k0 && z < k1 ? p.lid[ocol + sxy * z] : -1; };
    int32_t Lo = own_lid(ez0);
    for (int ez = ez0; ez <= ez1 + 1; ++ez) {
      // metadata of the owner's row in vertex layer ez (published with element layer ez),
      // then the LID of the next layer; both ahead of this layer's face loads
      RowMeta om = load_meta_flat(p, Lo);
      const int32_t Lo_next = own_lid(ez + 1);
      Hex8Out<Real, TRANS> o;
      double done = 0.0;   // set from the layer's last result
      if (ez <= ez1) {
        if (evalid) {
          // prefetch (FP32) of the face above the top one (layer ez+2), the LIDs beyond it and the
          // next element's coefficients, with the layers clamped to the mesh (what lies past it is
          // loaded again and never used).  No branch: a conditional load would be copied, and so
          // waited for, where the branch joins.  Issued after the element's quadrature loop, whose
          // geometry holds the most registers, and ahead of the rest of the element work.
          auto prefetch = [&]() {
            if constexpr (kPrefetch) {
              __builtin_amdgcn_sched_barrier(0);   // not hoisted into the quadrature loop either
              load_face(p, col0 + sxy * min(ez + 2, p.nz), sx, Ln, fn);
              load_face_lids(p, col0 + sxy * min(ez + 3, p.nz), sx, Ln);
              load_coef<KIND>(p, cell0 + sz * min(ez + 1, p.nz - 1), kn);
              // keep these loads ahead of the rest of the element work: the scheduler would
              // otherwise sink them toward their use at the end of the layer
              __builtin_amdgcn_sched_barrier(0);
            }
          };
          if (!kPrefetch) {
            load_coef<KIND>(p, cell0 + sz * ez, kc);
            if (ez > ez0) {
              // no register room for a prefetch (FP64): load the top face now
              int32_t L[4];
              load_face_lids(p, col0 + sxy * (ez + 1), sx, L);
              load_face(p, col0 + sxy * (ez + 1), sx, L, ft);
            }
          }
          double X[3][8], u[8], ud[8];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int d = 0; d < 3; ++d) { X[d][v] = fb.x[d][v]; X[d][v + 4] = ft.x[d][v]; }
            u[v] = fb.u[v]; u[v + 4] = ft.u[v];
            ud[v] = fb.ud[v]; ud[v + 4] = ft.ud[v];
          }
          cell_element<KIND, Real, TRANS, UNIT>(p, kc, X, u, ud, o, prefetch);
          fb = ft;
          if (kPrefetch) { ft = fn; kc = kn; }
        } else {
#pragma unroll
          for (int i = 0; i < 36; ++i) o.K[i] = Real(0);
#pragma unroll
          for (int i = 0; i < 8; ++i) o.R[i] = 0.0;
        }
        if constexpr (flux_residual<Real>()) {
          // x-pair sums of the residual rows with the +x neighbour (the next lane of the
          // element row; the last column's pairs are not read)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int v1 = 1 + 2 * (j & 1) + 4 * (j >> 1);
            o.R[j] = o.R[v1] + __shfl_down(o.R[v1 - 1], 1, TX);
          }
        }
        done = o.R[3];
      }
      // the metadata registers are first read here, after the element work (an input the
      // work produces orders it), so the wave waits for those loads only now; earlier
      // uses, or a register of the loads reused early, would wait for them before the work
      asm volatile("" : "+v"(om.base), "+v"(om.end), "+v"(om.w[0]), "+v"(om.w[1]), "+v"(om.w[2]), "+v"(om.w[3]),
                   "+v"(om.w[4]), "+v"(om.w[5]), "+v"(om.w[6]) : "v"(done));
      if constexpr (NSLOT == 1) __syncthreads();   // the owners are done with the stage
      const int sl = NSLOT == 2 ? (ez & 1) : 0;
      if (ez <= ez1) {
        Real *st = stage + sl * (NSTAGE * NE);
#pragma unroll
        for (int i = 0; i < 36; ++i) st[i * NE + t] = o.K[i];
#pragma unroll
        for (int i = 0; i < (flux_residual<Real>() ? 4 : 8); ++i)
          reinterpret_cast<double *>(st + 36 * NE)[i * NE + t] = o.R[i];
      }
      if (t < NOWN) put_meta<TXV>(meta + sl * MSLOT, t, om);
      Lo = Lo_next;
      __syncthreads();   // layer ez is in the stage (two slots: the owners are also done with layer ez-1)
    }
  } else {
    // ---------------- owner waves: vertex column (oi, oj) is corner (1,1) of element
    // column (tx, ty), corner (0,1) of (tx+1, ty), (1,0) of (tx, ty+1), (0,0) of (tx+1, ty+1)
    const int oi = x0 + 1 + tx, oj = y0 + 1 + ty;
    const bool owner = tx < TX - 1 && ty < TY - 1 && oi <= p.nx && oj <= p.ny;
    double acc[27], accn[18];
    double rc = 0.0, rn = 0.0;
#pragma unroll
    for (int s = 0; s < 27; ++s) acc[s] = 0.0;
#pragma unroll
    for (int s = 0; s < 18; ++s) accn[s] = 0.0;
    for (int ez = ez0; ez <= ez1 + 1; ++ez) {
      if (ez > ez0) {
        // sum element layer e = ez-1 from the stage
        const int e = ez - 1;
        const int sl = NSLOT == 2 ? (e & 1) : 0;
        const Real *st = stage + sl * (NSTAGE * NE);
        RowMeta m;
        if (owner) {
          m = get_meta<TXV>(meta + sl * MSLOT, t);
          take_corner<1, 1>(st, t, acc, accn, rc, rn);
          take_corner<0, 1>(st, t + 1, acc, accn, rc, rn);
          take_corner<1, 0>(st, t + TX, acc, accn, rc, rn);
          take_corner<0, 0>(st, t + TX + 1, acc, accn, rc, rn);
          take_pairs<TXV>(st, t, rc, rn);
        }
        if (e >= k0) emit_rows<TXV>(p, m, acc, rc, rowbuf, t / TX);
#pragma unroll
        for (int s = 0; s < 18; ++s) acc[s] = accn[s];
        rc = rn;
      }
      if constexpr (NSLOT == 1) __syncthreads();   // done with the stage
      __syncthreads();   // next layer is in the stage
    }
    // the top vertex layer of the mesh has no element layer above it
    if (k1 == p.nz + 1) {
      RowMeta m;
      if (owner) m = get_meta<TXV>(meta + (NSLOT == 2 ? ((ez1 + 1) & 1) : 0) * MSLOT, t);
      emit_rows<TXV>(p, m, acc, rc, rowbuf, t / TX);
    }
  }
}


template <int KIND, typename Real, bool TRANS, bool UNIT>
__global__ __launch_bounds__(NT) void k_q1_hex8_diffusion(Q1DiffParams p) {
  using T = Tile<kTX>;
  __shared__ __align__(16) Real stage[stage_slots<Real>() * NSTAGE * NE];
  __shared__ double rowbuf[(T::TY - 1) * T::ROWBUF];
  __shared__ uint32_t meta[stage_slots<Real>() * mslot<kTX>()];
  spec_block<kTX, KIND, Real, TRANS, UNIT>(p, blockIdx.x, blockIdx.y, blockIdx.z, stage, rowbuf, meta);
}

// ============================================================================
// Batched assembly: many meshes in one persistent launch
// ============================================================================
// Work item = one tile (workgroup's worth: a (TX-1) x (TY-1) patch of vertex columns and a
// z slab) of one mesh.  Workgroups pull items from a counter until the list is
// exhausted; the last workgroup to leave resets the counter, so the launch can
// be replayed (as a graph node) without a separate clear.
struct BatchQueue {
  unsigned next;
  unsigned done;
};

__device__ __forceinline__ int fetch_item(BatchQueue *q, int *slot) {
  if (threadIdx.x == 0) *slot = static_cast<int>(atomicAdd(&q->next, 1u));
  __syncthreads();
  const int it = *slot;
  __syncthreads();
  return it;
}

__device__ __forceinline__ void leave_queue(BatchQueue *q) {
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&q->done, 1u) == gridDim.x - 1) {
      atomicExch(&q->next, 0u);
      atomicExch(&q->done, 0u);
    }
  }
}

// One element kind per launch (FP64, transient or general-cell batches, and batches of one
// kind): a kernel holding both bodies spills there.  FP32 steady batches of both kinds with
// parallelepiped hex8 cells take k_q1_batch_mixed.
template <int TXV, int KIND, typename Real, bool TRANS>
__global__ __launch_bounds__(NT) void k_q1_batch_spec(const Q1DiffParams *__restrict__ prm,
                                                      const int4 *__restrict__ items, int nitems, BatchQueue *q) {
  using T = Tile<TXV>;
  __shared__ __align__(16) Real stage[stage_slots<Real>() * NSTAGE * NE];
  __shared__ double rowbuf[(T::TY - 1) * T::ROWBUF];
  __shared__ uint32_t meta[stage_slots<Real>() * mslot<TXV>()];
  __shared__ int slot;
  for (;;) {
    const int it = fetch_item(q, &slot);
    if (it >= nitems) break;
    const int4 w = items[it];
    const Q1DiffParams &p = prm[w.x];
    if (w.w * p.zc > p.nz) continue;   // slab past the mesh: nothing to do (uniform over the block)
    spec_block<TXV, KIND, Real, TRANS, false>(p, w.y, w.z, w.w, stage, rowbuf, meta);
    __syncthreads();
  }
  leave_queue(q);
}

// Both element kinds in one persistent launch: one queue of (mesh | kMixedTet for tet4
// meshes, tile x, tile y, slab) items over hex8 and tet4 meshes, longest first.
constexpr int kMixedTet = 1 << 30;

template <int TXV, int HEXKIND, typename Real, bool TRANS>
__global__ __launch_bounds__(NT) void k_q1_batch_mixed(const Q1DiffParams *__restrict__ prm,
                                                       const int4 *__restrict__ items, int nitems, BatchQueue *q) {
  using T = Tile<TXV>;
  __shared__ __align__(16) Real stage[stage_slots<Real>() * NSTAGE * NE];
  __shared__ double rowbuf[(T::TY - 1) * T::ROWBUF];
  __shared__ uint32_t meta[stage_slots<Real>() * mslot<TXV>()];
  __shared__ int slot;
  for (;;) {
    const int it = fetch_item(q, &slot);
    if (it >= nitems) break;
    const int4 w = items[it];
    const bool tet = (w.x & kMixedTet) != 0;
    const Q1DiffParams &p = prm[w.x & (kMixedTet - 1)];
    if (w.w * p.zc > p.nz) continue;   // slab past the mesh: nothing to do (uniform over the block)
    if (tet) spec_block<TXV, TXA_TET4, Real, TRANS, false>(p, w.y, w.z, w.w, stage, rowbuf, meta);
    else spec_block<TXV, HEXKIND, Real, TRANS, false>(p, w.y, w.z, w.w, stage, rowbuf, meta);
    __syncthreads();
  }
  leave_queue(q);
}

struct BcItem {
  const int32_t *lids;
  const double *g;
  const double *x;
  const int64_t *rowptr;
  const int32_t *col;
  double *f;
  double *vals;
  int64_t n, nrows;
  int64_t blk0;   // first block of this mesh's list entries
};

// list entries of the batch Dirichlet pass per 256-thread block
template <int G>
constexpr int64_t bc_rows_per_block() { return 4 * kDirU * (64 / G); }

// Dirichlet rows of every mesh of a batch: each block takes bc_rows_per_block consecutive
// entries of one mesh's list (blk_mesh: the mesh of each block), so the largest list is
// spread over as many blocks as it needs.  A LID outside a mesh's map touches no row and
// sets *err.
template <int G>
__global__ __launch_bounds__(256) void k_dirichlet_batch(const BcItem *__restrict__ items,
                                                         const int32_t *__restrict__ blk_mesh, int *err) {
  const BcItem &b = items[blk_mesh[blockIdx.x]];
  const int64_t i0 = (static_cast<int64_t>(blockIdx.x) - b.blk0) * bc_rows_per_block<G>() +
                     static_cast<int64_t>(threadIdx.x / 64) * kDirU * (64 / G);
  dirichlet_wave<G>(i0, b.n, b.nrows, b.lids, b.g, b.x, b.rowptr, b.col, b.f, b.vals, err);
}

// Whether a brick cell is a parallelepiped: the bilinear and trilinear coefficients of its
// trilinear map (Walsh coefficients 3, 5, 6, 7 of the vertex coordinates in binary order)
// are below 2^-47 of the cell's coordinate magnitude (its centroid and edge coefficients),
// i.e. at the level of the coordinates' own rounding.  One thread per cell; any cell that
// is not clears *flag.
__global__ void k_cell_affine(int nx, int ny, int nz, int64_t nv, const double *__restrict__ xyz, int *flag) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<int64_t>(nx) * ny * nz) return;
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (static_cast<int64_t>(nx) * ny);
  double nonaff = 0.0, mag = 0.0;
  for (int d = 0; d < 3; ++d) {
    double h[8];
#pragma unroll
    for (int v = 0; v < 8; ++v)
      h[v] = xyz[d * nv + (i + bit(v, 0)) + (nx + 1) * ((j + bit(v, 1)) + static_cast<int64_t>(ny + 1) * (k + bit(v, 2)))];
    walsh8(h);   // h[s] = sum_v prod_{d in s} sgn(v,d) X_v (walsh8 computes hi - lo: the same up to sign)
    nonaff += 0.125 * (fabs(h[3]) + fabs(h[5]) + fabs(h[6]) + fabs(h[7]));
    mag += 0.125 * (fabs(h[0]) + fabs(h[1]) + fabs(h[2]) + fabs(h[4]));
  }
  if (!(nonaff <= 0x1p-47 * mag)) __hip_atomic_store(flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Element and owner waves split in both precisions (the stage is double-buffered in FP32).
template <int KIND>
void launch_kind(const Q1DiffParams &p, dim3 grid, int precision, bool trans, hipStream_t s) {
  const bool unit = p.beta == 1.0;
  if (precision == TXA_FP64) {
    if (trans) hipLaunchKernelGGL((k_q1_hex8_diffusion<KIND, double, true, false>), grid, dim3(NT), 0, s, p);
    else if (unit) hipLaunchKernelGGL((k_q1_hex8_diffusion<KIND, double, false, true>), grid, dim3(NT), 0, s, p);
    else hipLaunchKernelGGL((k_q1_hex8_diffusion<KIND, double, false, false>), grid, dim3(NT), 0, s, p);
  } else {
    // FP32 scales by beta in every case: without the 36 multiplies the compiler's schedule
    // of the unit-beta instance spills (47 registers) where the scaled one fits in 240
    if (trans) hipLaunchKernelGGL((k_q1_hex8_diffusion<KIND, float, true, false>), grid, dim3(NT), 0, s, p);
    else hipLaunchKernelGGL((k_q1_hex8_diffusion<KIND, float, false, false>), grid, dim3(NT), 0, s, p);
  }
}

}  // namespace

namespace {

Q1DiffParams make_params(const Mesh &m, const txa_assembly_args &a, int zc) {
  Q1DiffParams p;
  p.nx = m.nx; p.ny = m.ny; p.nz = m.nz;
  p.zc = zc;
  p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
  p.lid = m.point_lid.p;
  p.coef = a.coef; p.x = a.x; p.xdot = a.xdot;
  p.alpha = a.alpha; p.beta = a.beta; p.rhoc = a.rho; p.src = a.src[0];
  p.f = a.f; p.vals = a.vals;
  p.rowptr = m.rowptr.p; p.perm = m.perm.p;
  return p;
}

// tiles of width tx (height 256 / tx) over a brick's vertex columns
inline int tiles_x(const Mesh &m, int tx = kTX) { return (m.nx + 1 + tx - 2) / (tx - 1); }
inline int tiles_y(const Mesh &m, int tx = kTX) {
  const int ty = NE / tx;
  return (m.ny + 1 + ty - 2) / (ty - 1);
}

}  // namespace

bool brick_cells_affine(const Mesh &m, hipStream_t s) {
  if (!m.structured || m.topo != TXA_HEX8) return false;
  const int64_t nc = static_cast<int64_t>(m.nx) * m.ny * m.nz;
  DevBuf<int> flag;
  flag.alloc(1);
  const int one = 1;
  TXA_HIP(hipMemcpyAsync(flag.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_cell_affine, dim3(static_cast<unsigned>((nc + 255) / 256)), dim3(256), 0, s, m.nx, m.ny, m.nz,
                     m.nverts, m.xyz.p, flag.p);
  TXA_HIP(hipGetLastError());
  int h = 0;
  TXA_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  TXA_HIP(hipStreamSynchronize(s));
  return h != 0;
}

void launch_q1_hex8_diffusion(Mesh &m, const txa_assembly_args &a, hipStream_t s) {
  // z slabs: enough workgroups to cover the 256 CUs several times over
  const int gx = tiles_x(m), gy = tiles_y(m);
  static_assert(NT == 512, "block shape");
  int nchunk = (kTargetBlocks + gx * gy - 1) / (gx * gy);
  nchunk = std::max(1, std::min(nchunk, (m.nz + 1 + 7) / 8));
  int zc = (m.nz + 1 + nchunk - 1) / nchunk;
  nchunk = (m.nz + 1 + zc - 1) / zc;
  const Q1DiffParams p = make_params(m, a, zc);
  dim3 grid(gx, gy, nchunk);
  const bool trans = a.transient != 0;
  if (m.topo == TXA_TET4) launch_kind<TXA_TET4>(p, grid, a.precision, trans, s);
  else if (m.affine) launch_kind<kHex8Affine>(p, grid, a.precision, trans, s);
  else launch_kind<TXA_HEX8>(p, grid, a.precision, trans, s);
  TXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------- batches
struct Batch {
  std::vector<int> ids;
  std::vector<Mesh *> meshes;
  int precision = TXA_FP64;
  bool transient = false;
  int slab = 0;   // vertex layers per work item (0: whole column)
  int tx = kTX;   // tile width (vertex columns along x; the tile is tx x 256/tx)
  int grid = 0;   // persistent workgroups
  std::vector<int> zc;
  std::vector<int4> items_h;   // hex8 items, then tet4 items
  size_t n_hex = 0;            // items of hex8 meshes (the first n_hex of items_h)
  // hex8 and tet4 meshes in one persistent launch over one queue (items of tet4 meshes marked
  // kMixedTet), where the combined body fits the register file unspilled: FP32 steady, every
  // hex8 mesh of parallelepiped cells
  bool mixed = false;
  bool hex_affine = false;     // every hex8 mesh of the batch has parallelepiped cells
  DevBuf<int4> items;
  DevBuf<Q1DiffParams> prm;
  DevBuf<BcItem> bc;
  DevBuf<BatchQueue> q;        // [2]: hex8 queue, tet4 queue
  Q1DiffParams *prm_h = nullptr;   // pinned: a graph's copy node reads it at replay
  BcItem *bc_h = nullptr;
  int bc_lanes = 32;              // lanes per Dirichlet row (the meshes' longest row)
  int64_t bc_blocks = 0;          // blocks of the Dirichlet pass for the current lists
  size_t bc_map_cap = 0;
  int32_t *bc_map_h = nullptr;    // pinned: the mesh of each Dirichlet block
  DevBuf<int32_t> bc_map;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  bool launched = false;
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (done) (void)hipEventDestroy(done);
    if (cap) (void)hipStreamDestroy(cap);
    if (prm_h) (void)hipHostFree(prm_h);
    if (bc_h) (void)hipHostFree(bc_h);
    if (bc_map_h) (void)hipHostFree(bc_map_h);
  }
};

namespace {

void batch_drop_graph(Batch &b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
}

// Work lists for a slab height, one per element kind: items of the tallest
// slabs first, so the queue hands out the longest items early and the short
// ones fill the tail.  The slab height is also written into the pinned
// argument blocks (their zc must match the items): no launch of the batch may
// be in flight.
void batch_build_items(Batch &b, int slab) {
  b.slab = slab;
  b.zc.resize(b.meshes.size());
  std::vector<int4> lists[2];
  for (size_t i = 0; i < b.meshes.size(); ++i) {
    const Mesh &m = *b.meshes[i];
    const int layers = m.nz + 1;
    const int zc = slab > 0 ? std::min(slab, layers) : layers;
    b.zc[i] = zc;
    b.prm_h[i].zc = zc;
    const int nch = (layers + zc - 1) / zc;
    std::vector<int4> &out = lists[m.topo == TXA_TET4 ? 1 : 0];
    for (int z = 0; z < nch; ++z)
      for (int y = 0; y < tiles_y(m, b.tx); ++y)
        for (int x = 0; x < tiles_x(m, b.tx); ++x) out.push_back(make_int4(static_cast<int>(i), x, y, z));
  }
  auto cost = [&](const int4 &w) {
    const Mesh &m = *b.meshes[w.x];
    const int k0 = w.w * b.zc[w.x];
    return std::min(k0 + b.zc[w.x], m.nz + 1) - std::max(k0 - 1, 0);
  };
  for (auto &l : lists)
    std::stable_sort(l.begin(), l.end(), [&](const int4 &u, const int4 &v) { return cost(u) > cost(v); });
  b.n_hex = lists[0].size();
  b.hex_affine = true;
  for (const Mesh *m : b.meshes)
    if (m->topo == TXA_HEX8 && !m->affine) b.hex_affine = false;
  b.items_h = lists[0];
  if (b.mixed) {
    // one list over both kinds, longest first (ties: hex8 before tet4, then build order)
    std::vector<int4> all(lists[0]);
    all.insert(all.end(), lists[1].begin(), lists[1].end());
    std::stable_sort(all.begin(), all.end(), [&](const int4 &u, const int4 &v) { return cost(u) > cost(v); });
    for (int4 &w : all)
      if (b.meshes[w.x]->topo == TXA_TET4) w.x |= kMixedTet;
    b.items_h = all;
  } else {
    b.items_h.insert(b.items_h.end(), lists[1].begin(), lists[1].end());
  }
  b.items.alloc(b.items_h.size());
  TXA_HIP(hipMemcpy(b.items.p, b.items_h.data(), sizeof(int4) * b.items_h.size(), hipMemcpyHostToDevice));
  batch_drop_graph(b);
}

void batch_fill(Batch &b, const txa_assembly_args *args) {
  for (size_t i = 0; i < b.meshes.size(); ++i) {
    const Mesh &m = *b.meshes[i];
    const txa_assembly_args &a = args[i];
    TXA_REQUIRE(a.physics == TXA_DIFFUSION, TXA_ERR_UNSUPPORTED, "batched assembly covers Q1 diffusion");
    TXA_REQUIRE(a.n_neumann == 0, TXA_ERR_UNSUPPORTED, "Neumann sides are assembled by txa_assemble, not in batches");
    TXA_REQUIRE(a.precision == b.precision, TXA_ERR_INVALID, "every batch item must use the batch precision");
    TXA_REQUIRE((a.transient != 0) == b.transient, TXA_ERR_INVALID, "every batch item must match the batch transient flag");
    TXA_REQUIRE(a.x && a.coef, TXA_ERR_INVALID, "x and coef are required");
    TXA_REQUIRE(!b.transient || a.xdot, TXA_ERR_INVALID, "transient assembly needs xdot");
    TXA_REQUIRE(a.n_bc == 0 || a.bc_lids, TXA_ERR_INVALID, "bc_lids missing");
    b.prm_h[i] = make_params(m, a, b.zc[i]);
    BcItem &c = b.bc_h[i];
    c.lids = a.bc_lids; c.g = a.bc_vals; c.x = a.x; c.rowptr = m.rowptr.p; c.col = m.col.p;
    c.f = a.f; c.vals = a.vals; c.n = a.n_bc; c.nrows = m.ndofs;
  }
  // Dirichlet blocks: each mesh's list in runs of bc_rows_per_block entries.  A change in the
  // block count changes the pass's grid and copy size: the captured graph is rebuilt.
  const int64_t rpb = b.bc_lanes == 8 ? bc_rows_per_block<8>() : b.bc_lanes == 16 ? bc_rows_per_block<16>()
                    : b.bc_lanes == 32 ? bc_rows_per_block<32>() : bc_rows_per_block<64>();
  int64_t nblk = 0;
  for (size_t i = 0; i < b.meshes.size(); ++i) {
    b.bc_h[i].blk0 = nblk;
    nblk += (b.bc_h[i].n + rpb - 1) / rpb;
  }
  TXA_REQUIRE(nblk < (int64_t(1) << 31), TXA_ERR_UNSUPPORTED, "too many Dirichlet rows for one batch pass");
  if (nblk != b.bc_blocks) batch_drop_graph(b);
  b.bc_blocks = nblk;
  if (static_cast<size_t>(nblk) > b.bc_map_cap) {
    if (b.bc_map_h) TXA_HIP(hipHostFree(b.bc_map_h));
    b.bc_map_h = nullptr;
    b.bc_map_cap = 0;
    TXA_HIP(hipHostMalloc(reinterpret_cast<void **>(&b.bc_map_h), sizeof(int32_t) * nblk, hipHostMallocDefault));
    b.bc_map_cap = static_cast<size_t>(nblk);
    b.bc_map.alloc(nblk);
  }
  for (size_t i = 0; i < b.meshes.size(); ++i) {
    const int64_t e = i + 1 < b.meshes.size() ? b.bc_h[i + 1].blk0 : nblk;
    for (int64_t k = b.bc_h[i].blk0; k < e; ++k) b.bc_map_h[k] = static_cast<int32_t>(i);
  }
}

template <int TXV, int KIND>
void launch_batch_tx(const Batch &b, dim3 g, const int4 *items, int n, BatchQueue *q, hipStream_t s) {
  if (b.precision == TXA_FP64) {
    if (b.transient) hipLaunchKernelGGL((k_q1_batch_spec<TXV, KIND, double, true>), g, dim3(NT), 0, s, b.prm.p, items, n, q);
    else hipLaunchKernelGGL((k_q1_batch_spec<TXV, KIND, double, false>), g, dim3(NT), 0, s, b.prm.p, items, n, q);
  } else {
    if (b.transient) hipLaunchKernelGGL((k_q1_batch_spec<TXV, KIND, float, true>), g, dim3(NT), 0, s, b.prm.p, items, n, q);
    else hipLaunchKernelGGL((k_q1_batch_spec<TXV, KIND, float, false>), g, dim3(NT), 0, s, b.prm.p, items, n, q);
  }
  TXA_HIP(hipGetLastError());
}

template <int KIND>
void launch_batch(const Batch &b, dim3 g, const int4 *items, int n, BatchQueue *q, hipStream_t s) {
  if (b.tx == 16) launch_batch_tx<16, KIND>(b, g, items, n, q, s);
  else launch_batch_tx<32, KIND>(b, g, items, n, q, s);
}

void batch_enqueue(Batch &b, hipStream_t s, int *err) {
  const size_t n = b.meshes.size();
  for (size_t i = 0; i < n; ++i) {
    const Q1DiffParams &p = b.prm_h[i];
    const Mesh &m = *b.meshes[i];
    TXA_REQUIRE(p.zc == b.zc[i] && p.nx == m.nx && p.ny == m.ny && p.nz == m.nz, TXA_ERR_LOGIC,
                "batch argument block out of step with its work list");
  }
  TXA_HIP(hipMemcpyAsync(b.prm.p, b.prm_h, sizeof(Q1DiffParams) * n, hipMemcpyHostToDevice, s));
  TXA_HIP(hipMemcpyAsync(b.bc.p, b.bc_h, sizeof(BcItem) * n, hipMemcpyHostToDevice, s));
  const int nk[2] = {static_cast<int>(b.n_hex), static_cast<int>(b.items_h.size() - b.n_hex)};
  const int4 *it[2] = {b.items.p, b.items.p + b.n_hex};
  if (b.mixed) {
    const int nall = static_cast<int>(b.items_h.size());
    const dim3 g(std::min(b.grid, nall));
    if (b.tx == 16)
      hipLaunchKernelGGL((k_q1_batch_mixed<16, kHex8Affine, float, false>), g, dim3(NT), 0, s, b.prm.p, b.items.p, nall, b.q.p);
    else
      hipLaunchKernelGGL((k_q1_batch_mixed<32, kHex8Affine, float, false>), g, dim3(NT), 0, s, b.prm.p, b.items.p, nall, b.q.p);
    TXA_HIP(hipGetLastError());
  }
  for (int k = 0; k < 2 && !b.mixed; ++k) {
    if (nk[k] == 0) continue;
    const dim3 g(std::min(b.grid, nk[k]));
    BatchQueue *q = b.q.p + k;
    if (k == 0 && b.hex_affine) launch_batch<kHex8Affine>(b, g, it[k], nk[k], q, s);
    else if (k == 0) launch_batch<TXA_HEX8>(b, g, it[k], nk[k], q, s);
    else launch_batch<TXA_TET4>(b, g, it[k], nk[k], q, s);
  }
  if (b.bc_blocks > 0) {
    TXA_HIP(hipMemcpyAsync(b.bc_map.p, b.bc_map_h, sizeof(int32_t) * b.bc_blocks, hipMemcpyHostToDevice, s));
    const dim3 g(static_cast<unsigned>(b.bc_blocks));
    if (b.bc_lanes == 8) hipLaunchKernelGGL(k_dirichlet_batch<8>, g, dim3(256), 0, s, b.bc.p, b.bc_map.p, err);
    else if (b.bc_lanes == 16) hipLaunchKernelGGL(k_dirichlet_batch<16>, g, dim3(256), 0, s, b.bc.p, b.bc_map.p, err);
    else if (b.bc_lanes == 32) hipLaunchKernelGGL(k_dirichlet_batch<32>, g, dim3(256), 0, s, b.bc.p, b.bc_map.p, err);
    else hipLaunchKernelGGL(k_dirichlet_batch<64>, g, dim3(256), 0, s, b.bc.p, b.bc_map.p, err);
    TXA_HIP(hipGetLastError());
  }
}

}  // namespace

Batch *batch_create(const std::vector<Mesh *> &meshes, const std::vector<int> &ids, int precision, bool transient) {
  TXA_REQUIRE(!meshes.empty(), TXA_ERR_INVALID, "empty batch");
  TXA_REQUIRE(precision == TXA_FP64 || precision == TXA_FP32, TXA_ERR_UNSUPPORTED,
              "batched assembly runs in FP64 or FP32 integration");
  std::unique_ptr<Batch> b(new Batch());
  b->ids = ids;
  b->meshes = meshes;
  b->precision = precision;
  b->transient = transient;
  for (Mesh *m : meshes) {
    TXA_REQUIRE(m->topo == TXA_HEX8 || m->topo == TXA_TET4, TXA_ERR_UNSUPPORTED, "batched assembly covers hex8 and tet4 meshes");
    TXA_REQUIRE(m->structured, TXA_ERR_UNSUPPORTED, "batched assembly covers bricks (txa_mesh_brick)");
    TXA_REQUIRE(m->has_graph && m->nfields == 1, TXA_ERR_LOGIC, "batch meshes need a scalar DOF map and a graph");
  }
  int max_row = 0, nhex = 0;
  bool affine = true;
  for (const Mesh *m : meshes) {
    max_row = std::max(max_row, m->max_row);
    nhex += m->topo == TXA_HEX8;
    affine = affine && (m->topo != TXA_HEX8 || m->affine);
  }
  b->mixed = precision == TXA_FP32 && !transient && affine && nhex > 0 && nhex < static_cast<int>(meshes.size());
  b->bc_lanes = dirichlet_lanes(max_row);
  b->prm.alloc(meshes.size());
  b->bc.alloc(meshes.size());
  b->q.alloc(2);
  TXA_HIP(hipMemset(b->q.p, 0, 2 * sizeof(BatchQueue)));
  TXA_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->prm_h), sizeof(Q1DiffParams) * meshes.size(), hipHostMallocDefault));
  TXA_HIP(hipHostMalloc(reinterpret_cast<void **>(&b->bc_h), sizeof(BcItem) * meshes.size(), hipHostMallocDefault));
  std::memset(b->prm_h, 0, sizeof(Q1DiffParams) * meshes.size());
  std::memset(b->bc_h, 0, sizeof(BcItem) * meshes.size());
  TXA_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  TXA_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  int dev = 0, ncu = 0;
  TXA_HIP(hipGetDevice(&dev));
  TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  b->grid = 2 * ncu;
  batch_build_items(*b, 32);
  return b.release();
}

void batch_free(Batch *b) { delete b; }

void batch_assemble(Batch &b, const txa_assembly_args *args, bool use_graph, hipStream_t s, int *err) {
  TXA_REQUIRE(args, TXA_ERR_INVALID, "null args");
  // the previous launch must be done reading the pinned parameter blocks
  if (b.launched) TXA_HIP(hipEventSynchronize(b.done));
  batch_fill(b, args);
  if (use_graph) {
    if (!b.exec) {
      TXA_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
      batch_enqueue(b, b.cap, err);
      TXA_HIP(hipStreamEndCapture(b.cap, &b.graph));
      TXA_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    }
    TXA_HIP(hipGraphLaunch(b.exec, s));
  } else {
    batch_enqueue(b, s, err);
  }
  TXA_HIP(hipEventRecord(b.done, s));
  b.launched = true;
}

// Autotuner: times the batch over tile shapes (32 x 8, 16 x 16), slab heights (the
// analogue of Panzer's workset size sweep) and persistent grid sizes, keeps the
// fastest.  out: best ms.
double batch_tune(Batch &b, const txa_assembly_args *args, hipStream_t s, int reps, int *err) {
  int dev = 0, ncu = 0;
  TXA_HIP(hipGetDevice(&dev));
  TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int slabs[] = {8, 16, 32, 64, 0};
  const int grids[] = {ncu, 2 * ncu, 4 * ncu};
  hipEvent_t e0, e1;
  TXA_HIP(hipEventCreate(&e0));
  TXA_HIP(hipEventCreate(&e1));
  double best = 1e30;
  int best_slab = b.slab, best_grid = b.grid, best_tx = b.tx;
  if (b.launched) TXA_HIP(hipEventSynchronize(b.done));
  batch_fill(b, args);   // the argument blocks stay as they are for every timed launch
  for (int tx : {32, 16}) {
    b.tx = tx;
    for (int sl : slabs) {
      TXA_HIP(hipStreamSynchronize(s));
      batch_build_items(b, sl);
      for (int g : grids) {
        b.grid = g;
        batch_enqueue(b, s, err);   // warm-up
        TXA_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) batch_enqueue(b, s, err);
        TXA_HIP(hipEventRecord(e1, s));
        TXA_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        TXA_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms / reps < best) { best = ms / reps; best_slab = sl; best_grid = g; best_tx = tx; }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  TXA_HIP(hipStreamSynchronize(s));
  b.tx = best_tx;
  batch_build_items(b, best_slab);
  b.grid = best_grid;
  TXA_HIP(hipEventRecord(b.done, s));
  b.launched = true;
  return best;
}

// The work shape set by hand (what the autotuner would otherwise pick): tile width,
// slab height and persistent grid.  No launch of the batch may be in flight.
void batch_shape(Batch &b, int tile_x, int slab, int grid) {
  TXA_REQUIRE(tile_x == 16 || tile_x == 32, TXA_ERR_INVALID, "tile width must be 16 or 32");
  TXA_REQUIRE(slab >= 0 && grid >= 1, TXA_ERR_INVALID, "slab must be >= 0 (0: whole column) and grid >= 1");
  if (b.launched) TXA_HIP(hipEventSynchronize(b.done));
  b.tx = tile_x;
  b.grid = grid;
  batch_build_items(b, slab);
}

void batch_info(const Batch &b, int64_t *out) {
  out[0] = static_cast<int64_t>(b.meshes.size());
  out[1] = static_cast<int64_t>(b.items_h.size());
  out[2] = b.slab;
  out[3] = b.grid;
  out[4] = b.tx;
  out[5] = b.mixed ? 1 : 2;
}

const std::vector<int> &batch_mesh_ids(const Batch &b) { return b.ids; }

}  // namespace txa
