// Linear elasticity, element-centric: hex8 (Q1) and hex27 (Q2), three fields.
//
// Replaces, for the displacement equation set, the chain
//   GatherSolution_Tpetra (+ seeds) -> DOFGradient -> stress -> Integrator_GradBasisDotVector,
//   Integrator_BasisTimesScalar (body force, inertia) -> ScatterResidual_Tpetra<Jacobian>
// (disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-650,
//  Panzer_DOFGradient_impl.hpp:60-160, Panzer_Integrator_GradBasisDotVector_impl.hpp:225-294,
//  Panzer_Integrator_BasisTimesScalar_impl.hpp:215-296,
//  Panzer_ScatterResidual_Tpetra_impl.hpp:446-520), with the Sacado-FAD element
// Jacobian replaced by the explicit tangent
//   J_e = beta K + alpha (a_R M + b_R K) + gamma2 M,   R_e = K u + C v + M a - F.
//
// One workgroup per element (grid-stride).  The reference element is the
// tensor product of 1D Gauss rules and Lagrange bases (tables in LDS), so no
// per-point tables are read from memory.  Geometry (J, det J, J^-1) is FP64 at
// every quadrature point.  The element stiffness is assembled from the
// gradient Gram matrix
//   T[(k,a),(l,b)] = sum_q w_q |det J_q| dN_a/dx_k dN_b/dx_l        (3 NB x 3 NB),
//   K[(a,i),(b,j)] = lambda T[(i,a),(j,b)] + mu T[(j,a),(i,b)] + delta_ij mu sum_k T[(k,a),(k,b)],
// i.e. B^T D B for isotropic D written as one symmetric product G^T W G.  T is
// computed on the matrix cores: T = A^T G with A = w det J G in FP64
// (v_mfma_f64_16x16x4f64) or FP32 (v_mfma_f32_16x16x4f32), or as T = O^T O with
// O = sqrt(w |det J|) G by bf16 MFMA (v_mfma_f32_16x16x32_bf16) on a three-way split of
// each FP32 operand (x = x0 + x1 + x2, every x_s in bf16) keeping the six cross products
// of order <= 2^-16, accumulated in FP32.
//
// Bricks of parallelepiped hex8 cells skip both phases: k_elastic8_affine forms each
// node's rows directly from closed-form cell blocks (one wave per node, no workspace).
//
// Assembly into the CRS is two-phase.  The element kernel writes each element's
// Jacobian block (FP32, FP64 in FP64 mode) and residual (FP64) to a workspace,
// coalesced; a node-centric gather then sums, for every node, the blocks of the
// elements around it into the node's three CRS rows in FP64, in a fixed element
// order, and writes each row once.  No global-memory atomics (the gather's LDS adds
// run in one wave's program order): results are bitwise reproducible.
// The row position of every (node, neighbour) pair comes from a per-node table
// built with the graph.
#include <map>
#include <type_traits>

#include "txa_internal.hpp"

namespace txa {
namespace {

constexpr int MODE_F64 = 0, MODE_F32 = 1, MODE_SPLIT = 2;

struct ElasticParams {
  int nq;
  int64_t nelem;
  int64_t e0, e1;               // elements [e0, e1) of this launch
  int nconn;
  const int64_t *conn;          // the first 8 entries of an element are its vertices
  const double *X, *Y, *Z;
  const int32_t *elem_lid;      // [E][3 NB]: node-major, field-minor
  const double *coef;           // Young's modulus per element
  double nu, rho, src[3], ra, rb;
  const double *x, *xdot, *xddot;
  double alpha, beta, gamma2;
  int transient;
  void *jbuf;        // element Jacobian blocks [E][NB][block_stride(NB)], node-major: node a holds the
                     // rows of DOFs (a, 0), (a, 1), (a, 2), each over the Gram columns j NB + b = DOF
                     // (b, j), then zero padding (FP32, FP64 in FP64 mode); null: none
  double *rbuf;      // element residuals [E][ND]; null: none
};

// Entries per node in the workspace block: the node's three rows (9 NB entries), padded for
// hex27 to 256 so every node's rows start on a 1 KiB boundary (FP32) and the gather reads
// whole cache lines; hex8's 72 entries stay unpadded.
__host__ __device__ constexpr int block_stride(int NB) { return NB == 27 ? 256 : 9 * NB; }

// lattice offsets of the element's nodes: shards vertex order (hex8), Q2 order (hex27)
__device__ const int8_t kHexV[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__device__ const int8_t kQ2[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2},
    {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
    {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// bf16 with round-to-nearest-even (finite inputs), returned as its bit pattern and its value
__device__ __forceinline__ short bf16_rne(float x, float &v) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  const uint32_t h = u & 0xffff0000u;
  v = __uint_as_float(h);
  return static_cast<short>(h >> 16);
}

// x = x0 + x1 + x2 with bf16 parts (exact up to ~2^-24 |x|)
__device__ __forceinline__ void split3(float x, short &s0, short &s1, short &s2) {
  float v0, v1, v2;
  s0 = bf16_rne(x, v0);
  const float r1 = x - v0;
  s1 = bf16_rne(r1, v1);
  const float r2 = r1 - v1;
  s2 = bf16_rne(r2, v2);
}

// One DPP lane move of an FP64 value (both halves); lanes of rows outside row_mask get 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Fixed-order sum over groups of GW (8 or 32) consecutive lanes, on the VALU (DPP):
// quad swaps (lanes ^1, ^2), the half-row mirror (8 lanes) and the row mirror (16
// lanes); for 32 lanes the first row's total is then broadcast into the second row.
// The group sum ends in every lane of the group (GW = 8) or in its second row (GW = 32).
template <int GW>
__device__ __forceinline__ double group_sum(double x) {
  x += dpp_f64<0xB1, 0xf>(x);    // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E, 0xf>(x);    // quad_perm [2,3,0,1]
  x += dpp_f64<0x141, 0xf>(x);   // row_half_mirror
  if constexpr (GW == 32) {
    x += dpp_f64<0x140, 0xf>(x);   // row_mirror
    x += dpp_f64<0x142, 0xa>(x);   // row_bcast15 into rows 1 and 3
  }
  return x;
}

template <int NB>
struct Dims {
  static constexpr int ND = 3 * NB;                    // element DOFs
  static constexpr int R = NB == 8 ? 1 : 2;            // lattice reach of a node's neighbours
  static constexpr int S = 2 * R + 1;
  static constexpr int MP = (ND + 15) / 16 * 16;       // padded operand rows (MFMA)
  static constexpr int KP = 32;                        // padded quadrature dimension (MFMA)
  // LDS row stride of the operand parts, in bf16: 40 (80 B) puts the 16 rows of an
  // MFMA tile on distinct bank pairs (a 64 B stride put them on two banks: 8-way
  // conflicts on every operand read) and keeps rows 16-byte aligned
  static constexpr int KS = 40;
  // FP64 form: quadrature dimension padded to the 4 of v_mfma_f64_16x16x4f64, and an odd
  // LDS row stride in doubles (the 16 rows of an operand read fall on distinct bank pairs)
  static constexpr int KP64 = NB == 8 ? 8 : 28;
  static constexpr int KS64 = KP64 + 1;
  // FP32 form: the same padding for v_mfma_f32_16x16x4f32; a row stride in floats whose 16
  // multiples are distinct multiples of 4 modulo the 64 banks, so the 16 rows x 4 points of an
  // operand read fall on 64 distinct banks
  static constexpr int KS32 = NB == 8 ? 12 : 36;
};

template <int NB, int NQ, int MODE>
struct Smem {
  using D = Dims<NB>;
  using TReal = typename std::conditional<MODE == MODE_F64, double, float>::type;
  // 1D tables on [-1, 1]: Gauss rule, field basis (2 or 3 nodes) and the linear
  // geometry basis, values and derivatives at each Gauss point
  double bL[3][3], bdL[3][3], gL[2][3], gdL[2][3], w1[3];
  int8_t off[NB][3];         // lattice offset of each element node
  double X[8][3];
  double Ji[NQ][9];
  double det[NQ];
  double wm[NQ];             // w det J (the sign of the first point is applied where used)
  double sw[NQ];             // sqrt(w |det J|)
  double u[D::ND], v[D::ND], a[D::ND], r[D::ND];
  TReal M[NB][NB];
  // VALU modes: gradients G and the Gram matrix T.  MFMA mode: the bf16 operand
  // parts, then T over the same bytes once the products are in registers.  The
  // operand rows are read as 8-byte vectors: the union is 16-byte aligned whatever
  // the fields before it add up to.
  // FP64 rows are not padded (81 doubles is already an odd bank stride), which keeps
  // the FP64 form under half of a CU's LDS: two workgroups per CU.
  static constexpr int TP = MODE == MODE_F64 ? D::ND : D::ND + 1;
  // (the bf16 MFMA form uses neither T nor G of `v`, the FP64 form no G: they shrink to one entry)
  static constexpr int VR = MODE == MODE_SPLIT ? 1 : D::ND;
  union alignas(16) {
    struct {
      TReal T[VR][TP];
      TReal G[1];
    } v;
    short ops[3 * D::MP * D::KS];
    float T32[D::ND][D::ND + 1];
    // FP64 form: the operands of T = A^T B, A = (w det sgn) G and B = G, rows (k, a) padded
    // to MP, points padded to KP64 (zero); T goes over them once the products are in registers
    double opd[MODE == MODE_F64 ? 2 * D::MP * D::KS64 : 1];
    // FP32 form: the same operands in FP32 for v_mfma_f32_16x16x4f32
    float opf[MODE == MODE_F32 ? 2 * D::MP * D::KS32 : 1];
  } w;
};

// The MFMA form is held to four waves per SIMD (at most 128 registers, AGPRs included),
// which its LDS also allows: four workgroups per CU (Q2 64^3 split form 9.21 -> 8.81 ms
// against three, at the cost of a 20-byte spill).
template <int NB, int NQ, int BS, int MODE>
__global__ __launch_bounds__(BS, MODE == MODE_SPLIT ? 4 : 1) void k_elastic(ElasticParams p) {
  using D = Dims<NB>;
  using TReal = typename Smem<NB, NQ, MODE>::TReal;
  constexpr int ND = D::ND;
  constexpr int Q1 = NQ == 27 ? 3 : 2;        // Gauss points per axis
  // final phase: lane (a, b) forms the 3x3 block of node pair (a, b); a group of GW
  // lanes shares a and runs over b
  constexpr int GW = NB == 27 ? 32 : 8;
  constexpr int GPW = 64 / GW;
  constexpr int AP = (BS / 64) * GPW;          // nodes a per pass
  constexpr int NPASS = (NB + AP - 1) / AP;
  __shared__ Smem<NB, NQ, MODE> sm;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int gb = lane % GW, ga = wave * GPW + lane / GW;
  const bool tr = p.transient && p.xdot;
  const bool tra = tr && p.xddot;
  const bool has_src = p.src[0] != 0.0 || p.src[1] != 0.0 || p.src[2] != 0.0;

  if (t == 0) {
    // Gauss-Legendre on [-1, 1]; Lagrange bases on the nodes {-1, 1} or {-1, 0, 1}
    const double g3 = 0.77459666924148337704;   // sqrt(3/5)
    const double g2 = 0.57735026918962576451;   // 1/sqrt(3)
    for (int q = 0; q < Q1; ++q) {
      const double xi = Q1 == 3 ? (q - 1) * g3 : (2 * q - 1) * g2;
      sm.w1[q] = Q1 == 3 ? (q == 1 ? 8.0 / 9.0 : 5.0 / 9.0) : 1.0;
      sm.gL[0][q] = 0.5 * (1.0 - xi);
      sm.gL[1][q] = 0.5 * (1.0 + xi);
      sm.gdL[0][q] = -0.5;
      sm.gdL[1][q] = 0.5;
      if (NB == 27) {
        sm.bL[0][q] = 0.5 * xi * (xi - 1.0);
        sm.bL[1][q] = 1.0 - xi * xi;
        sm.bL[2][q] = 0.5 * xi * (xi + 1.0);
        sm.bdL[0][q] = xi - 0.5;
        sm.bdL[1][q] = -2.0 * xi;
        sm.bdL[2][q] = xi + 0.5;
      } else {
        sm.bL[0][q] = sm.gL[0][q];
        sm.bL[1][q] = sm.gL[1][q];
        sm.bdL[0][q] = -0.5;
        sm.bdL[1][q] = 0.5;
      }
    }
  }
  for (int i = t; i < NB * 3; i += BS) sm.off[i / 3][i % 3] = NB == 8 ? kHexV[i / 3][i % 3] : kQ2[i / 3][i % 3];
  __syncthreads();

  // Gather data of the element after this one, fetched while this one is computed: the
  // coordinates, state and coefficient (thread t < 24: one vertex coordinate; t < ND:
  // one DOF) once the Gram matrix is done, from the connectivity and LIDs fetched one
  // element earlier; they are put in LDS after the pair phase, before the block store.
  // No global load is then pending when the block store goes out, so the next element
  // starts without waiting for those stores (one counter tracks both).  The loads have
  // no branch (indices clamped to a valid element and lane; the extra values are not
  // used): a conditional load would be copied, and so waited for, where the branch joins.
  static_assert(BS >= ND && BS >= 24, "one gather entry per thread");
  int64_t gn = 0;
  int32_t gL = 0;
  double gX = 0.0, gu = 0.0, gv = 0.0, gac = 0.0, gE = 0.0;
  auto fetch_index = [&](int64_t e) {
    const int64_t ec = e < p.e1 ? e : p.e1 - 1;
    gn = p.conn[ec * p.nconn + min(t, 23) / 3];
    gL = p.elem_lid[ec * ND + min(t, ND - 1)];
  };
  auto fetch_data = [&](int64_t e) {
    const int64_t ec = e < p.e1 ? e : p.e1 - 1;
    const int d = t % 3;
    const double *P = d == 0 ? p.X : (d == 1 ? p.Y : p.Z);
    gX = P[gn];
    gu = p.x[gL];
    gv = tr ? p.xdot[gL] : 0.0;
    gac = tra ? p.xddot[gL] : 0.0;
    gE = p.coef[ec];
  };
  auto put_data = [&]() {
    if (t < 24) sm.X[t / 3][t % 3] = gX;
    if (t < ND) {
      sm.u[t] = gu;
      sm.v[t] = gv;
      sm.a[t] = gac;
    }
    asm volatile("" : "+v"(gE));   // the coefficient is taken here too, not after the block store
  };
  fetch_index(p.e0 + blockIdx.x);
  fetch_data(p.e0 + blockIdx.x);
  put_data();
  double E = gE;
  fetch_index(p.e0 + blockIdx.x + gridDim.x);

  for (int64_t e = p.e0 + blockIdx.x; e < p.e1; e += gridDim.x) {
    __syncthreads();   // this element's vertices and state are in LDS
    // ---- geometry (FP64)
    if (t < NQ) {
      const int qx = t % Q1, qy = (t / Q1) % Q1, qz = t / (Q1 * Q1);
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
      for (int vtx = 0; vtx < 8; ++vtx) {
        const int ox = kHexV[vtx][0], oy = kHexV[vtx][1], oz = kHexV[vtx][2];
        const double gx = sm.gL[ox][qx], gy = sm.gL[oy][qy], gz = sm.gL[oz][qz];
        const double d[3] = {sm.gdL[ox][qx] * gy * gz, gx * sm.gdL[oy][qy] * gz, gx * gy * sm.gdL[oz][qz]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) J[i][j] += sm.X[vtx][i] * d[j];
      }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      double *Ji = sm.Ji[t];
      Ji[0] = c00 * id;
      Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[3] = c01 * id;
      Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[6] = c02 * id;
      Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      sm.det[t] = det;
      sm.wm[t] = sm.w1[qx] * sm.w1[qy] * sm.w1[qz] * det;
    }
    if constexpr (MODE == MODE_SPLIT) {
      // operand rows >= ND are zero (T covered them in the previous element)
      constexpr int PADV = (D::MP - ND) * D::KS / 8;   // 16-byte vectors per part
      for (int i = t; i < 3 * PADV; i += BS) {
        const int s = i / PADV, o = i % PADV;
        reinterpret_cast<uint4 *>(sm.w.ops + (s * D::MP + ND) * D::KS)[o] = make_uint4(0, 0, 0, 0);
      }
    }
    if constexpr (MODE == MODE_F64) {
      constexpr int PADD = (D::MP - ND) * D::KS64;   // doubles per operand
      for (int i = t; i < 2 * PADD; i += BS) sm.w.opd[((i / PADD) * D::MP + ND) * D::KS64 + i % PADD] = 0.0;
    }
    if constexpr (MODE == MODE_F32) {
      constexpr int PADF = (D::MP - ND) * D::KS32;   // floats per operand
      for (int i = t; i < 2 * PADF; i += BS) sm.w.opf[((i / PADF) * D::MP + ND) * D::KS32 + i % PADF] = 0.0f;
    }
    __syncthreads();
    // measure w |det J| with the sign of the first point (Intrepid2 computeCellMeasure)
    const double sg = sm.det[0] < 0.0 ? -1.0 : 1.0;
    if (t < NQ) sm.sw[t] = sqrt(sg * sm.wm[t]);
    // mass matrix (inertia and Rayleigh terms only)
    if (tr)
      for (int i = t; i < NB * NB; i += BS) {
        const int a = i / NB, b = i % NB;
        double s = 0.0;
        for (int q = 0; q < NQ; ++q) {
          const int qx = q % Q1, qy = (q / Q1) % Q1, qz = q / (Q1 * Q1);
          s += sm.wm[q] * (sm.bL[sm.off[a][0]][qx] * sm.bL[sm.off[a][1]][qy] * sm.bL[sm.off[a][2]][qz]) *
               (sm.bL[sm.off[b][0]][qx] * sm.bL[sm.off[b][1]][qy] * sm.bL[sm.off[b][2]][qz]);
        }
        sm.M[a][b] = static_cast<TReal>(sg * s);
      }
    __syncthreads();
    // ---- physical gradients at the quadrature points (MFMA mode: also the zero columns q >= NQ);
    // q runs fastest over the threads, so the operand writes of a wave are consecutive in LDS
    constexpr int QG = MODE == MODE_SPLIT ? D::KP : D::KP64;
    // when BS is a multiple of QG every thread keeps one q: its J^-1 and sqrt(w |det J|)
    // are read from LDS once
    constexpr bool QFIX = BS % QG == 0;
    double Jq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sqq = 1.0;
    if constexpr (QFIX) {
      const int q = t % QG;
      if (q < NQ) {
#pragma unroll
        for (int r = 0; r < 9; ++r) Jq[r] = sm.Ji[q][r];
        if constexpr (MODE == MODE_SPLIT) sqq = sm.sw[q];
      }
    }
    for (int i = t; i < QG * NB; i += BS) {
      const int q = i % QG, a = i / QG;
      if (MODE == MODE_SPLIT && q >= NQ) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int s = 0; s < 3; ++s) sm.w.ops[(s * D::MP + k * NB + a) * D::KS + q] = 0;
        continue;
      }
      if (MODE == MODE_F64 && q >= NQ) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          sm.w.opd[(0 * D::MP + k * NB + a) * D::KS64 + q] = 0.0;
          sm.w.opd[(1 * D::MP + k * NB + a) * D::KS64 + q] = 0.0;
        }
        continue;
      }
      if (MODE == MODE_F32 && q >= NQ) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          sm.w.opf[(0 * D::MP + k * NB + a) * D::KS32 + q] = 0.0f;
          sm.w.opf[(1 * D::MP + k * NB + a) * D::KS32 + q] = 0.0f;
        }
        continue;
      }
      const int qx = q % Q1, qy = (q / Q1) % Q1, qz = q / (Q1 * Q1);
      const int ax = sm.off[a][0], ay = sm.off[a][1], az = sm.off[a][2];
      const double lx = sm.bL[ax][qx], ly = sm.bL[ay][qy], lz = sm.bL[az][qz];
      const double dn[3] = {sm.bdL[ax][qx] * ly * lz, lx * sm.bdL[ay][qy] * lz, lx * ly * sm.bdL[az][qz]};
      const double *Ji = sm.Ji[q];
      auto ji = [&](int r) { return QFIX ? Jq[r] : Ji[r]; };
      const double sq = QFIX ? sqq : (MODE == MODE_SPLIT ? sm.sw[q] : 1.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double g = ji(0 * 3 + k) * dn[0] + ji(1 * 3 + k) * dn[1] + ji(2 * 3 + k) * dn[2];
        if constexpr (MODE == MODE_SPLIT) {
          // operand [m = (k,a)][q] = sqrt(w |det J|) g in three bf16 parts: T = O^T O
          short s0, s1, s2;
          split3(static_cast<float>(sq * g), s0, s1, s2);
          const int m = k * NB + a;
          sm.w.ops[(0 * D::MP + m) * D::KS + q] = s0;
          sm.w.ops[(1 * D::MP + m) * D::KS + q] = s1;
          sm.w.ops[(2 * D::MP + m) * D::KS + q] = s2;
        } else if constexpr (MODE == MODE_F64) {
          // T[(k,a),(l,b)] = sum_q (sgn w det g_ka) g_lb
          const int m = k * NB + a;
          sm.w.opd[(0 * D::MP + m) * D::KS64 + q] = sg * sm.wm[q] * g;
          sm.w.opd[(1 * D::MP + m) * D::KS64 + q] = g;
        } else {
          // T[(k,a),(l,b)] = sum_q (sgn w det g_ka) g_lb, operands rounded to FP32 once
          const int m = k * NB + a;
          sm.w.opf[(0 * D::MP + m) * D::KS32 + q] = static_cast<float>(sg * sm.wm[q] * g);
          sm.w.opf[(1 * D::MP + m) * D::KS32 + q] = static_cast<float>(g);
        }
      }
    }
    __syncthreads();
    // ---- gradient Gram matrix T
    if constexpr (MODE == MODE_SPLIT) {
      constexpr int NT = D::MP / 16;                 // 16x16 tiles per side
      constexpr int NTILE = NT * (NT + 1) / 2;       // upper triangle
      constexpr int TPW = (NTILE + BS / 64 - 1) / (BS / 64);
      const short *O = sm.w.ops;
      v4f acc[TPW];
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
        // gfx950's K = 32 bf16 product: lane l brings eight points of row l % 16, the eighth
        // 8 (l / 16) .. 8 (l / 16) + 7, of A and of B alike (whatever order the instruction gives
        // the eight, the sum over all 32 points is the same)
#pragma unroll
        for (int ks = 0; ks < D::KP / 32; ++ks) {
          const int k0 = ks * 32 + 8 * (lane >> 4);
          const int ma = tm * 16 + (lane & 15), nb = tn * 16 + (lane & 15);
          v8bf a_[3], b_[3];
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            a_[s] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8s *>(O + (s * D::MP + ma) * D::KS + k0));
            b_[s] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v8s *>(O + (s * D::MP + nb) * D::KS + k0));
          }
          // products of order 2^-16, 2^-8 and 1 (smallest first)
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[2], b_[0], acc[k], 0, 0, 0);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[1], b_[1], acc[k], 0, 0, 0);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0], b_[2], acc[k], 0, 0, 0);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[1], b_[0], acc[k], 0, 0, 0);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0], b_[1], acc[k], 0, 0, 0);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_[0], b_[0], acc[k], 0, 0, 0);
        }
      }
      __syncthreads();   // every wave has read the operands: T goes over them
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
        // D[4 (lane/16) + i][lane % 16] of the tile; off-diagonal tiles also fill the
        // mirrored tile (the summation order of the split products differs between
        // (m, n) and (n, m), so a diagonal tile keeps its own entries on both sides)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = tm * 16 + 4 * (lane >> 4) + i, cc = tn * 16 + (lane & 15);
          if (rr < ND && cc < ND) {
            sm.w.T32[rr][cc] = acc[k][i];
            if (tm != tn) sm.w.T32[cc][rr] = acc[k][i];
          }
        }
      }
    } else if constexpr (MODE == MODE_F64) {
      // T = A^T B on v_mfma_f64_16x16x4f64 over the upper-triangle 16x16 tiles: lane l holds
      // A and B element [row l % 16][point 4 ks + l / 16]; the tile's D[l/16 + 4 i][l % 16]
      // lands in component i (rows interleaved, unlike the FP32-accumulating 16x16 forms)
      constexpr int NT = D::MP / 16;
      constexpr int NTILE = NT * (NT + 1) / 2;
      constexpr int TPW = (NTILE + BS / 64 - 1) / (BS / 64);
      const double *A = sm.w.opd, *B = sm.w.opd + D::MP * D::KS64;
      v4d acc[TPW];
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        acc[k] = v4d{0.0, 0.0, 0.0, 0.0};
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
        const double *ar = A + (tm * 16 + (lane & 15)) * D::KS64 + (lane >> 4);
        const double *br = B + (tn * 16 + (lane & 15)) * D::KS64 + (lane >> 4);
#pragma unroll
        for (int ks = 0; ks < D::KP64 / 4; ++ks)
          acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[4 * ks], br[4 * ks], acc[k], 0, 0, 0);
      }
      __syncthreads();   // every wave has read the operands: T goes over them
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = tm * 16 + (lane >> 4) + 4 * i, cc = tn * 16 + (lane & 15);
          if (rr < ND && cc < ND) {
            sm.w.v.T[rr][cc] = acc[k][i];
            if (tm != tn) sm.w.v.T[cc][rr] = acc[k][i];
          }
        }
      }
    } else {
      // T = A^T B on v_mfma_f32_16x16x4f32 over the upper-triangle 16x16 tiles: lane l holds A
      // and B element [row l % 16][point 4 ks + l / 16]; D[4 (l / 16) + i][l % 16] in component i
      constexpr int NT = D::MP / 16;
      constexpr int NTILE = NT * (NT + 1) / 2;
      constexpr int TPW = (NTILE + BS / 64 - 1) / (BS / 64);
      const float *A = sm.w.opf, *B = sm.w.opf + D::MP * D::KS32;
      v4f acc[TPW];
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
        const float *ar = A + (tm * 16 + (lane & 15)) * D::KS32 + (lane >> 4);
        const float *br = B + (tn * 16 + (lane & 15)) * D::KS32 + (lane >> 4);
#pragma unroll
        for (int ks = 0; ks < D::KP64 / 4; ++ks)
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[4 * ks], br[4 * ks], acc[k], 0, 0, 0);
      }
      __syncthreads();   // every wave has read the operands: T goes over them
#pragma unroll
      for (int k = 0; k < TPW; ++k) {
        const int ti = wave + k * (BS / 64);
        if (ti >= NTILE) continue;
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = tm * 16 + 4 * (lane >> 4) + i, cc = tn * 16 + (lane & 15);
          if (rr < ND && cc < ND) {
            sm.w.v.T[rr][cc] = acc[k][i];
            if (tm != tn) sm.w.v.T[cc][rr] = acc[k][i];
          }
        }
      }
    }
    __syncthreads();
    fetch_data(e + gridDim.x);
    fetch_index(e + 2 * static_cast<int64_t>(gridDim.x));
    // ---- 3x3 blocks of node pairs: residual rows, Jacobian values written over the
    // pair's own T entries (the block of (a, b) depends only on those nine values)
    //   K_ij = lambda T[(i,a),(j,b)] + mu T[(j,a),(i,b)] + delta_ij mu sum_k T[(k,a),(k,b)]
    const double lam = E * p.nu / ((1.0 + p.nu) * (1.0 - 2.0 * p.nu));
    const double mu = E / (2.0 * (1.0 + p.nu));
    auto Tref = [&](int r, int c) -> TReal & {
      if constexpr (MODE == MODE_SPLIT) return sm.w.T32[r][c];
      else return sm.w.v.T[r][c];
    };
    const double cK = tr ? p.beta + p.alpha * p.rb : p.beta;
    const double cM = tr ? p.alpha * p.ra + (tra ? p.gamma2 : 0.0) : 0.0;
#pragma unroll
    for (int k = 0; k < NPASS; ++k) {
      const int a = ga + k * AP, b = gb;
      double rr[3] = {0.0, 0.0, 0.0};
      if (a < NB && b < NB) {
        double t9[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) t9[i][j] = static_cast<double>(Tref(i * NB + a, j * NB + b));
        const double trs = t9[0][0] + t9[1][1] + t9[2][2];
        const double Mab = tr ? p.rho * static_cast<double>(sm.M[a][b]) : 0.0;
        // with C = a_R M + b_R K:  J = (beta + alpha b_R) K + (alpha a_R + gamma2) M  and
        // R = K (u + b_R v) + M (a_R v + a), node b's state combined once for its nine entries
        double ut[3], mt[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          ut[j] = tr ? sm.u[3 * b + j] + p.rb * sm.v[3 * b + j] : sm.u[3 * b + j];
          mt[j] = tr ? p.ra * sm.v[3 * b + j] + (tra ? sm.a[3 * b + j] : 0.0) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double Kv = lam * t9[i][j] + mu * t9[j][i] + (i == j ? mu * trs : 0.0);
            rr[i] += Kv * ut[j];
            double jv = cK * Kv;
            if (i == j) {
              jv += cM * Mab;
              rr[i] += Mab * mt[j];
            }
            Tref(i * NB + a, j * NB + b) = static_cast<TReal>(jv);
          }
      }
      // row residuals: fixed-order reduction over the group's b (lanes b >= NB add 0)
      static_assert(GW == 8 || GW == 32, "group width");
#pragma unroll
      for (int i = 0; i < 3; ++i) rr[i] = group_sum<GW>(rr[i]);
      if (a < NB && b == (GW == 32 ? 16 : 0))
#pragma unroll
        for (int i = 0; i < 3; ++i) sm.r[3 * a + i] = rr[i];
    }
    __syncthreads();
    put_data();   // the next element's vertices and state (LDS X, u, v, a are free now)
    // ---- element block to the workspace, node-major (node a: rows i NB + a, i = 0..2, of
    // the Gram layout, then padding): consecutive lanes on consecutive entries
    if (p.jbuf) {
      constexpr int NP = block_stride(NB);
      for (int idx = t; idx < NB * NP; idx += BS) {
        const int a = idx / NP, q = idx % NP;
        const TReal v = q < 3 * ND ? TReal(Tref((q / ND) * NB + a, q % ND)) : TReal(0);
        TReal *dst = &static_cast<TReal *>(p.jbuf)[e * (NB * NP) + idx];
        // hex27 blocks stream past the caches (measured faster); hex8 blocks as plain stores
        if (NB == 27) __builtin_nontemporal_store(v, dst);
        else *dst = v;
      }
    }
    __syncthreads();
    if (p.rbuf)
      for (int i = t; i < ND; i += BS) {
        double F = 0.0;
        if (has_src) {
          const int a = i / 3;
          const int ax = sm.off[a][0], ay = sm.off[a][1], az = sm.off[a][2];
          for (int q = 0; q < NQ; ++q) {
            const int qx = q % Q1, qy = (q / Q1) % Q1, qz = q / (Q1 * Q1);
            F += sm.wm[q] * sm.bL[ax][qx] * sm.bL[ay][qy] * sm.bL[az][qz];
          }
          // a select, not an indexed read: src[i % 3] would be a vector load from the
          // parameters, and waiting for it would wait for the block stores
          const int c = i % 3;
          F *= sg * (c == 0 ? p.src[0] : (c == 1 ? p.src[1] : p.src[2]));
        }
        p.rbuf[e * ND + i] = sm.r[i] - F;
      }
    E = gE;
  }
}

// Second phase: node-centric gather.  One wave per lattice node assembles the
// node's three CRS rows (consecutive LIDs with one column structure) in LDS in
// FP64 from the blocks of the elements around it, in a fixed element order, then
// writes them once, coalesced, with the residual.
struct GatherParams {
  int nx, ny, nz;
  int64_t px, py, pz, npoints;
  int64_t p0, p1;             // lattice points [p0, p1) of this launch
  const int32_t *point_lid;   // [points][3]
  const int64_t *rowptr;
  const uint16_t *nperm;
  int nperm_w;
  const void *jbuf;
  const double *rbuf;
  double *f, *vals;
};

// One wave stores n consecutive FP64 values from LDS to dst, non-temporally (written once,
// not re-read here), in chunks aligned to 128-byte lines: only the first and the last line
// of the range are partial.
__device__ __forceinline__ void row_store(const double *src, double *dst, int n, int lane) {
  const int lead = static_cast<int>((reinterpret_cast<uintptr_t>(dst) >> 3) & 15);
  for (int k = lane - lead; k < n; k += 64)
    if (k >= 0) __builtin_nontemporal_store(src[k], dst + k);
}

__device__ __forceinline__ void gather_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NB, typename JT>
__global__ __launch_bounds__(256) void k_elastic_gather(GatherParams g) {
  constexpr int ND = 3 * NB;
  constexpr int R = NB == 8 ? 1 : 2, S = 2 * R + 1;
  constexpr int STEP = NB == 8 ? 1 : 2;        // lattice points per cell edge
  constexpr int MAXLEN = 3 * S * S * S;        // entries per row
  __shared__ double buf[4][3 * MAXLEN];
  __shared__ int8_t inv[3][3][3];
  // the wave index is uniform: everything derived from the node stays in scalar registers
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int NR = (3 * ND + 63) / 64;   // block entries per lane and element
  constexpr int NP = block_stride(NB);
  // per lane and r: entry idx = lane + 64 r of the node's three block rows (row field
  // idx / ND, Gram column idx % ND = field j, node bn), and the lattice offset index of
  // its column node plus R in each axis; the entry sits at idx in the node's part of the block
  int lterm[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int idx = lane + 64 * r, c = idx % ND;
    const int bn = c % NB;
    const int8_t *o = NB == 8 ? kHexV[bn] : kQ2[bn];
    lterm[r] = (o[0] + R) + S * (o[1] + R) + S * S * (o[2] + R);
  }
  for (int i = threadIdx.x; i < NB; i += 256) {
    const int8_t *o = NB == 8 ? kHexV[i] : kQ2[i];
    inv[o[0]][o[1]][o[2]] = static_cast<int8_t>(i);
  }
  __syncthreads();
  double *b = buf[w];
  const JT *J = static_cast<const JT *>(g.jbuf);
  // The node's LID and row pointers are wave-uniform: they are read through the scalar
  // cache (constant address space; neither array changes during the kernel), whose
  // counter is not the one the row stores use, and one node ahead.  Read as vector loads,
  // each node began by waiting for them and with them for every store of the node before.
  using cint = const __attribute__((address_space(4))) int32_t;
  using clong = const __attribute__((address_space(4))) int64_t;
  const cint *plid = (const cint *)g.point_lid;
  const clong *prow = (const clong *)g.rowptr;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 4;
  int64_t pt = g.p0 + static_cast<int64_t>(blockIdx.x) * 4 + w;
  int32_t Lc = plid[min(pt, g.p1 - 1) * 3];
  int64_t bc = prow[Lc], ec = prow[Lc + 1];
  // lattice coordinates of the node, advanced by the stride's own coordinates with
  // carries: the divisions happen once, not per node
  const int64_t pxy = g.px * g.py;
  int64_t I = pt % g.px, Jy = (pt / g.px) % g.py, K = pt / pxy;
  const int64_t sI = stride % g.px, sJ = (stride / g.px) % g.py, sK = stride / pxy;
  for (; pt < g.p1; pt += stride) {
    const int32_t L0 = Lc;
    const int64_t base = bc;
    // row length (at most S^3 neighbours x 3 fields; clamped so the LDS rows never overflow)
    const int len = min(static_cast<int>(ec - bc), MAXLEN);
    const int32_t Ln = plid[min(pt + stride, g.p1 - 1) * 3];   // the next node's LID
    const int64_t node = L0 / 3;
    const uint16_t *np = g.nperm + node * g.nperm_w;
    for (int k = lane; k < 3 * len; k += 64) b[k] = 0.0;
    gather_lds_sync();
    // cells containing the node (at most 8), per axis [c_lo, c_hi]
    const int64_t cx_lo = I % STEP == 0 ? I / STEP - 1 : I / STEP, cx_hi = I / STEP;
    const int64_t cy_lo = Jy % STEP == 0 ? Jy / STEP - 1 : Jy / STEP, cy_hi = Jy / STEP;
    const int64_t cz_lo = K % STEP == 0 ? K / STEP - 1 : K / STEP, cz_hi = K / STEP;
    int64_t ce[8];
    int ca[8], cs[8];
    int ne = 0;
    for (int64_t cz = cz_lo; cz <= cz_hi; ++cz)
      for (int64_t cy = cy_lo; cy <= cy_hi; ++cy)
        for (int64_t cx = cx_lo; cx <= cx_hi; ++cx)
          if (cx >= 0 && cx < g.nx && cy >= 0 && cy < g.ny && cz >= 0 && cz < g.nz) {
            const int ox = static_cast<int>(I - STEP * cx), oy = static_cast<int>(Jy - STEP * cy),
                      oz = static_cast<int>(K - STEP * cz);   // lattice offset of the node in the cell
            ce[ne] = cx + g.nx * (cy + g.ny * cz);
            ca[ne] = inv[ox][oy][oz];
            cs[ne] = ox + S * oy + S * S * oz;
            ++ne;
          }
    // every load first (one round trip per node), then the additions in element order
    JT val[8][NR];
    uint16_t pos[8][NR];
    double rr[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      rr[c] = 0.0;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        val[c][r] = JT(0);
        pos[c][r] = 0xffffu;
        const int idx = lane + 64 * r;
        if (c < ne && idx < 3 * ND && J) {
          val[c][r] = J[(ce[c] * NB + ca[c]) * NP + idx];   // node ca's rows: one contiguous run
          pos[c][r] = np[lterm[r] - cs[c]];
        }
      }
      if (c < ne && lane < 3 && g.rbuf) rr[c] = g.rbuf[ce[c] * ND + 3 * ca[c] + lane];
    }
    const int64_t bn = prow[Ln], en = prow[Ln + 1];   // the next node's row pointers
    // the values stay opaque until here, so no conversion pulls a wait in among the loads
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int r = 0; r < NR; ++r) asm volatile("" : "+v"(val[c][r]));
    // LDS atomic adds, issued back to back: the LDS executes one wave's operations in
    // order and no two lanes of one instruction share an entry, so every entry is summed
    // in cell order, as one read-add-write round per cell did, without a round trip per cell
    double racc = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (c < ne) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int idx = lane + 64 * r;
          const int i = idx / ND, j = (idx % ND) / NB;
          const uint16_t ps = pos[c][r];
          if (idx < 3 * ND && ps != 0xffffu && ps + j < len) atomicAdd(&b[i * len + ps + j], static_cast<double>(val[c][r]));
        }
        racc += rr[c];
      }
    }
    gather_lds_sync();   // all additions done before the rows are read
    if (g.vals)
      row_store(b, g.vals + base, 3 * len, lane);
    if (g.f && lane < 3) g.f[L0 + lane] = racc;
    gather_lds_sync();   // the buffer is reused by the next node
    Lc = Ln;
    bc = bn;
    ec = en;
    I += sI;
    Jy += sJ;
    K += sK;
    if (I >= g.px) { I -= g.px; ++Jy; }
    if (Jy >= g.py) { Jy -= g.py; ++K; }
  }
}

// The same gather on an uploaded hex8 / hex27 mesh of one block: node n (rows 3n, 3n + 1, 3n + 2:
// the first-touch numbering gives a node's three fields consecutive LIDs) takes its element
// instances (e, a) from the node's list, ascending, eight at a time, and the row position of
// each element node's field 0 from the instance's slot map.
struct GatherUParams {
  int64_t nnodes;
  const int64_t *occ_ptr;   // [nnodes + 1]
  const int32_t *occ;       // e * NB + a per instance
  const uint16_t *slot;     // [instances][NB] position of (element node b, field 0) in row 3n
  const int64_t *rowptr;
  const void *jbuf;
  const double *rbuf;
  double *f, *vals;
};

template <int NB, typename JT>
__global__ __launch_bounds__(256) void k_elastic_gather_u(GatherUParams g) {
  constexpr int ND = 3 * NB;
  constexpr int R = NB == 8 ? 1 : 2, S = 2 * R + 1;
  constexpr int MAXLEN = 3 * S * S * S;        // entries per row (checked when the lists are built)
  __shared__ double buf[4][3 * MAXLEN];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int NR = (3 * ND + 63) / 64;
  constexpr int NP = block_stride(NB);
  double *b = buf[w];
  const JT *J = static_cast<const JT *>(g.jbuf);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t n = static_cast<int64_t>(blockIdx.x) * 4 + w; n < g.nnodes; n += stride) {
    const int32_t L0 = static_cast<int32_t>(3 * n);
    const int64_t base = g.rowptr[L0];
    const int len = min(static_cast<int>(g.rowptr[L0 + 1] - base), MAXLEN);
    for (int k = lane; k < 3 * len; k += 64) b[k] = 0.0;
    gather_lds_sync();
    const int64_t o0 = g.occ_ptr[n], o1 = g.occ_ptr[n + 1];
    double racc = 0.0;
    for (int64_t k0 = o0; k0 < o1; k0 += 8) {
      const int ne = static_cast<int>(o1 - k0 < 8 ? o1 - k0 : 8);
      JT val[8][NR];
      uint16_t pos[8][NR];
      double rr[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        rr[c] = 0.0;
        const int32_t v = c < ne ? g.occ[k0 + c] : 0;
        const int64_t e = v / NB;
        const int a = v - static_cast<int>(e) * NB;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          val[c][r] = JT(0);
          pos[c][r] = 0xffffu;
          const int idx = lane + 64 * r;
          if (c < ne && idx < 3 * ND && J) {
            val[c][r] = J[(e * NB + a) * NP + idx];
            pos[c][r] = g.slot[(k0 + c) * NB + (idx % ND) % NB];
          }
        }
        if (c < ne && lane < 3 && g.rbuf) rr[c] = g.rbuf[e * ND + 3 * a + lane];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int r = 0; r < NR; ++r) asm volatile("" : "+v"(val[c][r]));
      // one wave's LDS operations execute in order and no two lanes of one instruction share
      // an entry: every entry is summed in instance order
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (c < ne) {
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int idx = lane + 64 * r;
            const int i = idx / ND, j = (idx % ND) / NB;
            const uint16_t ps = pos[c][r];
            if (idx < 3 * ND && ps != 0xffffu && ps + j < len)
              atomicAdd(&b[i * len + ps + j], static_cast<double>(val[c][r]));
          }
          racc += rr[c];
        }
      }
    }
    gather_lds_sync();
    if (g.vals)
      row_store(b, g.vals + base, 3 * len, lane);
    if (g.f && lane < 3) g.f[L0 + lane] = racc;
    gather_lds_sync();
  }
}

// Element instances of each node of an uploaded 3-field mesh (field-0 positions of elem_lid)
__global__ void k_uocc_count(int64_t npos0, int nb, const int32_t *__restrict__ elem_lid,
                             unsigned long long *__restrict__ cnt) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= npos0) return;   // t = e * nb + a
  atomicAdd(cnt + elem_lid[3 * t] / 3, 1ull);
}

__global__ void k_uocc_fill(int64_t npos0, int nb, const int32_t *__restrict__ elem_lid,
                            unsigned long long *__restrict__ cursor, int32_t *__restrict__ occ) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= npos0) return;
  occ[atomicAdd(cursor + elem_lid[3 * t] / 3, 1ull)] = static_cast<int32_t>(t);
}

// Sorts each node's list (ascending e * nb + a: a fixed summation order), the slot map of each
// instance, and the longest row and list (out[0], out[1]).
__global__ void k_uocc_finish(int64_t nnodes, int nb, const int64_t *__restrict__ optr, int32_t *__restrict__ occ,
                              const int32_t *__restrict__ elem_lid, const int64_t *__restrict__ rowptr,
                              const int32_t *__restrict__ col, uint16_t *__restrict__ slot, int *__restrict__ out) {
  const int64_t n = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int64_t lo = optr[n], hi = optr[n + 1];
  for (int64_t i = lo + 1; i < hi; ++i) {
    const int32_t v = occ[i];
    int64_t j = i;
    for (; j > lo && occ[j - 1] > v; --j) occ[j] = occ[j - 1];
    occ[j] = v;
  }
  const int64_t r0 = rowptr[3 * n], len = rowptr[3 * n + 1] - r0;
  atomicMax(out, static_cast<int>(len));
  atomicMax(out + 1, static_cast<int>(hi - lo));
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t e = occ[i] / nb;
    for (int b = 0; b < nb; ++b) {
      const int32_t key = elem_lid[(e * nb + b) * 3];
      int64_t l = 0, h = len;
      while (l < h) {
        const int64_t m = (l + h) >> 1;
        if (col[r0 + m] < key) l = m + 1; else h = m;
      }
      slot[i * nb + b] = l < len && col[r0 + l] == key && l < 0xffff ? static_cast<uint16_t>(l) : uint16_t(0xffff);
    }
  }
}

// Row position of (neighbour, field 0) for every node and lattice offset in
// [-R, R]^3 (0xffff: not a neighbour), by binary search in the node's first row.
__global__ void k_node_perm(int64_t px, int64_t py, int64_t pz, int R, int nf, int W,
                            const int32_t *__restrict__ point_lid, const int64_t *__restrict__ rowptr,
                            const int32_t *__restrict__ col, uint16_t *__restrict__ nperm) {
  const int64_t pt = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pt >= px * py * pz) return;
  const int64_t i = pt % px, j = (pt / px) % py, k = pt / (px * py);
  const int32_t row = point_lid[pt * nf];
  const int64_t lo0 = rowptr[row], hi0 = rowptr[row + 1];
  uint16_t *out = nperm + static_cast<int64_t>(row / nf) * W;
  const int S = 2 * R + 1;
  for (int s = 0; s < W; ++s) {
    uint16_t v = 0xffffu;
    if (s < S * S * S) {
      const int64_t ii = i + s % S - R, jj = j + (s / S) % S - R, kk = k + s / (S * S) - R;
      if (ii >= 0 && ii < px && jj >= 0 && jj < py && kk >= 0 && kk < pz) {
        const int32_t key = point_lid[(ii + px * (jj + py * kk)) * nf];
        int64_t lo = lo0, hi = hi0;
        while (lo < hi) {
          const int64_t m = (lo + hi) >> 1;
          if (col[m] < key) lo = m + 1; else hi = m;
        }
        if (lo < hi0 && col[lo] == key) v = static_cast<uint16_t>(lo - lo0);
      }
    }
    out[s] = v;
  }
}

// Affine HEX8 bricks (every cell a parallelepiped), three fields: node-centric assembly with
// no workspace.  One wave per lattice node.  Lane (c, b) = (cell c of the eight around the
// node, vertex b of that cell, binary order) forms the 3x3 block of the node's vertex a and b in
// closed form: with J = dx/dxi constant over the cell and C its cofactor matrix,
//   T_ab = |det J| J^-T S_ab J^-1 = C S_ab C^T / |det J|,
//   S_ab[m][n] = sum_q dN_a/dxi_m dN_b/dxi_n   (2x2x2 Gauss)
//              = s_a^m s_b^n f_d / 4 (m != n, d the third axis),  s_a^m s_b^m f_d1 f_d2 / 2 (m = n),
// with s the vertex's sign along an axis and f = 2/3 where a and b share it, 1/3 otherwise;
// the mass entry rho |det J| f_x f_y f_z, the source load src |det J| per vertex.  The rows are
// summed in LDS in eight steps, one per cell in a fixed order: in step c the cell's eight lanes
// (eight distinct neighbours) add their blocks at the neighbours' positions (from nperm); one
// wave's LDS operations execute in order, so every entry is summed in cell order and the result
// is bitwise reproducible.  The wave then writes the node's three rows once, coalesced.  The
// general (per-point) form is the element kernel plus the gather.
struct Elastic8Params {
  int nx, ny, nz;
  int64_t npoints;
  const int32_t *point_lid;   // [points][3]
  const int64_t *rowptr;
  const uint16_t *nperm;
  int nperm_w;
  const double *geom;         // [cells][10]: cofactors of J scaled by |det J|^-1/2, then |det J|
  const double *coef;         // Young's modulus per cell
  double nu, rho, src[3], ra, rb;
  const double *x, *xdot, *xddot;
  double alpha, beta, gamma2;
  int transient;
  double *f, *vals;
};

// Geometry of each parallelepiped cell, once per mesh: its constant Jacobian
// J[d][j] = (1/8) sum_b s_b^j X_b^d (s_b^j = -1 / +1 for vertex b's j-th bit 0 / 1), the
// cofactor matrix C (adj J = C^T) and |det J|, stored as C |det J|^-1/2 and |det J|, so that
// C S C^T / |det J| = Chat S Chat^T in the node kernel.  One thread per cell.
__global__ void k_cell_geom(int nx, int ny, int nz, int64_t nv, const double *__restrict__ xyz,
                            double *__restrict__ geom) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= static_cast<int64_t>(nx) * ny * nz) return;
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (static_cast<int64_t>(nx) * ny);
  double Jm[3][3] = {};
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
    const int64_t v = (i + bx) + (nx + 1) * ((j + by) + static_cast<int64_t>(ny + 1) * (k + bz));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double X = xyz[d * nv + v];
      Jm[d][0] += bx ? X : -X;
      Jm[d][1] += by ? X : -X;
      Jm[d][2] += bz ? X : -X;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) Jm[d][jj] *= 0.125;
  double C[9];
  C[0] = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
  C[1] = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
  C[2] = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
  C[3] = Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2];
  C[4] = Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0];
  C[5] = Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1];
  C[6] = Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1];
  C[7] = Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2];
  C[8] = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
  const double adet = fabs(Jm[0][0] * C[0] + Jm[0][1] * C[1] + Jm[0][2] * C[2]);
  const double sc = 1.0 / sqrt(adet);
  double *g = geom + 10 * c;
#pragma unroll
  for (int t = 0; t < 9; ++t) g[t] = C[t] * sc;
  g[9] = adet;
}

// Four waves per SIMD (125 VGPRs): the cell geometry is loaded with the node's second round
// rather than a node ahead (a prefetch held 20 more registers: three waves, 3% slower).
__global__ __launch_bounds__(256, 4) void k_elastic8_affine(Elastic8Params p) {
  constexpr int MAXLEN = 81;   // entries per row: 27 neighbours x 3 fields
  __shared__ double rows[4][3 * MAXLEN];
  // the wave index is uniform: the node's lattice arithmetic runs on the scalar unit
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane >> 3, b = lane & 7;
  const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;   // cell (I-1+ox, J-1+oy, K-1+oz)
  const int bxyz[3] = {b & 1, (b >> 1) & 1, b >> 2};
  const int oxyz[3] = {ox, oy, oz};
  // the node is vertex a = (1-ox, 1-oy, 1-oz) of cell c; vertex b lies at offset o - 1 + b
  double S[3][3], mf = 1.0;
  {
    double sa[3], sb[3], fd[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      sa[d] = oxyz[d] ? -1.0 : 1.0;
      sb[d] = bxyz[d] ? 1.0 : -1.0;
      fd[d] = sa[d] == sb[d] ? 2.0 / 3.0 : 1.0 / 3.0;
      mf *= fd[d];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 3; ++n)
        S[m][n] = m == n ? 0.5 * sa[m] * sb[m] * fd[(m + 1) % 3] * fd[(m + 2) % 3] : 0.25 * sa[m] * sb[n] * fd[3 - m - n];
  }
  const int dx = ox - 1 + bxyz[0], dy = oy - 1 + bxyz[1], dz = oz - 1 + bxyz[2];
  const int slot = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);   // neighbour slot of vertex b
  // 32-bit lattice arithmetic (the host takes this kernel only below 2^31 nodes)
  const int sx = p.nx + 1, sxy = sx * (p.ny + 1), np = static_cast<int>(p.npoints);
  const bool tr = p.transient && p.xdot;
  const bool tra = tr && p.xddot;
  const double lamf = p.nu / ((1.0 + p.nu) * (1.0 - 2.0 * p.nu)), muf = 0.5 / (1.0 + p.nu);
  const double cK = tr ? p.beta + p.alpha * p.rb : p.beta;
  const double cM = tr ? p.alpha * p.ra + (tra ? p.gamma2 : 0.0) : 0.0;
  const int stride = static_cast<int>(gridDim.x) * 4;
  // Loads come in two rounds per node (the lane's vertex LID and the node's LID; then the
  // state by that LID, the node's row pointers and slot map, and the cell's geometry and
  // modulus).  The first round of the next node is issued before this node's arithmetic,
  // the second round of this node as its iteration starts.
  struct Round1 {
    int I, J, K;
    int32_t Lb, L0;
  };
  auto round1 = [&](int q, Round1 &r) {
    const int qc = q < np ? q : np - 1;   // past the end: loaded, never used
    r.I = qc % sx;
    r.J = (qc / sx) % (p.ny + 1);
    r.K = qc / sxy;
    // this lane's vertex (clamped into the mesh for absent cells: loaded, never used)
    const int vx = min(max(r.I + dx, 0), p.nx), vy = min(max(r.J + dy, 0), p.ny), vz = min(max(r.K + dz, 0), p.nz);
    const int vi = vx + sx * vy + sxy * vz;
    r.Lb = p.point_lid[3 * vi];
    r.L0 = p.point_lid[3 * qc];
  };
  int pt = static_cast<int>(blockIdx.x) * 4 + w;
  Round1 cur, nxt;
  round1(pt, cur);
  for (; pt < np; pt += stride) {
    const int I = cur.I, Jy = cur.J, K = cur.K;
    const int cx = I - 1 + ox, cy = Jy - 1 + oy, cz = K - 1 + oz;
    const bool cell_ok = cx >= 0 && cx < p.nx && cy >= 0 && cy < p.ny && cz >= 0 && cz < p.nz;
    // second round of this node
    double ub[3], vb[3], ab[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ub[j] = p.x[cur.Lb + j];
      vb[j] = tr ? p.xdot[cur.Lb + j] : 0.0;
      ab[j] = tra ? p.xddot[cur.Lb + j] : 0.0;
    }
    const int32_t L0 = cur.L0;
    const int64_t base = p.rowptr[L0];
    const int len = min(static_cast<int>(p.rowptr[L0 + 1] - base), MAXLEN);
    const int node = L0 / 3;
    const uint16_t pos = p.nperm[static_cast<int64_t>(node) * p.nperm_w + slot];
    const int cell = cell_ok ? cx + p.nx * (cy + p.ny * cz) : 0;
    const double E = p.coef[cell];
    double2 G[5];
    const double2 *gp = reinterpret_cast<const double2 *>(p.geom + 10 * static_cast<int64_t>(cell));
#pragma unroll
    for (int t = 0; t < 5; ++t) G[t] = gp[t];
    // first round of the next node
    round1(pt + stride, nxt);
    __builtin_amdgcn_sched_barrier(0);
    // T = C S C^T / |det J| = Chat S Chat^T with the cell's scaled cofactors
    const double Ch[3][3] = {{G[0].x, G[0].y, G[1].x}, {G[1].y, G[2].x, G[2].y}, {G[3].x, G[3].y, G[4].x}};
    const double adet = G[4].y;
    double P[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int l = 0; l < 3; ++l) P[m][l] = S[m][0] * Ch[l][0] + S[m][1] * Ch[l][1] + S[m][2] * Ch[l][2];
    double T[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) T[k][l] = Ch[k][0] * P[0][l] + Ch[k][1] * P[1][l] + Ch[k][2] * P[2][l];
    const double lam = E * lamf, mu = E * muf;
    const double trs = T[0][0] + T[1][1] + T[2][2];
    const double Mab = tr ? p.rho * adet * mf : 0.0;
    double rr[3] = {0.0, 0.0, 0.0};
    double jb[9];
    // with C = a_R M + b_R K:  J = (beta + alpha b_R) K + (alpha a_R + gamma2) M  and
    // R = K (u + b_R v) + M (a_R v + a)
    double ut[3], mt[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      ut[j] = tr ? ub[j] + p.rb * vb[j] : ub[j];
      mt[j] = tr ? p.ra * vb[j] + (tra ? ab[j] : 0.0) : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double Kv = lam * T[i][j] + mu * T[j][i] + (i == j ? mu * trs : 0.0);
        rr[i] += Kv * ut[j];
        double jv = cK * Kv;
        if (i == j) {
          jv += cM * Mab;
          rr[i] += Mab * mt[j];
        }
        jb[3 * i + j] = jv;
      }
    // the cell's source load enters once (vertex b = a), F_a = src |det J|
    const bool ba = b == (1 - ox) + 2 * (1 - oy) + 4 * (1 - oz);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      // fixed-order sum over the wave: each half's total (DPP) ends in lanes 16 and 48
      const double v = group_sum<32>(cell_ok ? rr[i] - (ba ? p.src[i] * adet : 0.0) : 0.0);
      const double h0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16),
                                         __builtin_amdgcn_readlane(__double2loint(v), 16));
      const double h1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 48),
                                         __builtin_amdgcn_readlane(__double2loint(v), 48));
      rr[i] = h0 + h1;
    }
    // rows: zero, then the cells' blocks in cell order (every column of the rows is a neighbour
    // within one cell, so every entry is reached)
    for (int k = lane; k < 3 * len; k += 64) rows[w][k] = 0.0;
    const bool add = cell_ok && pos != 0xffffu && pos + 2 < len;
#pragma unroll
    for (int cc = 0; cc < 8; ++cc)
      if (c == cc && add)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) atomicAdd(&rows[w][i * len + pos + j], jb[3 * i + j]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (p.vals) row_store(rows[w], p.vals + base, 3 * len, lane);
    if (p.f && lane < 3) p.f[L0 + lane] = rr[lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    cur = nxt;
  }
}

template <int NB, int NQ, int BS, int MODE>
void launch(const ElasticParams &p, hipStream_t s) {
  int dev = 0, ncu = 0;
  TXA_HIP(hipGetDevice(&dev));
  TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // grid-stride over elements with exactly the resident workgroups: a workgroup that
  // only starts when another ends would stretch the tail by its whole share
  int per_cu = 0;
  TXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_elastic<NB, NQ, BS, MODE>, BS, 0));
  const int64_t want = static_cast<int64_t>(ncu) * std::max(per_cu, 1);
  const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(p.e1 - p.e0, want)));
  hipLaunchKernelGGL((k_elastic<NB, NQ, BS, MODE>), dim3(grid), dim3(BS), 0, s, p);
  TXA_HIP(hipGetLastError());
}

constexpr size_t kWsAlign = 256;

size_t ws_round(size_t b) { return (b + kWsAlign - 1) / kWsAlign * kWsAlign; }

// workspace bytes: residuals [E][ND] FP64 when f is wanted, blocks [E][ND][ND] when vals are
// affine HEX8 bricks take the node-centric kernel (no workspace) in every precision
bool affine_hex8(const Mesh &m) {
  return m.topo == TXA_HEX8 && m.affine && m.structured && m.nfields == 3 && m.npoints < (int64_t(1) << 31);
}

// Streams and events of the element / gather slab pipeline, one set per host thread and device
// (a context is driven by one thread at a time; the events are recorded and waited on within
// one call, so calls in sequence may reuse them).  Created on first use; kept until the
// process ends.
constexpr int kSlabs = 8;
struct SlabPipe {
  hipStream_t e = nullptr, g = nullptr;
  hipEvent_t ev[kSlabs] = {}, ev0 = nullptr, evg = nullptr;
};

SlabPipe &slab_pipe(int dev) {
  thread_local std::map<int, SlabPipe> pipes;
  SlabPipe &sp = pipes[dev];
  if (!sp.e) {
    TXA_HIP(hipStreamCreateWithFlags(&sp.g, hipStreamNonBlocking));
    for (int k = 0; k < kSlabs; ++k) TXA_HIP(hipEventCreateWithFlags(&sp.ev[k], hipEventDisableTiming));
    TXA_HIP(hipEventCreateWithFlags(&sp.ev0, hipEventDisableTiming));
    TXA_HIP(hipEventCreateWithFlags(&sp.evg, hipEventDisableTiming));
    TXA_HIP(hipStreamCreateWithFlags(&sp.e, hipStreamNonBlocking));   // last: marks the set complete
  }
  return sp;
}

void ws_parts(const Mesh &m, int precision, bool want_f, bool want_vals, size_t *rbytes, size_t *jbytes) {
  if (affine_hex8(m)) {
    *rbytes = *jbytes = 0;
    return;
  }
  const size_t nd = static_cast<size_t>(m.ndof_elem), ne = static_cast<size_t>(m.nelem);
  *rbytes = want_f ? ws_round(ne * nd * sizeof(double)) : 0;
  const size_t nb = nd / 3;
  *jbytes = want_vals ? ws_round(ne * nb * static_cast<size_t>(block_stride(static_cast<int>(nb))) *
                                 (precision == TXA_FP64 ? sizeof(double) : sizeof(float)))
                      : 0;
}

}  // namespace

bool elastic_supported(const Mesh &m, const txa_assembly_args &a) {
  if (a.physics != TXA_ELASTICITY || m.nfields != 3 || !(m.nperm.p || m.uocc.p)) return false;
  if (m.topo == TXA_HEX8) return a.precision == TXA_FP64 || a.precision == TXA_FP32;
  return m.topo == TXA_HEX27;
}

size_t elastic_workspace_bytes(const Mesh &m, int precision, bool want_f, bool want_vals) {
  size_t r = 0, j = 0;
  ws_parts(m, precision, want_f, want_vals, &r, &j);
  return r + j;
}

void node_perm_build(Mesh &m, hipStream_t s) {
  m.cell_geom.release();
  m.uocc_ptr.release();
  m.uocc.release();
  m.uslot.release();
  if (!m.structured && m.nfields == 3 && m.blocks.size() == 1 && (m.topo == TXA_HEX8 || m.topo == TXA_HEX27)) {
    // uploaded mesh: element instances of each node and their slot maps, for the general gather
    const int nb = m.nconn;
    const int64_t nnodes = m.ndofs / 3, npos0 = m.nelem * nb;
    if (npos0 >= (int64_t(1) << 31)) return;
    DevBuf<unsigned long long> cnt;
    cnt.alloc(nnodes + 1);
    TXA_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    const unsigned gp = static_cast<unsigned>((npos0 + 255) / 256), gn = static_cast<unsigned>((nnodes + 255) / 256);
    hipLaunchKernelGGL(k_uocc_count, dim3(gp), dim3(256), 0, s, npos0, nb, m.elem_lid.p, cnt.p);
    TXA_HIP(hipGetLastError());
    m.uocc_ptr.alloc(nnodes + 1);
    exclusive_scan_i64(reinterpret_cast<const int64_t *>(cnt.p), m.uocc_ptr.p, nnodes + 1, s);
    TXA_HIP(hipMemcpyAsync(cnt.p, m.uocc_ptr.p, sizeof(int64_t) * (nnodes + 1), hipMemcpyDeviceToDevice, s));
    m.uocc.alloc(npos0);
    hipLaunchKernelGGL(k_uocc_fill, dim3(gp), dim3(256), 0, s, npos0, nb, m.elem_lid.p, cnt.p, m.uocc.p);
    TXA_HIP(hipGetLastError());
    m.uslot.alloc(npos0 * nb);
    DevBuf<int> mx;
    mx.alloc(2);
    TXA_HIP(hipMemsetAsync(mx.p, 0, 2 * sizeof(int), s));
    hipLaunchKernelGGL(k_uocc_finish, dim3(gn), dim3(256), 0, s, nnodes, nb, m.uocc_ptr.p, m.uocc.p, m.elem_lid.p,
                       m.rowptr.p, m.col.p, m.uslot.p, mx.p);
    TXA_HIP(hipGetLastError());
    int h[2];
    TXA_HIP(hipMemcpyAsync(h, mx.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TXA_HIP(hipStreamSynchronize(s));
    const int S = nb == 8 ? 3 : 5;
    if (h[0] > 3 * S * S * S) {   // a row longer than the gather's LDS rows: the generic kernel
      m.uocc_ptr.release();
      m.uocc.release();
      m.uslot.release();
    }
    return;
  }
  if (!m.structured || m.nfields != 3 || (m.topo != TXA_HEX8 && m.topo != TXA_HEX27)) return;
  const int R = m.topo == TXA_HEX8 ? 1 : 2;
  const int64_t px = m.topo == TXA_HEX8 ? m.nx + 1 : 2 * static_cast<int64_t>(m.nx) + 1;
  const int64_t py = m.topo == TXA_HEX8 ? m.ny + 1 : 2 * static_cast<int64_t>(m.ny) + 1;
  const int64_t pz = m.topo == TXA_HEX8 ? m.nz + 1 : 2 * static_cast<int64_t>(m.nz) + 1;
  TXA_REQUIRE(px * py * pz == m.npoints, TXA_ERR_LOGIC, "lattice size mismatch");
  m.nperm_w = R == 1 ? 32 : 128;
  m.nperm.alloc(static_cast<size_t>(m.npoints) * m.nperm_w);
  const int64_t nt = m.npoints;
  hipLaunchKernelGGL(k_node_perm, dim3(static_cast<unsigned>((nt + 255) / 256)), dim3(256), 0, s, px, py, pz, R,
                     m.nfields, m.nperm_w, m.point_lid.p, m.rowptr.p, m.col.p, m.nperm.p);
  TXA_HIP(hipGetLastError());
  if (affine_hex8(m)) {
    const int64_t nc = static_cast<int64_t>(m.nx) * m.ny * m.nz;
    m.cell_geom.alloc(static_cast<size_t>(nc) * 10);
    hipLaunchKernelGGL(k_cell_geom, dim3(static_cast<unsigned>((nc + 255) / 256)), dim3(256), 0, s, m.nx, m.ny, m.nz,
                       m.nverts, m.xyz.p, m.cell_geom.p);
    TXA_HIP(hipGetLastError());
  }
}

void launch_elastic(Mesh &m, const txa_assembly_args &a, hipStream_t s, void *ws) {
  if (!a.f && !a.vals) return;
  if (affine_hex8(m)) {
    TXA_REQUIRE(m.cell_geom.p != nullptr, TXA_ERR_LOGIC, "cell geometry missing (graph_build)");
    Elastic8Params q;
    q.nx = m.nx; q.ny = m.ny; q.nz = m.nz;
    q.npoints = m.npoints;
    q.point_lid = m.point_lid.p;
    q.rowptr = m.rowptr.p;
    q.nperm = m.nperm.p;
    q.nperm_w = m.nperm_w;
    q.geom = m.cell_geom.p;
    q.coef = a.coef;
    q.nu = a.nu; q.rho = a.rho;
    q.src[0] = a.src[0]; q.src[1] = a.src[1]; q.src[2] = a.src[2];
    q.ra = a.rayleigh_a; q.rb = a.rayleigh_b;
    q.x = a.x; q.xdot = a.xdot; q.xddot = a.xddot;
    q.alpha = a.alpha; q.beta = a.beta; q.gamma2 = a.gamma2;
    q.transient = a.transient;
    q.f = a.f; q.vals = a.vals;
    int dev = 0, ncu = 0, per_cu = 0;
    TXA_HIP(hipGetDevice(&dev));
    TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TXA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_elastic8_affine, 256, 0));
    const int64_t want = static_cast<int64_t>(ncu) * std::max(per_cu, 1);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((m.npoints + 3) / 4, want)));
    hipLaunchKernelGGL(k_elastic8_affine, dim3(grid), dim3(256), 0, s, q);
    TXA_HIP(hipGetLastError());
    return;
  }
  size_t rbytes = 0, jbytes = 0;
  ws_parts(m, a.precision, a.f != nullptr, a.vals != nullptr, &rbytes, &jbytes);
  char *w = static_cast<char *>(ws);
  TXA_REQUIRE(w != nullptr, TXA_ERR_LOGIC, "elasticity assembly needs its workspace");
  ElasticParams p;
  p.nq = m.topo == TXA_HEX8 ? 8 : 27;
  p.nelem = m.nelem;
  p.e0 = 0;
  p.e1 = m.nelem;
  p.nconn = m.nconn;
  p.conn = m.conn.p;
  p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
  p.elem_lid = m.elem_lid.p;
  p.coef = a.coef;
  p.nu = a.nu; p.rho = a.rho;
  p.src[0] = a.src[0]; p.src[1] = a.src[1]; p.src[2] = a.src[2];
  p.ra = a.rayleigh_a; p.rb = a.rayleigh_b;
  p.x = a.x; p.xdot = a.xdot; p.xddot = a.xddot;
  p.alpha = a.alpha; p.beta = a.beta; p.gamma2 = a.gamma2;
  p.transient = a.transient;
  p.rbuf = rbytes ? reinterpret_cast<double *>(w) : nullptr;
  p.jbuf = jbytes ? static_cast<void *>(w + rbytes) : nullptr;
  auto elem_launch = [&](const ElasticParams &pp, hipStream_t st) {
    if (m.topo == TXA_HEX8) {
      if (a.precision == TXA_FP32) launch<8, 8, 64, MODE_F32>(pp, st);
      else launch<8, 8, 64, MODE_F64>(pp, st);
    } else {
      if (a.precision == TXA_SPLIT_BF16) launch<27, 27, 256, MODE_SPLIT>(pp, st);
      else if (a.precision == TXA_FP32) launch<27, 27, 256, MODE_F32>(pp, st);
      else launch<27, 27, 256, MODE_F64>(pp, st);
    }
  };
  if (!m.structured) elem_launch(p, s);   // bricks: element slabs interleaved with the gather below
  int dev = 0, ncu = 0;
  TXA_HIP(hipGetDevice(&dev));
  TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const bool wide = a.precision == TXA_FP64;
  if (!m.structured) {
    GatherUParams u;
    u.nnodes = m.ndofs / 3;
    u.occ_ptr = m.uocc_ptr.p;
    u.occ = m.uocc.p;
    u.slot = m.uslot.p;
    u.rowptr = m.rowptr.p;
    u.jbuf = p.jbuf;
    u.rbuf = p.rbuf;
    u.f = a.f;
    u.vals = a.vals;
    const unsigned ug = static_cast<unsigned>(std::min<int64_t>((u.nnodes + 3) / 4, static_cast<int64_t>(ncu) * 16));
    if (m.topo == TXA_HEX8) {
      if (wide) hipLaunchKernelGGL((k_elastic_gather_u<8, double>), dim3(ug), dim3(256), 0, s, u);
      else hipLaunchKernelGGL((k_elastic_gather_u<8, float>), dim3(ug), dim3(256), 0, s, u);
    } else {
      if (wide) hipLaunchKernelGGL((k_elastic_gather_u<27, double>), dim3(ug), dim3(256), 0, s, u);
      else hipLaunchKernelGGL((k_elastic_gather_u<27, float>), dim3(ug), dim3(256), 0, s, u);
    }
    TXA_HIP(hipGetLastError());
    return;
  }
  GatherParams g;
  g.nx = m.nx; g.ny = m.ny; g.nz = m.nz;
  const int64_t st = m.topo == TXA_HEX8 ? 1 : 2;
  g.px = st * m.nx + 1; g.py = st * m.ny + 1; g.pz = st * m.nz + 1;
  g.npoints = m.npoints;
  g.p0 = 0;
  g.p1 = m.npoints;
  g.point_lid = m.point_lid.p;
  g.rowptr = m.rowptr.p;
  g.nperm = m.nperm.p;
  g.nperm_w = m.nperm_w;
  g.jbuf = p.jbuf;
  g.rbuf = p.rbuf;
  g.f = a.f;
  g.vals = a.vals;
  auto gather_launch = [&](const GatherParams &gg, hipStream_t st) {
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>((gg.p1 - gg.p0 + 3) / 4, static_cast<int64_t>(ncu) * 16));
    if (m.topo == TXA_HEX8) {
      if (wide) hipLaunchKernelGGL((k_elastic_gather<8, double>), dim3(grid), dim3(256), 0, st, gg);
      else hipLaunchKernelGGL((k_elastic_gather<8, float>), dim3(grid), dim3(256), 0, st, gg);
    } else {
      if (wide) hipLaunchKernelGGL((k_elastic_gather<27, double>), dim3(grid), dim3(256), 0, st, gg);
      else hipLaunchKernelGGL((k_elastic_gather<27, float>), dim3(grid), dim3(256), 0, st, gg);
    }
    TXA_HIP(hipGetLastError());
  };
  // Slab pipeline: the elements in z slabs on one stream, the gather of each slab's node planes
  // on another once that slab (and the one below) is done, so the element kernel (bound by
  // its per-element latency and VALU) of slab k+1 runs beside the gather (bound by HBM) of
  // slab k.  Node plane st*z1 at a slab's top needs the next slab: the next gather takes it.
  // A stream under graph capture (or a brick one slab high) takes the two kernels in order
  // on that stream: no streams or events are created inside a capture.
  const int nslab = std::min(kSlabs, m.nz);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  TXA_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone || nslab < 2) {
    elem_launch(p, s);
    gather_launch(g, s);
    return;
  }
  SlabPipe &sp = slab_pipe(dev);
  const int64_t layer = static_cast<int64_t>(m.nx) * m.ny, pxy = g.px * g.py;
  TXA_HIP(hipEventRecord(sp.ev0, s));
  TXA_HIP(hipStreamWaitEvent(sp.e, sp.ev0, 0));
  TXA_HIP(hipStreamWaitEvent(sp.g, sp.ev0, 0));
  for (int k = 0; k < nslab; ++k) {
    const int64_t z0 = static_cast<int64_t>(m.nz) * k / nslab, z1 = static_cast<int64_t>(m.nz) * (k + 1) / nslab;
    ElasticParams pk = p;
    pk.e0 = z0 * layer;
    pk.e1 = z1 * layer;
    elem_launch(pk, sp.e);
    TXA_HIP(hipEventRecord(sp.ev[k], sp.e));
    TXA_HIP(hipStreamWaitEvent(sp.g, sp.ev[k], 0));
    GatherParams gk = g;
    gk.p0 = st * z0 * pxy;
    gk.p1 = k == nslab - 1 ? m.npoints : st * z1 * pxy;
    gather_launch(gk, sp.g);
  }
  TXA_HIP(hipEventRecord(sp.evg, sp.g));
  TXA_HIP(hipStreamWaitEvent(s, sp.evg, 0));
}

}  // namespace txa
