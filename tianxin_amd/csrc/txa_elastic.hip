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
// One workgroup per element (grid-stride).  Geometry (J, det J, J^-1) is FP64 at
// every quadrature point.  The element stiffness is assembled from the
// gradient Gram matrix
//   T[(k,a),(l,b)] = sum_q w_q |det J_q| dN_a/dx_k dN_b/dx_l        (3 NB x 3 NB),
//   K[(a,i),(b,j)] = lambda T[(i,a),(j,b)] + mu T[(j,a),(i,b)] + delta_ij mu sum_k T[(k,a),(k,b)],
// i.e. B^T D B for isotropic D written as one symmetric product G^T W G.  T is
// computed in FP64 or FP32 on the VALU, or with bf16 MFMA on a three-way split
// of each FP32 operand (x = x0 + x1 + x2, every x_s in bf16) keeping the six
// cross products of order <= 2^-16, accumulated in FP32.  The scatter adds FP64
// values into the CRS with hardware FP64 atomics; the row position of every
// (node, neighbour) pair comes from a per-node table built with the graph.
#include <type_traits>

#include "txa_internal.hpp"

namespace txa {
namespace {

constexpr int MODE_F64 = 0, MODE_F32 = 1, MODE_SPLIT = 2;

struct ElasticParams {
  int nq;
  int64_t nelem;
  int nconn;
  const int64_t *conn;          // the first 8 entries of an element are its vertices
  const double *X, *Y, *Z;
  const int32_t *elem_lid;      // [E][3 NB]: node-major, field-minor
  const uint16_t *nperm;        // [nodes][nperm_w]: position of (neighbour, field 0) in the node's rows
  int nperm_w;
  const double *tab;            // N[nb][nq] | dN[nb][nq][3] | gdN[8][nq][3] | qw[nq]
  const double *coef;           // Young's modulus per element
  double nu, rho, src[3], ra, rb;
  const double *x, *xdot, *xddot;
  double alpha, beta, gamma2;
  int transient;
  double *f, *vals;
  const int64_t *rowptr;
};

// lattice offsets of the element's nodes: shards vertex order (hex8), Q2 order (hex27)
__device__ const int8_t kHexV[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__device__ const int8_t kQ2[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2},
    {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
    {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

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

template <int NB>
struct Dims {
  static constexpr int ND = 3 * NB;                    // element DOFs
  static constexpr int R = NB == 8 ? 1 : 2;            // lattice reach of a node's neighbours
  static constexpr int S = 2 * R + 1;
  static constexpr int MP = (ND + 15) / 16 * 16;       // padded operand rows (MFMA)
  static constexpr int KP = 32;                        // padded quadrature dimension (MFMA)
};

template <int NB, int NQ, int MODE>
struct Smem {
  using D = Dims<NB>;
  using TReal = typename std::conditional<MODE == MODE_F64, double, float>::type;
  double X[8][3];
  double Ji[NQ][9];
  double det[NQ];
  double wm[NQ];
  double u[D::ND], v[D::ND], a[D::ND], r[D::ND];
  int32_t L[D::ND];
  int64_t row[D::ND];
  uint16_t pos[NB][NB];
  TReal M[NB][NB];
  TReal T[D::ND][D::ND + 1];
  // gradients (VALU modes) or the bf16 operand splits (MFMA mode)
  typename std::conditional<MODE == MODE_SPLIT, short, TReal>::type
      G[MODE == MODE_SPLIT ? 2 * 3 * D::MP * D::KP : NQ * 3 * NB];
};

template <int NB, int NQ, int BS, int MODE>
__global__ __launch_bounds__(BS) void k_elastic(ElasticParams p) {
  using D = Dims<NB>;
  using TReal = typename Smem<NB, NQ, MODE>::TReal;
  constexpr int ND = D::ND;
  __shared__ Smem<NB, NQ, MODE> sm;
  const double *tN = p.tab;
  const double *tdN = tN + NB * NQ;
  const double *tgdN = tdN + NB * NQ * 3;
  const double *tqw = tgdN + 8 * NQ * 3;
  const int t = threadIdx.x;
  const bool tr = p.transient && p.xdot;
  const bool tra = tr && p.xddot;

  if constexpr (MODE == MODE_SPLIT) {
    // zero the operand padding once: rows >= ND and quadrature columns >= NQ stay zero
    for (int i = t; i < 2 * 3 * D::MP * D::KP; i += BS) sm.G[i] = 0;
  }

  for (int64_t e = blockIdx.x; e < p.nelem; e += gridDim.x) {
    // ---- gather: vertices, LIDs, row starts, state
    if (t < 24) {
      const int vtx = t / 3, d = t % 3;
      const int64_t n = p.conn[e * p.nconn + vtx];
      sm.X[vtx][d] = d == 0 ? p.X[n] : (d == 1 ? p.Y[n] : p.Z[n]);
    }
    for (int i = t; i < ND; i += BS) {
      const int32_t L = p.elem_lid[e * ND + i];
      sm.L[i] = L;
      sm.row[i] = p.rowptr[L];
      sm.u[i] = p.x[L];
      sm.v[i] = tr ? p.xdot[L] : 0.0;
      sm.a[i] = tra ? p.xddot[L] : 0.0;
      sm.r[i] = 0.0;
    }
    __syncthreads();
    // ---- geometry (FP64) and the row positions of the element's node pairs
    if (t < NQ) {
      double J[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int vtx = 0; vtx < 8; ++vtx) s += sm.X[vtx][i] * tgdN[(vtx * NQ + t) * 3 + j];
          J[i][j] = s;
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
    }
    for (int i = t; i < NB * NB; i += BS) {
      const int a = i / NB, b = i % NB;
      const int8_t(*off)[3] = NB == 8 ? kHexV : kQ2;
      const int s = (off[b][0] - off[a][0] + D::R) + D::S * (off[b][1] - off[a][1] + D::R) +
                    D::S * D::S * (off[b][2] - off[a][2] + D::R);
      sm.pos[a][b] = p.nperm[static_cast<int64_t>(sm.L[3 * a] / 3) * p.nperm_w + s];
    }
    __syncthreads();
    if (t < NQ) sm.wm[t] = tqw[t] * sm.det[t] * (sm.det[0] < 0.0 ? -1.0 : 1.0);
    __syncthreads();
    // ---- physical gradients at the quadrature points, mass matrix
    for (int i = t; i < NQ * NB; i += BS) {
      const int q = i / NB, a = i % NB;
      const double *Ji = sm.Ji[q];
      const double *dn = tdN + (a * NQ + q) * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double g = Ji[0 * 3 + k] * dn[0] + Ji[1 * 3 + k] * dn[1] + Ji[2 * 3 + k] * dn[2];
        if constexpr (MODE == MODE_SPLIT) {
          // A operand [m = (k,a)][q] = g; B operand [n = (k,a)][q] = w_q g; three bf16 parts each
          short s0, s1, s2;
          const int m = k * NB + a;
          short *A = sm.G, *B = sm.G + 3 * D::MP * D::KP;
          split3(static_cast<float>(g), s0, s1, s2);
          A[(0 * D::MP + m) * D::KP + q] = s0;
          A[(1 * D::MP + m) * D::KP + q] = s1;
          A[(2 * D::MP + m) * D::KP + q] = s2;
          split3(static_cast<float>(sm.wm[q] * g), s0, s1, s2);
          B[(0 * D::MP + m) * D::KP + q] = s0;
          B[(1 * D::MP + m) * D::KP + q] = s1;
          B[(2 * D::MP + m) * D::KP + q] = s2;
        } else {
          sm.G[(q * 3 + k) * NB + a] = static_cast<TReal>(g);
        }
      }
    }
    for (int i = t; i < NB * NB; i += BS) {
      const int a = i / NB, b = i % NB;
      double s = 0.0;
      for (int q = 0; q < NQ; ++q) s += sm.wm[q] * tN[a * NQ + q] * tN[b * NQ + q];
      sm.M[a][b] = static_cast<TReal>(s);
    }
    __syncthreads();
    // ---- gradient Gram matrix T
    if constexpr (MODE == MODE_SPLIT) {
      constexpr int NT = D::MP / 16;                 // 16x16 tiles per side
      constexpr int NTILE = NT * (NT + 1) / 2;       // upper triangle
      const int lane = t & 63, wave = t >> 6;
      const short *A = sm.G, *B = sm.G + 3 * D::MP * D::KP;
      for (int ti = wave; ti < NTILE; ti += BS / 64) {
        int tm = 0, rem = ti;
        while (rem >= NT - tm) { rem -= NT - tm; ++tm; }
        const int tn = tm + rem;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < D::KP / 16; ++ks) {
          const int k0 = ks * 16 + 4 * (lane >> 4);
          const int ma = tm * 16 + (lane & 15), nb = tn * 16 + (lane & 15);
          v4s a_[3], b_[3];
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            a_[s] = *reinterpret_cast<const v4s *>(A + (s * D::MP + ma) * D::KP + k0);
            b_[s] = *reinterpret_cast<const v4s *>(B + (s * D::MP + nb) * D::KP + k0);
          }
          // products of order 1, 2^-8 and 2^-16
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[2], b_[0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[1], b_[1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[0], b_[2], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[1], b_[0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[0], b_[1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_[0], b_[0], acc, 0, 0, 0);
        }
        // D[4 (lane/16) + i][lane % 16] of the tile; off-diagonal tiles also fill the
        // mirrored tile (the split products are not exactly symmetric, so a diagonal
        // tile keeps its own entries on both sides)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = tm * 16 + 4 * (lane >> 4) + i, cc = tn * 16 + (lane & 15);
          if (rr < ND && cc < ND) {
            sm.T[rr][cc] = acc[i];
            if (tm != tn) sm.T[cc][rr] = acc[i];
          }
        }
      }
    } else {
      for (int i = t; i < ND * ND; i += BS) {
        const int rr = i / ND, cc = i % ND;
        const int kr = rr / NB, ar = rr % NB, kc = cc / NB, ac = cc % NB;
        TReal s = 0;
        for (int q = 0; q < NQ; ++q)
          s += static_cast<TReal>(sm.wm[q]) * sm.G[(q * 3 + kr) * NB + ar] * sm.G[(q * 3 + kc) * NB + ac];
        sm.T[rr][cc] = s;
      }
    }
    __syncthreads();
    // ---- element matrix entries: residual and CRS scatter
    const double E = p.coef[e];
    const double lam = E * p.nu / ((1.0 + p.nu) * (1.0 - 2.0 * p.nu));
    const double mu = E / (2.0 * (1.0 + p.nu));
    for (int idx = t; idx < ND * ND; idx += BS) {
      const int P = idx / ND, Q = idx % ND;
      const int a = P / 3, i = P % 3, b = Q / 3, j = Q % 3;
      double Kv = lam * static_cast<double>(sm.T[i * NB + a][j * NB + b]) +
                  mu * static_cast<double>(sm.T[j * NB + a][i * NB + b]);
      double Mv = 0.0;
      if (i == j) {
        Kv += mu * (static_cast<double>(sm.T[a][b]) + static_cast<double>(sm.T[NB + a][NB + b]) +
                    static_cast<double>(sm.T[2 * NB + a][2 * NB + b]));
        Mv = p.rho * static_cast<double>(sm.M[a][b]);
      }
      double rc = Kv * sm.u[Q];
      double jv = p.beta * Kv;
      if (tr) {
        const double Cv = p.ra * Mv + p.rb * Kv;
        rc += Cv * sm.v[Q];
        jv += p.alpha * Cv;
        if (tra) {
          rc += Mv * sm.a[Q];
          jv += p.gamma2 * Mv;
        }
      }
      unsafeAtomicAdd(&sm.r[P], rc);
      if (p.vals) {
        const uint16_t ps = sm.pos[a][b];
        if (ps != 0xffffu) unsafeAtomicAdd(p.vals + sm.row[P] + ps + j, jv);
      }
    }
    __syncthreads();
    if (p.f)
      for (int i = t; i < ND; i += BS) {
        const int a = i / 3, k = i % 3;
        double F = 0.0;
        for (int q = 0; q < NQ; ++q) F += sm.wm[q] * tN[a * NQ + q];
        unsafeAtomicAdd(p.f + sm.L[i], sm.r[i] - F * p.src[k]);
      }
    __syncthreads();
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

template <int NB, int NQ, int BS, int MODE>
void launch(const ElasticParams &p, hipStream_t s) {
  int dev = 0, ncu = 0;
  TXA_HIP(hipGetDevice(&dev));
  TXA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t want = static_cast<int64_t>(ncu) * (BS == 64 ? 16 : 4);
  const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(p.nelem, want)));
  hipLaunchKernelGGL((k_elastic<NB, NQ, BS, MODE>), dim3(grid), dim3(BS), 0, s, p);
  TXA_HIP(hipGetLastError());
}

}  // namespace

bool elastic_supported(const Mesh &m, const txa_assembly_args &a) {
  if (a.physics != TXA_ELASTICITY || m.nfields != 3 || !m.nperm.p) return false;
  if (m.topo == TXA_HEX8) return a.precision == TXA_FP64 || a.precision == TXA_FP32;
  return m.topo == TXA_HEX27;
}

void node_perm_build(Mesh &m, hipStream_t s) {
  if (m.nfields != 3 || (m.topo != TXA_HEX8 && m.topo != TXA_HEX27)) return;
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
}

void launch_elastic(Mesh &m, const txa_assembly_args &a, hipStream_t s) {
  ElasticParams p;
  p.nq = m.topo == TXA_HEX8 ? 8 : 27;
  p.nelem = m.nelem;
  p.nconn = m.nconn;
  p.conn = m.conn.p;
  p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
  p.elem_lid = m.elem_lid.p;
  p.nperm = m.nperm.p;
  p.nperm_w = m.nperm_w;
  p.tab = m.ref.p;
  p.coef = a.coef;
  p.nu = a.nu; p.rho = a.rho;
  p.src[0] = a.src[0]; p.src[1] = a.src[1]; p.src[2] = a.src[2];
  p.ra = a.rayleigh_a; p.rb = a.rayleigh_b;
  p.x = a.x; p.xdot = a.xdot; p.xddot = a.xddot;
  p.alpha = a.alpha; p.beta = a.beta; p.gamma2 = a.gamma2;
  p.transient = a.transient;
  p.f = a.f; p.vals = a.vals; p.rowptr = m.rowptr.p;
  if (a.f) TXA_HIP(hipMemsetAsync(a.f, 0, sizeof(double) * m.ndofs, s));
  if (a.vals) TXA_HIP(hipMemsetAsync(a.vals, 0, sizeof(double) * m.nnz, s));
  if (m.topo == TXA_HEX8) {
    if (a.precision == TXA_FP32) launch<8, 8, 64, MODE_F32>(p, s);
    else launch<8, 8, 64, MODE_F64>(p, s);
  } else {
    if (a.precision == TXA_SPLIT_BF16) launch<27, 27, 256, MODE_SPLIT>(p, s);
    else if (a.precision == TXA_FP32) launch<27, 27, 256, MODE_F32>(p, s);
    else launch<27, 27, 256, MODE_F64>(p, s);
  }
}

}  // namespace txa
