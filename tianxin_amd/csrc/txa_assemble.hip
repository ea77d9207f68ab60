// Residual + Jacobian assembly kernels for gfx950.
//
// Replaces, per workset, the reference chain
//   GatherSolution_Tpetra -> DOF / DOFGradient -> Integrator_GradBasisDotVector,
//   Integrator_BasisTimesScalar -> ScatterResidual_Tpetra<Jacobian>
// (disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-650,
//  Panzer_DOF_Functors.hpp:156-279, Panzer_DOFGradient_impl.hpp:60-160,
//  Panzer_Integrator_GradBasisDotVector_impl.hpp:225-294,
//  Panzer_Integrator_BasisTimesScalar_impl.hpp:215-270,
//  Panzer_ScatterResidual_Tpetra_impl.hpp:375-520)
// with geometry recomputed in registers (IntegrationValues2 / BasisValues2,
// disc-fe/src/Panzer_IntegrationValues2.cpp:949-1221,
// disc-fe/src/Panzer_BasisValues2_impl.hpp:1379-1460) and the Sacado-FAD
// Jacobian replaced by the explicit tangent beta*K + alpha*C + gamma2*M.
//
// Two kernel families:
//  * k_q1_hex8_diffusion: hex8 scalar diffusion on the structured brick.
//    Row-owner, z-marching: a workgroup owns a 31x7 patch of vertex columns,
//    computes the 32x8 element columns around it one z-layer at a time (one
//    element per thread), hands each element's rows to the owners of its
//    vertices through LDS, and every owner writes its complete CRS row once
//    with plain stores -- no atomics, no zero-fill, deterministic sums.
//  * k_generic: any topology / physics, one workgroup per element, FP64 atomic
//    scatter into a zeroed CRS (the unfused baseline and the coverage path).
#include "txa_internal.hpp"
#include "txa_refelem.hpp"

namespace txa {
namespace {

// ============================================================================
// Fast path: hex8 diffusion, row-owner z-marching
// ============================================================================
constexpr int TX = 32, TY = 8, NT = TX * TY;
constexpr double kG = 0.57735026918962576451;  // 1/sqrt(3)

struct Q1DiffParams {
  int nx, ny, nz;
  const double *X, *Y, *Z;     // SoA vertex coordinates
  const int32_t *lid;          // LID of each vertex (one field)
  const double *coef;          // k per element
  const double *x, *xdot;      // by LID
  double alpha, beta, rhoc, src;
  double *f, *vals;
  const int64_t *rowptr;
  const uint8_t *perm;         // [rows][32] stencil slot of the 27 neighbours, 255 = absent
};

__device__ __forceinline__ constexpr int sym8(int a, int b) {
  return a <= b ? a * 8 - (a * (a - 1)) / 2 + (b - a) : b * 8 - (b * (b - 1)) / 2 + (a - b);
}
__device__ __forceinline__ constexpr double sgnbit(int v, int d) { return ((v >> d) & 1) ? 1.0 : -1.0; }

// Element stiffness (and mass when TRANS) of a trilinear hex in binary vertex
// order v = bx + 2 by + 4 bz, 2x2x2 Gauss.  Geometry (J, cofactors, det) in
// FP64; the basis-gradient transform and the contraction in Real.
// Outputs: J_e = beta K (+ alpha M) packed symmetric, R_e = K u (+ M udot) - F.
template <typename Real, bool TRANS>
__device__ __forceinline__ void hex8_diffusion_element(const double (&Xc)[3][8], const double (&u)[8],
                                                       const double (&ud)[8], double k, double rhoc, double src,
                                                       double alpha, double beta, Real (&Je)[36], Real (&Re)[8]) {
  // x(xi) = c0 + c1 xi + c2 eta + c3 zeta + c4 xi eta + c5 eta zeta + c6 xi zeta + c7 xi eta zeta
  double c[8][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int m = 1; m < 8; ++m) c[m][d] = 0.0;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const double x = Xc[d][v];
      const double sx = sgnbit(v, 0), sy = sgnbit(v, 1), sz = sgnbit(v, 2);
      c[1][d] += sx * x; c[2][d] += sy * x; c[3][d] += sz * x;
      c[4][d] += sx * sy * x; c[5][d] += sy * sz * x; c[6][d] += sx * sz * x; c[7][d] += sx * sy * sz * x;
    }
#pragma unroll
    for (int m = 1; m < 8; ++m) c[m][d] *= 0.125;
  }
  Real K[36];
  Real M[TRANS ? 36 : 1];
  double F[8];
#pragma unroll
  for (int i = 0; i < 36; ++i) { K[i] = Real(0); if (TRANS) M[i] = Real(0); }
#pragma unroll
  for (int v = 0; v < 8; ++v) F[v] = 0.0;
  double sgn = 1.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double xi = (q & 1) ? kG : -kG, eta = (q & 2) ? kG : -kG, zeta = (q & 4) ? kG : -kG;
    double J[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      J[d][0] = c[1][d] + c[4][d] * eta + c[6][d] * zeta + c[7][d] * (eta * zeta);
      J[d][1] = c[2][d] + c[4][d] * xi + c[5][d] * zeta + c[7][d] * (xi * zeta);
      J[d][2] = c[3][d] + c[5][d] * eta + c[6][d] * xi + c[7][d] * (xi * eta);
    }
    // cofactor matrix: (J^-1)_{ji} = cof_{ij} / det
    double cof[3][3];
    cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    cof[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    cof[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    cof[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    cof[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
    if (q == 0) sgn = det < 0.0 ? -1.0 : 1.0;   // computeCellMeasure sign rule
    const double wm = det * sgn;                 // Gauss weights are 1
    const double idet = 1.0 / det;
    const Real cw = Real(k * wm * idet * idet);  // grad N = cof . dN / det
    Real cr[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) cr[i][j] = Real(cof[i][j]);
    Real g[8][3];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const double sx = sgnbit(v, 0), sy = sgnbit(v, 1), sz = sgnbit(v, 2);
      const Real dx = Real(0.125 * sx * (1.0 + sy * eta) * (1.0 + sz * zeta));
      const Real dy = Real(0.125 * sy * (1.0 + sx * xi) * (1.0 + sz * zeta));
      const Real dz = Real(0.125 * sz * (1.0 + sx * xi) * (1.0 + sy * eta));
#pragma unroll
      for (int i = 0; i < 3; ++i) g[v][i] = cr[i][0] * dx + cr[i][1] * dy + cr[i][2] * dz;
      const double Nv = 0.125 * (1.0 + sx * xi) * (1.0 + sy * eta) * (1.0 + sz * zeta);
      F[v] += (wm * src) * Nv;
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      Real ha[3] = {cw * g[a][0], cw * g[a][1], cw * g[a][2]};
#pragma unroll
      for (int b = a; b < 8; ++b) K[sym8(a, b)] += ha[0] * g[b][0] + ha[1] * g[b][1] + ha[2] * g[b][2];
    }
    if (TRANS) {
      const Real cm = Real(rhoc * wm);
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double Na = 0.125 * (1.0 + sgnbit(a, 0) * xi) * (1.0 + sgnbit(a, 1) * eta) * (1.0 + sgnbit(a, 2) * zeta);
#pragma unroll
        for (int b = a; b < 8; ++b) {
          const double Nb =
              0.125 * (1.0 + sgnbit(b, 0) * xi) * (1.0 + sgnbit(b, 1) * eta) * (1.0 + sgnbit(b, 2) * zeta);
          M[sym8(a, b)] += cm * Real(Na * Nb);
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    Real r = Real(0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      r += K[sym8(a, b)] * Real(u[b]);
      if (TRANS) r += M[sym8(a, b)] * Real(ud[b]);
    }
    Re[a] = Real(double(r) - F[a]);
  }
#pragma unroll
  for (int i = 0; i < 36; ++i) Je[i] = TRANS ? Real(beta) * K[i] + Real(alpha) * M[i] : Real(beta) * K[i];
}

// Stage rows a_lo = cx + 2cy and a_hi = a_lo + 4 (binary order) of one element.
template <typename Real>
__device__ __forceinline__ void stage_corner(Real *s, int slot0, int t, const Real (&Je)[36], const Real (&Re)[8],
                                             int alo) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    s[(slot0 + b) * NT + t] = Je[sym8(alo, b)];
    s[(slot0 + 8 + b) * NT + t] = Je[sym8(alo + 4, b)];
  }
  s[(slot0 + 16) * NT + t] = Re[alo];
  s[(slot0 + 17) * NT + t] = Re[alo + 4];
}

// Add the rows of the element at corner (cx, cy) of the owner's vertex column.
template <int CX, int CY, typename Get>
__device__ __forceinline__ void take_corner(double (&acc)[27], double (&accn)[18], double &rc, double &rn, Get get) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
    const int s = (bx - CX + 1) + 3 * (by - CY + 1);
    acc[s + 9 * (bz + 1)] += double(get(b));
    accn[s + 9 * bz] += double(get(8 + b));
  }
  rc += double(get(16));
  rn += double(get(17));
}

__device__ __forceinline__ void write_row(const Q1DiffParams &p, int64_t node, const double (&acc)[27], double r) {
  const int32_t L = p.lid[node];
  if (p.f) p.f[L] = r;
  if (p.vals) {
    const int64_t base = p.rowptr[L];
    const uint4 *pp = reinterpret_cast<const uint4 *>(p.perm + static_cast<int64_t>(L) * 32);
    const uint4 w0 = pp[0], w1 = pp[1];
    const uint32_t w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      const uint32_t o = (w[s >> 2] >> (8 * (s & 3))) & 0xffu;
      if (o != 0xffu) p.vals[base + o] = acc[s];
    }
  }
}

template <typename Real, bool TRANS>
__global__ __launch_bounds__(NT) void k_q1_hex8_diffusion(Q1DiffParams p) {
  __shared__ Real stage[36 * NT];
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int x0 = -1 + static_cast<int>(blockIdx.x) * (TX - 1);
  const int y0 = -1 + static_cast<int>(blockIdx.y) * (TY - 1);
  const int ex = x0 + tx, ey = y0 + ty;
  const bool evalid = ex >= 0 && ex < p.nx && ey >= 0 && ey < p.ny;
  const int oi = ex + 1, oj = ey + 1;  // owned vertex column: corner (1,1) of this thread's element
  const bool owner = tx < TX - 1 && ty < TY - 1 && oi <= p.nx && oj <= p.ny;
  const int64_t sx = p.nx + 1, sxy = static_cast<int64_t>(p.nx + 1) * (p.ny + 1);

  double acc[27], accn[18];
  double rc = 0.0, rn = 0.0;
#pragma unroll
  for (int s = 0; s < 27; ++s) acc[s] = 0.0;
#pragma unroll
  for (int s = 0; s < 18; ++s) accn[s] = 0.0;

  for (int ez = 0; ez < p.nz; ++ez) {
    Real Je[36], Re[8];
    if (evalid) {
      double Xc[3][8], u[8], ud[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int64_t n = (ex + (v & 1)) + sx * (ey + ((v >> 1) & 1)) + sxy * (ez + ((v >> 2) & 1));
        Xc[0][v] = p.X[n]; Xc[1][v] = p.Y[n]; Xc[2][v] = p.Z[n];
        const int32_t L = p.lid[n];
        u[v] = p.x[L];
        ud[v] = TRANS ? p.xdot[L] : 0.0;
      }
      const double k = p.coef[ex + static_cast<int64_t>(p.nx) * (ey + static_cast<int64_t>(p.ny) * ez)];
      hex8_diffusion_element<Real, TRANS>(Xc, u, ud, k, p.rhoc, p.src, p.alpha, p.beta, Je, Re);
    } else {
#pragma unroll
      for (int i = 0; i < 36; ++i) Je[i] = Real(0);
#pragma unroll
      for (int i = 0; i < 8; ++i) Re[i] = Real(0);
    }
    // round A: corners (0,1) [owner tx-1] and (1,0) [owner ty-1]
    stage_corner(stage, 0, t, Je, Re, 2);
    stage_corner(stage, 18, t, Je, Re, 1);
    __syncthreads();
    if (owner) {
      take_corner<0, 1>(acc, accn, rc, rn, [&](int v) { return stage[v * NT + t + 1]; });
      take_corner<1, 0>(acc, accn, rc, rn, [&](int v) { return stage[(18 + v) * NT + t + TX]; });
    }
    __syncthreads();
    // round B: corner (0,0) [owner (tx-1, ty-1)]
    stage_corner(stage, 0, t, Je, Re, 0);
    __syncthreads();
    if (owner) {
      take_corner<0, 0>(acc, accn, rc, rn, [&](int v) { return stage[v * NT + t + TX + 1]; });
      // own element: corner (1,1)
      take_corner<1, 1>(acc, accn, rc, rn, [&](int v) -> Real {
        if (v < 8) return Je[sym8(3, v)];
        if (v < 16) return Je[sym8(7, v - 8)];
        return v == 16 ? Re[3] : Re[7];
      });
      write_row(p, oi + sx * oj + sxy * ez, acc, rc);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 18; ++s) { acc[s] = accn[s]; accn[s] = 0.0; }
#pragma unroll
    for (int s = 18; s < 27; ++s) acc[s] = 0.0;
    rc = rn;
    rn = 0.0;
  }
  if (owner) write_row(p, oi + sx * oj + sxy * p.nz, acc, rc);
}

// ============================================================================
// Generic path: one workgroup per element, atomic scatter
// ============================================================================
constexpr int GB = 256;

struct GenericParams {
  int topo, physics, nb, nq, nv, nf, ndof, nconn;
  int64_t nelem;
  const int64_t *conn;             // first nv entries of an element are its vertices
  const double *X, *Y, *Z;
  const int32_t *elem_lid;
  const double *tab;               // N[nb][nq] | dN[nb][nq][3] | gdN[nv][nq][3] | qw[nq]
  const double *coef;
  double nu, rho, src[3], ra, rb;
  const double *x, *xdot, *xddot;
  double alpha, beta, gamma2;
  int transient;
  double *f, *vals;
  const int64_t *rowptr;
  const int32_t *col;
};

__device__ __forceinline__ int64_t find_slot(const int64_t *rowptr, const int32_t *col, int32_t row, int32_t key) {
  int64_t lo = rowptr[row], hi = rowptr[row + 1];
  while (lo < hi) {
    int64_t m = (lo + hi) >> 1;
    if (col[m] < key) lo = m + 1; else hi = m;
  }
  return lo;
}

__global__ __launch_bounds__(GB) void k_generic(GenericParams p) {
  __shared__ double sX[8][3], sDet[kMaxNQ], sJi[kMaxNQ][9], sWm[kMaxNQ];
  __shared__ double sG[kMaxNB * kMaxNQ * 3];
  __shared__ double sU[81], sV[81], sA[81], sR[81];
  const double *tN = p.tab;
  const double *tdN = tN + p.nb * p.nq;
  const double *tgdN = tdN + p.nb * p.nq * 3;
  const double *tqw = tgdN + p.nv * p.nq * 3;
  const int t = threadIdx.x;
  const bool elastic = p.physics == TXA_ELASTICITY;
  const bool tr = p.transient && p.xdot;
  for (int64_t e = blockIdx.x; e < p.nelem; e += gridDim.x) {
    const int32_t *lids = p.elem_lid + e * p.ndof;
    if (t < p.nv * 3) {
      const int v = t / 3, d = t % 3;
      const int64_t n = p.conn[e * p.nconn + v];
      sX[v][d] = d == 0 ? p.X[n] : (d == 1 ? p.Y[n] : p.Z[n]);
    }
    if (t < p.ndof) {
      const int32_t L = lids[t];
      sU[t] = p.x[L];
      sV[t] = tr ? p.xdot[L] : 0.0;
      sA[t] = (tr && p.xddot) ? p.xddot[L] : 0.0;
      sR[t] = 0.0;
    }
    __syncthreads();
    if (t < p.nq) {
      double J[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int v = 0; v < p.nv; ++v) s += sX[v][i] * tgdN[(v * p.nq + t) * 3 + j];
          J[i][j] = s;
        }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      double *Ji = sJi[t];
      Ji[0] = c00 * id;
      Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[3] = c01 * id;
      Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[6] = c02 * id;
      Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      sDet[t] = det;
    }
    __syncthreads();
    if (t < p.nq) sWm[t] = tqw[t] * sDet[t] * (sDet[0] < 0.0 ? -1.0 : 1.0);
    for (int idx = t; idx < p.nb * p.nq; idx += GB) {
      const int b = idx / p.nq, q = idx % p.nq;
      const double *Ji = sJi[q];
      const double *dn = tdN + (b * p.nq + q) * 3;
      for (int i = 0; i < 3; ++i) sG[idx * 3 + i] = Ji[0 * 3 + i] * dn[0] + Ji[1 * 3 + i] * dn[1] + Ji[2 * 3 + i] * dn[2];
    }
    __syncthreads();
    const double cf = p.coef[e];
    const double lam = elastic ? cf * p.nu / ((1.0 + p.nu) * (1.0 - 2.0 * p.nu)) : 0.0;
    const double mu = elastic ? cf / (2.0 * (1.0 + p.nu)) : 0.0;
    for (int idx = t; idx < p.ndof * p.ndof; idx += GB) {
      const int I = idx / p.ndof, Jd = idx % p.ndof;
      const int a = I / p.nf, i = I % p.nf, b = Jd / p.nf, j = Jd % p.nf;
      double Kv = 0.0, Mv = 0.0;
      for (int q = 0; q < p.nq; ++q) {
        const double *ga = &sG[(a * p.nq + q) * 3];
        const double *gb = &sG[(b * p.nq + q) * 3];
        const double w = sWm[q];
        const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        const double nn = tN[a * p.nq + q] * tN[b * p.nq + q];
        if (!elastic) {
          Kv += w * cf * gg;
          Mv += w * p.rho * nn;
        } else {
          Kv += w * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * gg : 0.0));
          if (i == j) Mv += w * p.rho * nn;
        }
      }
      const double Cv = elastic ? p.ra * Mv + p.rb * Kv : Mv;
      double rcontrib = Kv * sU[Jd];
      double jv = p.beta * Kv;
      if (tr) {
        rcontrib += Cv * sV[Jd];
        jv += p.alpha * Cv;
        if (elastic && p.xddot) { rcontrib += Mv * sA[Jd]; jv += p.gamma2 * Mv; }
      }
      atomicAdd(&sR[I], rcontrib);
      if (p.vals) {
        const int64_t slot = find_slot(p.rowptr, p.col, lids[I], lids[Jd]);
        if (slot < p.rowptr[lids[I] + 1] && p.col[slot] == lids[Jd]) atomicAdd(p.vals + slot, jv);
      }
    }
    __syncthreads();
    if (t < p.ndof && p.f) {
      const int a = t / p.nf, i = t % p.nf;
      double F = 0.0;
      for (int q = 0; q < p.nq; ++q) F += sWm[q] * tN[a * p.nq + q] * p.src[i];
      atomicAdd(p.f + lids[t], sR[t] - F);
    }
    __syncthreads();
  }
}

// ============================================================================
// Dirichlet rows: f_i = x_i - g_i; matrix row -> identity row
// (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:228-237, 300-318)
// ============================================================================
__global__ void k_dirichlet(int64_t n, int64_t nrows, const int32_t *__restrict__ lids, const double *__restrict__ g,
                            const double *__restrict__ x, const int64_t *__restrict__ rowptr,
                            const int32_t *__restrict__ col, double *__restrict__ f, double *__restrict__ vals) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= n) return;
  const int32_t r = lids[i];
  if (r < 0 || r >= nrows) return;  // LIDs outside the map are ignored
  if (lane == 0 && f) f[r] = x[r] - (g ? g[i] : 0.0);
  if (vals)
    for (int64_t k = rowptr[r] + lane; k < rowptr[r + 1]; k += 64) vals[k] = col[k] == r ? 1.0 : 0.0;
}

}  // namespace

void upload_refelem(Mesh &m, DevBuf<double> &tab, hipStream_t s) {
  RefElem r;
  refelem_build(r, m.topo);
  std::vector<double> h;
  for (int b = 0; b < r.nb; ++b) for (int q = 0; q < r.nq; ++q) h.push_back(r.N[b][q]);
  for (int b = 0; b < r.nb; ++b) for (int q = 0; q < r.nq; ++q) for (int d = 0; d < 3; ++d) h.push_back(r.dN[b][q][d]);
  for (int v = 0; v < r.nv; ++v) for (int q = 0; q < r.nq; ++q) for (int d = 0; d < 3; ++d) h.push_back(r.gdN[v][q][d]);
  for (int q = 0; q < r.nq; ++q) h.push_back(r.qw[q]);
  tab.alloc(h.size());
  TXA_HIP(hipMemcpyAsync(tab.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  TXA_HIP(hipStreamSynchronize(s));
}

void assemble(Mesh &m, const txa_assembly_args &a, hipStream_t s) {
  TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "assemble before graph_build");
  TXA_REQUIRE(a.x != nullptr, TXA_ERR_INVALID, "x is required");
  TXA_REQUIRE(a.coef != nullptr, TXA_ERR_INVALID, "coef is required");
  TXA_REQUIRE(a.physics == TXA_DIFFUSION || a.physics == TXA_ELASTICITY, TXA_ERR_INVALID, "unknown physics");
  TXA_REQUIRE(a.precision == TXA_FP64 || a.precision == TXA_FP32 || a.precision == TXA_SPLIT_BF16,
              TXA_ERR_INVALID, "unknown precision");
  const int want_nf = a.physics == TXA_ELASTICITY ? 3 : 1;
  TXA_REQUIRE(m.nfields == want_nf, TXA_ERR_INVALID,
              "DOF map has " + std::to_string(m.nfields) + " fields, physics needs " + std::to_string(want_nf));
  TXA_REQUIRE(!(a.transient && a.xdot == nullptr), TXA_ERR_INVALID, "transient assembly needs xdot");
  TXA_REQUIRE(a.n_bc == 0 || a.bc_lids != nullptr, TXA_ERR_INVALID, "bc_lids missing");
  const bool trans = a.transient != 0;

  if (m.topo == TXA_HEX8 && a.physics == TXA_DIFFUSION && a.precision != TXA_SPLIT_BF16) {
    Q1DiffParams p;
    p.nx = m.nx; p.ny = m.ny; p.nz = m.nz;
    p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
    p.lid = m.point_lid.p;
    p.coef = a.coef; p.x = a.x; p.xdot = a.xdot;
    p.alpha = a.alpha; p.beta = a.beta; p.rhoc = a.rho; p.src = a.src[0];
    p.f = a.f; p.vals = a.vals;
    p.rowptr = m.rowptr.p; p.perm = m.perm.p;
    dim3 grid((m.nx + 1 + TX - 2) / (TX - 1), (m.ny + 1 + TY - 2) / (TY - 1));
    if (a.precision == TXA_FP64) {
      if (trans) hipLaunchKernelGGL((k_q1_hex8_diffusion<double, true>), grid, dim3(NT), 0, s, p);
      else hipLaunchKernelGGL((k_q1_hex8_diffusion<double, false>), grid, dim3(NT), 0, s, p);
    } else {
      if (trans) hipLaunchKernelGGL((k_q1_hex8_diffusion<float, true>), grid, dim3(NT), 0, s, p);
      else hipLaunchKernelGGL((k_q1_hex8_diffusion<float, false>), grid, dim3(NT), 0, s, p);
    }
    TXA_HIP(hipGetLastError());
  } else {
    TXA_REQUIRE(a.precision == TXA_FP64, TXA_ERR_UNSUPPORTED,
                "reduced-precision integration is implemented for hex8 diffusion only");
    TXA_REQUIRE(m.ref.p != nullptr, TXA_ERR_LOGIC, "reference tables missing");
    RefElem r;
    refelem_build(r, m.topo);
    GenericParams p;
    p.topo = m.topo; p.physics = a.physics; p.nb = r.nb; p.nq = r.nq; p.nv = r.nv; p.nf = m.nfields;
    p.ndof = m.ndof_elem; p.nconn = m.nconn; p.nelem = m.nelem; p.conn = m.conn.p;
    p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
    p.elem_lid = m.elem_lid.p; p.tab = m.ref.p; p.coef = a.coef;
    p.nu = a.nu; p.rho = a.rho; p.src[0] = a.src[0]; p.src[1] = a.src[1]; p.src[2] = a.src[2];
    p.ra = a.rayleigh_a; p.rb = a.rayleigh_b;
    p.x = a.x; p.xdot = a.xdot; p.xddot = a.xddot;
    p.alpha = a.alpha; p.beta = a.beta; p.gamma2 = a.gamma2; p.transient = a.transient;
    p.f = a.f; p.vals = a.vals; p.rowptr = m.rowptr.p; p.col = m.col.p;
    if (a.f) TXA_HIP(hipMemsetAsync(a.f, 0, sizeof(double) * m.ndofs, s));
    if (a.vals) TXA_HIP(hipMemsetAsync(a.vals, 0, sizeof(double) * m.nnz, s));
    int64_t blocks = m.nelem < 65536 * 8 ? m.nelem : 65536 * 8;
    hipLaunchKernelGGL(k_generic, dim3(static_cast<unsigned>(blocks)), dim3(GB), 0, s, p);
    TXA_HIP(hipGetLastError());
  }
  if (a.n_bc > 0) {
    const int64_t per = 4;  // rows per 256-thread block
    hipLaunchKernelGGL(k_dirichlet, dim3(static_cast<unsigned>((a.n_bc + per - 1) / per)), dim3(256), 0, s, a.n_bc,
                       m.ndofs, a.bc_lids, a.bc_vals, a.x, m.rowptr.p, m.col.p, a.f, a.vals);
    TXA_HIP(hipGetLastError());
  }
}

}  // namespace txa
