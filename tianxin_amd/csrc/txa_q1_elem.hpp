// Q1 hex8 element routines shared by the lattice kernels (txa_q1_hex8.hip) and the
// uploaded-mesh patch kernels (txa_patch.hip): one trilinear cell with 2x2x2 Gauss, in
// the binary vertex order v = bx + 2by + 4bz.
//
// Element stiffness by sum factorisation.  At the 2x2x2 Gauss points the
// reference gradient of vertex a factors into 1D terms,
//   dN_a/dxi = (s_ax / 2) P(a_y, q_y) P(a_z, q_z),   P(bit, qbit) = (1 +- 1/sqrt3)/2,
// so K_ab = sum_q sum_jk A_q[jk] D_j(a,q) D_k(b,q) with the metric tensor
// A_q = k w |det J| J^-1 J^-T = k sgn / det * cof^T cof (cof = cofactor matrix of J)
// contracts axis by axis: three diagonal tables S_jj (3x3 classes each) and
// three cross tables X_jk (12 entries each) built from the 8x6 metric values,
// then each of the 36 symmetric entries is a signed sum of 9 table values.
// About 650 operations instead of ~1600 for the direct product.
//
// Geometry (trilinear-map coefficients, J, cofactors, det, 1/det) is FP64 in
// every precision; Real is the precision of the metric tensor and the
// contraction (the Jacobian); row sums over elements are FP64.
//
// Residual.  The FP64 form forms R = K u - F from its FP64 element matrix.  The
// FP32 form computes R in FP64 without the matrix, as the reference's
// DOFGradient -> GradBasisDotVector chain does: the reference gradient of u at
// each point from the Walsh coefficients of u (u(xi) = sum_s u^_s prod_{d in s} xi_d),
// the flux G_q grad u with the FP64 metric, and its transposed contraction back
// to the vertices in the same basis; the FP32 element matrix never touches R, so
// R keeps FP64 accuracy entry by entry (cancelling rows included).
#pragma once

#include <hip/hip_runtime.h>

namespace txa {
namespace q1elem {

constexpr double kG = 0.57735026918962576451;   // 1/sqrt(3)
constexpr double kP = 0.5 * (1.0 + kG);         // P(bit == qbit)
constexpr double kM = 0.5 * (1.0 - kG);         // P(bit != qbit)

__device__ __forceinline__ constexpr int sym8(int a, int b) {
  return a <= b ? a * 8 - (a * (a - 1)) / 2 + (b - a) : b * 8 - (b * (b - 1)) / 2 + (a - b);
}
__device__ __forceinline__ constexpr int bit(int v, int d) { return (v >> d) & 1; }
__device__ __forceinline__ constexpr double sgn(int v, int d) { return bit(v, d) ? 1.0 : -1.0; }
__device__ __forceinline__ constexpr double Pq(int b, int qb) { return b == qb ? kP : kM; }
// pair class of two bits: 0 = (0,0), 1 = (1,1), 2 = mixed
__device__ __forceinline__ constexpr int cls(int a, int b) { return a == b ? a : 2; }
__device__ __forceinline__ constexpr int cls_lo(int c) { return c == 1 ? 1 : 0; }
__device__ __forceinline__ constexpr int cls_hi(int c) { return c == 0 ? 0 : 1; }
__device__ __forceinline__ constexpr double Wq(int c, int qb) { return Pq(cls_lo(c), qb) * Pq(cls_hi(c), qb); }

// 1/x: hardware estimate refined by two Newton steps (correctly rounded to within an ulp
// for the normal range; det of a valid element is far from the denormal range).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  return r;
}

template <typename Real, bool TRANS>
struct Hex8Out {
  Real K[36];     // beta K + alpha M, packed symmetric, binary vertex order
  double R[8];    // K u + M udot - F, FP64 in every precision
};

// Whether the residual is formed without the element matrix (FP32 form: FP64 flux).
template <typename Real>
constexpr bool flux_residual() { return sizeof(Real) == 4; }

// In-place Walsh butterfly over the three vertex bits: v[s] <- sum_a prod_{d in s} sgn(a,d) v[a].
__device__ __forceinline__ void walsh8(double (&v)[8]) {
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (!bit(i, b)) {
        const double lo = v[i], hi = v[i | (1 << b)];
        v[i] = lo + hi;
        v[i | (1 << b)] = hi - lo;
      }
}

// Its transpose: v[a] <- sum_s prod_{d in s} sgn(a,d) v[s].
__device__ __forceinline__ void walsh8_t(double (&v)[8]) {
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (!bit(i, b)) {
        const double lo = v[i], hi = v[i | (1 << b)];
        v[i] = lo - hi;
        v[i | (1 << b)] = lo + hi;
      }
}

// Trilinear interpolation to the 2x2x2 Gauss points and its transpose (the matrix
// P(a_d, q_d) is symmetric): v[q] <- sum_a prod_d P(a_d, q_d) v[a].
__device__ __forceinline__ void interp8(double (&v)[8]) {
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (!bit(i, b)) {
        const double lo = v[i], hi = v[i | (1 << b)];
        v[i] = kP * lo + kM * hi;
        v[i | (1 << b)] = kM * lo + kP * hi;
      }
}

// One element.  X: vertex coordinates in binary order v = bx + 2by + 4bz.  mid() runs
// once the quadrature loop is done (the caller issues its prefetch loads there, where
// the geometry's registers are free).
template <typename Real, bool TRANS, bool UNIT, bool AFF, class Mid>
__device__ __forceinline__ void hex8_element(const double (&X)[3][8], const double (&u)[8], const double (&ud)[8],
                                             double k, double rhoc, double src, double alpha, double beta,
                                             Hex8Out<Real, TRANS> &o, Mid &&mid) {
  constexpr bool kFlux = flux_residual<Real>();
  // FP64 flux residual: Walsh coefficients of u (times 1/8; the xi-linear terms of the
  // gradient prescaled by the point coordinate 1/sqrt3, the bilinear one by 1/3) and the
  // accumulators rh[s] = sum_q sum_j d(phi_s)/d(xi_j)(q) z_j(q) of the flux z = A_q grad u
  double uh[8], rh[8];
  if constexpr (kFlux) {
#pragma unroll
    for (int v = 0; v < 8; ++v) { uh[v] = u[v]; rh[v] = 0.0; }
    walsh8(uh);
#pragma unroll
    for (int v = 1; v < 8; ++v) uh[v] *= 0.125;
    uh[3] *= kG; uh[5] *= kG; uh[6] *= kG; uh[7] *= kG * kG;
  }
  // Walsh transform of the vertex coordinates: h[s] = sum_v prod_{d in s} sgn(v,d) X_v,
  // the coefficients (times 8) of 1, xi, eta, xi eta, zeta, xi zeta, eta zeta, xi eta zeta
  double h[3][8];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int v = 0; v < 8; ++v) h[d][v] = X[d][v];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int v = 0; v < 8; ++v)
        if (!bit(v, b)) {
          const double lo = h[d][v], hi = h[d][v | (1 << b)];
          h[d][v] = lo + hi;
          h[d][v | (1 << b)] = hi - lo;
        }
  }
  // 1/8 of the trilinear map folded into the coefficients
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int v = 1; v < 8; ++v) h[d][v] *= 0.125;
  // Tables of the element matrix (sum-factorised over the points) and the point weights
  Real S[3][3][3];                                  // diagonal: S_dd[c1][c2]
  Real Xxy[2][2][3], Xyz[3][2][2], Xxz[2][3][2];    // cross terms
  double wm[8];
  // AFF: every cell is a parallelepiped (the bilinear and trilinear coefficients of the map
  // zero to within the rounding of its coordinates; cell_affine decides it once per mesh),
  // so there is one Jacobian and the metric, the tables and the flux residual take closed
  // forms: the point sums of the general form reduce to the weights w(0) = w(1) = 2/3,
  // w(2) = 1/3 of a pair class and to sum_q s_d(q) s_e(q) = 8 delta_de.
  if constexpr (AFF) {
    double c[3][3];   // cofactors of J = [h1 h2 h4]
    c[0][0] = h[1][2] * h[2][4] - h[1][4] * h[2][2];
    c[0][1] = h[1][4] * h[2][1] - h[1][1] * h[2][4];
    c[0][2] = h[1][1] * h[2][2] - h[1][2] * h[2][1];
    c[1][0] = h[0][4] * h[2][2] - h[0][2] * h[2][4];
    c[1][1] = h[0][1] * h[2][4] - h[0][4] * h[2][1];
    c[1][2] = h[0][2] * h[2][1] - h[0][1] * h[2][2];
    c[2][0] = h[0][2] * h[1][4] - h[0][4] * h[1][2];
    c[2][1] = h[0][4] * h[1][1] - h[0][1] * h[1][4];
    c[2][2] = h[0][1] * h[1][2] - h[0][2] * h[1][1];
    const double det = h[0][1] * c[0][0] + h[0][2] * c[0][1] + h[0][4] * c[0][2];
    const double sg = det < 0.0 ? -1.0 : 1.0;
    const double cw = 0.25 * k * sg * recip(det);
    const double mxx = cw * (c[0][0] * c[0][0] + c[1][0] * c[1][0] + c[2][0] * c[2][0]);
    const double myy = cw * (c[0][1] * c[0][1] + c[1][1] * c[1][1] + c[2][1] * c[2][1]);
    const double mzz = cw * (c[0][2] * c[0][2] + c[1][2] * c[1][2] + c[2][2] * c[2][2]);
    const double mxy = cw * (c[0][0] * c[0][1] + c[1][0] * c[1][1] + c[2][0] * c[2][1]);
    const double myz = cw * (c[0][1] * c[0][2] + c[1][1] * c[1][2] + c[2][1] * c[2][2]);
    const double mxz = cw * (c[0][0] * c[0][2] + c[1][0] * c[1][2] + c[2][0] * c[2][2]);
    if (!kFlux || TRANS)
#pragma unroll
      for (int q = 0; q < 8; ++q) wm[q] = det * sg;
    if constexpr (kFlux) {
      rh[0] = -2.0 * src * (det * sg);
      rh[1] = 8.0 * (mxx * uh[1] + mxy * uh[2] + mxz * uh[4]);
      rh[2] = 8.0 * (mxy * uh[1] + myy * uh[2] + myz * uh[4]);
      rh[4] = 8.0 * (mxz * uh[1] + myz * uh[2] + mzz * uh[4]);
      rh[3] = 8.0 * ((mxx + myy) * uh[3] + mxz * uh[6] + myz * uh[5]);
      rh[5] = 8.0 * ((mxx + mzz) * uh[5] + mxy * uh[6] + myz * uh[3]);
      rh[6] = 8.0 * ((myy + mzz) * uh[6] + mxy * uh[5] + mxz * uh[3]);
      rh[7] = 8.0 * (mxx + myy + mzz) * uh[7];
    }
    mid();
    const Real a[6] = {Real(mxx), Real(myy), Real(mzz), Real(mxy), Real(myz), Real(mxz)};
    constexpr double w[3] = {2.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0};
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
        for (int c2 = 0; c2 < 3; ++c2) S[d][c1][c2] = Real(2.0 * w[c1] * w[c2]) * a[d];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          Xxy[i][j][c3] = Real(w[c3]) * a[3];
          Xyz[c3][i][j] = Real(w[c3]) * a[4];
          Xxz[i][c3][j] = Real(w[c3]) * a[5];
        }
  } else {
    // metric tensor at the 8 points, scale 1/4 of the gradient products folded in
    Real A[8][6];     // xx yy zz xy yz xz
    double sg = 1.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double xi = bit(q, 0) ? kG : -kG, eta = bit(q, 1) ? kG : -kG, zeta = bit(q, 2) ? kG : -kG;
      double J[3][3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        J[d][0] = h[d][1] + h[d][3] * eta + h[d][5] * zeta + h[d][7] * (eta * zeta);
        J[d][1] = h[d][2] + h[d][3] * xi + h[d][6] * zeta + h[d][7] * (xi * zeta);
        J[d][2] = h[d][4] + h[d][5] * xi + h[d][6] * eta + h[d][7] * (xi * eta);
      }
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
      if (q == 0) sg = det < 0.0 ? -1.0 : 1.0;   // computeCellMeasure: sign of det at the first point
      if (!kFlux || TRANS) wm[q] = det * sg;     // Gauss weights are 1
      const double cw = 0.25 * k * sg * recip(det);   // k wm / det^2, times the 1/4 of (s/2)(s/2)
      // the metric in FP64, rounded once to Real (six conversions per point instead of ten)
      const double mxx = cw * (c[0][0] * c[0][0] + c[1][0] * c[1][0] + c[2][0] * c[2][0]);
      const double myy = cw * (c[0][1] * c[0][1] + c[1][1] * c[1][1] + c[2][1] * c[2][1]);
      const double mzz = cw * (c[0][2] * c[0][2] + c[1][2] * c[1][2] + c[2][2] * c[2][2]);
      const double mxy = cw * (c[0][0] * c[0][1] + c[1][0] * c[1][1] + c[2][0] * c[2][1]);
      const double myz = cw * (c[0][1] * c[0][2] + c[1][1] * c[1][2] + c[2][1] * c[2][2]);
      const double mxz = cw * (c[0][0] * c[0][2] + c[1][0] * c[1][2] + c[2][0] * c[2][2]);
      A[q][0] = Real(mxx); A[q][1] = Real(myy); A[q][2] = Real(mzz);
      A[q][3] = Real(mxy); A[q][4] = Real(myz); A[q][5] = Real(mxz);
      if constexpr (kFlux) {
        // reference gradient of u at the point, the flux z = A_q grad u (A_q = G_q / 4), and its
        // contraction with the gradients of the Walsh basis functions
        const double s0 = bit(q, 0) ? 1.0 : -1.0, s1 = bit(q, 1) ? 1.0 : -1.0, s2 = bit(q, 2) ? 1.0 : -1.0;
        const double g0 = uh[1] + s1 * uh[3] + s2 * uh[5] + (s1 * s2) * uh[7];
        const double g1 = uh[2] + s0 * uh[3] + s2 * uh[6] + (s0 * s2) * uh[7];
        const double g2 = uh[4] + s0 * uh[5] + s1 * uh[6] + (s0 * s1) * uh[7];
        const double z0 = mxx * g0 + mxy * g1 + mxz * g2;
        const double z1 = mxy * g0 + myy * g1 + myz * g2;
        const double z2 = mxz * g0 + myz * g1 + mzz * g2;
        rh[1] += z0;
        rh[2] += z1;
        rh[4] += z2;
        rh[3] += s1 * z0 + s0 * z1;
        rh[5] += s2 * z0 + s0 * z2;
        rh[6] += s2 * z1 + s1 * z2;
        rh[7] += (s1 * s2) * z0 + (s0 * s2) * z1 + (s0 * s1) * z2;
        // source: -F_a = -src sum_q wm_q N_a(q) enters as -(src/4) wm_q phi_s(q), with
        // phi_s(q) = 3^(-|s|/2) prod sgn, one power of 1/sqrt3 fewer than the later scaling supplies
        const double ws = -0.25 * src * (det * sg), wg = ws * kG;
        rh[0] += ws;
        rh[1] += s0 * wg;
        rh[2] += s1 * wg;
        rh[4] += s2 * wg;
        rh[3] += (s0 * s1) * wg;
        rh[5] += (s0 * s2) * wg;
        rh[6] += (s1 * s2) * wg;
        rh[7] += (s0 * s1 * s2) * wg;
      }
    }
    mid();
    // ---- diagonal tables S_dd[c1][c2]: sum over the two other axes after summing axis d
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;   // the two other axes, ascending
      Real Ab[2][2];
#pragma unroll
      for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
        for (int i2 = 0; i2 < 2; ++i2) {
          const int q0 = (i1 << d1) | (i2 << d2);
          Ab[i1][i2] = A[q0][d] + A[q0 | (1 << d)][d];
        }
      Real T[2][3];
#pragma unroll
      for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
        for (int c2 = 0; c2 < 3; ++c2) T[i1][c2] = Real(Wq(c2, 0)) * Ab[i1][0] + Real(Wq(c2, 1)) * Ab[i1][1];
#pragma unroll
      for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
        for (int c2 = 0; c2 < 3; ++c2) S[d][c1][c2] = Real(Wq(c1, 0)) * T[0][c2] + Real(Wq(c1, 1)) * T[1][c2];
    }
    // ---- cross tables
    // Xxy[al][be][cz] = sum_q A_xy P(al,qx) P(be,qy) W(cz,qz)
    // Xyz[cx][be][ga] = sum_q A_yz W(cx,qx) P(be,qy) P(ga,qz)
    // Xxz[al][cy][ga] = sum_q A_xz P(al,qx) W(cy,qy) P(ga,qz)
    {
      Real U[2][2][2];   // [al][qy][qz]
#pragma unroll
      for (int al = 0; al < 2; ++al)
#pragma unroll
        for (int qy = 0; qy < 2; ++qy)
#pragma unroll
          for (int qz = 0; qz < 2; ++qz)
            U[al][qy][qz] = Real(Pq(al, 0)) * A[0 | (qy << 1) | (qz << 2)][3] +
                            Real(Pq(al, 1)) * A[1 | (qy << 1) | (qz << 2)][3];
#pragma unroll
      for (int al = 0; al < 2; ++al)
#pragma unroll
        for (int be = 0; be < 2; ++be) {
          const Real V0 = Real(Pq(be, 0)) * U[al][0][0] + Real(Pq(be, 1)) * U[al][1][0];
          const Real V1 = Real(Pq(be, 0)) * U[al][0][1] + Real(Pq(be, 1)) * U[al][1][1];
#pragma unroll
          for (int c = 0; c < 3; ++c) Xxy[al][be][c] = Real(Wq(c, 0)) * V0 + Real(Wq(c, 1)) * V1;
        }
    }
    {
      Real U[2][2][2];   // [ga][qx][qy]: sum over qz with P(ga, qz)
#pragma unroll
      for (int ga = 0; ga < 2; ++ga)
#pragma unroll
        for (int qx = 0; qx < 2; ++qx)
#pragma unroll
          for (int qy = 0; qy < 2; ++qy)
            U[ga][qx][qy] = Real(Pq(ga, 0)) * A[qx | (qy << 1)][4] + Real(Pq(ga, 1)) * A[qx | (qy << 1) | 4][4];
#pragma unroll
      for (int ga = 0; ga < 2; ++ga)
#pragma unroll
        for (int be = 0; be < 2; ++be) {
          const Real V0 = Real(Pq(be, 0)) * U[ga][0][0] + Real(Pq(be, 1)) * U[ga][0][1];   // qx = 0
          const Real V1 = Real(Pq(be, 0)) * U[ga][1][0] + Real(Pq(be, 1)) * U[ga][1][1];   // qx = 1
#pragma unroll
          for (int c = 0; c < 3; ++c) Xyz[c][be][ga] = Real(Wq(c, 0)) * V0 + Real(Wq(c, 1)) * V1;
        }
    }
    {
      Real U[2][2][2];   // [al][qy][qz]: sum over qx with P(al, qx)
#pragma unroll
      for (int al = 0; al < 2; ++al)
#pragma unroll
        for (int qy = 0; qy < 2; ++qy)
#pragma unroll
          for (int qz = 0; qz < 2; ++qz)
            U[al][qy][qz] = Real(Pq(al, 0)) * A[0 | (qy << 1) | (qz << 2)][5] +
                            Real(Pq(al, 1)) * A[1 | (qy << 1) | (qz << 2)][5];
#pragma unroll
      for (int al = 0; al < 2; ++al)
#pragma unroll
        for (int ga = 0; ga < 2; ++ga) {
          const Real V0 = Real(Pq(ga, 0)) * U[al][0][0] + Real(Pq(ga, 1)) * U[al][0][1];   // qy = 0
          const Real V1 = Real(Pq(ga, 0)) * U[al][1][0] + Real(Pq(ga, 1)) * U[al][1][1];   // qy = 1
#pragma unroll
          for (int c = 0; c < 3; ++c) Xxz[al][c][ga] = Real(Wq(c, 0)) * V0 + Real(Wq(c, 1)) * V1;
        }
    }
  }
  // ---- entries
  Real K[36];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = a; b < 8; ++b) {
      const int ax = bit(a, 0), ay = bit(a, 1), az = bit(a, 2);
      const int bx = bit(b, 0), by = bit(b, 1), bz = bit(b, 2);
      const int cx = cls(ax, bx), cy = cls(ay, by), cz = cls(az, bz);
      Real v = Real(sgn(a, 0) * sgn(b, 0)) * S[0][cy][cz] + Real(sgn(a, 1) * sgn(b, 1)) * S[1][cx][cz] +
               Real(sgn(a, 2) * sgn(b, 2)) * S[2][cx][cy];
      v += Real(sgn(a, 0) * sgn(b, 1)) * Xxy[bx][ay][cz] + Real(sgn(a, 1) * sgn(b, 0)) * Xxy[ax][by][cz];
      v += Real(sgn(a, 1) * sgn(b, 2)) * Xyz[cx][by][az] + Real(sgn(a, 2) * sgn(b, 1)) * Xyz[cx][ay][bz];
      v += Real(sgn(a, 0) * sgn(b, 2)) * Xxz[bx][cy][az] + Real(sgn(a, 2) * sgn(b, 0)) * Xxz[ax][cy][bz];
      K[sym8(a, b)] = v;
    }
  // ---- source load F_a = src sum_q wm_q N_a(q) and mass classes (FP64), axis by axis
  double F[8];
  if constexpr (!kFlux) {
    double t1[2][2][2];   // [ax][qy][qz]
#pragma unroll
    for (int ax = 0; ax < 2; ++ax)
#pragma unroll
      for (int qy = 0; qy < 2; ++qy)
#pragma unroll
        for (int qz = 0; qz < 2; ++qz)
          t1[ax][qy][qz] = Pq(ax, 0) * wm[(qy << 1) | (qz << 2)] + Pq(ax, 1) * wm[1 | (qy << 1) | (qz << 2)];
#pragma unroll
    for (int ax = 0; ax < 2; ++ax)
#pragma unroll
      for (int ay = 0; ay < 2; ++ay) {
        const double t20 = Pq(ay, 0) * t1[ax][0][0] + Pq(ay, 1) * t1[ax][1][0];
        const double t21 = Pq(ay, 0) * t1[ax][0][1] + Pq(ay, 1) * t1[ax][1][1];
#pragma unroll
        for (int az = 0; az < 2; ++az) F[ax | (ay << 1) | (az << 2)] = src * (Pq(az, 0) * t20 + Pq(az, 1) * t21);
      }
  }
  Real M[TRANS ? 27 : 1];   // M_ab = rho_c sum_q wm N_a N_b, indexed by the three pair classes
  if (TRANS) {
    double m1[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int qy = 0; qy < 2; ++qy)
#pragma unroll
        for (int qz = 0; qz < 2; ++qz)
          m1[c][qy][qz] = Wq(c, 0) * wm[(qy << 1) | (qz << 2)] + Wq(c, 1) * wm[1 | (qy << 1) | (qz << 2)];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int cy = 0; cy < 3; ++cy) {
        const double m20 = Wq(cy, 0) * m1[c][0][0] + Wq(cy, 1) * m1[c][1][0];
        const double m21 = Wq(cy, 0) * m1[c][0][1] + Wq(cy, 1) * m1[c][1][1];
#pragma unroll
        for (int cz = 0; cz < 3; ++cz) M[c + 3 * (cy + 3 * cz)] = Real(rhoc * (Wq(cz, 0) * m20 + Wq(cz, 1) * m21));
      }
  }
  // ---- residual and tangent
  if constexpr (kFlux) {
    // R_a = 4 (1/8) sum_s w_s(a) rh_s (A = G/4, N_a = 1/8 sum_s w_s(a) phi_s), source included;
    // then M udot
    rh[3] *= kG; rh[5] *= kG; rh[6] *= kG; rh[7] *= kG * kG;
    walsh8_t(rh);
    double mr[8];
    if constexpr (TRANS) {
      // M udot = rho_c sum_q wm_q N_a(q) udot(q): udot at the points, weighted, back to the vertices
#pragma unroll
      for (int v = 0; v < 8; ++v) mr[v] = ud[v];
      interp8(mr);
#pragma unroll
      for (int q = 0; q < 8; ++q) mr[q] *= rhoc * wm[q];
      interp8(mr);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) o.R[a] = TRANS ? 0.5 * rh[a] + mr[a] : 0.5 * rh[a];
  } else {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      Real r = Real(0);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        r += K[sym8(a, b)] * Real(u[b]);
        if (TRANS) {
          const int mc = cls(bit(a, 0), bit(b, 0)) + 3 * (cls(bit(a, 1), bit(b, 1)) + 3 * cls(bit(a, 2), bit(b, 2)));
          r += M[mc] * Real(ud[b]);
        }
      }
      o.R[a] = double(r) - F[a];
    }
  }
  if constexpr (!TRANS && UNIT) {   // the steady Jacobian dR/du with beta = 1: K as it is
#pragma unroll
    for (int i = 0; i < 36; ++i) o.K[i] = K[i];
    return;
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = a; b < 8; ++b) {
      Real j = Real(beta) * K[sym8(a, b)];
      if (TRANS) {
        const int mc = cls(bit(a, 0), bit(b, 0)) + 3 * (cls(bit(a, 1), bit(b, 1)) + 3 * cls(bit(a, 2), bit(b, 2)));
        j += Real(alpha) * M[mc];
      }
      o.K[sym8(a, b)] = j;
    }
}

}  // namespace q1elem
}  // namespace txa
