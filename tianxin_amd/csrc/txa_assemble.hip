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
// Dispatch: diffusion on hex8 / tet4 bricks goes to the row-owner kernel
// (txa_q1_hex8.hip), three-field elasticity on bricks to txa_elastic.hip; one-field
// diffusion (steady or transient) on uploaded hex8 / tet4 meshes runs k_generic_lane here (one
// element per lane; its Jacobian through the row gather k_lane_gather, atomic-free)
// and every other combination k_generic (one workgroup per element, FP64 atomic
// scatter into a zeroed CRS).
#include "txa_internal.hpp"
#include "txa_dirichlet.hpp"
#include "txa_refelem.hpp"

#include <algorithm>
#include <vector>

namespace txa {
namespace {

// ============================================================================
// Generic path: one workgroup per element, atomic scatter
// ============================================================================
// Two shapes: 256 threads with tables for up to 27 basis functions and points
// (hex27), and one wave with tables for up to 8 (hex8, tet4) whose LDS is small
// enough for many elements in flight per CU.

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
  double *ke;                      // k_generic_lane: element matrices, packed upper triangles [nelem][nb(nb+1)/2], or null
};

__device__ __forceinline__ int64_t find_slot(const int64_t *rowptr, const int32_t *col, int32_t row, int32_t key) {
  int64_t lo = rowptr[row], hi = rowptr[row + 1];
  while (lo < hi) {
    int64_t m = (lo + hi) >> 1;
    if (col[m] < key) lo = m + 1; else hi = m;
  }
  return lo;
}

template <int GB, int MNB, int MNQ>
__global__ __launch_bounds__(GB) void k_generic(GenericParams p) {
  constexpr int MND = 3 * MNB;   // element DOFs, at most
  __shared__ double sX[8][3], sDet[MNQ], sJi[MNQ][9], sWm[MNQ];
  __shared__ double sG[MNB * MNQ * 3];
  __shared__ double sU[MND], sV[MND], sA[MND], sR[MND];
  __shared__ int32_t sL[MND];   // the element's LIDs
  // One-wave shape: the column lists of the element's rows are staged in LDS (up to RCAP
  // entries; longer ones fall back to the search in global memory), so the row search of
  // every entry runs in LDS instead of as a chain of dependent global loads.
  constexpr bool kStage = GB == 64;
  constexpr int RCAP = kStage ? 512 : 1;
  __shared__ int32_t sCol[RCAP];
  __shared__ int64_t sRB[kStage ? MND : 1];       // CRS start of each element row
  __shared__ int32_t sRO[kStage ? MND + 1 : 1];   // start of each element row in sCol
  const double *tN = p.tab;
  const double *tdN = tN + p.nb * p.nq;
  const double *tgdN = tdN + p.nb * p.nq * 3;
  const double *tqw = tgdN + p.nv * p.nq * 3;
  const int t = threadIdx.x;
  const bool elastic = p.physics == TXA_ELASTICITY;
  const bool tr = p.transient && p.xdot;
  for (int64_t e = blockIdx.x; e < p.nelem; e += gridDim.x) {
    const int32_t *lids = p.elem_lid + e * p.ndof;
    static_assert(GB >= 24, "one vertex coordinate per thread");
    if (t < p.nv * 3) {
      const int v = t / 3, d = t % 3;
      const int64_t n = p.conn[e * p.nconn + v];
      sX[v][d] = d == 0 ? p.X[n] : (d == 1 ? p.Y[n] : p.Z[n]);
    }
    for (int i = t; i < p.ndof; i += GB) {
      const int32_t L = lids[i];
      sL[i] = L;
      sU[i] = p.x[L];
      sV[i] = tr ? p.xdot[L] : 0.0;
      sA[i] = (tr && p.xddot) ? p.xddot[L] : 0.0;
      sR[i] = 0.0;
    }
    bool staged = false;   // uniform over the block
    if constexpr (kStage) {
      int64_t rb = 0;
      int rl = 0;
      if (t < p.ndof && p.vals) {
        const int32_t L = lids[t];
        rb = p.rowptr[L];
        rl = static_cast<int>(p.rowptr[L + 1] - rb);
      }
      int inc = rl;   // inclusive prefix sum of the row lengths over the wave
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o);
        if (t >= o) inc += v;
      }
      const int total = __shfl(inc, p.ndof - 1);
      staged = p.vals && total <= RCAP;
      if (t < p.ndof) { sRB[t] = rb; sRO[t + 1] = inc; }
      if (t == 0) sRO[0] = 0;
      __syncthreads();
      if (staged)
        for (int k0 = t; k0 < total; k0 += 4 * GB) {
          int32_t v[4];
          int kk[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {   // four independent loads before their LDS writes
            const int k = k0 + u * GB;
            kk[u] = k;
            v[u] = 0;
            if (k < total) {
              int I = 0;   // the element row of entry k: the last row starting at or before k
              for (int r = 1; r < p.ndof; ++r) I = sRO[r] <= k ? r : I;
              v[u] = p.col[sRB[I] + (k - sRO[I])];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (kk[u] < total) sCol[kk[u]] = v[u];
        }
    }
    __syncthreads();
    for (int q = t; q < p.nq; q += GB) {
      double J[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int v = 0; v < p.nv; ++v) s += sX[v][i] * tgdN[(v * p.nq + q) * 3 + j];
          J[i][j] = s;
        }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      double *Ji = sJi[q];
      Ji[0] = c00 * id;
      Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[3] = c01 * id;
      Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[6] = c02 * id;
      Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      sDet[q] = det;
    }
    __syncthreads();
    for (int q = t; q < p.nq; q += GB) sWm[q] = tqw[q] * sDet[q] * (sDet[0] < 0.0 ? -1.0 : 1.0);
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
        if (staged) {
          const int32_t key = sL[Jd];
          int lo = sRO[I], hi = sRO[I + 1];
          const int end = hi;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sCol[mid] < key) lo = mid + 1; else hi = mid;
          }
          if (lo < end && sCol[lo] == key) atomicAdd(p.vals + sRB[I] + (lo - sRO[I]), jv);
        } else {
          const int64_t slot = find_slot(p.rowptr, p.col, sL[I], sL[Jd]);
          if (slot < p.rowptr[sL[I] + 1] && p.col[slot] == sL[Jd]) atomicAdd(p.vals + slot, jv);
        }
      }
    }
    __syncthreads();
    for (int k = t; p.f && k < p.ndof; k += GB) {
      const int a = k / p.nf, i = k % p.nf;
      double F = 0.0;
      for (int q = 0; q < p.nq; ++q) F += sWm[q] * tN[a * p.nq + q] * p.src[i];
      atomicAdd(p.f + sL[k], sR[k] - F);
    }
    __syncthreads();
  }
}

// Generic diffusion on hex8 / tet4 meshes (one field): one element per lane.
// The reference tables sit in LDS; each lane integrates its element (the same
// formulas as k_generic, FP64) and adds its residual rows with atomics.  Steady (TR
// false): R = K u - F from the element matrix, J = beta K.  Transient (TR true): the
// quadrature loop is the same and keeps the point weights; afterwards each mass entry
// M_ab = rho sum_q w_q N_a N_b is formed once, added to R (M udot) and to the matrix,
// which becomes J_e = beta K + alpha M, so no second matrix is held.  The element
// matrix goes to the workspace for k_lane_gather when one is given (ke); otherwise its
// entries are added with atomics, each entry's CRS position found by a binary search
// whose steps depend only on the row length, so the NB searches of one row run in
// lockstep (NB independent loads per step instead of one dependent chain per entry).
template <int NB, int NQ, bool TR>
__global__ __launch_bounds__(256, 2) void k_generic_lane(GenericParams p) {
  __shared__ double tN[NB * NQ], tdN[NB * NQ * 3], tgdN[NB * NQ * 3], tqw[NQ];
  for (int i = threadIdx.x; i < NB * NQ; i += 256) tN[i] = p.tab[i];
  for (int i = threadIdx.x; i < NB * NQ * 3; i += 256) {
    tdN[i] = p.tab[NB * NQ + i];
    tgdN[i] = p.tab[NB * NQ * 4 + i];   // geometry vertices are the basis nodes (nv = nb)
  }
  for (int i = threadIdx.x; i < NQ; i += 256) tqw[i] = p.tab[NB * NQ * 7 + i];
  __syncthreads();
  const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= p.nelem) return;
  double X[NB][3], u[NB], ud[TR ? NB : 1];
  int32_t L[NB];
#pragma unroll
  for (int v = 0; v < NB; ++v) {
    const int64_t n = p.conn[e * p.nconn + v];
    X[v][0] = p.X[n]; X[v][1] = p.Y[n]; X[v][2] = p.Z[n];
    L[v] = p.elem_lid[e * NB + v];
  }
#pragma unroll
  for (int v = 0; v < NB; ++v) u[v] = p.x[L[v]];
  if constexpr (TR) {
#pragma unroll
    for (int v = 0; v < NB; ++v) ud[v] = p.xdot[L[v]];
  }
  const double cf = p.coef[e];
  double K[NB * (NB + 1) / 2], F[NB];
#pragma unroll
  for (int i = 0; i < NB * (NB + 1) / 2; ++i) K[i] = 0.0;
#pragma unroll
  for (int a = 0; a < NB; ++a) F[a] = 0.0;
  double sg = 1.0;
  double wq[TR ? NQ : 1];   // the point weights w detJ, for the mass entries
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < NB; ++v) s += X[v][i] * tgdN[(v * NQ + q) * 3 + j];
        J[i][j] = s;
      }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    double Ji[9];
    Ji[0] = c00 * id;
    Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[3] = c01 * id;
    Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[6] = c02 * id;
    Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    if (q == 0) sg = det < 0.0 ? -1.0 : 1.0;
    const double w = tqw[q] * det * sg;
    double G[NB][3];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double *dn = &tdN[(b * NQ + q) * 3];
#pragma unroll
      for (int i = 0; i < 3; ++i) G[b][i] = Ji[0 * 3 + i] * dn[0] + Ji[1 * 3 + i] * dn[1] + Ji[2 * 3 + i] * dn[2];
    }
    if constexpr (TR) wq[q] = w;
    int k = 0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      F[a] += w * tN[a * NQ + q] * p.src[0];
#pragma unroll
      for (int b = a; b < NB; ++b, ++k) K[k] += w * cf * (G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2]);
    }
  }
  auto kab = [&](int a, int b) -> double {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return K[lo * NB - (lo * (lo - 1)) / 2 + (hi - lo)];
  };
  // residual rows K u - F (then + M udot below, for the transient form)
  if (TR || p.f)
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      double r = 0.0;
#pragma unroll
      for (int b = 0; b < NB; ++b) r += kab(a, b) * u[b];
      F[a] = r - F[a];
    }
  if constexpr (TR) {
    int k = 0;
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = a; b < NB; ++b, ++k) {
        double mab = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) mab += wq[q] * tN[a * NQ + q] * tN[b * NQ + q];
        mab *= p.rho;
        F[a] += mab * ud[b];
        if (b != a) F[b] += mab * ud[a];
        K[k] = p.beta * K[k] + p.alpha * mab;
      }
  }
  if (p.f)
#pragma unroll
    for (int a = 0; a < NB; ++a) atomicAdd(p.f + L[a], F[a]);
  const double jb = TR ? 1.0 : p.beta;   // the transient matrix already holds beta and alpha
  if (!p.vals) return;
  if (p.ke) {
    // the packed upper triangle as accumulated: entry (lo, hi) at lo*NB - lo(lo-1)/2 + (hi - lo)
    double *ke = p.ke + e * (NB * (NB + 1) / 2);
#pragma unroll
    for (int i = 0; i < NB * (NB + 1) / 2; ++i) ke[i] = K[i];
    return;
  }
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    const int64_t r0 = p.rowptr[L[a]];
    const int64_t n = p.rowptr[L[a] + 1] - r0;
    int64_t step = 1;
    while (step * 2 <= n) step *= 2;
    int64_t pos[NB];   // lower bound of L[b] in the row: columns before it, in [0, n]
#pragma unroll
    for (int b = 0; b < NB; ++b) pos[b] = 0;
    for (; step > 0; step >>= 1) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (pos[b] + step <= n && p.col[r0 + pos[b] + step - 1] < L[b]) pos[b] += step;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (pos[b] < n && p.col[r0 + pos[b]] == L[b]) atomicAdd(p.vals + r0 + pos[b], jb * kab(a, b));
  }
}

// Row gather of the uploaded-mesh Jacobian (after k_generic_lane wrote the element
// matrices, packed upper triangles): 32 lanes per CRS row, lane j on columns j, j+32, ...
// of the row.  The row's items are the entries (L_e[b], K_e(a, b)) of its element rows p = e*NB + a, in
// ascending p and b.  The 32 lanes load 64 items at a time (two each, all loads in
// flight together) into their half-wave's LDS region; every lane then walks the 64 in
// item order (broadcast reads) and adds those whose column is its own.  Every entry of
// the row is stored once, with no atomics, so the values do not depend on the order in
// which elements ran.  A half-wave's region is written and read by that half-wave
// only, and one wave's LDS operations complete in program order, so no barrier is
// needed (the two halves of a wave may run different trip counts).
template <int NB>
__global__ __launch_bounds__(256) void k_lane_gather(int64_t nrows, const int64_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const int64_t *__restrict__ occ_ptr,
                                                     const int64_t *__restrict__ occ,
                                                     const int32_t *__restrict__ elem_lid,
                                                     const double *__restrict__ ke, double beta, int add,
                                                     double *__restrict__ vals) {
  __shared__ __align__(16) double sK[8][64];
  __shared__ __align__(16) int32_t sL[8][64];
  const int j = threadIdx.x % 32, hw = threadIdx.x / 32;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * (256 / 32);
  for (int64_t row = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) / 32; row < nrows; row += stride) {
    const int64_t r0 = rowptr[row], n = rowptr[row + 1] - r0;
    const int64_t o0 = occ_ptr[row];
    const int64_t nitem = (occ_ptr[row + 1] - o0) * NB;
    for (int64_t c0 = 0; c0 < n; c0 += 32) {
      const int64_t jj = c0 + j;
      const int32_t c = jj < n ? col[r0 + jj] : -1;
      double acc = 0.0;
      for (int64_t i0 = 0; i0 < nitem; i0 += 64) {
        int32_t Li[2];
        double Ki[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t it = i0 + h * 32 + j;
          Li[h] = -2;   // matches no column
          Ki[h] = 0.0;
          if (it < nitem) {
            const int64_t pe = occ[o0 + it / NB];   // e*NB + a
            const int64_t e = pe / NB;
            const int a = static_cast<int>(pe - e * NB), b = static_cast<int>(it % NB);
            const int lo = a < b ? a : b, hi = a < b ? b : a;   // K_e is symmetric, stored packed
            Li[h] = elem_lid[e * NB + b];
            Ki[h] = ke[e * (NB * (NB + 1) / 2) + lo * NB - (lo * (lo - 1)) / 2 + (hi - lo)];
          }
        }
        // the previous round's reads of this region are issued before these writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          sL[hw][h * 32 + j] = Li[h];
          sK[hw][h * 32 + j] = Ki[h];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // past the round's last item the slots hold column -2, which matches no lane; the
        // reads are unconditional (16-byte, eight items a step) and the sum takes a select
        const int cnt = nitem - i0 < 64 ? static_cast<int>(nitem - i0) : 64;
        const int cnt8 = (cnt + 7) & ~7;
        for (int k = 0; k < cnt8; k += 8) {
          const int4 la = *reinterpret_cast<const int4 *>(&sL[hw][k]);
          const int4 lb = *reinterpret_cast<const int4 *>(&sL[hw][k + 4]);
          const int32_t lv[8] = {la.x, la.y, la.z, la.w, lb.x, lb.y, lb.z, lb.w};
          double kv[8];
#pragma unroll
          for (int u = 0; u < 8; u += 2) {
            const double2 k2 = *reinterpret_cast<const double2 *>(&sK[hw][k + u]);
            kv[u] = k2.x;
            kv[u + 1] = k2.y;
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) acc = lv[u] == c ? acc + kv[u] : acc;
        }
      }
      if (jj < n) vals[r0 + jj] = add ? vals[r0 + jj] + beta * acc : beta * acc;
    }
  }
}

// The row gather over a mesh of several element blocks (hex8 / tet4 / wedge6): as k_lane_gather
// with every element row padded to 8 items (a row of a block with NB < 8 basis functions has
// NB real items; the others carry column -2, which matches no lane).  An element row is a
// position p of elem_lid; its block is the last one whose positions start at or before p.
constexpr int kMaxGatherBlocks = 8;
struct GatherBlocks {
  int n;
  int64_t pos_begin[kMaxGatherBlocks];   // first elem_lid position of the block
  int64_t ke_begin[kMaxGatherBlocks];    // first packed element-matrix entry of the block
  int nb[kMaxGatherBlocks];
};

__global__ __launch_bounds__(256) void k_lane_gather_blocks(int64_t nrows, const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ col,
                                                            const int64_t *__restrict__ occ_ptr,
                                                            const int64_t *__restrict__ occ,
                                                            const int32_t *__restrict__ elem_lid,
                                                            const double *__restrict__ ke, GatherBlocks gb,
                                                            int sel, double beta, int add,
                                                            double *__restrict__ vals) {
  constexpr int NBP = 8;
  __shared__ __align__(16) double sK[8][64];
  __shared__ __align__(16) int32_t sL[8][64];
  const int j = threadIdx.x % 32, hw = threadIdx.x / 32;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * (256 / 32);
  for (int64_t row = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) / 32; row < nrows; row += stride) {
    const int64_t r0 = rowptr[row], n = rowptr[row + 1] - r0;
    const int64_t o0 = occ_ptr[row];
    const int64_t nitem = (occ_ptr[row + 1] - o0) * NBP;
    for (int64_t c0 = 0; c0 < n; c0 += 32) {
      const int64_t jj = c0 + j;
      const int32_t c = jj < n ? col[r0 + jj] : -1;
      double acc = 0.0;
      for (int64_t i0 = 0; i0 < nitem; i0 += 64) {
        int32_t Li[2];
        double Ki[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t it = i0 + h * 32 + j;
          Li[h] = -2;
          Ki[h] = 0.0;
          if (it < nitem) {
            const int64_t pe = occ[o0 + it / NBP];
            int k = 0;
            for (int q = 1; q < gb.n; ++q) k = gb.pos_begin[q] <= pe ? q : k;
            const int nb = gb.nb[k];
            const int64_t loc = pe - gb.pos_begin[k], e = loc / nb;
            const int a = static_cast<int>(loc - e * nb), b = static_cast<int>(it % NBP);
            if (b < nb && (sel < 0 || k == sel)) {   // rows of the selected block only (sel >= 0)
              const int lo = a < b ? a : b, hi = a < b ? b : a;
              Li[h] = elem_lid[gb.pos_begin[k] + e * nb + b];
              Ki[h] = ke[gb.ke_begin[k] + e * (nb * (nb + 1) / 2) + lo * nb - (lo * (lo - 1)) / 2 + (hi - lo)];
            }
          }
        }
        // the previous round's reads of this region are issued before these writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          sL[hw][h * 32 + j] = Li[h];
          sK[hw][h * 32 + j] = Ki[h];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int cnt = nitem - i0 < 64 ? static_cast<int>(nitem - i0) : 64;
        const int cnt8 = (cnt + 7) & ~7;
        for (int k = 0; k < cnt8; k += 8) {
          const int4 la = *reinterpret_cast<const int4 *>(&sL[hw][k]);
          const int4 lb = *reinterpret_cast<const int4 *>(&sL[hw][k + 4]);
          const int32_t lv[8] = {la.x, la.y, la.z, la.w, lb.x, lb.y, lb.z, lb.w};
          double kv[8];
#pragma unroll
          for (int u = 0; u < 8; u += 2) {
            const double2 k2 = *reinterpret_cast<const double2 *>(&sK[hw][k + u]);
            kv[u] = k2.x;
            kv[u + 1] = k2.y;
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) acc = lv[u] == c ? acc + kv[u] : acc;
        }
      }
      if (jj < n) vals[r0 + jj] = add ? vals[r0 + jj] + beta * acc : beta * acc;
    }
  }
}

// y += x over n entries (accumulating a brick's assembly into the caller's container).
__global__ void k_add(double *__restrict__ y, const double *__restrict__ x, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] += x[i];
}

void add_async(double *y, const double *x, int64_t n, hipStream_t s) {
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_add, dim3(static_cast<unsigned>(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, y, x, n);
  TXA_HIP(hipGetLastError());
}

// y = A x on a CRS matrix: 16 lanes per row, lane-strided over the row, then a fixed-order
// lane reduction (one wave holds four rows).
__global__ __launch_bounds__(256) void k_spmv(int64_t nrows, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ col, const double *__restrict__ vals,
                                              const double *__restrict__ x, double *__restrict__ y) {
  const int l = threadIdx.x % 16;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * (256 / 16);
  for (int64_t row = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) / 16; row < nrows; row += stride) {
    double acc = 0.0;
    for (int64_t k = rowptr[row] + l; k < rowptr[row + 1]; k += 16) acc += vals[k] * x[col[k]];
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 16);
    if (l == 0) y[row] = acc;
  }
}

// Fill of an output (zero before the atomic kernels add into it; a kernel rather than
// hipMemsetAsync, so the assembly can be recorded in a HIP graph).
__global__ void k_fill(double *__restrict__ p, int64_t n, double v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

void zero_async(double *p, int64_t n, hipStream_t s) { fill_async(p, n, 0.0, s); }

// ============================================================================
// Dirichlet rows: f_i = x_i - g_i; matrix row -> identity row
// (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:228-237, 300-318)
// ============================================================================
// one wave per kDirU * 64/G consecutive list entries
template <int G>
__global__ __launch_bounds__(256) void k_dirichlet(int64_t n, int64_t nrows, const int32_t *__restrict__ lids,
                                                   const double *__restrict__ g, const double *__restrict__ x,
                                                   const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                   double *__restrict__ f, double *__restrict__ vals, int *err) {
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64;
  dirichlet_wave<G>(wave * kDirU * (64 / G), n, nrows, lids, g, x, rowptr, col, f, vals, err);
}

// Neumann flux on one side of the brick: one thread per lattice point of the
// side sums, in a fixed face order, its share of the faces around it,
//   f(point, i) += val_i * sum_q w_q L_a(xi_q) L_b(eta_q) |x_xi x x_eta|(q),
// with the bilinear face map of the four corner vertices, Gauss points of the
// volume degree (2 for hex8, 3 for hex27) and the face basis of the element
// restricted to the face (a, b: the point's position in the face, 0..STEP).
struct NeumannParams {
  int step;                       // lattice points per cell edge: 1 (hex8), 2 (hex27)
  int nx, ny, nz, nf;
  int axis, at_max;               // side normal axis, and whether the side is at its upper end
  double val[3];
  const double *X, *Y, *Z;        // vertex coordinates
  const int32_t *point_lid;       // [points][nf]
  double *f;
};

__device__ __forceinline__ double lagrange(int step, int a, double t) {
  if (step == 1) return a == 0 ? 0.5 * (1.0 - t) : 0.5 * (1.0 + t);
  return a == 0 ? 0.5 * t * (t - 1.0) : (a == 1 ? 1.0 - t * t : 0.5 * t * (t + 1.0));
}

__global__ void k_neumann(NeumannParams p) {
  const int n[3] = {p.nx, p.ny, p.nz};
  const int d = p.axis, u = d == 0 ? 1 : 0, v = d == 2 ? 1 : 2;   // in-plane axes, ascending
  const int64_t pu = static_cast<int64_t>(p.step) * n[u] + 1, pv = static_cast<int64_t>(p.step) * n[v] + 1;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= pu * pv) return;
  const int I = static_cast<int>(t % pu), J = static_cast<int>(t / pu);
  const int nq = p.step == 1 ? 2 : 3;
  const double g2 = 0.57735026918962576451, g3 = 0.77459666924148337704;   // 1/sqrt(3), sqrt(3/5)
  double acc = 0.0;
  const int cu_lo = I % p.step == 0 ? I / p.step - 1 : I / p.step, cv_lo = J % p.step == 0 ? J / p.step - 1 : J / p.step;
  for (int cv = cv_lo; cv <= J / p.step; ++cv)
    for (int cu = cu_lo; cu <= I / p.step; ++cu) {
      if (cu < 0 || cu >= n[u] || cv < 0 || cv >= n[v]) continue;
      const int a = I - p.step * cu, b = J - p.step * cv;
      double C[2][2][3];   // corners [du][dv], coordinates
      for (int dv = 0; dv < 2; ++dv)
        for (int du = 0; du < 2; ++du) {
          int c[3];
          c[d] = p.at_max ? n[d] : 0;
          c[u] = cu + du;
          c[v] = cv + dv;
          const int64_t vid = c[0] + static_cast<int64_t>(p.nx + 1) * (c[1] + static_cast<int64_t>(p.ny + 1) * c[2]);
          C[du][dv][0] = p.X[vid]; C[du][dv][1] = p.Y[vid]; C[du][dv][2] = p.Z[vid];
        }
      for (int qv = 0; qv < nq; ++qv)
        for (int qu = 0; qu < nq; ++qu) {
          const double xi = nq == 2 ? (2 * qu - 1) * g2 : (qu - 1) * g3;
          const double eta = nq == 2 ? (2 * qv - 1) * g2 : (qv - 1) * g3;
          const double w = nq == 2 ? 1.0 : (qu == 1 ? 8.0 / 9.0 : 5.0 / 9.0) * (qv == 1 ? 8.0 / 9.0 : 5.0 / 9.0);
          double tx[3], ty[3];
          for (int k = 0; k < 3; ++k) {
            tx[k] = 0.25 * ((1.0 - eta) * (C[1][0][k] - C[0][0][k]) + (1.0 + eta) * (C[1][1][k] - C[0][1][k]));
            ty[k] = 0.25 * ((1.0 - xi) * (C[0][1][k] - C[0][0][k]) + (1.0 + xi) * (C[1][1][k] - C[1][0][k]));
          }
          const double nx_ = tx[1] * ty[2] - tx[2] * ty[1], ny_ = tx[2] * ty[0] - tx[0] * ty[2],
                       nz_ = tx[0] * ty[1] - tx[1] * ty[0];
          acc += w * lagrange(p.step, a, xi) * lagrange(p.step, b, eta) * sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
        }
    }
  int64_t c[3];
  c[d] = p.at_max ? static_cast<int64_t>(p.step) * n[d] : 0;
  c[u] = I;
  c[v] = J;
  const int64_t px = static_cast<int64_t>(p.step) * p.nx + 1, py = static_cast<int64_t>(p.step) * p.ny + 1;
  const int64_t pt = c[0] + px * (c[1] + py * c[2]);
  for (int i = 0; i < p.nf; ++i) p.f[p.point_lid[pt * p.nf + i]] += p.val[i] * acc;
}

// ---- Neumann flux on listed element sides, any mesh (the side worksets of a side set)
// Side node lists in shards order: Hexahedron<8>, Tetrahedron<4>, Wedge<6> (quads, then the
// two triangles).  Quads in cyclic order: corner i at (xi, eta) = (-1,-1), (1,-1), (1,1), (-1,1).
__device__ const int8_t kHexSide[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                          {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
__device__ const int8_t kTetSide[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
__device__ const int8_t kWedgeQuad[3][4] = {{0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}};
__device__ const int8_t kWedgeTri[2][3] = {{0, 2, 1}, {3, 4, 5}};
// lattice offsets of the shards hex vertices and of the Q2 nodes (subcell order)
__device__ const int8_t kHexOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                         {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__device__ const int8_t kQ2Off[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2},
    {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
    {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};
__device__ const int8_t kSideAxis[6] = {1, 0, 1, 0, 2, 2}, kSideMax[6] = {0, 1, 1, 0, 0, 1};

struct SideBlocks {
  int n;
  int64_t elem_begin[kMaxGatherBlocks + 1];   // n + 1 entries: the last is the element count
  int64_t conn_begin[kMaxGatherBlocks];
  int64_t pos_begin[kMaxGatherBlocks];
  int nconn[kMaxGatherBlocks], topo[kMaxGatherBlocks];
};

struct SideParams {
  int64_t n;
  const int64_t *elems;   // [n] mesh element index (blocks in order)
  const int *ords;        // [n] shards side ordinal of the element's topology
  const double *values;   // [n][nf] flux per field
  int nf;
  const double *X, *Y, *Z;
  const int64_t *conn;
  const int32_t *elem_lid;
  SideBlocks sb;
  double *f;
  int *err;               // bit 1 set for a side that does not exist
};

// integral over a cyclic quad (corners c[0..3], bilinear map) of each corner's bilinear basis
// times the face measure, 2x2 Gauss
__device__ void quad_loads(const double (&c)[4][3], double (&out)[4]) {
  const double g = 0.57735026918962576451;
  for (int i = 0; i < 4; ++i) out[i] = 0.0;
  for (int qe = 0; qe < 2; ++qe)
    for (int qx = 0; qx < 2; ++qx) {
      const double xi = qx ? g : -g, eta = qe ? g : -g;
      double tx[3], ty[3];
      for (int k = 0; k < 3; ++k) {
        tx[k] = 0.25 * ((1.0 - eta) * (c[1][k] - c[0][k]) + (1.0 + eta) * (c[2][k] - c[3][k]));
        ty[k] = 0.25 * ((1.0 - xi) * (c[3][k] - c[0][k]) + (1.0 + xi) * (c[2][k] - c[1][k]));
      }
      const double nx_ = tx[1] * ty[2] - tx[2] * ty[1], ny_ = tx[2] * ty[0] - tx[0] * ty[2],
                   nz_ = tx[0] * ty[1] - tx[1] * ty[0];
      const double J = sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
      out[0] += 0.25 * (1.0 - xi) * (1.0 - eta) * J;
      out[1] += 0.25 * (1.0 + xi) * (1.0 - eta) * J;
      out[2] += 0.25 * (1.0 + xi) * (1.0 + eta) * J;
      out[3] += 0.25 * (1.0 - xi) * (1.0 + eta) * J;
    }
}

// one third of a triangle's area (the integral of each linear vertex basis over it)
__device__ double tri_third(const double (&a)[3], const double (&b)[3], const double (&c)[3]) {
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2], n2 = u[0] * v[1] - u[1] * v[0];
  return sqrt(n0 * n0 + n1 * n1 + n2 * n2) / 6.0;
}

// One thread per side; the loads add into f with FP64 atomics (sides share nodes).
__global__ void k_neumann_sides(SideParams p) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  const int64_t e = p.elems[k];
  const int ord = p.ords[k];
  if (e < 0 || e >= p.sb.elem_begin[p.sb.n]) {
    if (p.err) atomicOr(p.err, 2);
    return;
  }
  int b = 0;
  while (b + 1 < p.sb.n && p.sb.elem_begin[b + 1] <= e) ++b;
  const int topo = p.sb.topo[b], nconn = p.sb.nconn[b];
  const int64_t le = e - p.sb.elem_begin[b];
  const int64_t *cn = p.conn + p.sb.conn_begin[b] + le * nconn;
  const int32_t *el = p.elem_lid + p.sb.pos_begin[b] + le * nconn * p.nf;
  const int nsides = topo == TXA_TET4 ? 4 : (topo == TXA_WEDGE6 ? 5 : 6);
  if (ord < 0 || ord >= nsides) {
    if (p.err) atomicOr(p.err, 2);
    return;
  }
  auto vtx = [&](int v, double (&x)[3]) {
    const int64_t id = cn[v];
    x[0] = p.X[id]; x[1] = p.Y[id]; x[2] = p.Z[id];
  };
  auto add = [&](int node, double w) {
    for (int i = 0; i < p.nf; ++i) atomicAdd(&p.f[el[node * p.nf + i]], p.values[k * p.nf + i] * w);
  };
  const bool tri = topo == TXA_TET4 || (topo == TXA_WEDGE6 && ord >= 3);
  if (tri) {
    const int8_t *t = topo == TXA_TET4 ? kTetSide[ord] : kWedgeTri[ord - 3];
    double a[3], bb[3], c[3];
    vtx(t[0], a); vtx(t[1], bb); vtx(t[2], c);
    const double w = tri_third(a, bb, c);
    for (int i = 0; i < 3; ++i) add(t[i], w);
  } else if (topo != TXA_HEX27) {
    const int8_t *q = topo == TXA_WEDGE6 ? kWedgeQuad[ord] : kHexSide[ord];
    double c[4][3], w[4];
    for (int i = 0; i < 4; ++i) vtx(q[i], c[i]);
    quad_loads(c, w);
    for (int i = 0; i < 4; ++i) add(q[i], w[i]);
  } else {
    // 9 face nodes with the biquadratic face basis over the bilinear face map, 3x3 Gauss
    const int d = kSideAxis[ord], am = kSideMax[ord], u = d == 0 ? 1 : 0, v = d == 2 ? 1 : 2;
    double C[2][2][3];
    for (int vv = 0; vv < 8; ++vv)
      if (kHexOff[vv][d] == am) {
        double x[3];
        vtx(vv, x);
        for (int i = 0; i < 3; ++i) C[kHexOff[vv][u]][kHexOff[vv][v]][i] = x[i];
      }
    const double g3 = 0.77459666924148337704;
    for (int nd = 0; nd < 27; ++nd) {
      if (kQ2Off[nd][d] != 2 * am) continue;
      const int a = kQ2Off[nd][u], bq = kQ2Off[nd][v];
      double acc = 0.0;
      for (int qv = 0; qv < 3; ++qv)
        for (int qu = 0; qu < 3; ++qu) {
          const double xi = (qu - 1) * g3, eta = (qv - 1) * g3;
          const double wq = (qu == 1 ? 8.0 / 9.0 : 5.0 / 9.0) * (qv == 1 ? 8.0 / 9.0 : 5.0 / 9.0);
          double tx[3], ty[3];
          for (int kk = 0; kk < 3; ++kk) {
            tx[kk] = 0.25 * ((1.0 - eta) * (C[1][0][kk] - C[0][0][kk]) + (1.0 + eta) * (C[1][1][kk] - C[0][1][kk]));
            ty[kk] = 0.25 * ((1.0 - xi) * (C[0][1][kk] - C[0][0][kk]) + (1.0 + xi) * (C[1][1][kk] - C[1][0][kk]));
          }
          const double n0 = tx[1] * ty[2] - tx[2] * ty[1], n1 = tx[2] * ty[0] - tx[0] * ty[2],
                       n2 = tx[0] * ty[1] - tx[1] * ty[0];
          acc += wq * lagrange(2, a, xi) * lagrange(2, bq, eta) * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        }
      add(nd, acc);
    }
  }
}

// Brick sides of a TET4 brick: the boundary triangles of the six-tet split.  One thread per
// (cell on the side, tet); each of the tet's faces with all three vertices on the side's
// plane adds a third of its area to its vertices' rows (FP64 atomics).
struct TetSideParams {
  int nx, ny, nz, nf, axis, at_max;
  double val[3];
  const double *X, *Y, *Z;
  const int64_t *conn;
  const int32_t *elem_lid;
  double *f;
};

__global__ void k_neumann_tet_brick(TetSideParams p) {
  const int n[3] = {p.nx, p.ny, p.nz};
  const int d = p.axis, u = d == 0 ? 1 : 0, v = d == 2 ? 1 : 2;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(n[u]) * n[v] * 6) return;
  const int tet = static_cast<int>(t % 6);
  const int64_t cf = t / 6;
  int cc[3];
  cc[d] = p.at_max ? n[d] - 1 : 0;
  cc[u] = static_cast<int>(cf % n[u]);
  cc[v] = static_cast<int>(cf / n[u]);
  const int64_t cell = cc[0] + static_cast<int64_t>(p.nx) * (cc[1] + static_cast<int64_t>(p.ny) * cc[2]);
  const int64_t e = 6 * cell + tet;
  const int64_t *cn = p.conn + e * 4;
  const int32_t *el = p.elem_lid + e * 4 * p.nf;
  const int64_t sx = p.nx + 1, sxy = sx * (p.ny + 1);
  const int64_t plane = p.at_max ? n[d] : 0;
  bool on[4];
  for (int i = 0; i < 4; ++i) {
    const int64_t id = cn[i];
    const int64_t c = d == 0 ? id % sx : (d == 1 ? (id / sx) % (p.ny + 1) : id / sxy);
    on[i] = c == plane;
  }
  for (int fs = 0; fs < 4; ++fs) {
    const int8_t *tr = kTetSide[fs];
    if (!(on[tr[0]] && on[tr[1]] && on[tr[2]])) continue;
    double x[3][3];
    for (int i = 0; i < 3; ++i) {
      const int64_t id = cn[tr[i]];
      x[i][0] = p.X[id]; x[i][1] = p.Y[id]; x[i][2] = p.Z[id];
    }
    const double w = tri_third(x[0], x[1], x[2]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < p.nf; ++j) atomicAdd(&p.f[el[tr[i] * p.nf + j]], p.val[j] * w);
  }
}

}  // namespace

void neumann_sides(const Mesh &m, int64_t n, const int64_t *elems, const int *ords, const double *values, double *f,
                   hipStream_t s, int *err) {
  if (n == 0) return;
  TXA_REQUIRE(m.nfields >= 1 && m.elem_lid.p, TXA_ERR_LOGIC, "DOF map not built");
  TXA_REQUIRE(elems && ords && values && f, TXA_ERR_INVALID, "null argument");
  TXA_REQUIRE(static_cast<int>(m.blocks.size()) <= kMaxGatherBlocks, TXA_ERR_UNSUPPORTED,
              "Neumann sides on meshes of more than 8 element blocks");
  SideParams p;
  p.n = n; p.elems = elems; p.ords = ords; p.values = values; p.nf = m.nfields;
  p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
  p.conn = m.conn.p; p.elem_lid = m.elem_lid.p;
  p.sb.n = static_cast<int>(m.blocks.size());
  for (int b = 0; b < p.sb.n; ++b) {
    const MeshBlock &mb = m.blocks[b];
    p.sb.elem_begin[b] = mb.elem_begin;
    p.sb.conn_begin[b] = mb.conn_begin;
    p.sb.pos_begin[b] = m.pos_begin(mb);
    p.sb.nconn[b] = mb.nconn;
    p.sb.topo[b] = mb.topo;
  }
  p.sb.elem_begin[p.sb.n] = m.nelem;
  p.f = f;
  p.err = err;
  hipLaunchKernelGGL(k_neumann_sides, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p);
  TXA_HIP(hipGetLastError());
}

void upload_refelem(int topo, DevBuf<double> &tab, hipStream_t s) {
  RefElem r;
  refelem_build(r, topo);
  std::vector<double> h;
  for (int b = 0; b < r.nb; ++b) for (int q = 0; q < r.nq; ++q) h.push_back(r.N[b][q]);
  for (int b = 0; b < r.nb; ++b) for (int q = 0; q < r.nq; ++q) for (int d = 0; d < 3; ++d) h.push_back(r.dN[b][q][d]);
  for (int v = 0; v < r.nv; ++v) for (int q = 0; q < r.nq; ++q) for (int d = 0; d < 3; ++d) h.push_back(r.gdN[v][q][d]);
  for (int q = 0; q < r.nq; ++q) h.push_back(r.qw[q]);
  tab.alloc(h.size());
  TXA_HIP(hipMemcpyAsync(tab.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  TXA_HIP(hipStreamSynchronize(s));
}

void upload_block_refs(Mesh &m, hipStream_t s) {
  if (m.topo != 0) upload_refelem(m.topo, m.ref, s);
  for (const MeshBlock &b : m.blocks)
    if (!m.block_ref[b.topo].p) upload_refelem(b.topo, m.block_ref[b.topo], s);
}

static bool uses_q1_path(const Mesh &m, int physics, int precision) {
  return m.structured && (m.topo == TXA_HEX8 || m.topo == TXA_TET4) && physics == TXA_DIFFUSION &&
         precision != TXA_SPLIT_BF16;
}

// Basis functions of a block's topology that runs one element per lane (0: none).
static int lane_nb(int topo) { return topo == TXA_HEX8 ? 8 : topo == TXA_TET4 ? 4 : topo == TXA_WEDGE6 ? 6 : 0; }

// Uploaded one-field hex8 / tet4 / wedge6 diffusion, steady or transient: the element matrices
// go to the workspace for the row gather (the mesh's element-row lists exist; a mesh of more
// blocks than the gather's table holds adds its matrices with atomics instead).
static bool lane_gather_mesh(const Mesh &m, int physics, int precision) {
  if (m.structured || m.nfields != 1 || physics != TXA_DIFFUSION || precision != TXA_FP64 || !m.occ.p) return false;
  if (m.blocks.size() > static_cast<size_t>(kMaxGatherBlocks)) return false;
  for (const MeshBlock &b : m.blocks)
    if (!lane_nb(b.topo)) return false;
  return true;
}

// Offset (in doubles, 256-byte aligned per block) of each block's packed element matrices.
static size_t gather_ke_doubles(const Mesh &m, std::vector<int64_t> *begin) {
  size_t off = 0;
  for (const MeshBlock &b : m.blocks) {
    if (begin) begin->push_back(static_cast<int64_t>(off));
    const int nb = lane_nb(b.topo);
    off += (static_cast<size_t>(b.nelem) * (nb * (nb + 1) / 2) + 31) / 32 * 32;
  }
  return off;
}

// Workspace of the structured kernels' own parts (the elasticity blocks), before the
// staging arrays of an accumulating call.
static size_t kernel_workspace_bytes(const Mesh &m, int physics, int precision, bool want_f, bool want_vals) {
  if (patch_supported(m, physics, precision)) return patch_workspace_bytes(m);
  if (lane_gather_mesh(m, physics, precision)) return want_vals ? sizeof(double) * gather_ke_doubles(m, nullptr) : 0;
  txa_assembly_args a{};
  a.physics = physics;
  a.precision = precision;
  if (uses_q1_path(m, physics, precision) || !elastic_supported(m, a)) return 0;
  return elastic_workspace_bytes(m, precision, want_f, want_vals);
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// An accumulating call on a brick stages the kernels' rows (they store each row once) in the
// workspace and adds them to f and vals afterwards.
static bool stages_rows(const Mesh &m, int physics, int precision, bool accumulate) {
  txa_assembly_args a{};
  a.physics = physics;
  a.precision = precision;
  return accumulate && (uses_q1_path(m, physics, precision) || elastic_supported(m, a));
}

size_t assemble_workspace_bytes(const Mesh &m, int physics, int precision, bool want_f, bool want_vals,
                                bool accumulate) {
  size_t b = round256(kernel_workspace_bytes(m, physics, precision, want_f, want_vals));
  if (stages_rows(m, physics, precision, accumulate))
    b += (want_f ? round256(sizeof(double) * m.ndofs) : 0) + (want_vals ? round256(sizeof(double) * m.nnz) : 0);
  return b;
}

void fill_async(double *p, int64_t n, double v, hipStream_t s) {
  TXA_REQUIRE(n >= 0 && (n == 0 || p), TXA_ERR_INVALID, "null array or negative length");
  if (n == 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n, v);
  TXA_HIP(hipGetLastError());
}

void spmv(const Mesh &m, const double *vals, const double *x, double *y, hipStream_t s) {
  TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "spmv before graph_build");
  TXA_REQUIRE(vals && x && y, TXA_ERR_INVALID, "null argument");
  const int64_t blocks = std::min<int64_t>((m.ndofs + 15) / 16, int64_t(1) << 20);
  hipLaunchKernelGGL(k_spmv, dim3(static_cast<unsigned>(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, m.ndofs,
                     m.rowptr.p, m.col.p, vals, x, y);
  TXA_HIP(hipGetLastError());
}

void apply_dirichlet(const Mesh &m, const int32_t *lids, const double *g, int64_t n, const double *x, double *f,
                     double *vals, hipStream_t s, int *err) {
  TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "Dirichlet rows before graph_build");
  TXA_REQUIRE(n >= 0 && (n == 0 || lids), TXA_ERR_INVALID, "lids missing");
  TXA_REQUIRE(!f || x, TXA_ERR_INVALID, "x is required for the residual rows");
  if (n == 0) return;
  const int G = dirichlet_lanes(m.max_row);
  const int64_t per = 4 * (64 / G) * kDirU;   // rows per 256-thread block
  const dim3 grid(static_cast<unsigned>((n + per - 1) / per));
  if (G == 8) hipLaunchKernelGGL(k_dirichlet<8>, grid, dim3(256), 0, s, n, m.ndofs, lids, g, x, m.rowptr.p, m.col.p, f, vals, err);
  else if (G == 16) hipLaunchKernelGGL(k_dirichlet<16>, grid, dim3(256), 0, s, n, m.ndofs, lids, g, x, m.rowptr.p, m.col.p, f, vals, err);
  else if (G == 32) hipLaunchKernelGGL(k_dirichlet<32>, grid, dim3(256), 0, s, n, m.ndofs, lids, g, x, m.rowptr.p, m.col.p, f, vals, err);
  else hipLaunchKernelGGL(k_dirichlet<64>, grid, dim3(256), 0, s, n, m.ndofs, lids, g, x, m.rowptr.p, m.col.p, f, vals, err);
  TXA_HIP(hipGetLastError());
}

void assemble(Mesh &m, const txa_assembly_args &a_in, hipStream_t s, void *ws, int *err) {
  TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "assemble before graph_build");
  TXA_REQUIRE(a_in.block >= 0 && a_in.block <= static_cast<int>(m.blocks.size()), TXA_ERR_INVALID,
              "block " + std::to_string(a_in.block) + ": the mesh has " + std::to_string(m.blocks.size()) +
                  " element block(s) (0 = all, b = block b - 1)");
  txa_assembly_args a = a_in;
  const bool add = a_in.accumulate != 0;
  // bricks accumulate through staging rows in the workspace, added to f and vals after the kernels
  const bool stage = stages_rows(m, a.physics, a.precision, add);
  char *wsc = static_cast<char *>(ws);
  double *stage_f = nullptr, *stage_v = nullptr;
  if (stage) {
    size_t off = round256(kernel_workspace_bytes(m, a.physics, a.precision, a.f != nullptr, a.vals != nullptr));
    TXA_REQUIRE(wsc != nullptr, TXA_ERR_LOGIC, "an accumulating assembly needs its workspace");
    if (a.f) { stage_f = reinterpret_cast<double *>(wsc + off); off += round256(sizeof(double) * m.ndofs); }
    if (a.vals) stage_v = reinterpret_cast<double *>(wsc + off);
    a.f = stage_f ? stage_f : a.f;
    a.vals = stage_v ? stage_v : a.vals;
  }
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
  TXA_REQUIRE(a.n_neumann >= 0 && (a.n_neumann == 0 || (a.neumann_sides && a.neumann_values)), TXA_ERR_INVALID,
              "neumann_sides / neumann_values missing");
  if (a.n_neumann > 0) {
    TXA_REQUIRE(m.structured, TXA_ERR_UNSUPPORTED,
                "brick sides (TXA_SIDE_*) need a brick; other meshes list their sides (txa_neumann_sides)");
    for (int k = 0; k < a.n_neumann; ++k)
      TXA_REQUIRE(a.neumann_sides[k] >= 0 && a.neumann_sides[k] <= 5, TXA_ERR_INVALID,
                  "side " + std::to_string(a.neumann_sides[k]) + " is not a side of the brick (0..5)");
  }

  if (uses_q1_path(m, a.physics, a.precision)) {
    launch_q1_hex8_diffusion(m, a, s);
  } else if (elastic_supported(m, a)) {
    launch_elastic(m, a, s, ws);
  } else if (patch_supported(m, a.physics, a.precision)) {
    launch_patch(m, a, s, ws);
  } else {
    TXA_REQUIRE(a.precision == TXA_FP64, TXA_ERR_UNSUPPORTED,
                "reduced-precision integration is implemented for hex8 / tet4 diffusion (bricks and single-block "
                "uploaded meshes) and 3-field elasticity on bricks");
    // the Jacobian of the lane kernels by the row gather: element matrices in ws, rows stored whole
    const bool gather = a.vals != nullptr && ws != nullptr && lane_gather_mesh(m, a.physics, a.precision);
    std::vector<int64_t> ke_begin;
    if (gather) gather_ke_doubles(m, &ke_begin);
    if (a.f && !add) zero_async(a.f, m.ndofs, s);
    if (a.vals && !gather && !add) zero_async(a.vals, m.nnz, s);
    // each element block with its own basis and cubature, all adding into the same rows
    for (size_t bi = 0; bi < m.blocks.size(); ++bi) {
      if (a.block > 0 && static_cast<int>(bi) != a.block - 1) continue;
      const MeshBlock &b = m.blocks[bi];
      TXA_REQUIRE(m.block_ref[b.topo].p != nullptr, TXA_ERR_LOGIC, "reference tables missing");
      RefElem r;
      refelem_build(r, b.topo);
      GenericParams p;
      p.topo = b.topo; p.physics = a.physics; p.nb = r.nb; p.nq = r.nq; p.nv = r.nv; p.nf = m.nfields;
      p.ndof = b.nconn * m.nfields; p.nconn = b.nconn; p.nelem = b.nelem; p.conn = m.conn.p + b.conn_begin;
      p.X = m.xyz.p; p.Y = m.xyz.p + m.nverts; p.Z = m.xyz.p + 2 * m.nverts;
      p.elem_lid = m.elem_lid.p + m.pos_begin(b); p.tab = m.block_ref[b.topo].p; p.coef = a.coef + b.elem_begin;
      p.nu = a.nu; p.rho = a.rho; p.src[0] = a.src[0]; p.src[1] = a.src[1]; p.src[2] = a.src[2];
      p.ra = a.rayleigh_a; p.rb = a.rayleigh_b;
      p.x = a.x; p.xdot = a.xdot; p.xddot = a.xddot;
      p.alpha = a.alpha; p.beta = a.beta; p.gamma2 = a.gamma2; p.transient = a.transient;
      p.f = a.f; p.vals = a.vals; p.rowptr = m.rowptr.p; p.col = m.col.p;
      const int64_t blocks = b.nelem < 65536 * 8 ? b.nelem : 65536 * 8;
      const int nbl = a.physics == TXA_DIFFUSION && m.nfields == 1 ? lane_nb(b.topo) : 0;
      p.ke = gather ? static_cast<double *>(ws) + ke_begin[bi] : nullptr;
      const unsigned lblocks = static_cast<unsigned>((b.nelem + 255) / 256);
      const bool tr = a.transient != 0;
      if (nbl == 8 && tr) hipLaunchKernelGGL((k_generic_lane<8, 8, true>), dim3(lblocks), dim3(256), 0, s, p);
      else if (nbl == 8) hipLaunchKernelGGL((k_generic_lane<8, 8, false>), dim3(lblocks), dim3(256), 0, s, p);
      else if (nbl == 6 && tr) hipLaunchKernelGGL((k_generic_lane<6, 6, true>), dim3(lblocks), dim3(256), 0, s, p);
      else if (nbl == 6) hipLaunchKernelGGL((k_generic_lane<6, 6, false>), dim3(lblocks), dim3(256), 0, s, p);
      else if (nbl == 4 && tr) hipLaunchKernelGGL((k_generic_lane<4, 4, true>), dim3(lblocks), dim3(256), 0, s, p);
      else if (nbl == 4) hipLaunchKernelGGL((k_generic_lane<4, 4, false>), dim3(lblocks), dim3(256), 0, s, p);
      else if (r.nb <= 8 && r.nq <= 8)
        hipLaunchKernelGGL((k_generic<64, 8, 8>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, p);
      else
        hipLaunchKernelGGL((k_generic<256, kMaxNB, kMaxNQ>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p);
      TXA_HIP(hipGetLastError());
    }
    if (gather) {
      const double gbeta = a.transient ? 1.0 : a.beta;   // the transient element matrices hold beta already
      const int64_t gblocks = std::min<int64_t>((m.ndofs + 7) / 8, int64_t(1) << 20);   // 8 rows per block
      const double *ke = static_cast<const double *>(ws);
      if (m.blocks.size() == 1 && m.topo == TXA_HEX8)
        hipLaunchKernelGGL((k_lane_gather<8>), dim3(static_cast<unsigned>(gblocks)), dim3(256), 0, s, m.ndofs,
                           m.rowptr.p, m.col.p, m.occ_ptr.p, m.occ.p, m.elem_lid.p, ke, gbeta, int(add), a.vals);
      else if (m.blocks.size() == 1 && m.topo == TXA_TET4)
        hipLaunchKernelGGL((k_lane_gather<4>), dim3(static_cast<unsigned>(gblocks)), dim3(256), 0, s, m.ndofs,
                           m.rowptr.p, m.col.p, m.occ_ptr.p, m.occ.p, m.elem_lid.p, ke, gbeta, int(add), a.vals);
      else {
        GatherBlocks gb{};
        gb.n = static_cast<int>(m.blocks.size());
        for (int k = 0; k < gb.n; ++k) {
          gb.pos_begin[k] = m.pos_begin(m.blocks[k]);
          gb.ke_begin[k] = ke_begin[k];
          gb.nb[k] = lane_nb(m.blocks[k].topo);
        }
        hipLaunchKernelGGL(k_lane_gather_blocks, dim3(static_cast<unsigned>(gblocks)), dim3(256), 0, s, m.ndofs,
                           m.rowptr.p, m.col.p, m.occ_ptr.p, m.occ.p, m.elem_lid.p, ke, gb, a.block - 1, gbeta,
                           int(add), a.vals);
      }
      TXA_HIP(hipGetLastError());
    }
  }
  if (stage) {
    if (stage_f) add_async(a_in.f, stage_f, m.ndofs, s);
    if (stage_v) add_async(a_in.vals, stage_v, m.nnz, s);
    a.f = a_in.f;
    a.vals = a_in.vals;
  }
  // Neumann sides, one launch each (stream order keeps the sums of shared edge points in a fixed order)
  for (int k = 0; a.f && k < a.n_neumann && m.topo == TXA_TET4; ++k) {
    TetSideParams tp;
    tp.nx = m.nx; tp.ny = m.ny; tp.nz = m.nz; tp.nf = m.nfields;
    const int side = a.neumann_sides[k];
    static const int kAx[6] = {1, 0, 1, 0, 2, 2}, kMx[6] = {0, 1, 1, 0, 0, 1};
    tp.axis = kAx[side];
    tp.at_max = kMx[side];
    for (int i = 0; i < 3; ++i) tp.val[i] = i < m.nfields ? a.neumann_values[k * m.nfields + i] : 0.0;
    tp.X = m.xyz.p; tp.Y = m.xyz.p + m.nverts; tp.Z = m.xyz.p + 2 * m.nverts;
    tp.conn = m.conn.p;
    tp.elem_lid = m.elem_lid.p;
    tp.f = a.f;
    const int n3[3] = {m.nx, m.ny, m.nz};
    const int u = tp.axis == 0 ? 1 : 0, v = tp.axis == 2 ? 1 : 2;
    const int64_t nt = static_cast<int64_t>(n3[u]) * n3[v] * 6;
    hipLaunchKernelGGL(k_neumann_tet_brick, dim3(static_cast<unsigned>((nt + 255) / 256)), dim3(256), 0, s, tp);
    TXA_HIP(hipGetLastError());
  }
  for (int k = 0; a.f && k < a.n_neumann && m.topo != TXA_TET4; ++k) {
    NeumannParams np;
    np.step = m.topo == TXA_HEX27 ? 2 : 1;
    np.nx = m.nx; np.ny = m.ny; np.nz = m.nz; np.nf = m.nfields;
    const int side = a.neumann_sides[k];
    static const int kAxis[6] = {1, 0, 1, 0, 2, 2}, kMax[6] = {0, 1, 1, 0, 0, 1};
    np.axis = kAxis[side];
    np.at_max = kMax[side];
    for (int i = 0; i < 3; ++i) np.val[i] = i < m.nfields ? a.neumann_values[k * m.nfields + i] : 0.0;
    np.X = m.xyz.p; np.Y = m.xyz.p + m.nverts; np.Z = m.xyz.p + 2 * m.nverts;
    np.point_lid = m.point_lid.p;
    np.f = a.f;
    const int n3[3] = {m.nx, m.ny, m.nz};
    const int u = np.axis == 0 ? 1 : 0, v = np.axis == 2 ? 1 : 2;
    const int64_t npts = (static_cast<int64_t>(np.step) * n3[u] + 1) * (static_cast<int64_t>(np.step) * n3[v] + 1);
    hipLaunchKernelGGL(k_neumann, dim3(static_cast<unsigned>((npts + 255) / 256)), dim3(256), 0, s, np);
    TXA_HIP(hipGetLastError());
  }
  if (a.n_bc > 0) apply_dirichlet(m, a.bc_lids, a.bc_vals, a.n_bc, a.x, a.f, a.vals, s, err);
}

}  // namespace txa
