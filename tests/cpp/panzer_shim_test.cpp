// Drives the GPU path through the Panzer-shaped C++ layer (include/txa_panzer.hpp)
// the way AssemblyEngine drives an element block: DOF manager, linear-object
// container, evaluator lifecycle, Residual then Jacobian evaluation with
// transient seeds and Dirichlet rows.  Checks the results against the CPU
// oracle (graph bit-exact, values to 1e-12) and prints "TEST PASSED", the
// reference harness's pass marker.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "../../include/txa_panzer.hpp"
#include "../../oracle/oracle.h"

namespace {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  hip_check(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())), "hipMalloc");
  hip_check(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return d;
}

template <class T>
std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> h(n);
  hip_check(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return h;
}

double rel_err(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num = std::fmax(num, std::fabs(a[i] - b[i]));
    den = std::fmax(den, std::fabs(b[i]));
  }
  return num / den;
}

// deterministic pseudo-random values in [-1, 1]
double pr(uint64_t i, uint64_t seed) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return 2.0 * static_cast<double>(z >> 11) / 9007199254740992.0 - 1.0;
}

}  // namespace

int main() {
  using namespace txa::panzer;
  const int nx = 12, ny = 10, nz = 9;
  const int64_t nv = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
  std::vector<double> X(3 * nv);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const int64_t n = i + (nx + 1) * (j + static_cast<int64_t>(ny + 1) * k);
        const bool inner = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
        const double h[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) X[3 * n + d] = ijk[d] * h[d] + (inner ? 0.1 * h[d] * pr(3 * n + d, 7) : 0.0);
      }
  BrickDOFManager dm(0, TXA_HEX8, nx, ny, nz, X, 1);
  std::vector<int64_t> owned, rowptr;
  std::vector<int32_t> lids, col;
  dm.getOwnedIndices(owned);
  dm.getElementLIDs(lids);
  dm.getGraph(rowptr, col);

  // oracle: same mesh, DOF map, graph
  int64_t sz[6];
  orc_mesh_sizes(ORC_HEX8, nx, ny, nz, sz);
  std::vector<int64_t> conn(sz[0] * sz[1]), ev(sz[0] * sz[2]), eg(sz[0] * 8), oowned(sz[3]);
  std::vector<int32_t> olids(sz[0] * 8);
  orc_mesh_build(ORC_HEX8, nx, ny, nz, conn.data(), ev.data());
  const int64_t nd = orc_dofmap(sz[0], 8, conn.data(), 1, eg.data(), oowned.data(), olids.data());
  std::vector<int64_t> orp(nd + 1);
  const int64_t nnz = orc_graph_rowptr(sz[0], 8, olids.data(), nd, orp.data());
  std::vector<int32_t> ocol(nnz);
  orc_graph_cols(sz[0], 8, olids.data(), nd, orp.data(), ocol.data());
  if (owned != oowned || lids != olids || rowptr != orp || col != ocol) {
    std::printf("DOF map or graph differs from the oracle\nTEST FAILED\n");
    return 1;
  }

  // closure data, state, Dirichlet rows (vertices at x = 0)
  const int64_t ne = dm.numElements();
  std::vector<double> k(ne), x(nd), xd(nd);
  for (int64_t e = 0; e < ne; ++e) k[e] = std::exp(0.5 * pr(e, 11));
  for (int64_t i = 0; i < nd; ++i) { x[i] = pr(i, 12); xd[i] = pr(i, 13); }
  std::vector<int32_t> bc;
  for (int64_t e = 0; e < ne; ++e)
    for (int v = 0; v < 8; ++v)
      if (conn[e * 8 + v] % (nx + 1) == 0) bc.push_back(olids[e * 8 + v]);
  std::sort(bc.begin(), bc.end());
  bc.erase(std::unique(bc.begin(), bc.end()), bc.end());

  VolumeParams prm;
  prm.coef = to_device(k);
  prm.rho = 2.0;
  prm.src[0] = 1.0;
  prm.dirichlet_lids = to_device(bc);
  prm.n_dirichlet = static_cast<int64_t>(bc.size());
  LinearObjContainer c;
  c.x = to_device(x);
  c.dxdt = to_device(xd);
  double *f = nullptr, *A = nullptr;
  hip_check(hipMalloc(&f, sizeof(double) * nd), "hipMalloc f");
  hip_check(hipMalloc(&A, sizeof(double) * nnz), "hipMalloc A");
  c.f = f;
  c.A_vals = A;

  Workset ws;
  ws.num_cells = static_cast<int>(ne);
  ws.alpha = 3.0;
  ws.beta = 1.0;
  ws.evaluate_transient_terms = true;

  GpuVolumeAssembly<Traits::Residual> res(dm, prm);
  GpuVolumeAssembly<Traits::Jacobian> jac(dm, prm);
  bool threw = false;
  try {
    res.evaluateFields(ws, c);   // before postRegistrationSetup: must throw std::logic_error
  } catch (const std::logic_error &) {
    threw = true;
  }
  res.postRegistrationSetup();
  jac.postRegistrationSetup();
  res.evaluateFields(ws, c);
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f_res = to_host(f, nd);
  jac.evaluateFields(ws, c);
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f_jac = to_host(f, nd), vals = to_host(A, nnz);

  orc_problem op{};
  op.topo = ORC_HEX8; op.physics = ORC_DIFFUSION; op.nelem = ne; op.nvert = 8;
  op.elem_vert = ev.data(); op.coords = X.data(); op.ndof = 8; op.elem_lid = olids.data();
  op.coef = k.data(); op.rho = 2.0; op.src[0] = 1.0; op.workset_size = 1000; op.nthreads = 1;
  op.rowptr = orp.data(); op.col = ocol.data();
  orc_ctx *oc = orc_setup(&op);
  orc_state st{};
  st.x = x.data(); st.xdot = xd.data(); st.alpha = 3.0; st.beta = 1.0; st.transient = 1;
  st.mode = ORC_MODE_FAD; st.want_jac = 1;
  std::vector<double> fr(nd, 0.0), vr(nnz, 0.0);
  orc_evaluate(oc, &st, fr.data(), vr.data());
  orc_dirichlet(static_cast<int64_t>(bc.size()), bc.data(), nullptr, x.data(), orp.data(), ocol.data(), fr.data(),
                vr.data());
  orc_free(oc);

  const double e1 = rel_err(f_res, fr), e2 = rel_err(f_jac, fr), e3 = rel_err(vals, vr);
  std::printf("residual eval: %.3e  jacobian eval: f %.3e  A %.3e  lifecycle error thrown: %d\n", e1, e2, e3,
              static_cast<int>(threw));
  const bool ok = threw && e1 <= 1e-12 && e2 <= 1e-12 && e3 <= 1e-12;
  std::printf(ok ? "TEST PASSED\n" : "TEST FAILED\n");
  return ok ? 0 : 1;
}
