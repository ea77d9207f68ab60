// Drives the GPU path through the Panzer-shaped C++ layer (include/txa_panzer.hpp) the way
// Panzer's AssemblyEngine and ModelEvaluator drive it, and checks every result against the CPU
// oracle (DOF map and graph bit-exact, values to 1e-12):
//  1. the Q1 Poisson 32^3 unit cube (k lognormal, f = 1, u = 0 on x = 0) as 33 worksets of
//     1000 cells, Residual then Jacobian, evaluators built from validated parameter lists,
//     containers found by key, Dirichlet rows from the BC field manager;
//  2. an uploaded mesh of two element blocks (one volume field manager each, worksets of 100
//     cells, transient seeds): both blocks add into one container;
//  3. the lifecycle and parameter errors the reference throws for;
//  4. one Newton step through the ModelEvaluator (txa_model_eval, txa_spmv): W against the
//     oracle's matrix, the linear problem solved to 1e-9 of its first residual.
// Prints "TEST PASSED", the reference harness's pass marker.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "../../include/txa_panzer.hpp"
#include "../../oracle/oracle.h"

using namespace txa::panzer;

namespace {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  hip_check(hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size())), "hipMalloc");
  if (!h.empty()) hip_check(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return d;
}

template <class T>
std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> h(n);
  hip_check(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return h;
}

double *device_zeros(size_t n) {
  double *d = nullptr;
  hip_check(hipMalloc(&d, sizeof(double) * (n ? n : 1)), "hipMalloc");
  hip_check(hipMemset(d, 0, sizeof(double) * (n ? n : 1)), "hipMemset");
  return d;
}

double rel_err(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num = std::fmax(num, std::fabs(a[i] - b[i]));
    den = std::fmax(den, std::fabs(b[i]));
  }
  return num / den;
}

// deterministic pseudo-random values in [-1, 1] (SplitMix64)
double pr(uint64_t i, uint64_t seed) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return 2.0 * static_cast<double>(z >> 11) / 9007199254740992.0 - 1.0;
}

std::vector<double> brick_coords(int nx, int ny, int nz, double jitter, uint64_t seed) {
  const int64_t nv = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
  std::vector<double> X(3 * nv);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const int64_t n = i + (nx + 1) * (j + static_cast<int64_t>(ny + 1) * k);
        const bool inner = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
        const double h[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) X[3 * n + d] = ijk[d] * h[d] + (inner ? jitter * h[d] * pr(3 * n + d, seed) : 0.0);
      }
  return X;
}

// The oracle on a hex8 connectivity: DOF map, graph, and a (block-by-block) assembly.
struct OracleSystem {
  std::vector<int64_t> owned, rowptr;
  std::vector<int32_t> lids, col;
  int64_t nd = 0, nnz = 0;
  OracleSystem(const std::vector<int64_t> &conn) {
    const int64_t ne = static_cast<int64_t>(conn.size()) / 8;
    std::vector<int64_t> eg(ne * 8);
    owned.resize(ne * 8);
    lids.resize(ne * 8);
    nd = orc_dofmap(ne, 8, conn.data(), 1, eg.data(), owned.data(), lids.data());
    owned.resize(nd);
    rowptr.resize(nd + 1);
    nnz = orc_graph_rowptr(ne, 8, lids.data(), nd, rowptr.data());
    col.resize(nnz);
    orc_graph_cols(ne, 8, lids.data(), nd, rowptr.data(), col.data());
  }
  // adds the assembly of elements [e0, e1) of conn into f / vals
  void assemble(const std::vector<int64_t> &conn, const std::vector<double> &X, const std::vector<double> &k,
                int64_t e0, int64_t e1, const orc_state &st, double rho, std::vector<double> &f,
                std::vector<double> &vals) const {
    orc_problem op{};
    op.topo = ORC_HEX8; op.physics = ORC_DIFFUSION; op.nelem = e1 - e0; op.nvert = 8;
    op.elem_vert = conn.data() + 8 * e0; op.coords = X.data(); op.ndof = 8; op.elem_lid = lids.data() + 8 * e0;
    op.coef = k.data() + e0; op.rho = rho; op.src[0] = 1.0; op.workset_size = 1000; op.nthreads = 1;
    op.rowptr = rowptr.data(); op.col = col.data();
    orc_ctx *oc = orc_setup(&op);
    if (!oc || orc_evaluate(oc, &st, f.data(), vals.data())) throw std::runtime_error("oracle evaluation failed");
    orc_free(oc);
  }
};

template <class E>
bool throws_logic(E &&fn) {
  try {
    fn();
  } catch (const std::logic_error &) {
    return true;
  }
  return false;
}

template <class EvalT>
std::shared_ptr<GpuVolumeAssembly<EvalT>> volume(const std::shared_ptr<const DOFManager> &dm, const std::string &block,
                                                const DeviceValues &k, double rho) {
  ParameterList p;
  p.set("DOF Manager", dm);
  p.set("Element Block ID", block);
  p.set("Physics", std::string("Diffusion"));
  p.set("Coefficient", k);
  p.set("Density", rho);
  p.set("Source", std::vector<double>{1.0});
  return std::make_shared<GpuVolumeAssembly<EvalT>>(p);
}

template <class EvalT>
std::shared_ptr<GpuDirichletRows<EvalT>> dirichlet(const std::shared_ptr<const DOFManager> &dm, const DeviceLIDs &l) {
  ParameterList p;
  p.set("DOF Manager", dm);
  p.set("Dirichlet LIDs", l);
  return std::make_shared<GpuDirichletRows<EvalT>>(p);
}

// LIDs of the vertices with x index 0 of an nx-wide brick, sorted
std::vector<int32_t> left_lids(const std::vector<int64_t> &conn, const std::vector<int32_t> &lids, int nx) {
  std::vector<int32_t> bc;
  for (size_t i = 0; i < conn.size(); ++i)
    if (conn[i] % (nx + 1) == 0) bc.push_back(lids[i]);
  std::sort(bc.begin(), bc.end());
  bc.erase(std::unique(bc.begin(), bc.end()), bc.end());
  return bc;
}

// ------------------------------------------------------------------ 1. Cfg1 in worksets of 1000
bool cfg1_worksets() {
  const int n = 32;
  std::vector<double> X = brick_coords(n, n, n, 0.0, 1);
  auto dm = std::make_shared<const DOFManager>(0, TXA_HEX8, n, n, n, X, 1);
  int64_t sz[6];
  orc_mesh_sizes(ORC_HEX8, n, n, n, sz);
  std::vector<int64_t> conn(sz[0] * 8), ev(sz[0] * 8);
  orc_mesh_build(ORC_HEX8, n, n, n, conn.data(), ev.data());
  OracleSystem o(conn);
  std::vector<int64_t> owned, rowptr;
  std::vector<int32_t> lids, col;
  dm->getOwnedIndices(owned);
  dm->getElementLIDs(lids);
  dm->getGraph(rowptr, col);
  const bool graph_ok = owned == o.owned && lids == o.lids && rowptr == o.rowptr && col == o.col &&
                        o.nd == 35937 && o.nnz == 912673;

  const int64_t ne = dm->numElements();
  std::vector<double> k(ne), x(o.nd);
  for (int64_t e = 0; e < ne; ++e) k[e] = std::exp(0.5 * pr(e, 11));
  for (int64_t i = 0; i < o.nd; ++i) x[i] = pr(i, 12);
  const std::vector<int32_t> bc = left_lids(conn, o.lids, n);
  const DeviceValues dk{to_device(k), ne};
  const DeviceLIDs dbc{to_device(bc), static_cast<int64_t>(bc.size())};

  std::vector<std::string> ids;
  dm->getElementBlockIds(ids);
  auto fm = std::make_shared<FieldManager>();
  fm->registerEvaluator<Traits::Residual>(volume<Traits::Residual>(dm, ids[0], dk, 1.0));
  fm->registerEvaluator<Traits::Jacobian>(volume<Traits::Jacobian>(dm, ids[0], dk, 1.0));
  auto ws = buildWorksets(*dm, ids[0], 1000);
  Traits::SD sd;
  sd.worksets_ = ws;
  fm->postRegistrationSetup(sd);
  auto bcs = std::make_shared<FieldManager>();
  bcs->registerEvaluator<Traits::Residual>(dirichlet<Traits::Residual>(dm, dbc));
  bcs->registerEvaluator<Traits::Jacobian>(dirichlet<Traits::Jacobian>(dm, dbc));
  bcs->postRegistrationSetup(Traits::SD());
  AssemblyEngine ae(dm, {{fm, ws}}, bcs);

  auto c = std::make_shared<LinearObjContainer>();
  c->x = to_device(x);
  c->f = device_zeros(o.nd);
  c->A = device_zeros(o.nnz);
  AssemblyEngineInArgs in;
  in.container = c;
  ae.evaluate<Traits::Residual>(in);
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f_res = to_host(c->f, o.nd);
  ae.evaluate<Traits::Jacobian>(in);
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f_jac = to_host(c->f, o.nd), A = to_host(c->A, o.nnz);

  orc_state st{};
  st.x = x.data(); st.beta = 1.0; st.mode = ORC_MODE_FAD; st.want_jac = 1;
  std::vector<double> fr(o.nd, 0.0), vr(o.nnz, 0.0);
  o.assemble(conn, X, k, 0, ne, st, 1.0, fr, vr);
  orc_dirichlet(static_cast<int64_t>(bc.size()), bc.data(), nullptr, x.data(), o.rowptr.data(), o.col.data(),
                fr.data(), vr.data());
  const double e1 = rel_err(f_res, fr), e2 = rel_err(f_jac, fr), e3 = rel_err(A, vr);
  std::printf("[1] 32^3 in %zu worksets of 1000: graph %s, residual eval %.2e, jacobian eval f %.2e A %.2e\n",
              ws->size(), graph_ok ? "bit-exact" : "DIFFERS", e1, e2, e3);
  return graph_ok && ws->size() == 33 && e1 <= 1e-12 && e2 <= 1e-12 && e3 <= 1e-12;
}

// ------------------------------------------------------------------ 2. two element blocks
bool two_blocks() {
  const int n = 10;
  std::vector<double> X = brick_coords(n, n, n, 0.1, 2);
  int64_t sz[6];
  orc_mesh_sizes(ORC_HEX8, n, n, n, sz);
  std::vector<int64_t> brick(sz[0] * 8), ev(sz[0] * 8);
  orc_mesh_build(ORC_HEX8, n, n, n, brick.data(), ev.data());
  // block "lower": cells with z < n/2 in brick order; block "upper": the rest, in reverse order
  std::vector<ElementBlock> blocks(2);
  blocks[0] = {"lower", TXA_HEX8, {}};
  blocks[1] = {"upper", TXA_HEX8, {}};
  const int64_t ne = sz[0], half = ne / 2;
  for (int64_t e = 0; e < half; ++e) blocks[0].conn.insert(blocks[0].conn.end(), &brick[8 * e], &brick[8 * e + 8]);
  for (int64_t e = ne - 1; e >= half; --e)
    blocks[1].conn.insert(blocks[1].conn.end(), &brick[8 * e], &brick[8 * e + 8]);
  std::vector<int64_t> conn = blocks[0].conn;
  conn.insert(conn.end(), blocks[1].conn.begin(), blocks[1].conn.end());
  auto dm = std::make_shared<const DOFManager>(0, blocks, X, 1);
  OracleSystem o(conn);   // first touch over the blocks in order = over the concatenation
  std::vector<int64_t> owned, rowptr;
  std::vector<int32_t> lids, col;
  dm->getOwnedIndices(owned);
  dm->getElementLIDs(lids);
  dm->getGraph(rowptr, col);
  const bool graph_ok = owned == o.owned && lids == o.lids && rowptr == o.rowptr && col == o.col;

  std::vector<double> k(ne), x(o.nd), xd(o.nd);
  for (int64_t e = 0; e < ne; ++e) k[e] = std::exp(0.5 * pr(e, 21));
  for (int64_t i = 0; i < o.nd; ++i) { x[i] = pr(i, 22); xd[i] = pr(i, 23); }
  const std::vector<int32_t> bc = left_lids(conn, o.lids, n);
  const DeviceValues dk{to_device(k), ne};
  const DeviceLIDs dbc{to_device(bc), static_cast<int64_t>(bc.size())};
  std::vector<AssemblyEngine::VolumeBlock> vol;
  for (const char *id : {"lower", "upper"}) {
    auto fm = std::make_shared<FieldManager>();
    fm->registerEvaluator<Traits::Residual>(volume<Traits::Residual>(dm, id, dk, 2.0));
    fm->registerEvaluator<Traits::Jacobian>(volume<Traits::Jacobian>(dm, id, dk, 2.0));
    auto ws = buildWorksets(*dm, id, 100);
    Traits::SD sd;
    sd.worksets_ = ws;
    fm->postRegistrationSetup(sd);
    vol.push_back({fm, ws});
  }
  auto bcs = std::make_shared<FieldManager>();
  bcs->registerEvaluator<Traits::Jacobian>(dirichlet<Traits::Jacobian>(dm, dbc));
  AssemblyEngine ae(dm, vol, bcs);
  auto c = std::make_shared<LinearObjContainer>();
  c->x = to_device(x);
  c->dxdt = to_device(xd);
  c->f = device_zeros(o.nd);
  c->A = device_zeros(o.nnz);
  AssemblyEngineInArgs in;
  in.container = c;
  in.alpha = 3.0;
  in.beta = 0.5;
  in.evaluate_transient_terms = true;
  ae.evaluate<Traits::Jacobian>(in);
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f = to_host(c->f, o.nd), A = to_host(c->A, o.nnz);

  orc_state st{};
  st.x = x.data(); st.xdot = xd.data(); st.alpha = 3.0; st.beta = 0.5; st.transient = 1;
  st.mode = ORC_MODE_EXPLICIT; st.want_jac = 1;
  std::vector<double> fr(o.nd, 0.0), vr(o.nnz, 0.0);
  o.assemble(conn, X, k, 0, half, st, 2.0, fr, vr);
  o.assemble(conn, X, k, half, ne, st, 2.0, fr, vr);
  orc_dirichlet(static_cast<int64_t>(bc.size()), bc.data(), nullptr, x.data(), o.rowptr.data(), o.col.data(),
                fr.data(), vr.data());
  const double e1 = rel_err(f, fr), e2 = rel_err(A, vr);
  std::printf("[2] two element blocks, worksets of 100: graph %s, f %.2e A %.2e\n", graph_ok ? "bit-exact" : "DIFFERS",
              e1, e2);
  return graph_ok && e1 <= 1e-12 && e2 <= 1e-12;
}

// ------------------------------------------------------------------ 3. contract errors
bool contract_errors() {
  const int n = 4;
  auto dm = std::make_shared<const DOFManager>(0, TXA_HEX8, n, n, n, brick_coords(n, n, n, 0.0, 3), 1);
  std::vector<std::string> ids;
  dm->getElementBlockIds(ids);
  const DeviceValues dk{to_device(std::vector<double>(dm->numElements(), 1.0)), dm->numElements()};
  bool ok = true;
  // an unknown parameter name, a parameter of the wrong type
  ok &= throws_logic([&] {
    ParameterList p;
    p.set("DOF Manager", dm);
    p.set("Element Block ID", ids[0]);
    p.set("Coefficient", dk);
    p.set("Bogus Entry", 1.0);
    GpuVolumeAssembly<Traits::Residual> e(p);
  });
  ok &= throws_logic([&] {
    ParameterList p;
    p.set("DOF Manager", dm);
    p.set("Element Block ID", ids[0]);
    p.set("Coefficient", dk);
    p.set("Density", 1);   // int, not double
    GpuVolumeAssembly<Traits::Residual> e(p);
  });
  auto ev = volume<Traits::Residual>(dm, ids[0], dk, 1.0);
  auto c = std::make_shared<LinearObjContainer>();
  Traits::PED ped;
  ped.gedc->addDataObject("Residual Scatter Container", c);
  ped.gedc->addDataObject("Solution Gather Container", c);
  auto ws = buildWorksets(*dm, ids[0], 10);
  // preEvaluate before postRegistrationSetup; evaluateFields outside preEvaluate/postEvaluate
  ok &= throws_logic([&] { ev->preEvaluate(ped); });
  FieldManager fm;
  Traits::SD sd;
  sd.worksets_ = ws;
  ev->postRegistrationSetup(sd, fm);
  ok &= throws_logic([&] { ev->evaluateFields((*ws)[0]); });
  // worksets that leave cells out, or list one twice
  ev->preEvaluate(ped);
  for (size_t i = 0; i + 1 < ws->size(); ++i) ev->evaluateFields((*ws)[i]);
  ok &= throws_logic([&] { ev->postEvaluate(nullptr); });
  ev->preEvaluate(ped);
  ev->evaluateFields((*ws)[0]);
  ok &= throws_logic([&] { ev->evaluateFields((*ws)[0]); });
  // a container key that is missing
  Traits::PED empty;
  ok &= throws_logic([&] { ev->preEvaluate(empty); });
  std::printf("[3] contract errors thrown: %s\n", ok ? "all" : "MISSING");
  return ok;
}

// ------------------------------------------------------------------ 4. Newton step through the ModelEvaluator
// Jacobi-preconditioned BiCGStab on the host for the Newton update (the matrix from the device)
std::vector<double> bicgstab(const std::vector<int64_t> &rp, const std::vector<int32_t> &col,
                             const std::vector<double> &A, const std::vector<double> &b, double tol) {
  const size_t n = b.size();
  std::vector<double> dinv(n, 1.0);
  for (size_t i = 0; i < n; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == static_cast<int32_t>(i) && A[k] != 0.0) dinv[i] = 1.0 / A[k];
  auto mv = [&](const std::vector<double> &v) {
    std::vector<double> y(n, 0.0);
    for (size_t i = 0; i < n; ++i)
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) y[i] += A[k] * v[col[k]];
    return y;
  };
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &c) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  std::vector<double> x(n, 0.0), r = b, r0 = b, p(n, 0.0), v(n, 0.0), ph(n), sh(n), s(n), t;
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  const double nb = std::sqrt(dot(b, b));
  for (int it = 0; it < 5000; ++it) {
    const double rn = dot(r0, r), beta = (rn / rho) * (alpha / omega);
    for (size_t i = 0; i < n; ++i) { p[i] = r[i] + beta * (p[i] - omega * v[i]); ph[i] = dinv[i] * p[i]; }
    v = mv(ph);
    alpha = rn / dot(r0, v);
    for (size_t i = 0; i < n; ++i) { s[i] = r[i] - alpha * v[i]; sh[i] = dinv[i] * s[i]; }
    t = mv(sh);
    omega = dot(t, s) / std::fmax(dot(t, t), 1e-300);
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * ph[i] + omega * sh[i]; r[i] = s[i] - omega * t[i]; }
    rho = rn;
    if (std::sqrt(dot(r, r)) <= tol * nb) break;
  }
  return x;
}

bool newton_step() {
  const int n = 14;
  std::vector<double> X = brick_coords(n, n, n, 0.1, 4);
  auto dm = std::make_shared<const DOFManager>(0, TXA_HEX8, n, n, n, X, 1);
  int64_t sz[6];
  orc_mesh_sizes(ORC_HEX8, n, n, n, sz);
  std::vector<int64_t> conn(sz[0] * 8), ev(sz[0] * 8);
  orc_mesh_build(ORC_HEX8, n, n, n, conn.data(), ev.data());
  OracleSystem o(conn);
  const int64_t ne = dm->numElements();
  std::vector<double> k(ne), x0(o.nd);
  for (int64_t e = 0; e < ne; ++e) k[e] = std::exp(0.5 * pr(e, 31));
  for (int64_t i = 0; i < o.nd; ++i) x0[i] = pr(i, 32);
  const std::vector<int32_t> bc = left_lids(conn, o.lids, n);
  txa_model_desc d{};
  d.physics = TXA_DIFFUSION;
  d.precision = TXA_FP64;
  d.coef = to_device(k);
  d.rho = 1.0;
  d.src[0] = 1.0;
  d.bc_lids = to_device(bc);
  d.n_bc = static_cast<int64_t>(bc.size());
  ModelEvaluator me(dm, d);
  double *dx = to_device(x0), *df = device_zeros(o.nd), *dW = device_zeros(o.nnz), *dy = device_zeros(o.nd);
  ModelEvaluator::InArgs in;
  in.x = dx;
  me.evalModel(in, {df, dW});
  // W * x0 through txa_spmv against the host product
  check(dm->context(), txa_spmv(dm->context(), dm->mesh(), dW, dx, dy, nullptr));
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f0 = to_host(df, o.nd), W = to_host(dW, o.nnz), y = to_host(dy, o.nd);
  std::vector<double> yh(o.nd, 0.0);
  for (int64_t i = 0; i < o.nd; ++i)
    for (int64_t q = o.rowptr[i]; q < o.rowptr[i + 1]; ++q) yh[i] += W[q] * x0[o.col[q]];
  orc_state st{};
  st.x = x0.data(); st.beta = 1.0; st.mode = ORC_MODE_EXPLICIT; st.want_jac = 1;
  std::vector<double> fr(o.nd, 0.0), vr(o.nnz, 0.0);
  o.assemble(conn, X, k, 0, ne, st, 1.0, fr, vr);
  orc_dirichlet(static_cast<int64_t>(bc.size()), bc.data(), nullptr, x0.data(), o.rowptr.data(), o.col.data(),
                fr.data(), vr.data());
  const double ew = rel_err(W, vr), ef = rel_err(f0, fr), ey = rel_err(y, yh);
  // x1 = x0 - W^-1 f(x0); the problem is linear, so f(x1) is the linear solver's residual
  std::vector<double> rhs(o.nd);
  for (int64_t i = 0; i < o.nd; ++i) rhs[i] = -f0[i];
  const std::vector<double> dxh = bicgstab(o.rowptr, o.col, W, rhs, 1e-13);
  std::vector<double> x1(o.nd);
  for (int64_t i = 0; i < o.nd; ++i) x1[i] = x0[i] + dxh[i];
  double *dx1 = to_device(x1);
  in.x = dx1;
  me.evalModel(in, {df, nullptr});   // the Residual evaluation alone
  hip_check(hipDeviceSynchronize(), "sync");
  const std::vector<double> f1 = to_host(df, o.nd);
  double r0 = 0.0, r1 = 0.0;
  for (int64_t i = 0; i < o.nd; ++i) { r0 = std::fmax(r0, std::fabs(f0[i])); r1 = std::fmax(r1, std::fabs(f1[i])); }
  // a model without transient support refuses x_dot
  in.x_dot = dx;
  const bool refused = throws_logic([&] { me.evalModel(in, {df, nullptr}); });   // std::invalid_argument
  std::printf("[4] ModelEvaluator: W %.2e f %.2e vs oracle, txa_spmv %.2e; Newton step |f| %.3e -> %.3e; x_dot "
              "refused: %d\n", ew, ef, ey, r0, r1, static_cast<int>(refused));
  return ew <= 1e-12 && ef <= 1e-12 && ey <= 1e-14 && r1 <= 1e-9 * r0 && refused;
}

}  // namespace

int main() {
  bool ok = true;
  try {
    ok &= cfg1_worksets();
    ok &= two_blocks();
    ok &= contract_errors();
    ok &= newton_step();
  } catch (const std::exception &e) {
    std::printf("exception: %s\n", e.what());
    ok = false;
  }
  std::printf(ok ? "TEST PASSED\n" : "TEST FAILED\n");
  return ok ? 0 : 1;
}
