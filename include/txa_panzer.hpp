// txa_panzer.hpp -- C++ layer above the C ABI that mirrors the Panzer evaluator
// contract for the volume residual + Jacobian path, so the GPU path drops in
// beside the plain evaluators:
//
//   evaluation types     panzer::Traits::Residual / Jacobian      (disc-fe/src/Panzer_Traits.hpp:95-106)
//   per-workset inputs   panzer::WorksetDetails (num_cells, block_id, alpha, beta, time,
//                        step_size, evaluate_transient_terms)       (disc-fe/src/Panzer_Workset.hpp:100-136, 361-367)
//   global inputs        panzer::AssemblyEngineInArgs              (disc-fe/src/Panzer_AssemblyEngine_InArgs.hpp:95-106)
//   DOF numbering        panzer::GlobalIndexer                     (dof-mgr/src/Panzer_GlobalIndexer.hpp:60-602)
//   linear objects       panzer::TpetraLinearObjContainer get_x/get_dxdt/get_f/get_A
//                                                                   (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:110-123)
//   evaluator lifecycle  ctor(parameters), postRegistrationSetup, evaluateFields(workset)
//                                                                   (disc-fe/src/Panzer_Traits.hpp:125-142)
//
// One GpuVolumeAssembly<EvalT> stands in for the Gather -> DOF/DOFGradient ->
// Integrator_* -> ScatterResidual_Tpetra chain of one element block; the
// workset it receives covers the whole block (one launch per block).  Errors
// from the C ABI are rethrown as std::invalid_argument / std::logic_error /
// std::runtime_error, as the reference's TEUCHOS_TEST_FOR_EXCEPTION does.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "txa.h"

namespace txa {
namespace panzer {

struct Traits {
  struct Residual {};
  struct Jacobian {};
};

inline void check(txa_ctx ctx, int rc) {
  if (rc == TXA_OK) return;
  const std::string msg = std::string("txa: ") + txa_last_error(ctx);
  if (rc == TXA_ERR_INVALID) throw std::invalid_argument(msg);
  if (rc == TXA_ERR_LOGIC || rc == TXA_ERR_UNSUPPORTED) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

// Members of panzer::Workset read by the path.
struct Workset {
  int num_cells = 0;
  std::string block_id = "eblock-0_0_0";
  double alpha = 0.0;       // d/d(xdot) seed
  double beta = 1.0;        // d/dx seed
  double gamma2 = 0.0;      // d/d(xddot) coefficient (Tianxin's second time derivative)
  double time = 0.0;
  double step_size = 0.0;
  bool evaluate_transient_terms = false;
};

// Owns the context and one structured element block with its DOF map and CRS graph.
class BrickDOFManager {
 public:
  BrickDOFManager(int device, int topo, int nx, int ny, int nz, const std::vector<double> &vertex_coords,
                  int nfields) {
    check(nullptr, txa_create(device, &ctx_));
    try {
      check(ctx_, txa_mesh_brick(ctx_, topo, nx, ny, nz, vertex_coords.data(), &mesh_));
      check(ctx_, txa_dofmap_build(ctx_, mesh_, nfields, &ndofs_));
      check(ctx_, txa_graph_build(ctx_, mesh_, &nnz_));
      int64_t info[3];
      check(ctx_, txa_mesh_info(ctx_, mesh_, info));
      nelem_ = info[0];
      ndof_elem_ = static_cast<int>(info[1]) * nfields;
    } catch (...) {
      txa_destroy(ctx_);
      throw;
    }
  }
  ~BrickDOFManager() { txa_destroy(ctx_); }
  BrickDOFManager(const BrickDOFManager &) = delete;
  BrickDOFManager &operator=(const BrickDOFManager &) = delete;

  txa_ctx context() const { return ctx_; }
  int mesh() const { return mesh_; }
  int64_t numOwned() const { return ndofs_; }
  int64_t numElements() const { return nelem_; }
  int elementDofCount() const { return ndof_elem_; }
  int64_t nnz() const { return nnz_; }

  // GlobalIndexer::getOwnedIndices (GIDs in LID order)
  void getOwnedIndices(std::vector<int64_t> &gids) const {
    gids.resize(ndofs_);
    check(ctx_, txa_dofmap_copy(ctx_, mesh_, gids.data(), nullptr));
  }
  // GlobalIndexer::getElementLIDs for every element, [elements][ndof_elem]
  void getElementLIDs(std::vector<int32_t> &lids) const {
    lids.resize(static_cast<size_t>(nelem_) * ndof_elem_);
    check(ctx_, txa_dofmap_copy(ctx_, mesh_, nullptr, lids.data()));
  }
  // the CRS graph of TpetraLinearObjFactory::buildGraph, host copy
  void getGraph(std::vector<int64_t> &rowptr, std::vector<int32_t> &col) const {
    rowptr.resize(ndofs_ + 1);
    col.resize(nnz_);
    check(ctx_, txa_graph_copy(ctx_, mesh_, rowptr.data(), col.data()));
  }

 private:
  txa_ctx ctx_ = nullptr;
  int mesh_ = 0;
  int64_t ndofs_ = 0, nnz_ = 0, nelem_ = 0;
  int ndof_elem_ = 0;
};

// Device arrays of TpetraLinearObjContainer (x, dxdt, d2xdt2, f, local matrix values).
struct LinearObjContainer {
  const double *x = nullptr;
  const double *dxdt = nullptr;
  const double *d2xdt2 = nullptr;
  double *f = nullptr;
  double *A_vals = nullptr;
};

// Closure data the equation set would attach (per-element coefficient field,
// material constants, source) and the Dirichlet rows of the block.
struct VolumeParams {
  int physics = TXA_DIFFUSION;
  int precision = TXA_FP64;
  const double *coef = nullptr;   // device [elements]
  double nu = 0.3, rho = 1.0;
  double src[3] = {0.0, 0.0, 0.0};
  double rayleigh_a = 0.0, rayleigh_b = 0.0;
  const int32_t *dirichlet_lids = nullptr;   // device
  const double *dirichlet_vals = nullptr;    // device or null (= 0)
  int64_t n_dirichlet = 0;
};

template <typename EvalT>
class GpuVolumeAssembly {
  static_assert(std::is_same<EvalT, Traits::Residual>::value || std::is_same<EvalT, Traits::Jacobian>::value,
                "EvalT must be Traits::Residual or Traits::Jacobian");

 public:
  GpuVolumeAssembly(const BrickDOFManager &dm, const VolumeParams &params) : dm_(dm), prm_(params) {
    if (!params.coef) throw std::invalid_argument("GpuVolumeAssembly: coefficient field is required");
  }

  // Nothing to bind lazily: the DOF map and graph are resident in the manager.
  void postRegistrationSetup() { ready_ = true; }

  void evaluateFields(const Workset &ws, LinearObjContainer &c, void *stream = nullptr) {
    if (!ready_) throw std::logic_error("GpuVolumeAssembly: evaluateFields before postRegistrationSetup");
    if (ws.num_cells != dm_.numElements())
      throw std::invalid_argument("GpuVolumeAssembly: the workset must cover the element block");
    txa_assembly_args a{};
    a.physics = prm_.physics;
    a.precision = prm_.precision;
    a.coef = prm_.coef;
    a.nu = prm_.nu;
    a.rho = prm_.rho;
    for (int i = 0; i < 3; ++i) a.src[i] = prm_.src[i];
    a.rayleigh_a = prm_.rayleigh_a;
    a.rayleigh_b = prm_.rayleigh_b;
    a.x = c.x;
    a.xdot = c.dxdt;
    a.xddot = c.d2xdt2;
    a.alpha = ws.alpha;
    a.beta = ws.beta;
    a.gamma2 = ws.gamma2;
    a.transient = ws.evaluate_transient_terms ? 1 : 0;
    a.f = c.f;
    a.vals = std::is_same<EvalT, Traits::Jacobian>::value ? c.A_vals : nullptr;
    a.bc_lids = prm_.dirichlet_lids;
    a.bc_vals = prm_.dirichlet_vals;
    a.n_bc = prm_.n_dirichlet;
    check(dm_.context(), txa_assemble(dm_.context(), dm_.mesh(), &a, stream));
  }

 private:
  const BrickDOFManager &dm_;
  VolumeParams prm_;
  bool ready_ = false;
};

}  // namespace panzer
}  // namespace txa
