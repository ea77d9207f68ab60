// txa_panzer.hpp -- C++ layer above the C ABI that mirrors the Panzer/Phalanx contract
// for the volume residual + Jacobian path, so the GPU path drops in beside the plain
// evaluators:
//
//   evaluation types   panzer::Traits::Residual / Jacobian, SetupData, PreEvalData, EvalData,
//                      PostEvalData                           (disc-fe/src/Panzer_Traits.hpp:95-142)
//   evaluator life     ctor(ParameterList) validated against getValidParameters(),
//                      postRegistrationSetup(SetupData, FieldManager&), preEvaluate(PreEvalData),
//                      evaluateFields(EvalData) per workset, postEvaluate(PostEvalData)
//                      (disc-fe/src/evaluators/Panzer_Integrator_GradBasisDotVector_impl.hpp:148-177,
//                       469-495; disc-fe/src/Panzer_AssemblyEngine_impl.hpp:134-180)
//   containers         GlobalEvaluationDataContainer looked up by "Residual Scatter Container" /
//                      "Solution Gather Container"          (Panzer_AssemblyEngine_impl.hpp:145-146)
//                      holding a TpetraLinearObjContainer's get_x/get_dxdt/get_d2xdt2/get_f/get_A
//                                                           (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:110-123)
//   DOF numbering      panzer::DOFManager over a ConnManager's element blocks
//                                                           (dof-mgr/src/Panzer_DOFManager.cpp:449-715)
//   orchestration      AssemblyEngine::evaluate: zero the container, every block's volume field
//                      manager over its worksets, then the Dirichlet field manager
//                                                           (Panzer_AssemblyEngine_impl.hpp:67-177)
//   solver hand-off    ModelEvaluator::evalModel (x, x_dot, alpha, beta in; f, W out)
//                                                           (disc-fe/src/Panzer_ModelEvaluator_impl.hpp:725-780, 1715-1840)
//
// One GpuVolumeAssembly<EvalT> stands in for the Gather -> DOF/DOFGradient -> Integrator_* ->
// ScatterResidual_Tpetra chain of one element block's field manager.  evaluateFields accounts
// for the cells of each workset (any workset size); the block is assembled by one launch in
// postEvaluate, once the worksets have covered every cell of the block exactly once, and its
// sums are added into the container as the reference's scatter adds.  Errors from the C ABI
// are rethrown as std::invalid_argument / std::logic_error / std::runtime_error, as the
// reference's TEUCHOS_TEST_FOR_EXCEPTION does; ParameterList validation errors are
// std::logic_error (Teuchos' InvalidParameter derives from it).
#pragma once

#include <any>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "txa.h"

namespace txa {
namespace panzer {

inline void check(txa_ctx ctx, int rc) {
  if (rc == TXA_OK) return;
  const std::string msg = std::string("txa: ") + txa_last_error(ctx);
  if (rc == TXA_ERR_INVALID) throw std::invalid_argument(msg);
  if (rc == TXA_ERR_LOGIC || rc == TXA_ERR_UNSUPPORTED) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

// ---------------------------------------------------------------- parameter lists
// Typed named values (Teuchos::ParameterList's get / set / isType / validateParameters).
class ParameterList {
 public:
  template <class T>
  ParameterList &set(const std::string &name, T value) {
    v_[name] = std::any(std::move(value));
    return *this;
  }
  bool isParameter(const std::string &name) const { return v_.count(name) != 0; }
  template <class T>
  bool isType(const std::string &name) const {
    auto it = v_.find(name);
    return it != v_.end() && it->second.type() == typeid(T);
  }
  template <class T>
  T get(const std::string &name) const {
    auto it = v_.find(name);
    if (it == v_.end()) throw std::logic_error("ParameterList: no parameter \"" + name + "\"");
    if (it->second.type() != typeid(T)) throw std::logic_error("ParameterList: \"" + name + "\" has another type");
    return std::any_cast<T>(it->second);
  }
  template <class T>
  T get(const std::string &name, T def) const {
    return isParameter(name) ? get<T>(name) : def;
  }
  // Every entry must be a valid name with the valid entry's type.
  void validateParameters(const ParameterList &valid) const {
    for (const auto &kv : v_) {
      auto it = valid.v_.find(kv.first);
      if (it == valid.v_.end())
        throw std::logic_error("ParameterList: \"" + kv.first + "\" is not a valid parameter");
      if (it->second.type() != kv.second.type())
        throw std::logic_error("ParameterList: \"" + kv.first + "\" has the wrong type");
    }
  }

 private:
  std::map<std::string, std::any> v_;
};

// Device arrays handed in through parameter lists.
struct DeviceValues {
  const double *p = nullptr;
  int64_t n = 0;
};
struct DeviceLIDs {
  const int32_t *p = nullptr;
  int64_t n = 0;
};

// ---------------------------------------------------------------- worksets and traits
// Members of panzer::Workset read by the path (disc-fe/src/Panzer_Workset.hpp:100-136, 361-367).
struct Workset {
  int num_cells = 0;
  std::string block_id;
  std::vector<int64_t> cell_local_ids;   // cells of the workset: element indices within the block
  double alpha = 0.0;       // d/d(xdot) seed
  double beta = 1.0;        // d/dx seed
  double gamma2 = 0.0;      // d/d(xddot) coefficient (Tianxin's second time derivative)
  double time = 0.0;
  double step_size = 0.0;
  bool evaluate_transient_terms = false;
};

class GlobalEvaluationDataContainer;
class FieldManager;

struct Traits {
  struct Residual {};
  struct Jacobian {};
  struct SD {
    std::shared_ptr<const std::vector<Workset>> worksets_;
  };
  using SetupData = const SD &;
  using EvalData = const Workset &;
  struct PED {
    PED();
    std::shared_ptr<GlobalEvaluationDataContainer> gedc;
    std::string first_sensitivities_name, second_sensitivities_name;
  };
  using PreEvalData = const PED &;
  using PostEvalData = void *;
};

// ---------------------------------------------------------------- global evaluation data
class GlobalEvaluationData {
 public:
  virtual ~GlobalEvaluationData() = default;
};

// The device arrays of a TpetraLinearObjContainer: x, dx/dt, d2x/dt2, f and the values of A
// on the DOF manager's graph.
class LinearObjContainer : public GlobalEvaluationData {
 public:
  const double *x = nullptr, *dxdt = nullptr, *d2xdt2 = nullptr;
  double *f = nullptr, *A = nullptr;
  const double *get_x() const { return x; }
  const double *get_dxdt() const { return dxdt; }
  const double *get_d2xdt2() const { return d2xdt2; }
  double *get_f() const { return f; }
  double *get_A() const { return A; }
};

class GlobalEvaluationDataContainer {
 public:
  void addDataObject(const std::string &key, std::shared_ptr<GlobalEvaluationData> ged) { d_[key] = std::move(ged); }
  bool containsDataObject(const std::string &key) const { return d_.count(key) != 0; }
  std::shared_ptr<GlobalEvaluationData> getDataObject(const std::string &key) const {
    auto it = d_.find(key);
    if (it == d_.end()) throw std::logic_error("GlobalEvaluationDataContainer: no data object \"" + key + "\"");
    return it->second;
  }

 private:
  std::map<std::string, std::shared_ptr<GlobalEvaluationData>> d_;
};

inline Traits::PED::PED() : gedc(std::make_shared<GlobalEvaluationDataContainer>()) {}

// ---------------------------------------------------------------- DOF manager
// One element block of a ConnManager: its id, topology (TXA_HEX8, TXA_TET4, TXA_WEDGE6,
// TXA_HEX27) and connectivity [elements][nodes per element] of entity ids.
struct ElementBlock {
  std::string id;
  int topo = TXA_HEX8;
  std::vector<int64_t> conn;
};

// Owns the context and one mesh with its DOF map and CRS graph on the device
// (DOFManager::buildGlobalUnknowns + TpetraLinearObjFactory's graph).
class DOFManager {
 public:
  // a structured brick, one element block "eblock-0_0_0" (CubeHexMeshFactory's block name)
  DOFManager(int device, int topo, int nx, int ny, int nz, const std::vector<double> &vertex_coords, int nfields) {
    init(device, [&](txa_ctx c, int *id) { return txa_mesh_brick(c, topo, nx, ny, nz, vertex_coords.data(), id); },
         nfields);
    ids_ = {"eblock-0_0_0"};
  }
  // element blocks of a ConnManager over vertex_coords [nverts][3] (nentities: bound of the
  // entity ids, 0 = the vertex count)
  DOFManager(int device, const std::vector<ElementBlock> &blocks, const std::vector<double> &vertex_coords,
             int nfields, int64_t nentities = 0) {
    std::vector<txa_element_block> b(blocks.size());
    for (size_t k = 0; k < blocks.size(); ++k) {
      const int npe = blocks[k].topo == TXA_HEX8 ? 8 : blocks[k].topo == TXA_TET4 ? 4
                                                       : blocks[k].topo == TXA_WEDGE6 ? 6 : 27;
      b[k].topo = blocks[k].topo;
      b[k].nelem = static_cast<int64_t>(blocks[k].conn.size()) / npe;
      b[k].conn = blocks[k].conn.data();
      ids_.push_back(blocks[k].id);
    }
    const int64_t nv = static_cast<int64_t>(vertex_coords.size()) / 3;
    init(device, [&](txa_ctx c, int *id) {
      return txa_mesh_upload_blocks(c, static_cast<int>(b.size()), b.data(), nentities > 0 ? nentities : nv, nv,
                                    vertex_coords.data(), id);
    }, nfields);
  }
  ~DOFManager() { txa_destroy(ctx_); }
  DOFManager(const DOFManager &) = delete;
  DOFManager &operator=(const DOFManager &) = delete;

  txa_ctx context() const { return ctx_; }
  int mesh() const { return mesh_; }
  int64_t numOwned() const { return ndofs_; }
  int64_t numElements() const { return nelem_; }
  int64_t nnz() const { return nnz_; }
  // the element block ids, in block order (DOFManager::getElementBlockIds)
  void getElementBlockIds(std::vector<std::string> &ids) const { ids = ids_; }
  int blockIndex(const std::string &id) const {
    for (size_t k = 0; k < ids_.size(); ++k)
      if (ids_[k] == id) return static_cast<int>(k);
    throw std::invalid_argument("DOFManager: no element block \"" + id + "\"");
  }
  int64_t blockElements(const std::string &id) const { return info_[2 + 3 * blockIndex(id)]; }
  int64_t blockFirstElement(const std::string &id) const { return info_[3 + 3 * blockIndex(id)]; }
  int blockTopology(const std::string &id) const { return static_cast<int>(info_[1 + 3 * blockIndex(id)]); }

  // GlobalIndexer::getOwnedIndices (GIDs in LID order)
  void getOwnedIndices(std::vector<int64_t> &gids) const {
    gids.resize(ndofs_);
    check(ctx_, txa_dofmap_copy(ctx_, mesh_, gids.data(), nullptr));
  }
  // GlobalIndexer::getElementLIDs of every element, blocks one after another
  void getElementLIDs(std::vector<int32_t> &lids) const {
    lids.resize(npos_);
    check(ctx_, txa_dofmap_copy(ctx_, mesh_, nullptr, lids.data()));
  }
  // the CRS graph of TpetraLinearObjFactory::buildGraph, host copy
  void getGraph(std::vector<int64_t> &rowptr, std::vector<int32_t> &col) const {
    rowptr.resize(ndofs_ + 1);
    col.resize(nnz_);
    check(ctx_, txa_graph_copy(ctx_, mesh_, rowptr.data(), col.data()));
  }

 private:
  template <class Make>
  void init(int device, Make make, int nfields) {
    check(nullptr, txa_create(device, &ctx_));
    try {
      check(ctx_, make(ctx_, &mesh_));
      check(ctx_, txa_dofmap_build(ctx_, mesh_, nfields, &ndofs_));
      check(ctx_, txa_graph_build(ctx_, mesh_, &nnz_));
      int64_t mi[3];
      check(ctx_, txa_mesh_info(ctx_, mesh_, mi));
      nelem_ = mi[0];
      info_.assign(1 + 3 * 64, 0);
      check(ctx_, txa_mesh_blocks(ctx_, mesh_, info_.data(), static_cast<int64_t>(info_.size())));
      npos_ = 0;
      for (int64_t k = 0; k < info_[0]; ++k) {
        const int t = static_cast<int>(info_[1 + 3 * k]);
        const int npe = t == TXA_HEX8 ? 8 : t == TXA_TET4 ? 4 : t == TXA_WEDGE6 ? 6 : 27;
        npos_ += info_[2 + 3 * k] * npe * nfields;
      }
    } catch (...) {
      txa_destroy(ctx_);
      throw;
    }
  }
  txa_ctx ctx_ = nullptr;
  int mesh_ = 0;
  int64_t ndofs_ = 0, nnz_ = 0, nelem_ = 0, npos_ = 0;
  std::vector<int64_t> info_;
  std::vector<std::string> ids_;
};

// ---------------------------------------------------------------- evaluators
template <typename EvalT>
class Evaluator {
 public:
  virtual ~Evaluator() = default;
  virtual void postRegistrationSetup(Traits::SetupData sd, FieldManager &fm) = 0;
  virtual void preEvaluate(Traits::PreEvalData) {}
  virtual void evaluateFields(Traits::EvalData workset) = 0;
  virtual void postEvaluate(Traits::PostEvalData) {}
};

// The evaluators of one field manager (PHX::FieldManager<panzer::Traits>), by evaluation type.
class FieldManager {
 public:
  template <typename EvalT>
  void registerEvaluator(std::shared_ptr<Evaluator<EvalT>> e) {
    list<EvalT>().push_back(std::move(e));
  }
  void postRegistrationSetup(Traits::SetupData sd) {
    for (auto &e : res_) e->postRegistrationSetup(sd, *this);
    for (auto &e : jac_) e->postRegistrationSetup(sd, *this);
  }
  template <typename EvalT>
  void preEvaluate(Traits::PreEvalData d) {
    for (auto &e : list<EvalT>()) e->preEvaluate(d);
  }
  template <typename EvalT>
  void evaluateFields(Traits::EvalData w) {
    for (auto &e : list<EvalT>()) e->evaluateFields(w);
  }
  template <typename EvalT>
  void postEvaluate(Traits::PostEvalData d) {
    for (auto &e : list<EvalT>()) e->postEvaluate(d);
  }

 private:
  template <typename EvalT>
  std::vector<std::shared_ptr<Evaluator<EvalT>>> &list() {
    if constexpr (std::is_same<EvalT, Traits::Residual>::value) return res_;
    else return jac_;
  }
  std::vector<std::shared_ptr<Evaluator<Traits::Residual>>> res_;
  std::vector<std::shared_ptr<Evaluator<Traits::Jacobian>>> jac_;
};

// Volume residual (+ Jacobian) of one element block on the GPU.
template <typename EvalT>
class GpuVolumeAssembly : public Evaluator<EvalT> {
  static_assert(std::is_same<EvalT, Traits::Residual>::value || std::is_same<EvalT, Traits::Jacobian>::value,
                "EvalT must be Traits::Residual or Traits::Jacobian");

 public:
  explicit GpuVolumeAssembly(const ParameterList &p) {
    p.validateParameters(getValidParameters());
    dm_ = p.get<std::shared_ptr<const DOFManager>>("DOF Manager");
    if (!dm_) throw std::invalid_argument("GpuVolumeAssembly: \"DOF Manager\" is required");
    block_ = p.get<std::string>("Element Block ID");
    bidx_ = dm_->blockIndex(block_);
    const std::string phys = p.get<std::string>("Physics", "Diffusion");
    if (phys != "Diffusion" && phys != "Elasticity")
      throw std::invalid_argument("GpuVolumeAssembly: \"Physics\" is Diffusion or Elasticity");
    physics_ = phys == "Diffusion" ? TXA_DIFFUSION : TXA_ELASTICITY;
    const std::string prec = p.get<std::string>("Precision", "FP64");
    precision_ = prec == "FP64" ? TXA_FP64 : prec == "FP32" ? TXA_FP32 : prec == "Split BF16" ? TXA_SPLIT_BF16 : -1;
    if (precision_ < 0) throw std::invalid_argument("GpuVolumeAssembly: \"Precision\" is FP64, FP32 or Split BF16");
    coef_ = p.get<DeviceValues>("Coefficient");
    if (!coef_.p || coef_.n != dm_->numElements())
      throw std::invalid_argument("GpuVolumeAssembly: \"Coefficient\" needs one value per element of the mesh");
    nu_ = p.get<double>("Poisson Ratio", 0.3);
    rho_ = p.get<double>("Density", 1.0);
    const std::vector<double> src = p.get<std::vector<double>>("Source", std::vector<double>());
    for (size_t i = 0; i < src.size() && i < 3; ++i) src_[i] = src[i];
    ra_ = p.get<double>("Rayleigh a", 0.0);
    rb_ = p.get<double>("Rayleigh b", 0.0);
    scatter_ = p.get<std::string>("Scatter Container Name", "Residual Scatter Container");
    gather_ = p.get<std::string>("Gather Container Name", "Solution Gather Container");
    stream_ = p.get<void *>("Stream", nullptr);
  }

  ParameterList getValidParameters() const {
    ParameterList p;
    p.set("DOF Manager", std::shared_ptr<const DOFManager>());
    p.set("Element Block ID", std::string("?"));
    p.set("Physics", std::string("Diffusion"));
    p.set("Precision", std::string("FP64"));
    p.set("Coefficient", DeviceValues());
    p.set("Poisson Ratio", 0.3);
    p.set("Density", 1.0);
    p.set("Source", std::vector<double>());
    p.set("Rayleigh a", 0.0);
    p.set("Rayleigh b", 0.0);
    p.set("Scatter Container Name", std::string("Residual Scatter Container"));
    p.set("Gather Container Name", std::string("Solution Gather Container"));
    p.set("Stream", static_cast<void *>(nullptr));
    return p;
  }

  // The field manager's worksets must be this block's and cover each of its cells once.
  void postRegistrationSetup(Traits::SetupData sd, FieldManager &) override {
    if (sd.worksets_) {
      reset();
      for (const Workset &w : *sd.worksets_) account(w, false);
      require_covered("postRegistrationSetup");
    }
    ready_ = true;
  }

  void preEvaluate(Traits::PreEvalData d) override {
    if (!ready_) throw std::logic_error("GpuVolumeAssembly: preEvaluate before postRegistrationSetup");
    if (!d.gedc) throw std::invalid_argument("GpuVolumeAssembly: no GlobalEvaluationDataContainer");
    scatter_c_ = std::dynamic_pointer_cast<LinearObjContainer>(d.gedc->getDataObject(scatter_));
    gather_c_ = std::dynamic_pointer_cast<LinearObjContainer>(d.gedc->getDataObject(gather_));
    if (!scatter_c_ || !gather_c_)
      throw std::logic_error("GpuVolumeAssembly: the containers are not LinearObjContainers");
    reset();
    open_ = true;
  }

  void evaluateFields(Traits::EvalData w) override {
    if (!open_) throw std::logic_error("GpuVolumeAssembly: evaluateFields outside preEvaluate / postEvaluate");
    account(w, true);
  }

  // One launch for the whole block, its sums added into the scatter container.
  void postEvaluate(Traits::PostEvalData) override {
    if (!open_) throw std::logic_error("GpuVolumeAssembly: postEvaluate without preEvaluate");
    open_ = false;
    require_covered("postEvaluate");
    txa_assembly_args a{};
    a.physics = physics_;
    a.precision = precision_;
    a.coef = coef_.p;
    a.nu = nu_;
    a.rho = rho_;
    for (int i = 0; i < 3; ++i) a.src[i] = src_[i];
    a.rayleigh_a = ra_;
    a.rayleigh_b = rb_;
    a.x = gather_c_->get_x();
    a.xdot = gather_c_->get_dxdt();
    a.xddot = gather_c_->get_d2xdt2();
    a.alpha = seeds_.alpha;
    a.beta = seeds_.beta;
    a.gamma2 = seeds_.gamma2;
    a.transient = seeds_.evaluate_transient_terms ? 1 : 0;
    a.f = scatter_c_->get_f();
    a.vals = std::is_same<EvalT, Traits::Jacobian>::value ? scatter_c_->get_A() : nullptr;
    a.block = bidx_ + 1;
    a.accumulate = 1;
    check(dm_->context(), txa_assemble(dm_->context(), dm_->mesh(), &a, stream_));
  }

 private:
  void reset() {
    seen_.assign(static_cast<size_t>(dm_->blockElements(block_)), 0);
    nseen_ = 0;
    have_seeds_ = false;
  }
  void account(const Workset &w, bool seeds) {
    if (w.block_id != block_)
      throw std::invalid_argument("GpuVolumeAssembly: workset of block \"" + w.block_id + "\" given to block \"" +
                                  block_ + "\"");
    if (static_cast<int64_t>(w.cell_local_ids.size()) != w.num_cells)
      throw std::invalid_argument("GpuVolumeAssembly: num_cells differs from the cell_local_ids it lists");
    for (int64_t c : w.cell_local_ids) {
      if (c < 0 || c >= static_cast<int64_t>(seen_.size()))
        throw std::invalid_argument("GpuVolumeAssembly: cell " + std::to_string(c) + " is not in the block");
      if (seen_[c]) throw std::logic_error("GpuVolumeAssembly: cell " + std::to_string(c) + " in two worksets");
      seen_[c] = 1;
      ++nseen_;
    }
    if (!seeds) return;
    if (!have_seeds_) {
      seeds_ = w;
      have_seeds_ = true;
    } else if (w.alpha != seeds_.alpha || w.beta != seeds_.beta || w.gamma2 != seeds_.gamma2 ||
               w.evaluate_transient_terms != seeds_.evaluate_transient_terms) {
      throw std::logic_error("GpuVolumeAssembly: worksets of one evaluation carry different seeds");
    }
  }
  void require_covered(const char *where) const {
    if (nseen_ != static_cast<int64_t>(seen_.size()))
      throw std::logic_error(std::string("GpuVolumeAssembly::") + where + ": the worksets cover " +
                             std::to_string(nseen_) + " of the " + std::to_string(seen_.size()) +
                             " cells of block \"" + block_ + "\"");
  }

  std::shared_ptr<const DOFManager> dm_;
  std::string block_, scatter_, gather_;
  int bidx_ = 0, physics_ = TXA_DIFFUSION, precision_ = TXA_FP64;
  DeviceValues coef_;
  double nu_ = 0.3, rho_ = 1.0, src_[3] = {0.0, 0.0, 0.0}, ra_ = 0.0, rb_ = 0.0;
  void *stream_ = nullptr;
  bool ready_ = false, open_ = false, have_seeds_ = false;
  std::vector<char> seen_;
  int64_t nseen_ = 0;
  Workset seeds_;
  std::shared_ptr<LinearObjContainer> scatter_c_, gather_c_;
};

// Dirichlet rows of a nodeset, applied after the volume assembly (the BC field manager's
// evaluator: TianXin_Dirichlet_impl.hpp:50-82 with the container's row replacement).
template <typename EvalT>
class GpuDirichletRows : public Evaluator<EvalT> {
 public:
  explicit GpuDirichletRows(const ParameterList &p) {
    ParameterList valid;
    valid.set("DOF Manager", std::shared_ptr<const DOFManager>());
    valid.set("Dirichlet LIDs", DeviceLIDs());
    valid.set("Dirichlet Values", DeviceValues());
    valid.set("Scatter Container Name", std::string("Residual Scatter Container"));
    valid.set("Gather Container Name", std::string("Solution Gather Container"));
    valid.set("Stream", static_cast<void *>(nullptr));
    p.validateParameters(valid);
    dm_ = p.get<std::shared_ptr<const DOFManager>>("DOF Manager");
    lids_ = p.get<DeviceLIDs>("Dirichlet LIDs");
    vals_ = p.get<DeviceValues>("Dirichlet Values", DeviceValues());
    if (!dm_ || (lids_.n > 0 && !lids_.p)) throw std::invalid_argument("GpuDirichletRows: DOF manager and LIDs");
    if (vals_.p && vals_.n != lids_.n) throw std::invalid_argument("GpuDirichletRows: one value per LID");
    scatter_ = p.get<std::string>("Scatter Container Name", "Residual Scatter Container");
    gather_ = p.get<std::string>("Gather Container Name", "Solution Gather Container");
    stream_ = p.get<void *>("Stream", nullptr);
  }
  void postRegistrationSetup(Traits::SetupData, FieldManager &) override {}
  void preEvaluate(Traits::PreEvalData d) override {
    scatter_c_ = std::dynamic_pointer_cast<LinearObjContainer>(d.gedc->getDataObject(scatter_));
    gather_c_ = std::dynamic_pointer_cast<LinearObjContainer>(d.gedc->getDataObject(gather_));
    if (!scatter_c_ || !gather_c_) throw std::logic_error("GpuDirichletRows: the containers are not LinearObjContainers");
  }
  void evaluateFields(Traits::EvalData) override {}
  void postEvaluate(Traits::PostEvalData) override {
    double *A = std::is_same<EvalT, Traits::Jacobian>::value ? scatter_c_->get_A() : nullptr;
    check(dm_->context(), txa_apply_dirichlet(dm_->context(), dm_->mesh(), lids_.p, vals_.p, lids_.n,
                                              gather_c_->get_x(), scatter_c_->get_f(), A, stream_));
  }

 private:
  std::shared_ptr<const DOFManager> dm_;
  DeviceLIDs lids_;
  DeviceValues vals_;
  std::string scatter_, gather_;
  void *stream_ = nullptr;
  std::shared_ptr<LinearObjContainer> scatter_c_, gather_c_;
};

// ---------------------------------------------------------------- assembly engine
// AssemblyEngineInArgs (disc-fe/src/Panzer_AssemblyEngine_InArgs.hpp:95-106).
struct AssemblyEngineInArgs {
  std::shared_ptr<LinearObjContainer> container;   // read (x, dxdt, d2xdt2) and written (f, A)
  double alpha = 0.0, beta = 1.0, gamma2 = 0.0, time = 0.0, step_size = 0.0;
  bool evaluate_transient_terms = false;
};

// AssemblyEngine::evaluate (Panzer_AssemblyEngine_impl.hpp:67-177): zero f (and A for the
// Jacobian), run each element block's volume field manager over its worksets with the
// in-arguments' seeds, then the Dirichlet field manager.
class AssemblyEngine {
 public:
  struct VolumeBlock {
    std::shared_ptr<FieldManager> fm;
    std::shared_ptr<std::vector<Workset>> worksets;
  };
  AssemblyEngine(std::shared_ptr<const DOFManager> dm, std::vector<VolumeBlock> volume,
                 std::shared_ptr<FieldManager> bcs, void *stream = nullptr)
      : dm_(std::move(dm)), volume_(std::move(volume)), bcs_(std::move(bcs)), stream_(stream) {}

  template <typename EvalT>
  void evaluate(const AssemblyEngineInArgs &in) {
    const bool jac = std::is_same<EvalT, Traits::Jacobian>::value;
    txa_ctx c = dm_->context();
    if (in.container->get_f()) check(c, txa_fill(c, in.container->get_f(), dm_->numOwned(), 0.0, stream_));
    if (jac && in.container->get_A()) check(c, txa_fill(c, in.container->get_A(), dm_->nnz(), 0.0, stream_));
    Traits::PED ped;
    ped.gedc->addDataObject("Solution Gather Container", in.container);
    ped.gedc->addDataObject("Residual Scatter Container", in.container);
    for (VolumeBlock &vb : volume_) {
      vb.fm->template preEvaluate<EvalT>(ped);
      for (Workset &w : *vb.worksets) {
        w.alpha = in.alpha;
        w.beta = in.beta;
        w.gamma2 = in.gamma2;
        w.time = in.time;
        w.step_size = in.step_size;
        w.evaluate_transient_terms = in.evaluate_transient_terms;
        vb.fm->template evaluateFields<EvalT>(w);
      }
      vb.fm->template postEvaluate<EvalT>(nullptr);
    }
    if (bcs_) {
      bcs_->template preEvaluate<EvalT>(ped);
      bcs_->template postEvaluate<EvalT>(nullptr);
    }
  }

 private:
  std::shared_ptr<const DOFManager> dm_;
  std::vector<VolumeBlock> volume_;
  std::shared_ptr<FieldManager> bcs_;
  void *stream_ = nullptr;
};

// Worksets of one element block in Panzer's way: consecutive chunks of workset_size cells.
inline std::shared_ptr<std::vector<Workset>> buildWorksets(const DOFManager &dm, const std::string &block,
                                                          int workset_size) {
  if (workset_size < 1) throw std::invalid_argument("buildWorksets: workset size must be >= 1");
  auto out = std::make_shared<std::vector<Workset>>();
  const int64_t n = dm.blockElements(block);
  for (int64_t c0 = 0; c0 < n; c0 += workset_size) {
    Workset w;
    w.block_id = block;
    for (int64_t c = c0; c < n && c < c0 + workset_size; ++c) w.cell_local_ids.push_back(c);
    w.num_cells = static_cast<int>(w.cell_local_ids.size());
    out->push_back(std::move(w));
  }
  return out;
}

// ---------------------------------------------------------------- model evaluator
// panzer::ModelEvaluator's evalModel for the volume path (txa_model_*).
class ModelEvaluator {
 public:
  struct InArgs {
    const double *x = nullptr, *x_dot = nullptr, *x_dot_dot = nullptr;
    double alpha = 0.0, beta = 1.0, gamma2 = 0.0, t = 0.0, step_size = 0.0;
  };
  struct OutArgs {
    double *f = nullptr;   // residual, or null
    double *W = nullptr;   // Jacobian values on the DOF manager's graph, or null
  };
  ModelEvaluator(std::shared_ptr<const DOFManager> dm, const txa_model_desc &desc, void *stream = nullptr)
      : dm_(std::move(dm)), stream_(stream) {
    check(dm_->context(), txa_model_create(dm_->context(), dm_->mesh(), &desc, &id_));
  }
  ~ModelEvaluator() { (void)txa_model_destroy(dm_->context(), id_); }
  ModelEvaluator(const ModelEvaluator &) = delete;
  ModelEvaluator &operator=(const ModelEvaluator &) = delete;
  void evalModel(const InArgs &in, const OutArgs &out) const {
    txa_model_in i{in.x, in.x_dot, in.x_dot_dot, in.alpha, in.beta, in.gamma2, in.t, in.step_size};
    txa_model_out o{out.f, out.W};
    check(dm_->context(), txa_model_eval(dm_->context(), id_, &i, &o, stream_));
  }
  const DOFManager &dofManager() const { return *dm_; }

 private:
  std::shared_ptr<const DOFManager> dm_;
  int id_ = 0;
  void *stream_ = nullptr;
};

}  // namespace panzer
}  // namespace txa
