// C ABI of the assembly engine: status codes + per-context last-error string,
// no exception crosses the boundary.
#include <cstring>
#include <mutex>
#include <new>

#include "txa_internal.hpp"

using txa::Context;
using txa::Error;
using txa::Mesh;

struct txa_context : Context {};

namespace {

thread_local std::string g_create_error;

// Workspace of txa_assemble_noalloc: one static device array, so that entry point allocates
// nothing.  Each device the library runs on holds its own copy until the process ends.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

// The device address of this device's copy, looked up once per device with the device
// current (host code must not use the array's name as an address).
void *scratch_address(int device) {
  static std::mutex mu;
  static std::map<int, void *> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void *p = nullptr;
  TXA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  TXA_REQUIRE(p != nullptr, TXA_ERR_DEVICE, "no device address for the static workspace");
  addr[device] = p;
  return p;
}

template <class F>
int guarded(txa_ctx ctx, F &&f) {
  if (!ctx) return TXA_ERR_INVALID;
  try {
    TXA_HIP(hipSetDevice(ctx->device));
    f();
    ctx->last_error.clear();
    return TXA_OK;
  } catch (const Error &e) {
    ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    ctx->last_error = "host allocation failed";
    return TXA_ERR_NOMEM;
  } catch (const std::exception &e) {
    ctx->last_error = e.what();
    return TXA_ERR_LOGIC;
  }
}

// Rebuilding a mesh's DOF map or graph reallocates arrays that batches and servers
// launch on (and can change its DOF count): refused while one holds the mesh.
void require_unshared(txa_ctx ctx, int mesh_id, const char *what) {
  for (auto &kv : ctx->batches)
    for (int id : txa::batch_mesh_ids(*kv.second))
      TXA_REQUIRE(id != mesh_id, TXA_ERR_LOGIC,
                  std::string(what) + ": mesh is part of batch " + std::to_string(kv.first));
  for (auto &kv : ctx->servers)
    for (int id : txa::serve_mesh_ids(*kv.second))
      TXA_REQUIRE(id != mesh_id, TXA_ERR_LOGIC,
                  std::string(what) + ": mesh is served by server " + std::to_string(kv.first));
}

}  // namespace

extern "C" {

const char *txa_version(void) { return "tianxin_amd 0.1 (gfx950)"; }

int txa_create(int device, txa_ctx *out) {
  if (!out) return TXA_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    g_create_error = "no HIP device";
    return TXA_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    g_create_error = "device index out of range";
    return TXA_ERR_INVALID;
  }
  if (hipSetDevice(device) != hipSuccess) {
    g_create_error = "hipSetDevice failed";
    return TXA_ERR_DEVICE;
  }
  txa_context *c = new (std::nothrow) txa_context();
  if (!c) return TXA_ERR_NOMEM;
  c->device = device;
  try {
    c->err.alloc(1);
    TXA_HIP(hipMemset(c->err.p, 0, sizeof(int)));
    c->scratch = scratch_address(device);
  } catch (const Error &e) {
    g_create_error = e.what();
    delete c;
    return e.code;
  }
  *out = c;
  return TXA_OK;
}

int txa_destroy(txa_ctx ctx) {
  if (!ctx) return TXA_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  delete ctx;
  return TXA_OK;
}

const char *txa_last_error(txa_ctx ctx) { return ctx ? ctx->last_error.c_str() : g_create_error.c_str(); }

int txa_sync(txa_ctx ctx) {
  return guarded(ctx, [&] {
    TXA_HIP(hipDeviceSynchronize());
    int flag = 0;
    TXA_HIP(hipMemcpy(&flag, ctx->err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
      TXA_HIP(hipMemset(ctx->err.p, 0, sizeof(int)));
      std::string msg;
      if (flag & 1) msg = "a Dirichlet LID of an assembly since the last txa_sync lies outside the DOF map "
                          "(its row was not applied)";
      if (flag & 2) msg += std::string(msg.empty() ? "" : "; ") + "a Neumann side since the last txa_sync does not "
                           "exist (element index or side ordinal out of range; it was not applied)";
      throw Error(TXA_ERR_INVALID, msg);
    }
  });
}

int txa_mesh_brick(txa_ctx ctx, int topo, int nx, int ny, int nz, const double *vertex_coords, int *mesh_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(mesh_id && vertex_coords, TXA_ERR_INVALID, "null argument");
    TXA_REQUIRE(topo == TXA_HEX8 || topo == TXA_TET4 || topo == TXA_HEX27, TXA_ERR_INVALID, "unknown topology");
    TXA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, TXA_ERR_INVALID, "brick dimensions must be >= 1");
    TXA_REQUIRE(static_cast<int64_t>(nx + 1) * (ny + 1) < (1LL << 31), TXA_ERR_UNSUPPORTED, "brick too wide");
    auto m = std::make_unique<Mesh>();
    m->topo = topo; m->nx = nx; m->ny = ny; m->nz = nz;
    txa::mesh_generate(*m, vertex_coords, nullptr);
    txa::upload_block_refs(*m, nullptr);
    int id = ctx->next_id++;
    ctx->meshes[id] = std::move(m);
    *mesh_id = id;
  });
}

int txa_mesh_upload(txa_ctx ctx, int topo, int64_t nelem, const int64_t *conn, int64_t nentities, int64_t nverts,
                    const double *vertex_coords, int *mesh_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(mesh_id && conn && vertex_coords, TXA_ERR_INVALID, "null argument");
    auto m = std::make_unique<Mesh>();
    m->topo = topo;
    txa::mesh_upload(*m, nelem, conn, nentities, nverts, vertex_coords, nullptr);
    txa::upload_block_refs(*m, nullptr);
    int id = ctx->next_id++;
    ctx->meshes[id] = std::move(m);
    *mesh_id = id;
  });
}

int txa_mesh_upload_blocks(txa_ctx ctx, int nblocks, const txa_element_block *blocks, int64_t nentities,
                           int64_t nverts, const double *vertex_coords, int *mesh_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(mesh_id && blocks && vertex_coords, TXA_ERR_INVALID, "null argument");
    auto m = std::make_unique<Mesh>();
    txa::mesh_upload_blocks(*m, nblocks, blocks, nentities, nverts, vertex_coords, nullptr);
    txa::upload_block_refs(*m, nullptr);
    int id = ctx->next_id++;
    ctx->meshes[id] = std::move(m);
    *mesh_id = id;
  });
}

int txa_mesh_blocks(txa_ctx ctx, int mesh_id, int64_t *out, int64_t cap) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(out, TXA_ERR_INVALID, "null argument");
    Mesh &m = ctx->mesh(mesh_id);
    const int64_t need = 1 + 3 * static_cast<int64_t>(m.blocks.size());
    TXA_REQUIRE(cap >= need, TXA_ERR_INVALID, "out holds " + std::to_string(cap) + " entries, " +
                                                  std::to_string(need) + " needed");
    out[0] = static_cast<int64_t>(m.blocks.size());
    for (size_t k = 0; k < m.blocks.size(); ++k) {
      out[1 + 3 * k] = m.blocks[k].topo;
      out[2 + 3 * k] = m.blocks[k].nelem;
      out[3 + 3 * k] = m.blocks[k].elem_begin;
    }
  });
}

int txa_mesh_destroy(txa_ctx ctx, int mesh_id) {
  return guarded(ctx, [&] {
    auto it = ctx->meshes.find(mesh_id);
    TXA_REQUIRE(it != ctx->meshes.end(), TXA_ERR_INVALID, "unknown mesh id");
    require_unshared(ctx, mesh_id, "txa_mesh_destroy");
    TXA_HIP(hipDeviceSynchronize());
    ctx->meshes.erase(it);
  });
}

int txa_mesh_info(txa_ctx ctx, int mesh_id, int64_t *out) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(out, TXA_ERR_INVALID, "null argument");
    Mesh &m = ctx->mesh(mesh_id);
    out[0] = m.nelem; out[1] = m.nconn; out[2] = m.nverts;
  });
}

int txa_mesh_affine(txa_ctx ctx, int mesh_id, int *affine) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(affine, TXA_ERR_INVALID, "null argument");
    *affine = ctx->mesh(mesh_id).affine ? 1 : 0;
  });
}

int txa_mesh_patches(txa_ctx ctx, int mesh_id, int64_t *out) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(out, TXA_ERR_INVALID, "null argument");
    const Mesh &m = ctx->mesh(mesh_id);
    const bool on = m.npatch > 0;
    out[0] = on ? m.patch_P : 0;
    out[1] = m.npatch;
    out[2] = on ? m.patch_ninst : 0;
    out[3] = on ? m.patch_NC : 0;
    out[4] = on ? m.patch_RS : 0;
    out[5] = on && m.patch_bypos ? 1 : 0;
    out[6] = on && m.patch_affine ? 1 : 0;
  });
}

int txa_dofmap_build(txa_ctx ctx, int mesh_id, int nfields, int64_t *ndofs) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    require_unshared(ctx, mesh_id, "txa_dofmap_build");
    txa::dofmap_build(m, nfields, nullptr);
    if (ndofs) *ndofs = m.ndofs;
  });
}

int txa_dofmap_copy(txa_ctx ctx, int mesh_id, int64_t *owned_gids, int32_t *elem_lids) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    TXA_REQUIRE(m.ndofs > 0, TXA_ERR_LOGIC, "DOF map not built");
    if (owned_gids) TXA_HIP(hipMemcpy(owned_gids, m.owned.p, m.owned.bytes(), hipMemcpyDeviceToHost));
    if (elem_lids) TXA_HIP(hipMemcpy(elem_lids, m.elem_lid.p, sizeof(int32_t) * m.npos(), hipMemcpyDeviceToHost));
  });
}

int txa_dofmap_point_lids(txa_ctx ctx, int mesh_id, int32_t *point_lids, int64_t *npoints) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    TXA_REQUIRE(m.ndofs > 0, TXA_ERR_LOGIC, "DOF map not built");
    if (npoints) *npoints = m.npoints;
    if (point_lids)
      TXA_HIP(hipMemcpy(point_lids, m.point_lid.p, sizeof(int32_t) * m.npoints * m.nfields, hipMemcpyDeviceToHost));
  });
}

int txa_graph_build(txa_ctx ctx, int mesh_id, int64_t *nnz) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    require_unshared(ctx, mesh_id, "txa_graph_build");
    txa::graph_build(m, nullptr);
    if (nnz) *nnz = m.nnz;
  });
}

int txa_graph_copy(txa_ctx ctx, int mesh_id, int64_t *rowptr, int32_t *col) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "graph not built");
    if (rowptr) TXA_HIP(hipMemcpy(rowptr, m.rowptr.p, sizeof(int64_t) * (m.ndofs + 1), hipMemcpyDeviceToHost));
    if (col) TXA_HIP(hipMemcpy(col, m.col.p, sizeof(int32_t) * m.nnz, hipMemcpyDeviceToHost));
  });
}

int txa_graph_device(txa_ctx ctx, int mesh_id, const int64_t **rowptr, const int32_t **col) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "graph not built");
    if (rowptr) *rowptr = m.rowptr.p;
    if (col) *col = m.col.p;
  });
}

int txa_write_matrix_market(txa_ctx ctx, int mesh_id, const double *vals, const double *f, const double *x,
                            const char *a_path, const char *b_path, const char *x_path) {
  return guarded(ctx, [&] {
    Mesh &m = ctx->mesh(mesh_id);
    TXA_HIP(hipDeviceSynchronize());   // the arrays may still be written by queued work
    txa::write_matrix_market(m, vals, f, x, a_path, b_path, x_path);
  });
}

int txa_read_vector_mm(txa_ctx ctx, const char *path, double *dst, int64_t n) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(path && (dst || n == 0) && n >= 0, TXA_ERR_INVALID, "null argument or negative length");
    txa::read_vector_mm(path, dst, n);
  });
}

int txa_assemble(txa_ctx ctx, int mesh_id, const txa_assembly_args *args, void *stream) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(args, TXA_ERR_INVALID, "null args");
    Mesh &m = ctx->mesh(mesh_id);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t wsb = txa::assemble_workspace_bytes(m, args->physics, args->precision, args->f != nullptr,
                                                     args->vals != nullptr, args->accumulate != 0);
    void *ws = wsb ? ctx->ws_alloc(wsb, s) : nullptr;
    try {
      txa::assemble(m, *args, s, ws, ctx->err.p);
    } catch (...) {
      if (ws) (void)hipFreeAsync(ws, s);
      throw;
    }
    if (ws) TXA_HIP(hipFreeAsync(ws, s));
  });
}

int txa_assemble_noalloc(txa_ctx ctx, int mesh_id, const txa_assembly_args *args, void *stream) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(args, TXA_ERR_INVALID, "null args");
    Mesh &m = ctx->mesh(mesh_id);
    const size_t wsb = txa::assemble_workspace_bytes(m, args->physics, args->precision, args->f != nullptr,
                                                     args->vals != nullptr, args->accumulate != 0);
    TXA_REQUIRE(wsb <= kScratchBytes, TXA_ERR_NOMEM,
                "txa_assemble_noalloc: this assembly needs " + std::to_string(wsb) +
                    " bytes of workspace; the static workspace holds " + std::to_string(kScratchBytes) +
                    " (use txa_assemble)");
    TXA_REQUIRE(ctx->scratch != nullptr, TXA_ERR_LOGIC, "static workspace not set up");
    txa::assemble(m, *args, static_cast<hipStream_t>(stream), wsb ? ctx->scratch : nullptr, ctx->err.p);
  });
}

int txa_apply_dirichlet(txa_ctx ctx, int mesh_id, const int32_t *lids, const double *values, int64_t n,
                        const double *x, double *f, double *vals, void *stream) {
  return guarded(ctx, [&] {
    txa::apply_dirichlet(ctx->mesh(mesh_id), lids, values, n, x, f, vals, static_cast<hipStream_t>(stream),
                         ctx->err.p);
  });
}

int txa_neumann_sides(txa_ctx ctx, int mesh_id, int64_t n, const int64_t *elems, const int *ords,
                      const double *values, double *f, void *stream) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(n >= 0, TXA_ERR_INVALID, "negative side count");
    Mesh &m = ctx->mesh(mesh_id);
    txa::neumann_sides(m, n, elems, ords, values, f, static_cast<hipStream_t>(stream), ctx->err.p);
  });
}

int txa_fill(txa_ctx ctx, double *p, int64_t n, double value, void *stream) {
  return guarded(ctx, [&] { txa::fill_async(p, n, value, static_cast<hipStream_t>(stream)); });
}

int txa_spmv(txa_ctx ctx, int mesh_id, const double *vals, const double *x, double *y, void *stream) {
  return guarded(ctx, [&] { txa::spmv(ctx->mesh(mesh_id), vals, x, y, static_cast<hipStream_t>(stream)); });
}

int txa_model_create(txa_ctx ctx, int mesh_id, const txa_model_desc *desc, int *model_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(desc && model_id, TXA_ERR_INVALID, "null argument");
    Mesh &m = ctx->mesh(mesh_id);
    TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "a model needs the mesh's DOF map and graph");
    TXA_REQUIRE(desc->coef, TXA_ERR_INVALID, "coef is required");
    TXA_REQUIRE(desc->n_bc == 0 || desc->bc_lids, TXA_ERR_INVALID, "bc_lids missing");
    txa::Model md;
    md.mesh_id = mesh_id;
    md.desc = *desc;
    const int id = ctx->next_id++;
    ctx->models[id] = md;
    *model_id = id;
  });
}

// panzer::ModelEvaluator::evalModelImpl_basic (disc-fe/src/Panzer_ModelEvaluator_impl.hpp:1715-1840)
// with setupAssemblyInArgs (:725-780): a model without transient support fills the steady
// system (alpha 0, beta 1); with it, alpha / beta / gamma2 come from the in-arguments and the
// transient terms are evaluated when x_dot is given.
int txa_model_eval(txa_ctx ctx, int model_id, const txa_model_in *in, const txa_model_out *out, void *stream) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(in && out, TXA_ERR_INVALID, "null argument");
    auto it = ctx->models.find(model_id);
    TXA_REQUIRE(it != ctx->models.end(), TXA_ERR_INVALID, "unknown model id " + std::to_string(model_id));
    const txa::Model &md = it->second;
    if (!out->f && !out->W) return;                 // nothing asked for
    const txa_model_desc &d = md.desc;
    TXA_REQUIRE(in->x, TXA_ERR_INVALID, "x is required");
    TXA_REQUIRE(!in->x_dot || d.transient_support, TXA_ERR_INVALID,
                "x_dot given to a model built without transient support");
    Mesh &m = ctx->mesh(md.mesh_id);
    txa_assembly_args a{};
    a.physics = d.physics;
    a.precision = d.precision;
    a.coef = d.coef;
    a.nu = d.nu;
    a.rho = d.rho;
    for (int i = 0; i < 3; ++i) a.src[i] = d.src[i];
    a.rayleigh_a = d.rayleigh_a;
    a.rayleigh_b = d.rayleigh_b;
    a.x = in->x;
    if (d.transient_support) {
      a.alpha = in->alpha;
      a.beta = in->beta;
      a.gamma2 = in->gamma2;
      a.transient = in->x_dot != nullptr;
      a.xdot = in->x_dot;
      a.xddot = in->x_dot_dot;
    } else {
      a.alpha = 0.0;
      a.beta = 1.0;
    }
    a.f = out->f;
    a.vals = out->W;
    a.bc_lids = d.bc_lids;
    a.bc_vals = d.bc_vals;
    a.n_bc = d.n_bc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t wsb = txa::assemble_workspace_bytes(m, a.physics, a.precision, a.f != nullptr, a.vals != nullptr);
    void *ws = wsb ? ctx->ws_alloc(wsb, s) : nullptr;
    try {
      txa::assemble(m, a, s, ws, ctx->err.p);
    } catch (...) {
      if (ws) (void)hipFreeAsync(ws, s);
      throw;
    }
    if (ws) TXA_HIP(hipFreeAsync(ws, s));
  });
}

int txa_model_destroy(txa_ctx ctx, int model_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(ctx->models.erase(model_id) == 1, TXA_ERR_INVALID, "unknown model id " + std::to_string(model_id));
  });
}

int txa_batch_create(txa_ctx ctx, const int *mesh_ids, int n, int precision, int transient, int *batch_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(mesh_ids && batch_id && n > 0, TXA_ERR_INVALID, "null argument or empty batch");
    std::vector<Mesh *> ms;
    std::vector<int> ids(mesh_ids, mesh_ids + n);
    for (int id : ids) ms.push_back(&ctx->mesh(id));
    txa::Batch *b = txa::batch_create(ms, ids, precision, transient != 0);
    const int id = ctx->next_id++;
    ctx->batches[id] = b;
    *batch_id = id;
  });
}

int txa_batch_destroy(txa_ctx ctx, int batch_id) {
  return guarded(ctx, [&] {
    auto it = ctx->batches.find(batch_id);
    TXA_REQUIRE(it != ctx->batches.end(), TXA_ERR_INVALID, "unknown batch id");
    TXA_HIP(hipDeviceSynchronize());
    txa::batch_free(it->second);
    ctx->batches.erase(it);
  });
}

int txa_batch_assemble(txa_ctx ctx, int batch_id, const txa_assembly_args *args, int use_graph, void *stream) {
  return guarded(ctx, [&] {
    txa::batch_assemble(ctx->batch(batch_id), args, use_graph != 0, static_cast<hipStream_t>(stream), ctx->err.p);
  });
}

int txa_batch_tune(txa_ctx ctx, int batch_id, const txa_assembly_args *args, int reps, void *stream,
                   double *best_ms) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(args, TXA_ERR_INVALID, "null args");
    TXA_REQUIRE(reps >= 1 && reps <= 1000, TXA_ERR_INVALID, "reps must be in [1, 1000]");
    const double ms = txa::batch_tune(ctx->batch(batch_id), args, static_cast<hipStream_t>(stream), reps,
                                      ctx->err.p);
    if (best_ms) *best_ms = ms;
  });
}

int txa_batch_shape(txa_ctx ctx, int batch_id, int tile_x, int slab, int grid) {
  return guarded(ctx, [&] { txa::batch_shape(ctx->batch(batch_id), tile_x, slab, grid); });
}

int txa_batch_info(txa_ctx ctx, int batch_id, int64_t *out) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(out, TXA_ERR_INVALID, "null argument");
    txa::batch_info(ctx->batch(batch_id), out);
  });
}

int txa_serve_create(txa_ctx ctx, const int *mesh_ids, int n, int64_t hbm_budget, int slots, int *server_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(mesh_ids && server_id && n > 0, TXA_ERR_INVALID, "null argument or no meshes");
    std::vector<int> ids(mesh_ids, mesh_ids + n);
    std::vector<Mesh *> ms;
    for (int id : ids) ms.push_back(&ctx->mesh(id));
    txa::Server *s = txa::serve_create(ctx->device, ms, ids, hbm_budget, slots);
    const int id = ctx->next_id++;
    ctx->servers[id] = s;
    *server_id = id;
  });
}

int txa_serve_destroy(txa_ctx ctx, int server_id) {
  return guarded(ctx, [&] {
    auto it = ctx->servers.find(server_id);
    TXA_REQUIRE(it != ctx->servers.end(), TXA_ERR_INVALID, "unknown server id");
    txa::serve_free(it->second);
    ctx->servers.erase(it);
    TXA_HIP(hipDeviceSynchronize());
  });
}

int txa_serve_submit(txa_ctx ctx, int server_id, const txa_serve_request *req, int64_t *request_id) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(req && request_id, TXA_ERR_INVALID, "null argument");
    *request_id = txa::serve_submit(ctx->server(server_id), *req);
  });
}

int txa_serve_poll(txa_ctx ctx, int server_id, int64_t request_id, double timeout_s, int *status, const double **f,
                   const double **vals) {
  return guarded(ctx, [&] { txa::serve_poll(ctx->server(server_id), request_id, timeout_s, status, f, vals); });
}

int txa_serve_release(txa_ctx ctx, int server_id, int64_t request_id) {
  return guarded(ctx, [&] { txa::serve_release(ctx->server(server_id), request_id); });
}

int txa_serve_stats(txa_ctx ctx, int server_id, double *out) {
  return guarded(ctx, [&] {
    TXA_REQUIRE(out, TXA_ERR_INVALID, "null argument");
    txa::serve_stats(ctx->server(server_id), out);
  });
}

}  // extern "C"
