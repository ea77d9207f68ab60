// Internal declarations of the MI355X assembly engine (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/txa.h"

namespace txa {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define TXA_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::txa::Error(TXA_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TXA_REQUIRE(cond, code, msg)                 \
  do {                                               \
    if (!(cond)) throw ::txa::Error((code), (msg));  \
  } while (0)

// Owning device buffer.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(e == hipErrorOutOfMemory ? TXA_ERR_NOMEM : TXA_ERR_DEVICE,
                  std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes: " +
                      hipGetErrorString(e));
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// One element block of a mesh: elements of one topology, numbered globally from
// elem_begin, connectivity at conn[conn_begin .. conn_begin + nelem * nconn).
struct MeshBlock {
  int topo = 0, nconn = 0, nvert_elem = 0;
  int64_t nelem = 0, elem_begin = 0, conn_begin = 0;
};

// A mesh with its DOF map and CRS graph, resident on the device: a structured brick,
// or uploaded connectivity of one or several element blocks.
struct Mesh {
  int topo = 0, nx = 0, ny = 0, nz = 0;   // topo: that of every block (0 when the blocks differ)
  bool structured = true;  // a brick (txa_mesh_brick): lattice kernels; false: uploaded connectivity
  bool affine = false;     // HEX8 brick whose cells are all parallelepipeds (brick_cells_affine)
  std::vector<MeshBlock> blocks;   // one for bricks and single-topology uploads
  int64_t nelem = 0;     // elements of all blocks (6 per cell for a TET4 brick)
  int64_t nconn_tot = 0; // connectivity entries of all blocks
  int nconn = 0;         // connectivity entities per element (0 when the blocks differ)
  int nvert_elem = 0;    // geometry vertices per element
  int64_t nverts = 0;    // vertex nodes
  int64_t nentities = 0; // entity-id bound
  DevBuf<double> xyz;    // SoA vertex coordinates: x[nverts], y[nverts], z[nverts]
  DevBuf<int64_t> conn;  // [nelem][nconn] entity ids
  // DOF map
  int nfields = 0;
  int ndof_elem = 0;          // element DOFs (0 when the blocks differ)
  int64_t ndofs = 0;
  DevBuf<int64_t> owned;      // [ndofs] GID of each LID
  DevBuf<int32_t> elem_lid;   // per block [nelem][nconn * nfields], blocks one after another
  int64_t npoints = 0;        // lattice points carrying DOFs (vertices; (2n+1)^3 for HEX27)
  DevBuf<int32_t> point_lid;  // [npoints][nfields] LID of each lattice point and field
  // graph
  bool has_graph = false;
  int64_t nnz = 0;
  int max_row = 0;            // entries in the longest CRS row
  DevBuf<int64_t> rowptr;
  DevBuf<int32_t> col;
  int perm_width = 0;         // bytes per row of the stencil slot map (27*nfields rounded up to 16)
  DevBuf<uint8_t> perm;       // [ndofs][perm_width] position of each stencil neighbour in the row, 255 = absent
  int nperm_w = 0;            // entries per node of nperm
  DevBuf<uint16_t> nperm;     // 3-field HEX8 / HEX27: [nodes][nperm_w] position of (neighbour, field 0) in the node's rows
  DevBuf<double> cell_geom;   // 3-field HEX8 bricks of parallelepiped cells: [cells][10] scaled cofactors, |det J|
  // uploaded 3-field HEX8 / HEX27 meshes of one block: the element instances (e * nb + a) of each
  // node, ascending, and each instance's slot map (position of element node b's field 0 in row 3n)
  DevBuf<int64_t> uocc_ptr;
  DevBuf<int32_t> uocc;
  DevBuf<uint16_t> uslot;
  // uploaded one-field HEX8 / TET4 / WEDGE6 blocks: the element rows of each LID, as positions p
  // in elem_lid (element e of a block, row a: block's position base + e*ndof + a), ascending
  // (occ[occ_ptr[r] .. occ_ptr[r+1]) for row r), for the row gather of the assembly
  DevBuf<int64_t> occ_ptr, occ;
  // node patches of uploaded one-field HEX8 / TET4 meshes (txa_patch.hip): rows along a Morton
  // curve cut into patches of patch_P, the elements touching each patch by (colour, sweep order)
  int patch_P = 0, patch_RS = 0, patch_NC = 0;   // rows per patch, LDS row stride, colours
  int64_t npatch = 0, patch_ninst = 0;
  DevBuf<int64_t> patch_cstart;   // [npatch*NC + 1] first instance of each (patch, colour)
  DevBuf<int32_t> patch_inst;     // element of each instance
  DevBuf<uint8_t> patch_icol;     // colour of each instance
  DevBuf<int32_t> patch_rows;     // [npatch*P] row at each curve position, -1 past the last row
  DevBuf<double> patch_rec;       // [rows][4] vertex x, y, z and the row's curve position, by LID
  bool patch_affine = false;      // every hex8 element a parallelepiped: the one-metric element form
  bool patch_bypos = false;       // the LIDs do not follow the curve: state and coordinates read in curve order
  DevBuf<int32_t> patch_ivtx;     // [instances][nb] vertices of each instance's element: curve positions
                                  // (patch_bypos) or LIDs
  DevBuf<double> patch_xyz;       // [npatch*P][4] vertex x, y, z in curve order (patch_bypos)
  DevBuf<int64_t> patch_r0;       // [npatch*P] first CRS position of the row at each curve position
  DevBuf<int32_t> patch_len;      // [npatch*P] its length
  DevBuf<uint8_t> patch_islot;    // [instances][nb][nb] position of LID b in the CRS row of LID a
  DevBuf<double> ref;         // reference-element tables of the topology (single-topology meshes)
  DevBuf<double> block_ref[5];   // reference-element tables by topology (TXA_HEX8 .. TXA_WEDGE6)
  bool mixed() const { return blocks.size() > 1 && topo == 0; }
  int64_t npos() const { return nconn_tot * nfields; }   // element DOF positions of all blocks
  int64_t pos_begin(const MeshBlock &b) const { return b.conn_begin * nfields; }
};

struct Batch;
void batch_free(Batch *b);
struct Server;
void serve_free(Server *s);

// A ModelEvaluator binding (txa_model_create): the mesh and the physics of its evaluations.
struct Model {
  int mesh_id = 0;
  txa_model_desc desc{};
};

struct Context {
  int device = 0;
  std::string last_error;
  std::map<int, std::unique_ptr<Mesh>> meshes;
  std::map<int, Model> models;
  std::map<int, Batch *> batches;
  std::map<int, Server *> servers;
  int next_id = 1;
  // stream-ordered pool for assembly workspaces: freed blocks stay mapped for the
  // next call (release threshold unbounded) until the context is destroyed
  hipMemPool_t ws_pool = nullptr;
  // device flag the assembly kernels set when a Dirichlet LID lies outside the DOF map;
  // txa_sync reports and clears it (the reference throws at that point)
  DevBuf<int> err;
  // this device's copy of the static workspace of txa_assemble_noalloc (its device address)
  void *scratch = nullptr;
  ~Context() {
    for (auto &kv : servers) serve_free(kv.second);   // joins the worker threads first
    for (auto &kv : batches) batch_free(kv.second);
    if (ws_pool) {
      (void)hipSetDevice(device);
      (void)hipDeviceSynchronize();
      (void)hipMemPoolDestroy(ws_pool);
    }
  }
  // device workspace of `bytes` for work on stream s; give it back with hipFreeAsync on s
  void *ws_alloc(size_t bytes, hipStream_t s) {
    if (!ws_pool) {
      hipMemPoolProps props{};
      props.allocType = hipMemAllocationTypePinned;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = device;
      TXA_HIP(hipMemPoolCreate(&ws_pool, &props));
      uint64_t keep = UINT64_MAX;
      TXA_HIP(hipMemPoolSetAttribute(ws_pool, hipMemPoolAttrReleaseThreshold, &keep));
    }
    void *p = nullptr;
    TXA_HIP(hipMallocFromPoolAsync(&p, bytes, ws_pool, s));
    return p;
  }
  Server &server(int id) {
    auto it = servers.find(id);
    TXA_REQUIRE(it != servers.end(), TXA_ERR_INVALID, "unknown server id " + std::to_string(id));
    return *it->second;
  }
  Batch &batch(int id) {
    auto it = batches.find(id);
    TXA_REQUIRE(it != batches.end(), TXA_ERR_INVALID, "unknown batch id " + std::to_string(id));
    return *it->second;
  }
  Mesh &mesh(int id) {
    auto it = meshes.find(id);
    TXA_REQUIRE(it != meshes.end(), TXA_ERR_INVALID, "unknown mesh id " + std::to_string(id));
    return *it->second;
  }
};

// setup kernels (txa_setup.hip)
void mesh_generate(Mesh &m, const double *host_coords, hipStream_t s);
void mesh_upload(Mesh &m, int64_t nelem, const int64_t *host_conn, int64_t nentities, int64_t nverts,
                 const double *host_coords, hipStream_t s);
void mesh_upload_blocks(Mesh &m, int nblocks, const txa_element_block *blocks, int64_t nentities, int64_t nverts,
                        const double *host_coords, hipStream_t s);
void dofmap_build(Mesh &m, int nfields, hipStream_t s);
void graph_build(Mesh &m, hipStream_t s);
void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);

// MatrixMarket export / import (txa_io.cpp)
void write_matrix_market(Mesh &m, const double *vals, const double *f, const double *x, const char *a_path,
                         const char *b_path, const char *x_path);
void read_vector_mm(const char *path, double *dst, int64_t n);

// assembly (txa_assemble.hip)
void upload_refelem(int topo, DevBuf<double> &tab, hipStream_t s);
void upload_block_refs(Mesh &m, hipStream_t s);   // tables of the mesh's topology and of each block's
// ws: device workspace of assemble_workspace_bytes(...) bytes, 256-byte aligned, for
// the work on s (may be null when that size is 0)
// err: device (or host-mapped) int set to 1 when a Dirichlet LID lies outside the map (may be null)
void assemble(Mesh &m, const txa_assembly_args &a, hipStream_t s, void *ws, int *err = nullptr);
size_t assemble_workspace_bytes(const Mesh &m, int physics, int precision, bool want_f, bool want_vals,
                                bool accumulate = false);
void apply_dirichlet(const Mesh &m, const int32_t *lids, const double *g, int64_t n, const double *x, double *f,
                     double *vals, hipStream_t s, int *err);
void spmv(const Mesh &m, const double *vals, const double *x, double *y, hipStream_t s);
void fill_async(double *p, int64_t n, double v, hipStream_t s);
// element-centric elasticity (txa_elastic.hip)
void node_perm_build(Mesh &m, hipStream_t s);
bool elastic_supported(const Mesh &m, const txa_assembly_args &a);
size_t elastic_workspace_bytes(const Mesh &m, int precision, bool want_f, bool want_vals);
void launch_elastic(Mesh &m, const txa_assembly_args &a, hipStream_t s, void *ws);
// Q1 diffusion fast path and batches (txa_q1_hex8.hip)
void launch_q1_hex8_diffusion(Mesh &m, const txa_assembly_args &a, hipStream_t s);
// uploaded one-field HEX8 / TET4 diffusion by node patches (txa_patch.hip)
void patch_build(Mesh &m, hipStream_t s);
bool patch_supported(const Mesh &m, int physics, int precision);
size_t patch_workspace_bytes(const Mesh &m);   // the state in curve order
void launch_patch(const Mesh &m, const txa_assembly_args &a, hipStream_t s, void *ws);
// Neumann flux on listed element sides (device arrays); a side that does not exist sets bit 1 of *err.
void neumann_sides(const Mesh &m, int64_t n, const int64_t *elems, const int *ords, const double *values, double *f,
                   hipStream_t s, int *err);
// Whether every cell of a HEX8 brick is a parallelepiped (to within its coordinates' rounding).
bool brick_cells_affine(const Mesh &m, hipStream_t s);
Batch *batch_create(const std::vector<Mesh *> &meshes, const std::vector<int> &ids, int precision, bool transient);
void batch_assemble(Batch &b, const txa_assembly_args *args, bool use_graph, hipStream_t s, int *err);
double batch_tune(Batch &b, const txa_assembly_args *args, hipStream_t s, int reps, int *err);
void batch_info(const Batch &b, int64_t *out);
void batch_shape(Batch &b, int tile_x, int slab, int grid);
const std::vector<int> &batch_mesh_ids(const Batch &b);
// serving loop (txa_serve.cpp)
Server *serve_create(int device, const std::vector<Mesh *> &meshes, const std::vector<int> &ids, int64_t budget,
                     int slots);
const std::vector<int> &serve_mesh_ids(const Server &s);
int64_t serve_submit(Server &s, const txa_serve_request &r);
void serve_poll(Server &s, int64_t id, double timeout_s, int *status, const double **f, const double **vals);
void serve_release(Server &s, int64_t id);
void serve_stats(Server &s, double *out);

}  // namespace txa
