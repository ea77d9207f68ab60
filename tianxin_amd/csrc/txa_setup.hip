// Device-side setup: structured connectivity, DOF map (GUN GIDs, first-touch
// LIDs), CRS graph and the per-row stencil slot map used by the assembly kernels.
//
// Numbering rules follow the reference (paths relative to its root):
//   vertices/elements  adapters-stk/src/stk_interface/Panzer_STK_CubeHexMeshFactory.cpp:425-456
//   6-tet split        dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:228-235
//   edge/face/cell ids dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:744-890
//   GIDs (one rank)    dof-mgr/src/Panzer_DOFManager.cpp:717-865, 1294-1349
//   owned_ / LIDs      dof-mgr/src/Panzer_DOFManager.cpp:580-640, Panzer_GlobalIndexer.hpp:605-640
//   CRS graph          disc-fe/src/lof/Panzer_TpetraLinearObjFactory_impl.hpp:559-650
// The algorithms are data-parallel re-formulations: GID ranks by a flag scan
// over entity ids, first touch by an atomic minimum over element positions,
// rows by a counting sort plus a per-row bitonic sort in LDS.
#include "txa_internal.hpp"

#include <vector>

namespace txa {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int per = kBlock) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

// ---------------------------------------------------------------- scan
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void k_scan_tiles(const int64_t *__restrict__ in, int64_t *__restrict__ out, int64_t n,
                             int64_t *__restrict__ tile_sums) {
  __shared__ int64_t s[kBlock];
  int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + i;
    v[i] = idx < n ? in[idx] : 0;
    sum += v[i];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t idx = base + i;
    if (idx < n) out[idx] = run;
    run += v[i];
  }
  if (threadIdx.x == kBlock - 1) tile_sums[blockIdx.x] = s[kBlock - 1];
}

__global__ void k_scan_add(int64_t *__restrict__ out, int64_t n, const int64_t *__restrict__ tile_off) {
  int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += tile_off[idx / kScanTile];
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- connectivity
__device__ const int kHexV[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__device__ const int kTets[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                                    {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};

struct BrickIds {
  int64_t nx, ny, nz, nn, ne, nf;
  __device__ int64_t node(int64_t x, int64_t y, int64_t z) const { return x + (nx + 1) * (y + (ny + 1) * z); }
  __device__ int64_t eks() const { return (nx + 1) * ny + nx * (ny + 1) + (nx + 1) * (ny + 1); }
  __device__ int64_t fks() const { return nx * ny + nx * (ny + 1) + (nx + 1) * ny; }
  // edges: per z level, rows of [x-edges | y-edges], then the vertical edges of the level
  __device__ int64_t ex(int64_t x, int64_t y, int64_t z) const { return nn + x + y * (2 * nx + 1) + z * eks(); }
  __device__ int64_t ey(int64_t x, int64_t y, int64_t z) const { return nn + nx + x + y * (2 * nx + 1) + z * eks(); }
  __device__ int64_t ez(int64_t x, int64_t y, int64_t z) const {
    return nn + z * eks() + (nx + 1) * ny + nx * (ny + 1) + x + y * (nx + 1);
  }
  // faces: per z level, the xy faces, then rows of [xz faces | yz faces]
  __device__ int64_t fxy(int64_t x, int64_t y, int64_t z) const { return nn + ne + x + y * nx + z * fks(); }
  __device__ int64_t fxz(int64_t x, int64_t y, int64_t z) const {
    return nn + ne + z * fks() + nx * ny + y * (2 * nx + 1) + x;
  }
  __device__ int64_t fyz(int64_t x, int64_t y, int64_t z) const {
    return nn + ne + z * fks() + nx * ny + y * (2 * nx + 1) + nx + x;
  }
};

__global__ void k_brick_conn(int topo, BrickIds b, int64_t ncell, int64_t *__restrict__ conn) {
  int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int64_t x = c % b.nx, y = (c / b.nx) % b.ny, z = c / (b.nx * b.ny);
  int64_t hv[8];
  for (int v = 0; v < 8; ++v) hv[v] = b.node(x + kHexV[v][0], y + kHexV[v][1], z + kHexV[v][2]);
  if (topo == TXA_HEX8) {
    for (int v = 0; v < 8; ++v) conn[c * 8 + v] = hv[v];
  } else if (topo == TXA_TET4) {
    for (int t = 0; t < 6; ++t)
      for (int v = 0; v < 4; ++v) conn[(6 * c + t) * 4 + v] = hv[kTets[t][v]];
  } else {
    int64_t *cc = conn + c * 27;
    for (int v = 0; v < 8; ++v) cc[v] = hv[v];
    cc[8] = b.ex(x, y, z);
    cc[9] = b.ey(x + 1, y, z);
    cc[10] = b.ex(x, y + 1, z);
    cc[11] = b.ey(x, y, z);
    cc[12] = b.ex(x, y, z + 1);
    cc[13] = b.ey(x + 1, y, z + 1);
    cc[14] = b.ex(x, y + 1, z + 1);
    cc[15] = b.ey(x, y, z + 1);
    cc[16] = b.ez(x, y, z);
    cc[17] = b.ez(x + 1, y, z);
    cc[18] = b.ez(x + 1, y + 1, z);
    cc[19] = b.ez(x, y + 1, z);
    cc[20] = b.fxz(x, y, z);
    cc[21] = b.fyz(x + 1, y, z);
    cc[22] = b.fxz(x, y + 1, z);
    cc[23] = b.fyz(x, y, z);
    cc[24] = b.fxy(x, y, z);
    cc[25] = b.fxy(x, y, z + 1);
    cc[26] = b.nn + b.ne + b.nf + c;
  }
}

// ---------------------------------------------------------------- DOF map
__global__ void k_mark_entities(const int64_t *__restrict__ conn, int64_t n, int64_t *__restrict__ flag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) flag[conn[i]] = 1;  // benign race: every writer stores 1
}

// element GIDs: per connectivity entity (subcell order), per field
__global__ void k_elem_gid(const int64_t *__restrict__ conn, int64_t nconn_tot, int nfields,
                           const int64_t *__restrict__ rank, int64_t *__restrict__ elem_gid,
                           unsigned long long *__restrict__ first) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nconn_tot) return;
  int64_t r = rank[conn[i]];
  for (int f = 0; f < nfields; ++f) {
    int64_t p = i * nfields + f;  // element-major, entity, field: the element DOF order
    int64_t g = r * nfields + f;
    elem_gid[p] = g;
    atomicMin(first + g, static_cast<unsigned long long>(p));
  }
}

__global__ void k_first_flag(const int64_t *__restrict__ elem_gid, int64_t n,
                             const unsigned long long *__restrict__ first, int64_t *__restrict__ flag) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < n) flag[p] = first[elem_gid[p]] == static_cast<unsigned long long>(p) ? 1 : 0;
}

__global__ void k_assign_lids(const int64_t *__restrict__ elem_gid, int64_t n,
                              const unsigned long long *__restrict__ first, const int64_t *__restrict__ pos,
                              int64_t *__restrict__ owned, int32_t *__restrict__ elem_lid) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int64_t g = elem_gid[p];
  int64_t fp = static_cast<int64_t>(first[g]);
  int64_t lid = pos[fp];
  elem_lid[p] = static_cast<int32_t>(lid);
  if (fp == p) owned[lid] = g;
}

// lattice point of connectivity entry c of element e: vertices for HEX8/TET4,
// the (2nx+1)(2ny+1)(2nz+1) lattice of Q2 nodes for HEX27
__device__ const int kQ2Node[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2},
    {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
    {1, 0, 1}, {2, 1, 1}, {1, 2, 1}, {0, 1, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};

__global__ void k_point_lid(int topo, int64_t nx, int64_t ny, const int64_t *__restrict__ conn, int nconn,
                            int64_t nelem, int nfields, const int32_t *__restrict__ elem_lid,
                            int32_t *__restrict__ point_lid) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nelem * nconn) return;
  int64_t e = i / nconn;
  int c = static_cast<int>(i % nconn);
  int64_t pt;
  if (topo == TXA_HEX27) {
    int64_t x = e % nx, y = (e / nx) % ny, z = e / (nx * ny);
    int64_t X = 2 * x + kQ2Node[c][0], Y = 2 * y + kQ2Node[c][1], Z = 2 * z + kQ2Node[c][2];
    pt = X + (2 * nx + 1) * (Y + (2 * ny + 1) * Z);
  } else {
    pt = conn[i];
  }
  for (int f = 0; f < nfields; ++f) point_lid[pt * nfields + f] = elem_lid[i * nfields + f];
}

// ---------------------------------------------------------------- graph
__global__ void k_row_count(const int32_t *__restrict__ elem_lid, int64_t n, int ndof, int64_t *__restrict__ cnt) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < n) atomicAdd(reinterpret_cast<unsigned long long *>(cnt + elem_lid[p]), static_cast<unsigned long long>(ndof));
}

__global__ void k_row_fill(const int32_t *__restrict__ elem_lid, int64_t n, int ndof,
                           int64_t *__restrict__ cursor, int32_t *__restrict__ cand) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int64_t e = p / ndof;
  int32_t row = elem_lid[p];
  int64_t base = static_cast<int64_t>(
      atomicAdd(reinterpret_cast<unsigned long long *>(cursor + row), static_cast<unsigned long long>(ndof)));
  for (int j = 0; j < ndof; ++j) cand[base + j] = elem_lid[e * ndof + j];
}

constexpr int kMaxCand = 1024;

// One wavefront per row: bitonic sort of the candidate columns in LDS, then
// in-place compaction of the unique values; ulen[row] = unique count.
__global__ __launch_bounds__(64) void k_row_sort_unique(int64_t nrows, const int64_t *__restrict__ cptr,
                                                        int32_t *__restrict__ cand, int64_t *__restrict__ ulen,
                                                        int *__restrict__ overflow) {
  __shared__ int32_t s[kMaxCand];
  __shared__ int32_t cnts[64];
  int64_t row = blockIdx.x;
  if (row >= nrows) return;
  int lane = threadIdx.x;
  int64_t b0 = cptr[row];
  int n = static_cast<int>(cptr[row + 1] - b0);
  if (n > kMaxCand) {
    if (lane == 0) atomicOr(overflow, 1);
    return;
  }
  int P = 64;
  while (P < n) P <<= 1;
  for (int i = lane; i < P; i += 64) s[i] = i < n ? cand[b0 + i] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        int l = i ^ j;
        if (l > i) {
          int32_t a = s[i], b = s[l];
          bool up = (i & k) == 0;
          if ((a > b) == up) { s[i] = b; s[l] = a; }
        }
      }
      __syncthreads();
    }
  // compaction: lane owns a contiguous chunk
  int per = (n + 63) / 64;
  int lo = lane * per, hi = min(n, lo + per);
  int c = 0;
  for (int i = lo; i < hi; ++i) c += (i == 0 || s[i] != s[i - 1]);
  cnts[lane] = c;
  __syncthreads();
  int off = 0;
  for (int l = 0; l < lane; ++l) off += cnts[l];
  for (int i = lo; i < hi; ++i)
    if (i == 0 || s[i] != s[i - 1]) cand[b0 + off++] = s[i];
  if (lane == 63) ulen[row] = off;
}

// longest row: each wave's maximum by shuffles, one atomic per wave
__global__ void k_len_max(int64_t n, const int64_t *__restrict__ len, int *__restrict__ out) {
  int v = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    v = max(v, static_cast<int>(len[i]));
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

__global__ void k_row_copy(int64_t nrows, const int64_t *__restrict__ cptr, const int32_t *__restrict__ cand,
                           const int64_t *__restrict__ rowptr, int32_t *__restrict__ col) {
  int64_t row = blockIdx.x;
  if (row >= nrows) return;
  int64_t n = rowptr[row + 1] - rowptr[row];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) col[rowptr[row] + i] = cand[cptr[row] + i];
}

// Element rows of each LID (uploaded one-field meshes): p = e*ndof + a lands in the list
// of row elem_lid[p]; the lists are then sorted so that the row gather sums in a fixed order.
__global__ void k_occ_fill(const int32_t *__restrict__ elem_lid, int64_t n, int64_t *__restrict__ cursor,
                           int64_t *__restrict__ occ) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t slot = static_cast<int64_t>(
      atomicAdd(reinterpret_cast<unsigned long long *>(cursor + elem_lid[p]), 1ull));
  occ[slot] = p;
}

__global__ void k_occ_sort(int64_t nrows, const int64_t *__restrict__ optr, int64_t *__restrict__ occ) {
  int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const int64_t lo = optr[row], hi = optr[row + 1];
  for (int64_t i = lo + 1; i < hi; ++i) {   // insertion sort: a row has a few element rows
    const int64_t v = occ[i];
    int64_t j = i;
    for (; j > lo && occ[j - 1] > v; --j) occ[j] = occ[j - 1];
    occ[j] = v;
  }
}

// Per-row stencil slot map for nodal fields on the vertex lattice: slot
// s*nf + g holds the position, within the row, of field g at the neighbour
// point offset s = (dx+1) + 3(dy+1) + 9(dz+1), or 255 when that column is
// not in the graph (outside the mesh or not connected, as for tets).
__global__ void k_stencil_perm(int64_t nx, int64_t ny, int64_t nz, int nf, int W, const int32_t *__restrict__ point_lid,
                               const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                               uint8_t *__restrict__ perm) {
  int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t npts = (nx + 1) * (ny + 1) * (nz + 1);
  if (t >= npts * nf) return;
  int64_t pt = t / nf;
  int f = static_cast<int>(t % nf);
  int64_t i = pt % (nx + 1), j = (pt / (nx + 1)) % (ny + 1), k = pt / ((nx + 1) * (ny + 1));
  int32_t row = point_lid[pt * nf + f];
  int64_t lo0 = rowptr[row], hi0 = rowptr[row + 1];
  uint8_t *pr = perm + static_cast<int64_t>(row) * W;
  for (int s = 27 * nf; s < W; ++s) pr[s] = 255;
  for (int s = 0; s < 27; ++s) {
    int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
    int64_t ii = i + dx, jj = j + dy, kk = k + dz;
    bool in = ii >= 0 && ii <= nx && jj >= 0 && jj <= ny && kk >= 0 && kk <= nz;
    for (int g = 0; g < nf; ++g) {
      uint8_t slot = 255;
      if (in) {
        int32_t key = point_lid[(ii + (nx + 1) * (jj + (ny + 1) * kk)) * nf + g];
        int64_t lo = lo0, hi = hi0;
        while (lo < hi) {
          int64_t m = (lo + hi) >> 1;
          if (col[m] < key) lo = m + 1; else hi = m;
        }
        if (lo < hi0 && col[lo] == key) slot = static_cast<uint8_t>(lo - lo0);
      }
      pr[s * nf + g] = slot;
    }
  }
}

}  // namespace

void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  DevBuf<int64_t> sums, offs;
  sums.alloc(tiles);
  offs.alloc(tiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s, in, out, n, sums.p);
  TXA_HIP(hipGetLastError());
  if (tiles > 1) {
    exclusive_scan_i64(sums.p, offs.p, tiles, s);
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n)), dim3(kBlock), 0, s, out, n, offs.p);
    TXA_HIP(hipGetLastError());
  }
  TXA_HIP(hipStreamSynchronize(s));  // temporaries are freed on return
}

void mesh_generate(Mesh &m, const double *host_coords, hipStream_t s) {
  int64_t NX = m.nx, NY = m.ny, NZ = m.nz;
  int64_t nn = (NX + 1) * (NY + 1) * (NZ + 1);
  int64_t ne = NX * (NY + 1) * (NZ + 1) + (NX + 1) * NY * (NZ + 1) + (NX + 1) * (NY + 1) * NZ;
  int64_t nf = NX * NY * (NZ + 1) + (NX * (NY + 1) + (NX + 1) * NY) * NZ;
  int64_t nc = NX * NY * NZ;
  m.nverts = nn;
  switch (m.topo) {
    case TXA_HEX8: m.nelem = nc; m.nconn = 8; m.nvert_elem = 8; m.nentities = nn; break;
    case TXA_TET4: m.nelem = 6 * nc; m.nconn = 4; m.nvert_elem = 4; m.nentities = nn; break;
    case TXA_HEX27: m.nelem = nc; m.nconn = 27; m.nvert_elem = 8; m.nentities = nn + ne + nf + nc; break;
    default: throw Error(TXA_ERR_INVALID, "unknown topology " + std::to_string(m.topo));
  }
  // SoA coordinates
  std::vector<double> soa(static_cast<size_t>(3 * nn));
  for (int64_t i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) soa[d * nn + i] = host_coords[3 * i + d];
  m.xyz.alloc(3 * nn);
  TXA_HIP(hipMemcpyAsync(m.xyz.p, soa.data(), sizeof(double) * 3 * nn, hipMemcpyHostToDevice, s));
  m.nconn_tot = m.nelem * m.nconn;
  m.blocks.assign(1, MeshBlock{m.topo, m.nconn, m.nvert_elem, m.nelem, 0, 0});
  m.conn.alloc(m.nelem * m.nconn);
  BrickIds b{NX, NY, NZ, nn, ne, nf};
  hipLaunchKernelGGL(k_brick_conn, dim3(grid_for(nc)), dim3(kBlock), 0, s, m.topo, b, nc, m.conn.p);
  TXA_HIP(hipGetLastError());
  TXA_HIP(hipStreamSynchronize(s));
  m.affine = brick_cells_affine(m, s);
}

// Connectivity width and geometry vertices of an uploadable topology.
static void topo_sizes(int topo, int *nconn, int *nvert) {
  switch (topo) {
    case TXA_HEX8: *nconn = 8; *nvert = 8; break;
    case TXA_TET4: *nconn = 4; *nvert = 4; break;
    case TXA_WEDGE6: *nconn = 6; *nvert = 6; break;
    case TXA_HEX27: *nconn = 27; *nvert = 8; break;
    default: throw Error(TXA_ERR_INVALID, "unknown topology " + std::to_string(topo));
  }
}

// A general mesh of one or several element blocks: connectivity as entity ids (the first
// nvert of an element are its vertices, in [0, nverts); HEX27 adds its edges, faces and
// cell as further ids below nentities), vertex coordinates [nverts][3].  Validated on the
// host; the blocks' connectivity is stored one block after another.
void mesh_upload_blocks(Mesh &m, int nblocks, const txa_element_block *blocks, int64_t nentities, int64_t nverts,
                        const double *host_coords, hipStream_t s) {
  TXA_REQUIRE(nblocks >= 1 && blocks, TXA_ERR_INVALID, "a mesh needs at least one element block");
  TXA_REQUIRE(nverts >= 1 && nentities >= nverts, TXA_ERR_INVALID, "need vertices and nentities >= nverts");
  m.blocks.clear();
  int64_t elem = 0, pos = 0;
  for (int k = 0; k < nblocks; ++k) {
    const txa_element_block &in = blocks[k];
    MeshBlock b;
    b.topo = in.topo;
    topo_sizes(in.topo, &b.nconn, &b.nvert_elem);
    TXA_REQUIRE(in.nelem >= 1 && in.conn, TXA_ERR_INVALID,
                "block " + std::to_string(k) + ": need elements and their connectivity");
    TXA_REQUIRE(b.nvert_elem <= nverts, TXA_ERR_INVALID, "fewer vertices than an element has");
    for (int64_t e = 0; e < in.nelem; ++e)
      for (int c = 0; c < b.nconn; ++c) {
        const int64_t id = in.conn[e * b.nconn + c];
        const int64_t hi = c < b.nvert_elem ? nverts : nentities;
        TXA_REQUIRE(id >= 0 && id < hi, TXA_ERR_INVALID,
                    (nblocks > 1 ? "block " + std::to_string(k) + " " : std::string()) + "element " +
                        std::to_string(e) + " entry " + std::to_string(c) + ": entity id " + std::to_string(id) +
                        (c < b.nvert_elem ? " is not a vertex" : " is out of range"));
      }
    b.nelem = in.nelem;
    b.elem_begin = elem;
    b.conn_begin = pos;
    elem += in.nelem;
    pos += in.nelem * b.nconn;
    m.blocks.push_back(b);
  }
  bool same = true;
  for (const MeshBlock &b : m.blocks) same = same && b.topo == m.blocks[0].topo;
  m.topo = same ? m.blocks[0].topo : 0;
  m.nconn = same ? m.blocks[0].nconn : 0;
  m.nvert_elem = same ? m.blocks[0].nvert_elem : 0;
  m.structured = false;
  m.nelem = elem;
  m.nconn_tot = pos;
  m.nverts = nverts;
  m.nentities = nentities;
  std::vector<double> soa(static_cast<size_t>(3 * nverts));
  for (int64_t i = 0; i < nverts; ++i)
    for (int d = 0; d < 3; ++d) soa[d * nverts + i] = host_coords[3 * i + d];
  m.xyz.alloc(3 * nverts);
  TXA_HIP(hipMemcpyAsync(m.xyz.p, soa.data(), sizeof(double) * 3 * nverts, hipMemcpyHostToDevice, s));
  m.conn.alloc(pos);
  for (const MeshBlock &b : m.blocks)
    TXA_HIP(hipMemcpyAsync(m.conn.p + b.conn_begin, blocks[&b - m.blocks.data()].conn,
                           sizeof(int64_t) * b.nelem * b.nconn, hipMemcpyHostToDevice, s));
  TXA_HIP(hipStreamSynchronize(s));
}

void mesh_upload(Mesh &m, int64_t nelem, const int64_t *host_conn, int64_t nentities, int64_t nverts,
                 const double *host_coords, hipStream_t s) {
  txa_element_block b{m.topo, nelem, host_conn};
  mesh_upload_blocks(m, 1, &b, nentities, nverts, host_coords, s);
}

void dofmap_build(Mesh &m, int nfields, hipStream_t s) {
  TXA_REQUIRE(nfields >= 1 && nfields <= 3, TXA_ERR_INVALID, "nfields must be 1..3");
  m.nfields = nfields;
  m.ndof_elem = m.nconn * nfields;
  int64_t ncon_tot = m.nconn_tot;
  int64_t npos = ncon_tot * nfields;
  // GUN on one rank: rank of each entity id among the ids present
  DevBuf<int64_t> flag, rank;
  flag.alloc(m.nentities);
  rank.alloc(m.nentities);
  TXA_HIP(hipMemsetAsync(flag.p, 0, flag.bytes(), s));
  hipLaunchKernelGGL(k_mark_entities, dim3(grid_for(ncon_tot)), dim3(kBlock), 0, s, m.conn.p, ncon_tot, flag.p);
  TXA_HIP(hipGetLastError());
  exclusive_scan_i64(flag.p, rank.p, m.nentities, s);
  int64_t last_rank = 0, last_flag = 0;
  TXA_HIP(hipMemcpy(&last_rank, rank.p + m.nentities - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  TXA_HIP(hipMemcpy(&last_flag, flag.p + m.nentities - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t nuniq = last_rank + last_flag;
  m.ndofs = nuniq * nfields;
  TXA_REQUIRE(m.ndofs < 0x7fffffffLL, TXA_ERR_UNSUPPORTED, "more than 2^31-1 DOFs: LIDs are int32");
  flag.release();
  // element GIDs and first touch
  DevBuf<int64_t> gid, fpflag, pos;
  DevBuf<unsigned long long> first;
  gid.alloc(npos);
  first.alloc(m.ndofs);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(m.ndofs)), dim3(kBlock), 0, s,
                     reinterpret_cast<int64_t *>(first.p), m.ndofs, static_cast<int64_t>(0x7fffffffffffffffLL));
  hipLaunchKernelGGL(k_elem_gid, dim3(grid_for(ncon_tot)), dim3(kBlock), 0, s, m.conn.p, ncon_tot, nfields, rank.p,
                     gid.p, first.p);
  TXA_HIP(hipGetLastError());
  rank.release();
  fpflag.alloc(npos);
  pos.alloc(npos);
  hipLaunchKernelGGL(k_first_flag, dim3(grid_for(npos)), dim3(kBlock), 0, s, gid.p, npos, first.p, fpflag.p);
  TXA_HIP(hipGetLastError());
  exclusive_scan_i64(fpflag.p, pos.p, npos, s);
  fpflag.release();
  m.owned.alloc(m.ndofs);
  m.elem_lid.alloc(npos);
  hipLaunchKernelGGL(k_assign_lids, dim3(grid_for(npos)), dim3(kBlock), 0, s, gid.p, npos, first.p, pos.p, m.owned.p,
                     m.elem_lid.p);
  TXA_HIP(hipGetLastError());
  // LID of every lattice point and field (bricks only)
  m.npoints = 0;
  if (m.structured) {
    m.npoints = m.topo == TXA_HEX27
                    ? static_cast<int64_t>(2 * m.nx + 1) * (2 * m.ny + 1) * (2 * m.nz + 1)
                    : m.nverts;
    m.point_lid.alloc(m.npoints * nfields);
    hipLaunchKernelGGL(k_point_lid, dim3(grid_for(ncon_tot)), dim3(kBlock), 0, s, m.topo, static_cast<int64_t>(m.nx),
                       static_cast<int64_t>(m.ny), m.conn.p, m.nconn, m.nelem, nfields, m.elem_lid.p, m.point_lid.p);
    TXA_HIP(hipGetLastError());
  }
  TXA_HIP(hipStreamSynchronize(s));
  m.has_graph = false;
}

void graph_build(Mesh &m, hipStream_t s) {
  TXA_REQUIRE(m.ndofs > 0, TXA_ERR_LOGIC, "graph_build before dofmap_build");
  int nd = m.ndof_elem;
  int64_t npos = m.npos();
  int64_t nrows = m.ndofs;
  // a brick's rows have at most valence x nd candidates; an uploaded mesh's rows are
  // checked against the capacity while they are sorted
  int valence = m.topo == TXA_TET4 ? 24 : 8;
  TXA_REQUIRE(!m.structured || static_cast<int64_t>(valence) * nd <= kMaxCand, TXA_ERR_UNSUPPORTED,
              "row candidate count exceeds the per-row sort capacity");
  DevBuf<int64_t> cnt, cptr, cursor, ulen;
  cnt.alloc(nrows + 1);
  cptr.alloc(nrows + 1);
  TXA_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
  // candidates: every element row of every block contributes the element's LIDs
  for (const MeshBlock &b : m.blocks) {
    const int bnd = b.nconn * m.nfields;
    const int64_t bn = b.nelem * bnd;
    hipLaunchKernelGGL(k_row_count, dim3(grid_for(bn)), dim3(kBlock), 0, s, m.elem_lid.p + m.pos_begin(b), bn, bnd,
                       cnt.p);
    TXA_HIP(hipGetLastError());
  }
  exclusive_scan_i64(cnt.p, cptr.p, nrows + 1, s);
  int64_t ncand = 0;
  TXA_HIP(hipMemcpy(&ncand, cptr.p + nrows, sizeof(int64_t), hipMemcpyDeviceToHost));
  cnt.release();
  cursor.alloc(nrows + 1);
  TXA_HIP(hipMemcpyAsync(cursor.p, cptr.p, sizeof(int64_t) * (nrows + 1), hipMemcpyDeviceToDevice, s));
  DevBuf<int32_t> cand;
  cand.alloc(ncand);
  for (const MeshBlock &b : m.blocks) {
    const int bnd = b.nconn * m.nfields;
    const int64_t bn = b.nelem * bnd;
    hipLaunchKernelGGL(k_row_fill, dim3(grid_for(bn)), dim3(kBlock), 0, s, m.elem_lid.p + m.pos_begin(b), bn, bnd,
                       cursor.p, cand.p);
    TXA_HIP(hipGetLastError());
  }
  cursor.release();
  ulen.alloc(nrows + 1);
  TXA_HIP(hipMemsetAsync(ulen.p, 0, ulen.bytes(), s));
  DevBuf<int> overflow;
  overflow.alloc(1);
  TXA_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
  TXA_REQUIRE(nrows < 0x7fffffffLL, TXA_ERR_UNSUPPORTED, "too many rows for one launch");
  hipLaunchKernelGGL(k_row_sort_unique, dim3(static_cast<unsigned>(nrows)), dim3(64), 0, s, nrows, cptr.p, cand.p,
                     ulen.p, overflow.p);
  TXA_HIP(hipGetLastError());
  int ovf = 0;
  TXA_HIP(hipMemcpy(&ovf, overflow.p, sizeof(int), hipMemcpyDeviceToHost));
  TXA_REQUIRE(ovf == 0, TXA_ERR_UNSUPPORTED, "a graph row has more than 1024 candidate columns (element valence)");
  TXA_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_len_max, dim3(std::min(grid_for(nrows), 2048u)), dim3(kBlock), 0, s, nrows, ulen.p, overflow.p);
  TXA_HIP(hipGetLastError());
  TXA_HIP(hipMemcpy(&m.max_row, overflow.p, sizeof(int), hipMemcpyDeviceToHost));
  m.rowptr.alloc(nrows + 1);
  exclusive_scan_i64(ulen.p, m.rowptr.p, nrows + 1, s);
  TXA_HIP(hipMemcpy(&m.nnz, m.rowptr.p + nrows, sizeof(int64_t), hipMemcpyDeviceToHost));
  m.col.alloc(m.nnz);
  hipLaunchKernelGGL(k_row_copy, dim3(static_cast<unsigned>(nrows)), dim3(64), 0, s, nrows, cptr.p, cand.p,
                     m.rowptr.p, m.col.p);
  TXA_HIP(hipGetLastError());
  // stencil slot map for the vertex-lattice kernels
  if (m.structured && (m.topo == TXA_HEX8 || m.topo == TXA_TET4)) {
    m.perm_width = (27 * m.nfields + 15) / 16 * 16;
    m.perm.alloc(static_cast<size_t>(nrows) * m.perm_width);
    int64_t nt = m.nverts * m.nfields;
    hipLaunchKernelGGL(k_stencil_perm, dim3(grid_for(nt)), dim3(kBlock), 0, s, static_cast<int64_t>(m.nx),
                       static_cast<int64_t>(m.ny), static_cast<int64_t>(m.nz), m.nfields, m.perm_width, m.point_lid.p,
                       m.rowptr.p, m.col.p, m.perm.p);
    TXA_HIP(hipGetLastError());
  }
  node_perm_build(m, s);   // neighbour positions for the 3-field elasticity kernels
  // element rows of each LID for the row gather of uploaded one-field hex8 / tet4 meshes
  m.occ_ptr.release();
  m.occ.release();
  bool lane_blocks = !m.structured && m.nfields == 1;
  for (const MeshBlock &b : m.blocks)
    lane_blocks = lane_blocks && (b.topo == TXA_HEX8 || b.topo == TXA_TET4 || b.topo == TXA_WEDGE6);
  if (lane_blocks) {
    DevBuf<int64_t> ocnt, ocur;
    ocnt.alloc(nrows + 1);
    TXA_HIP(hipMemsetAsync(ocnt.p, 0, ocnt.bytes(), s));
    hipLaunchKernelGGL(k_row_count, dim3(grid_for(npos)), dim3(kBlock), 0, s, m.elem_lid.p, npos, 1, ocnt.p);
    TXA_HIP(hipGetLastError());
    m.occ_ptr.alloc(nrows + 1);
    exclusive_scan_i64(ocnt.p, m.occ_ptr.p, nrows + 1, s);
    ocnt.release();
    ocur.alloc(nrows + 1);
    TXA_HIP(hipMemcpyAsync(ocur.p, m.occ_ptr.p, sizeof(int64_t) * (nrows + 1), hipMemcpyDeviceToDevice, s));
    m.occ.alloc(npos);
    hipLaunchKernelGGL(k_occ_fill, dim3(grid_for(npos)), dim3(kBlock), 0, s, m.elem_lid.p, npos, ocur.p, m.occ.p);
    TXA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_occ_sort, dim3(grid_for(nrows)), dim3(kBlock), 0, s, nrows, m.occ_ptr.p, m.occ.p);
    TXA_HIP(hipGetLastError());
    TXA_HIP(hipStreamSynchronize(s));   // ocur is freed at scope end
  }
  patch_build(m, s);   // node patches of single-block hex8 / tet4 meshes
  TXA_HIP(hipStreamSynchronize(s));
  m.has_graph = true;
}

}  // namespace txa
