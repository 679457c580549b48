// tc_abi.hip -- extern "C" entry points of the tessellation stages.
//
// Host code only validates arguments, assembles the kernel argument block
// and enqueues kernels on the caller's stream.  No call synchronises the
// stream; nothing here allocates except tc_table_upload.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "tc_sort.h"
#include "tc_stage.h"

namespace tc {

#include "tess_table_nv_data.inc"

constexpr uint32_t kMaxInstances = 1u << 16;

static uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }

WsLayout ws_layout(const tc_config& cfg, uint32_t) {
  WsLayout L{};
  uint32_t tilesSplit = (cfg.maxSplitTriangles + kWave - 1) / kWave;
  uint32_t tilesPart = (cfg.maxPartTriangles + kWave - 1) / kWave;
  L.maxTiles = std::max(std::max(cfg.maxVisibleClusters, tilesSplit), std::max(tilesPart, 1u));
  L.sortCapacity = std::max(cfg.maxGeneratedClusters, 1u);
  // tiles of the larger of the CLAS sort and the TLAS instance sort
  L.sortMaxTiles = (std::max(L.sortCapacity, kMaxInstances) + kSortTileKeys - 1) / kSortTileKeys;
  L.instanceCapacity = kMaxInstances;
  uint64_t o = kWsHeaderBytes;
  // the radix sort's state follows the stage header at a fixed offset
  L.sortState = o;
  o = align256(o + sizeof(SortState));
  L.status = o;
  o = align256(o + (uint64_t)L.maxTiles * kMaxCounters * 8);
  for (int i = 0; i < 2; i++) {
    L.sortKeys[i] = o;
    o = align256(o + (uint64_t)L.sortCapacity * 8);
  }
  for (int i = 0; i < 2; i++) {
    L.sortVals[i] = o;
    o = align256(o + (uint64_t)L.sortCapacity * 4);
  }
  L.sortStatus = o;
  o = align256(o + (uint64_t)kSortRadix * L.sortMaxTiles * 8);
  L.sortAux = o;
  o = align256(o + (uint64_t)kMaxInstances * 4);
  L.cullOffsets = o;
  o = align256(o + (uint64_t)kMaxInstances * 4);
  L.total = o;
  return L;
}

static int g_numCU = 0;
static std::once_flag g_cuOnce;

static int num_cu() {
  std::call_once(g_cuOnce, [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      g_numCU = n;
    if (g_numCU <= 0) g_numCU = 256;
  });
  return g_numCU;
}

static bool config_ok(const tc_config& c) {
  if (c.clusterTriangleCount == 0 || c.clusterTriangleCount > kMaxClusterTris) return false;
  if (c.clusterVertexCount == 0 || c.clusterVertexCount > kMaxClusterVerts) return false;
  if (c.maxSplitFactor < 2 || c.maxSplitFactor > TC_TESSTABLE_SIZE) return false;
  if (c.maxVisibleClusters == 0 || c.maxPartTriangles == 0 || c.maxSplitTriangles == 0) return false;
  if (c.maxGeneratedClusters == 0 || c.maxGeneratedVertices == 0) return false;
  return true;
}

static constexpr uint32_t kSupportedFlags =
    TC_FLAG_DISPLACEMENT | TC_FLAG_CULLING | TC_FLAG_PN | TC_FLAG_VERTEX_TAPE | TC_FLAG_1X_TRANSIENT |
    TC_FLAG_2X_TRANSIENT | TC_FLAG_ANIMATION | TC_FLAG_PERSISTENT_SPLIT | TC_FLAG_HIZ;

// useWorkspace false: the caller's workspace is not used (nor required);
// the argument block's ws / status / maxTiles are left for the caller.
static tc_status make_args(const tc_bindings* b, const tc_config& cfg, KArgs& a, WsLayout* layout = nullptr,
                           bool useWorkspace = true) {
  if (!b || !b->view || !b->tessTable || !b->readback || !b->instances || !b->build) return TC_ERROR_INVALID_ARGUMENT;
  if (useWorkspace && !b->workspace) return TC_ERROR_INVALID_ARGUMENT;
  if (!config_ok(cfg)) return TC_ERROR_INVALID_ARGUMENT;
  if (cfg.flags & ~kSupportedFlags) return TC_ERROR_UNSUPPORTED;
  if (b->numRenderInstances == 0 || b->numRenderInstances > kMaxInstances) return TC_ERROR_INVALID_ARGUMENT;
  if ((cfg.flags & TC_FLAG_DISPLACEMENT) && b->numDisplacementTextures && !b->displacementTextures) return TC_ERROR_INVALID_ARGUMENT;
  if ((cfg.flags & TC_FLAG_VERTEX_TAPE) && !b->vertexTape) return TC_ERROR_INVALID_ARGUMENT;
  // the depth adjoint differentiates displacement only (not the ripple) on
  // cluster and part vertices (not on 2X batch vertices)
  if ((cfg.flags & TC_FLAG_VERTEX_TAPE) && (cfg.flags & (TC_FLAG_ANIMATION | TC_FLAG_2X_TRANSIENT))) return TC_ERROR_UNSUPPORTED;
  WsLayout L = ws_layout(cfg, b->numRenderInstances);
  if (useWorkspace && b->workspaceBytes < L.total) return TC_ERROR_INVALID_ARGUMENT;
  a.view = b->view;
  a.tessTable = b->tessTable;
  a.readback = b->readback;
  a.instances = b->instances;
  a.build = b->build;
  a.textures = b->displacementTextures;
  a.numInstances = b->numRenderInstances;
  a.numTextures = (cfg.flags & TC_FLAG_DISPLACEMENT) ? b->numDisplacementTextures : 0;
  a.ws = useWorkspace ? reinterpret_cast<WsHeader*>(b->workspace) : nullptr;
  a.status = useWorkspace ? reinterpret_cast<uint64_t*>(static_cast<char*>(b->workspace) + L.status) : nullptr;
  a.maxTiles = useWorkspace ? L.maxTiles : 0;
  a._pad = 0;
  a.tape = (cfg.flags & TC_FLAG_VERTEX_TAPE) ? b->vertexTape : nullptr;
  a.cfg = cfg;
  a.viewLast = b->viewLast ? b->viewLast : b->view;
  a.hiz = b->hiz;
  // a bound HiZ must be a whole pyramid
  if ((cfg.flags & TC_FLAG_HIZ) && a.hiz.texels &&
      (a.hiz.size == 0 || (a.hiz.size & (a.hiz.size - 1)) || a.hiz.mipLevels == 0 || a.hiz.mipLevels > 16 ||
       (a.hiz.size >> (a.hiz.mipLevels - 1)) != 1))
    return TC_ERROR_INVALID_ARGUMENT;
  if (!(cfg.flags & TC_FLAG_HIZ)) a.hiz.texels = nullptr;
  if (layout) *layout = L;
  return TC_OK;
}

static SortBuffers sort_buffers(const tc_bindings* b, const WsLayout& L) {
  char* w = static_cast<char*>(b->workspace);
  SortBuffers sb;
  for (int i = 0; i < 2; i++) {
    sb.keys[i] = reinterpret_cast<uint64_t*>(w + L.sortKeys[i]);
    sb.vals[i] = reinterpret_cast<uint32_t*>(w + L.sortVals[i]);
  }
  sb.status = reinterpret_cast<uint64_t*>(w + L.sortStatus);
  sb.state = reinterpret_cast<SortState*>(w + L.sortState);
  sb.capacity = L.sortCapacity;
  sb.maxTiles = L.sortMaxTiles;
  return sb;
}

tc_status make_args_public(const tc_bindings* b, const tc_config& cfg, KArgs& a) { return make_args(b, cfg, a); }
int num_cu_public() { return num_cu(); }

// defaults from the autotuner on the 64-instance grid frame: traversal with
// about one ray per thread (256-thread workgroups, 32 per CU) hides its
// memory latency best
// The last field is the sleeps (s_sleep 2, about 128 clocks each) before a
// persistent-split wait gives up: seconds, far beyond any frame's split time.
static constexpr uint32_t kSplitWaitLimitDefault = 1u << 24;
static constexpr tc_tuning kDefaultTuning = {5, 8, 2, 64, 512, 64, 8, kSplitWaitLimitDefault};
static std::mutex g_tuningMutex;
static tc_tuning g_tuning = kDefaultTuning;

tc_tuning tuning_snapshot() {
  std::lock_guard<std::mutex> lock(g_tuningMutex);
  return g_tuning;
}

extern "C" tc_status tc_get_tuning(tc_tuning* out) {
  if (!out) return TC_ERROR_INVALID_ARGUMENT;
  *out = tuning_snapshot();
  return TC_OK;
}

extern "C" tc_status tc_set_tuning(const tc_tuning* t) {
  if (t) {
    const uint32_t v[7] = {t->classifyBlocksPerCU, t->splitBlocksPerCU, t->instantiateBlocksPerCU, t->clasWavesPerCU,
                           t->traceBlocksPerCU, t->traceThreads, t->bvhBlocksPerCU};
    for (uint32_t x : v)
      if (x == 0 || x > 4096) return TC_ERROR_INVALID_ARGUMENT;
    if (t->traceThreads != 64 && t->traceThreads != 128 && t->traceThreads != 256) return TC_ERROR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_tuningMutex);
  g_tuning = t ? *t : kDefaultTuning;
  if (g_tuning.splitWaitLimit == 0) g_tuning.splitWaitLimit = kSplitWaitLimitDefault;
  return TC_OK;
}

static tc_status launched() { return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH; }

// ------------------------------------------------------ table templates

static uint32_t round16(uint32_t v) { return (v + 15u) & ~15u; }

// 30-bit Morton code of a point in [0,1]^3 (10 bits per axis)
static uint32_t morton3(float x, float y, float z) {
  auto q = [](float f) {
    float c = std::min(std::max(f, 0.0f), 1.0f);
    uint32_t v = (uint32_t)(c * 1023.0f + 0.5f);
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  return q(x) | (q(y) << 1) | (q(z) << 2);
}

// Orders triangles along a Morton curve of their centroids (normalised to
// the cluster's bounds) so runs of TC_CLAS_GROUP_TRIANGLES are compact.
static void order_triangles(uint32_t nt, const uint32_t* words, const float* pos3, uint32_t* out) {
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  std::vector<float> cen(nt * 3);
  for (uint32_t t = 0; t < nt; t++)
    for (int k = 0; k < 3; k++) {
      float c = 0;
      for (int v = 0; v < 3; v++) c += pos3[((words[t] >> (8 * v)) & 0xFF) * 3 + k];
      c /= 3.0f;
      cen[t * 3 + k] = c;
      lo[k] = std::min(lo[k], c);
      hi[k] = std::max(hi[k], c);
    }
  std::vector<std::pair<uint32_t, uint32_t>> keys(nt);
  for (uint32_t t = 0; t < nt; t++) {
    float n[3];
    for (int k = 0; k < 3; k++) n[k] = hi[k] > lo[k] ? (cen[t * 3 + k] - lo[k]) / (hi[k] - lo[k]) : 0.5f;
    keys[t] = {morton3(n[0], n[1], n[2]), t};
  }
  std::stable_sort(keys.begin(), keys.end(), [](auto& a, auto& b) { return a.first < b.first; });
  for (uint32_t i = 0; i < nt; i++) out[i] = words[keys[i].second];
}

struct TableHost {
  tc_tess_table_entry lookup[TC_TESSTABLE_NUM_LOOKUP];
  std::vector<uint32_t> templWords;                      // all template triangle words
  std::vector<tc_clas_template> templ;                   // per lookup index (device addresses patched)
  std::vector<uint32_t> templWordOffset;                 // per lookup index
  uint32_t sizes[TC_TESSTABLE_NUM_LOOKUP];

  TableHost() {
    memset(lookup, 0, sizeof(lookup));
    memset(sizes, 0, sizeof(sizes));
    // tessellation_table.cpp:53-80
    uint32_t idx = 0;
    for (uint32_t x = 1; x <= TC_TABLE_MAX_EDGE_SEGMENTS; x++)
      for (uint32_t y = 1; y <= x; y++)
        for (uint32_t z = 1; z <= y; z++, idx++) {
          tc_tess_table_entry e{tc_table_nv_configs[idx * 4 + 0], tc_table_nv_configs[idx * 4 + 1], tc_table_nv_configs[idx * 4 + 2],
                                tc_table_nv_configs[idx * 4 + 3]};
          lookup[x + y * 16 + z * 256 - 273] = e;
          if (z != y && x > 1) lookup[x + z * 16 + y * 256 - 273] = e;
        }
    // per-pattern templates: one per used lookup entry; the (x,z,y) alias
    // (z > y) is the flipped pattern (tessellation_table.cpp:116-153)
    templ.assign(TC_TESSTABLE_NUM_LOOKUP, tc_clas_template{});
    templWordOffset.assign(TC_TESSTABLE_NUM_LOOKUP, 0);
    for (uint32_t li = 0; li < TC_TESSTABLE_NUM_LOOKUP; li++) {
      const tc_tess_table_entry& e = lookup[li];
      if (e.numTriangles == 0) continue;
      uint32_t f = li + 273;
      uint32_t fy = (f >> 4) & 15, fz = (f >> 8) & 15;
      bool flipped = fz > fy;
      std::vector<float> uv(e.numVertices * 3);
      for (uint32_t v = 0; v < e.numVertices; v++) {
        uint32_t raw = tc_table_nv_vertices[e.firstVertex + v];
        float u = float(raw & 0xFFFF) / 32768.0f, w = float(raw >> 16) / 32768.0f;
        float x = (1.0f - u) - w;
        if (flipped) std::swap(x, u);  // barycentric (w,u,v) -> (u,w,v)
        uv[v * 3 + 0] = u;
        uv[v * 3 + 1] = w;
        uv[v * 3 + 2] = 0.0f;
      }
      std::vector<uint32_t> words(e.numTriangles), ordered(e.numTriangles);
      for (uint32_t t = 0; t < e.numTriangles; t++) {
        uint32_t packed = tc_table_nv_triangles[e.firstTriangle + t];
        uint32_t i0 = packed & 0xFF, i1 = (packed >> 8) & 0xFF, i2 = (packed >> 16) & 0xFF;
        if (flipped) std::swap(i1, i2);
        words[t] = i0 | (i1 << 8) | (i2 << 16) | (t << 24);
      }
      order_triangles(e.numTriangles, words.data(), uv.data(), ordered.data());
      templWordOffset[li] = (uint32_t)templWords.size();
      templWords.insert(templWords.end(), ordered.begin(), ordered.end());
      tc_clas_template& T = templ[li];
      T.numTriangles = e.numTriangles;
      T.numVertices = e.numVertices;
      T.numGroups = (e.numTriangles + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
      T.clusterID = 0;
      sizes[li] = tc_clas_size(e.numTriangles, e.numVertices);
    }
  }
};

static const TableHost& table_host() {
  static TableHost t;
  return t;
}

// device allocation bundle behind one uploaded table
struct TableDev {
  tc_tessellation_table hdr;
  void* mem;
};

}  // namespace tc

using namespace tc;

extern "C" {

tc_status tc_device_check(int* deviceCount) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    if (deviceCount) *deviceCount = 0;
    return TC_ERROR_NO_DEVICE;
  }
  if (deviceCount) *deviceCount = n;
  return TC_OK;
}

uint64_t tc_workspace_size(const tc_config* config) {
  if (!config || !config_ok(*config)) return 0;
  return ws_layout(*config, 1).total;
}

tc_status tc_workspace_init(const tc_bindings* b, const tc_config* config, tc_stream stream) {
  if (!b || !config || !b->workspace) return TC_ERROR_INVALID_ARGUMENT;
  if (!config_ok(*config)) return TC_ERROR_INVALID_ARGUMENT;
  WsLayout L = ws_layout(*config, 1);
  if (b->workspaceBytes < L.total) return TC_ERROR_INVALID_ARGUMENT;
  if (hipMemsetAsync(b->workspace, 0, L.total, (hipStream_t)stream) != hipSuccess) return TC_ERROR_LAUNCH;
  return TC_OK;
}

uint32_t tc_clas_size(uint32_t numTriangles, uint32_t numVertices) {
  uint32_t groups = (numTriangles + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
  return (uint32_t)sizeof(tc_clas_header) + round16(groups * (uint32_t)sizeof(tc_aabb)) + round16(numTriangles * 4u) +
         round16(numVertices * 12u);
}

tc_status tc_cluster_template_order(uint32_t numTriangles, uint32_t numVertices, const uint8_t* localTriangles,
                                    const float* positions, uint32_t* trianglesOut) {
  if (!localTriangles || !positions || !trianglesOut || numTriangles == 0 || numTriangles > 256 || numVertices > 256)
    return TC_ERROR_INVALID_ARGUMENT;
  std::vector<uint32_t> words(numTriangles);
  for (uint32_t t = 0; t < numTriangles; t++) {
    uint32_t i0 = localTriangles[t * 3], i1 = localTriangles[t * 3 + 1], i2 = localTriangles[t * 3 + 2];
    if (i0 >= numVertices || i1 >= numVertices || i2 >= numVertices) return TC_ERROR_INVALID_ARGUMENT;
    words[t] = i0 | (i1 << 8) | (i2 << 16) | ((t & 0xFF) << 24);
  }
  order_triangles(numTriangles, words.data(), positions, trianglesOut);
  return TC_OK;
}

tc_status tc_table_lookup_host(tc_tess_table_entry* out4096) {
  if (!out4096) return TC_ERROR_INVALID_ARGUMENT;
  memcpy(out4096, table_host().lookup, sizeof(table_host().lookup));
  return TC_OK;
}

tc_status tc_table_templates_host(tc_clas_template* out4096, uint32_t* sizes4096, uint32_t* words, uint32_t* numWords) {
  const TableHost& H = table_host();
  if (numWords) {
    if (words && *numWords < H.templWords.size()) return TC_ERROR_INVALID_ARGUMENT;
    *numWords = (uint32_t)H.templWords.size();
  }
  if (words) memcpy(words, H.templWords.data(), H.templWords.size() * 4);
  if (out4096)
    for (uint32_t li = 0; li < TC_TESSTABLE_NUM_LOOKUP; li++) {
      out4096[li] = H.templ[li];
      out4096[li].triangles = H.templWordOffset[li];
    }
  if (sizes4096) memcpy(sizes4096, H.sizes, sizeof(H.sizes));
  return TC_OK;
}

tc_status tc_memcpy(void* dst, const void* src, uint64_t bytes) {
  if ((!dst || !src) && bytes) return TC_ERROR_INVALID_ARGUMENT;
  if (!bytes) return TC_OK;
  // after all queued device work, on every stream (readbacks follow frames
  // queued on non-blocking streams)
  if (hipDeviceSynchronize() != hipSuccess) return TC_ERROR_LAUNCH;
  return hipMemcpy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

tc_status tc_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH; }

tc_status tc_table_upload(tc_tessellation_table** tableDev) {
  if (!tableDev) return TC_ERROR_INVALID_ARGUMENT;
  const TableHost& H = table_host();
  // one allocation: header | vertices | triangles | lookup | templates |
  // template words | template addresses | sizes
  uint64_t o = 256;
  uint64_t oV = o;
  o = align256(o + sizeof(tc_table_nv_vertices));
  uint64_t oT = o;
  o = align256(o + sizeof(tc_table_nv_triangles));
  uint64_t oL = o;
  o = align256(o + sizeof(H.lookup));
  uint64_t oTm = o;
  o = align256(o + H.templ.size() * sizeof(tc_clas_template));
  uint64_t oW = o;
  o = align256(o + H.templWords.size() * 4);
  uint64_t oA = o;
  o = align256(o + TC_TESSTABLE_NUM_LOOKUP * 8);
  uint64_t oS = o;
  o = align256(o + sizeof(H.sizes));
  void* mem = nullptr;
  if (hipMalloc(&mem, o) != hipSuccess) return TC_ERROR_OUT_OF_MEMORY;
  uint64_t base = reinterpret_cast<uint64_t>(mem);
  std::vector<char> host(o, 0);
  tc_tessellation_table hdr;
  hdr.vertices = base + oV;
  hdr.triangles = base + oT;
  hdr.entries = base + oL;
  hdr.templateAddresses = base + oA;
  hdr.templateInstantiationSizes = base + oS;
  memcpy(host.data(), &hdr, sizeof(hdr));
  memcpy(host.data() + oV, tc_table_nv_vertices, sizeof(tc_table_nv_vertices));
  memcpy(host.data() + oT, tc_table_nv_triangles, sizeof(tc_table_nv_triangles));
  memcpy(host.data() + oL, H.lookup, sizeof(H.lookup));
  std::vector<tc_clas_template> templ = H.templ;
  std::vector<uint64_t> addrs(TC_TESSTABLE_NUM_LOOKUP, 0);
  for (uint32_t li = 0; li < TC_TESSTABLE_NUM_LOOKUP; li++) {
    if (templ[li].numTriangles == 0) continue;
    templ[li].triangles = base + oW + (uint64_t)H.templWordOffset[li] * 4;
    addrs[li] = base + oTm + (uint64_t)li * sizeof(tc_clas_template);
  }
  memcpy(host.data() + oTm, templ.data(), templ.size() * sizeof(tc_clas_template));
  memcpy(host.data() + oW, H.templWords.data(), H.templWords.size() * 4);
  memcpy(host.data() + oA, addrs.data(), addrs.size() * 8);
  memcpy(host.data() + oS, H.sizes, sizeof(H.sizes));
  if (hipMemcpy(mem, host.data(), o, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return TC_ERROR_LAUNCH;
  }
  *tableDev = reinterpret_cast<tc_tessellation_table*>(mem);
  return TC_OK;
}

tc_status tc_table_free(tc_tessellation_table* tableDev) {
  if (!tableDev) return TC_ERROR_INVALID_ARGUMENT;
  return hipFree(tableDev) == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

uint32_t tc_num_split_passes(uint32_t splitFactor) {
  // renderer_raytrace_clusters_tess.cpp:516-520
  uint32_t coord = TC_TESSTABLE_COORD_MAX, passes = 0;
  splitFactor = std::max(2u, std::min(splitFactor, TC_TESSTABLE_SIZE));
  while (coord > splitFactor) {
    coord /= splitFactor;
    passes++;
  }
  return passes;
}

tc_status tc_build_setup(const tc_bindings* b, tc_config config, int mode, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  if (mode < TC_BUILD_SETUP_CLASSIFY || mode > TC_BUILD_SETUP_BUILD_BLAS) return TC_ERROR_INVALID_ARGUMENT;
  hipLaunchKernelGGL(k_build_setup, dim3(1), dim3(1), 0, (hipStream_t)stream, a, mode);
  return launched();
}

tc_status tc_instances_classify(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  hipLaunchKernelGGL(k_instances_classify, dim3((a.numInstances + 63) / 64), dim3(64), 0, (hipStream_t)stream, a);
  return launched();
}

tc_status tc_clusters_cull(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  WsLayout L;
  tc_status s = make_args(b, config, a, &L);
  if (s != TC_OK) return s;
  uint32_t* offsets = reinterpret_cast<uint32_t*>(static_cast<char*>(b->workspace) + L.cullOffsets);
  hipLaunchKernelGGL(k_clusters_cull_scan, dim3(1), dim3(kWave), 0, (hipStream_t)stream, a, offsets);
  hipLaunchKernelGGL(k_clusters_cull_fill, dim3(a.numInstances), dim3(256), 0, (hipStream_t)stream, a, (const uint32_t*)offsets);
  return launched();
}

// The stage's prologue, fused with the build_setup of `setupMode` before it
// when setupMode >= 0 (the frame sequence), else alone (the public entries).
static void launch_prologue(const KArgs& a, int setupMode, uint32_t stage, hipStream_t st) {
  if (setupMode >= 0)
    hipLaunchKernelGGL(k_setup_prologue, dim3(1), dim3(256), 0, st, a, setupMode, stage);
  else
    hipLaunchKernelGGL(k_prologue, dim3(1), dim3(256), 0, st, a, stage);
}

static tc_status launch_classify(const KArgs& a, tc_config config, int setupMode, hipStream_t st) {
  launch_prologue(a, setupMode, 0, st);
  uint32_t tiles = (config.maxVisibleClusters + kClassifyTile - 1) / kClassifyTile;
  uint32_t grid = std::min<uint32_t>(tiles, (uint32_t)num_cu() * tuning_snapshot().classifyBlocksPerCU);
  if (config.flags & TC_FLAG_2X_TRANSIENT) {
    hipLaunchKernelGGL((k_cluster_classify<true, true>), dim3(grid), dim3(kClassifyWaves * kWave), 0, st, a);
    // the 2X batch vertices, flat over the batches
    const uint32_t mgrid = std::min<uint32_t>((config.maxGeneratedClusters + kClassifyWaves - 1) / kClassifyWaves,
                                              (uint32_t)num_cu() * 8u);
    hipLaunchKernelGGL(k_mini_vertices, dim3(mgrid), dim3(kClassifyWaves * kWave), 0, st, a);
  } else if (config.flags & TC_FLAG_1X_TRANSIENT) {
    hipLaunchKernelGGL((k_cluster_classify<false, true>), dim3(grid), dim3(kClassifyWaves * kWave), 0, st, a);
  } else {
    hipLaunchKernelGGL((k_cluster_classify<false, false>), dim3(grid), dim3(kClassifyWaves * kWave), 0, st, a);
  }
  return launched();
}

tc_status tc_cluster_classify(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  return launch_classify(a, config, -1, (hipStream_t)stream);
}

// One split pass, or with TC_FLAG_PERSISTENT_SPLIT all `passes` passes in
// one persistent launch (renderer_raytrace_clusters_tess.cpp:510-513).
static tc_status launch_split(const KArgs& a, tc_config config, uint32_t passes, hipStream_t st, int setupMode = -1) {
  if ((config.flags & TC_FLAG_PERSISTENT_SPLIT) && passes > kMaxSplitPasses) return TC_ERROR_UNSUPPORTED;
  launch_prologue(a, setupMode, 1, st);
  uint32_t tiles = (config.maxSplitTriangles + kWave - 1) / kWave;
  uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, (uint32_t)num_cu() * tuning_snapshot().splitBlocksPerCU));
  if (config.flags & TC_FLAG_PERSISTENT_SPLIT) {
    hipLaunchKernelGGL(k_split_queue_init, dim3(1), dim3(256), 0, st, a, passes, tuning_snapshot().splitWaitLimit);
    hipLaunchKernelGGL(k_triangle_split_persistent, dim3(grid), dim3(64), 0, st, a);
  } else {
    hipLaunchKernelGGL(k_triangle_split, dim3(grid), dim3(64), 0, st, a);
  }
  return launched();
}

tc_status tc_triangle_split(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  return launch_split(a, config, tc_num_split_passes(config.maxSplitFactor), (hipStream_t)stream);
}

static tc_status launch_instantiate(const KArgs& a, tc_config config, int setupMode, hipStream_t st) {
  launch_prologue(a, setupMode, 2, st);
  uint32_t tiles = (config.maxPartTriangles + kInstTile - 1) / kInstTile;
  uint32_t grid = std::min<uint32_t>(tiles, (uint32_t)num_cu() * tuning_snapshot().instantiateBlocksPerCU);
  hipLaunchKernelGGL(k_instantiate, dim3(grid), dim3(kInstWaves * kWave), 0, st, a);
  // vertices: four waves per workgroup, each taking 64 parts at a time
  const uint32_t vgrid = std::min<uint32_t>((config.maxPartTriangles + kInstTile - 1) / kInstTile, (uint32_t)num_cu() * 8u);
  if (config.flags & TC_FLAG_PN)
    hipLaunchKernelGGL(k_instantiate_vertices<true>, dim3(vgrid), dim3(kInstWaves * kWave), 0, st, a);
  else
    hipLaunchKernelGGL(k_instantiate_vertices<false>, dim3(vgrid), dim3(kInstWaves * kWave), 0, st, a);
  return launched();
}

tc_status tc_triangle_tess_template_instantiate(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  return launch_instantiate(a, config, -1, (hipStream_t)stream);
}

tc_status tc_blas_setup_insertion(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  hipLaunchKernelGGL(k_blas_setup_insertion, dim3(1), dim3(1024), 0, (hipStream_t)stream, a);
  return launched();
}

tc_status tc_blas_clusters_insert(const tc_bindings* b, tc_config config, int transient, tc_stream stream) {
  KArgs a;
  WsLayout L;
  tc_status s = make_args(b, config, a, &L);
  if (s != TC_OK) return s;
  if (transient && !transient_builds(config.flags)) return TC_OK;  // no transient builds configured
  const int tr = transient ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  SortBuffers sb = sort_buffers(b, L);
  uint32_t* segStart = reinterpret_cast<uint32_t*>(static_cast<char*>(b->workspace) + L.sortAux);
  uint32_t grid = std::min<uint32_t>((sb.capacity + 255) / 256, (uint32_t)num_cu() * 4);
  uint32_t bits = 1;
  while ((1u << bits) < a.numInstances) bits++;
  uint32_t res = 0;
  // the sort reads the list's instance IDs itself (the field holding the
  // list's address lies in the device-side building record)
  const uint64_t* ids = tr ? &b->build->transInstanceIDs : &b->build->tempInstanceIDs;
  radix_sort_pairs(sb, tr ? &b->build->transBuildCounter : &b->build->tempInstantiateCounter, 0, bits, st, &res, ids);
  const uint32_t* starts = segStart;
  if (!tr && bits <= 8) {
    // one digit pass: an instance's first sorted position is its digit base
    starts = sb.state->bases[0];
  } else {
    hipLaunchKernelGGL(k_insert_seg_start, dim3(grid), dim3(256), 0, st, a, (const uint64_t*)sb.keys[res], segStart,
                       sb.capacity, tr);
  }
  // 1024-thread workgroups, at most one per CU: each ends with one
  // device-scope atomic on the readback's size total, and 1024 workgroups'
  // atomics on that one word took 9 of the launch's 17 us at grid64
  const uint32_t sgrid = std::min<uint32_t>((sb.capacity + 1023) / 1024, (uint32_t)num_cu());
  hipLaunchKernelGGL(k_insert_scatter, dim3(sgrid), dim3(1024), 0, st, a, (const uint64_t*)sb.keys[res],
                     (const uint32_t*)sb.vals[res], starts, sb.capacity, tr);
  return launched();
}

// renderer_raytrace_clusters_tess.cpp:467-582 on a prepared argument block;
// each build_setup that precedes a stage runs in that stage's prologue
// launch
static tc_status frame_tessellate(const KArgs& a, tc_config config, uint32_t numSplitPasses, hipStream_t st) {
  tc_status s;
  if (numSplitPasses == 0) numSplitPasses = tc_num_split_passes(config.maxSplitFactor);
  if ((s = launch_classify(a, config, TC_BUILD_SETUP_CLASSIFY, st)) != TC_OK) return s;
  if (config.flags & TC_FLAG_PERSISTENT_SPLIT) {
    // every pass, with its SPLIT_PASS setup, in one launch
    if ((s = launch_split(a, config, numSplitPasses, st, TC_BUILD_SETUP_SPLIT)) != TC_OK) return s;
  } else {
    for (uint32_t p = 0; p < numSplitPasses; p++)
      if ((s = launch_split(a, config, numSplitPasses, st, p == 0 ? TC_BUILD_SETUP_SPLIT : TC_BUILD_SETUP_SPLIT_PASS)) != TC_OK)
        return s;
  }
  if ((s = launch_instantiate(a, config, TC_BUILD_SETUP_INSTANTIATE_TESS, st)) != TC_OK) return s;
  hipLaunchKernelGGL(k_build_setup, dim3(1), dim3(1), 0, st, a, (int)TC_BUILD_SETUP_BUILD_BLAS);
  return launched();
}

tc_status tc_frame_tessellate(const tc_bindings* b, tc_config config, uint32_t numSplitPasses, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a);
  if (s != TC_OK) return s;
  return frame_tessellate(a, config, numSplitPasses, (hipStream_t)stream);
}

// ---------------------------------------------- allocation-free variant
//
// tc_frame_tessellate_noalloc runs the same chain with its look-back state
// (the workspace header with the tile ticket, epochs, counter snapshots and
// the persistent split's pass table, then the tiles' status words) in this
// file-scope array instead of the caller's workspace.  Each device holds
// its own copy from module load to process end.  It starts zeroed, as
// tc_workspace_init leaves a workspace; after that nothing needs clearing
// between calls: every stage's prologue resets the ticket and opens a new
// epoch, and status words of earlier epochs read as unpublished.  Calls
// that overlap on one device race on it.
constexpr uint32_t kScratchTiles = 16384;
constexpr uint64_t kScratchBytes = kWsHeaderBytes + (uint64_t)kMaxCounters * kScratchTiles * 8;
static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratchMutex;
static void* g_scratchPtr[kMaxScratchDevices];

// Device address of `scratch` on the current device, looked up once per
// device.  Host code uses only this address, never the symbol itself.
static tc_status scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TC_ERROR_NO_DEVICE;
  if (dev < 0 || dev >= kMaxScratchDevices) return TC_ERROR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TC_ERROR_LAUNCH;
    g_scratchPtr[dev] = p;
  }
  *out = g_scratchPtr[dev];
  return TC_OK;
}

tc_status tc_frame_tessellate_noalloc(const tc_bindings* b, tc_config config, uint32_t numSplitPasses, tc_stream stream) {
  KArgs a;
  tc_status s = make_args(b, config, a, nullptr, false);
  if (s != TC_OK) return s;
  // every stage's tile count must fit the status words of the array
  if ((config.maxVisibleClusters + kClassifyTile - 1) / kClassifyTile > kScratchTiles ||
      (config.maxSplitTriangles + kWave - 1) / kWave > kScratchTiles ||
      (config.maxPartTriangles + kInstTile - 1) / kInstTile > kScratchTiles)
    return TC_ERROR_CAPACITY;
  void* base = nullptr;
  if ((s = scratch_address(&base)) != TC_OK) return s;
  a.ws = reinterpret_cast<WsHeader*>(base);
  a.status = reinterpret_cast<uint64_t*>(static_cast<char*>(base) + kWsHeaderBytes);
  a.maxTiles = kScratchTiles;
  return frame_tessellate(a, config, numSplitPasses, (hipStream_t)stream);
}

}  // extern "C"
