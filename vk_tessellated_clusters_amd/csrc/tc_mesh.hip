// tc_mesh.hip -- host-side mesh clustering (no device code).
//
// Stands for the reference's geometry processing (scene.cpp:365-552:
// meshopt_buildMeshletsSpatial, the cluster bounding boxes and the
// per-cluster vertex duplication) with a clusterizer of our own: recursive
// bisection of the triangle centroids.  A range of n triangles that is too
// large (more than maxTriangles triangles or more than maxVertices distinct
// vertices) is cut along the longest axis of its centroid bounds into
// k = ceil(n / maxTriangles) parts' worth, floor(k / 2) parts to the left,
// with the cut at the proportional rank (ties broken by triangle index).
// Leaves become clusters in depth-first order, so neighbouring clusters are
// spatially close.  Within a cluster triangles keep their mesh order and
// vertices are numbered in first use.  Everything is deterministic.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "tc_abi.h"

namespace {

struct Clusterizer {
  const float* pos;
  const uint32_t* idx;
  uint32_t maxT, maxV;
  std::vector<float> centroid;  // 3 per triangle
  std::vector<uint32_t> tris;   // triangle order being partitioned
  std::vector<uint32_t> stamp;  // per vertex: last leaf test that saw it
  uint32_t stampValue = 0;
  std::vector<std::pair<uint32_t, uint32_t>> leaves;  // [begin, end) of tris

  uint32_t distinct_vertices(uint32_t b, uint32_t e) {
    ++stampValue;
    uint32_t n = 0;
    for (uint32_t i = b; i < e; i++)
      for (int k = 0; k < 3; k++) {
        uint32_t v = idx[3 * tris[i] + k];
        if (stamp[v] != stampValue) stamp[v] = stampValue, n++;
      }
    return n;
  }

  void split(uint32_t b, uint32_t e) {
    const uint32_t n = e - b;
    if (n <= maxT && distinct_vertices(b, e) <= maxV) {
      std::sort(tris.begin() + b, tris.begin() + e);
      leaves.emplace_back(b, e);
      return;
    }
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = b; i < e; i++)
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], centroid[3 * tris[i] + k]);
        hi[k] = std::max(hi[k], centroid[3 * tris[i] + k]);
      }
    int axis = 0;
    for (int k = 1; k < 3; k++)
      if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
    const uint32_t parts = std::max(2u, (n + maxT - 1) / maxT);
    const uint32_t m = b + (uint32_t)((uint64_t)n * (parts / 2) / parts);
    std::nth_element(tris.begin() + b, tris.begin() + m, tris.begin() + e, [&](uint32_t x, uint32_t y) {
      const float cx = centroid[3 * x + axis], cy = centroid[3 * y + axis];
      return cx < cy || (cx == cy && x < y);
    });
    split(b, m);
    split(m, e);
  }
};

}  // namespace

extern "C" tc_status tc_clusterize_mesh(const float* positions, uint32_t numVertices, const uint32_t* indices,
                                        uint32_t numTriangles, uint32_t maxTriangles, uint32_t maxVertices,
                                        tc_cluster* clusters, uint32_t* clusterLocalVertices,
                                        uint8_t* clusterLocalTriangles, tc_bbox* clusterBboxes, uint32_t* numClusters,
                                        uint32_t* numClusterVertices) {
  if (!positions || !indices || !clusters || !clusterLocalVertices || !clusterLocalTriangles || !clusterBboxes ||
      !numClusters || !numClusterVertices)
    return TC_ERROR_INVALID_ARGUMENT;
  // clusters hold at most 128 triangles (the classify tile) and 8-bit
  // local vertex indices
  if (maxTriangles < 1 || maxTriangles > 128 || maxVertices < 3 || maxVertices > 256) return TC_ERROR_INVALID_ARGUMENT;
  for (uint64_t i = 0; i < 3ull * numTriangles; i++)
    if (indices[i] >= numVertices) return TC_ERROR_INVALID_ARGUMENT;

  Clusterizer cz;
  cz.pos = positions;
  cz.idx = indices;
  cz.maxT = maxTriangles;
  cz.maxV = maxVertices;
  cz.centroid.resize(3ull * numTriangles);
  for (uint32_t t = 0; t < numTriangles; t++)
    for (int k = 0; k < 3; k++)
      cz.centroid[3ull * t + k] = (positions[3ull * indices[3ull * t] + k] + positions[3ull * indices[3ull * t + 1] + k] +
                                   positions[3ull * indices[3ull * t + 2] + k]) /
                                  3.0f;
  cz.tris.resize(numTriangles);
  std::iota(cz.tris.begin(), cz.tris.end(), 0u);
  cz.stamp.assign(numVertices, 0u);
  if (numTriangles) cz.split(0, numTriangles);

  // scene.cpp:519-552 per-cluster vertices; :463-517 bounds and edge lengths
  std::vector<uint32_t> local(numVertices, 0xFFFFFFFFu);
  uint32_t vOut = 0, tOut = 0, c = 0;
  for (const auto& leaf : cz.leaves) {
    tc_cluster& cl = clusters[c];
    cl.firstTriangle = tOut;
    cl.firstLocalTriangle = 3 * tOut;
    cl.firstLocalVertex = vOut;
    uint32_t nv = 0;
    for (uint32_t i = leaf.first; i < leaf.second; i++, tOut++) {
      const uint32_t t = cz.tris[i];
      for (int k = 0; k < 3; k++) {
        const uint32_t v = indices[3ull * t + k];
        if (local[v] == 0xFFFFFFFFu) {
          local[v] = nv++;
          clusterLocalVertices[vOut + local[v]] = v;
        }
        clusterLocalTriangles[3ull * tOut + k] = (uint8_t)local[v];
      }
    }
    cl.numVertices = (uint16_t)nv;
    cl.numTriangles = (uint16_t)(leaf.second - leaf.first);
    tc_bbox bb = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}, FLT_MAX, -FLT_MAX};
    for (uint32_t v = 0; v < nv; v++) {
      const float* p = positions + 3ull * clusterLocalVertices[vOut + v];
      for (int k = 0; k < 3; k++) bb.lo[k] = std::min(bb.lo[k], p[k]), bb.hi[k] = std::max(bb.hi[k], p[k]);
    }
    for (uint32_t t = cl.firstTriangle; t < tOut; t++) {
      const float* q[3];
      for (int k = 0; k < 3; k++) q[k] = positions + 3ull * clusterLocalVertices[vOut + clusterLocalTriangles[3ull * t + k]];
      for (int e = 0; e < 3; e++) {
        const float* a = q[e];
        const float* b = q[(e + 1) % 3];
        const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
        bb.shortestEdge = std::min(bb.shortestEdge, d);
        bb.longestEdge = std::max(bb.longestEdge, d);
      }
    }
    clusterBboxes[c] = bb;
    for (uint32_t v = 0; v < nv; v++) local[clusterLocalVertices[vOut + v]] = 0xFFFFFFFFu;
    vOut += nv;
    c++;
  }
  *numClusters = c;
  *numClusterVertices = vOut;
  return TC_OK;
}
