// Dumps the reference's tessellation table exactly as its own header
// defines it (compiled against src/tessellation_table_nv_raw.hpp of the
// upstream tree), so tests can compare it with the data the build uses.
// Output: text lines "max_edge_segments N", then "V <hex>", "T <hex>",
// "C <dec>" for every vertex, triangle and config word.
#include <cstdio>

#include "tessellation_table_nv_raw.hpp"

int main() {
  using namespace tessellation_table;
  std::printf("max_edge_segments %u\n", max_edge_segments);
  for (unsigned i = 0; i < max_vertices; i++) std::printf("V %08x\n", vertices[i]);
  for (unsigned i = 0; i < max_triangles; i++) std::printf("T %08x\n", triangles[i]);
  for (unsigned i = 0; i < max_configs * 4; i++) std::printf("C %u\n", configs[i]);
  return 0;
}
