// On-disk hex/quad mesh with its node restriction (any numbering), read and
// written by the library.
//
// File layout (little-endian):
//   char[8]  "FEMHMSH1"
//   int32    dim (1..3), num_elem, P (nodes per direction of the node
//            restriction), num_nodes, num_vertices
//   int32    offsets[num_elem][P^dim]       node restriction (L index of local node
//                                           i = (iz*P + iy)*P + ix, x fastest)
//   int32    vertex_offsets[num_elem][2^dim] vertices of each element, same order
//   float64  coords[dim][num_vertices]      vertex coordinates, component-blocked
//   uint64   checksum: FNV-1a of every byte before it
// Reading validates the header, sizes, index ranges and the checksum.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace femhip {

struct MeshData {
  int dim = 0, num_elem = 0, P = 0, num_nodes = 0, num_vertices = 0;
  std::vector<int> offsets, vertex_offsets;
  std::vector<double> coords;
};

void mesh_write(const std::string& path, const MeshData& m);
MeshData mesh_read(const std::string& path);
// Throws std::invalid_argument unless sizes and index ranges are consistent.
void mesh_validate(const MeshData& m);

}  // namespace femhip
