// On-disk mesh format (see mesh_io.hpp).
#include "mesh_io.hpp"

#include <cstdint>
#include <cstring>
#include <fstream>

namespace femhip {

namespace {

constexpr char kMagic[8] = {'F', 'E', 'M', 'H', 'M', 'S', 'H', '1'};

struct Fnv1a {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
};

long long ipow_ll(int b, int e) {
  long long r = 1;
  while (e-- > 0) r *= b;
  return r;
}

}  // namespace

void mesh_validate(const MeshData& m) {
  require(m.dim >= 1 && m.dim <= 3, "mesh: dim must be 1, 2 or 3");
  require(m.num_elem > 0 && m.P >= 2 && m.num_nodes > 0 && m.num_vertices > 0, "mesh: sizes must be positive");
  const long long es = ipow_ll(m.P, m.dim), vs = ipow_ll(2, m.dim);
  require(static_cast<long long>(m.offsets.size()) == m.num_elem * es, "mesh: offsets size != num_elem * P^dim");
  require(static_cast<long long>(m.vertex_offsets.size()) == m.num_elem * vs,
          "mesh: vertex_offsets size != num_elem * 2^dim");
  require(static_cast<long long>(m.coords.size()) == static_cast<long long>(m.dim) * m.num_vertices,
          "mesh: coords size != dim * num_vertices");
  for (int o : m.offsets) require(o >= 0 && o < m.num_nodes, "mesh: node offset out of range");
  for (int o : m.vertex_offsets) require(o >= 0 && o < m.num_vertices, "mesh: vertex offset out of range");
}

void mesh_write(const std::string& path, const MeshData& m) {
  mesh_validate(m);
  std::ofstream out(path, std::ios::binary);
  require(static_cast<bool>(out), "mesh_write: cannot open " + path);
  Fnv1a sum;
  auto put = [&](const void* p, size_t n) {
    out.write(static_cast<const char*>(p), static_cast<std::streamsize>(n));
    sum.add(p, n);
  };
  const int32_t hdr[5] = {m.dim, m.num_elem, m.P, m.num_nodes, m.num_vertices};
  put(kMagic, sizeof(kMagic));
  put(hdr, sizeof(hdr));
  put(m.offsets.data(), m.offsets.size() * sizeof(int32_t));
  put(m.vertex_offsets.data(), m.vertex_offsets.size() * sizeof(int32_t));
  put(m.coords.data(), m.coords.size() * sizeof(double));
  const uint64_t h = sum.h;
  out.write(reinterpret_cast<const char*>(&h), sizeof(h));
  require(static_cast<bool>(out), "mesh_write: write failed: " + path);
}

MeshData mesh_read(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  require(static_cast<bool>(in), "mesh_read: cannot open " + path);
  Fnv1a sum;
  auto get = [&](void* p, size_t n) {
    in.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
    require(static_cast<size_t>(in.gcount()) == n, "mesh_read: file is truncated: " + path);
    sum.add(p, n);
  };
  char magic[8];
  get(magic, sizeof(magic));
  require(std::memcmp(magic, kMagic, sizeof(kMagic)) == 0, "mesh_read: not a FEMHMSH1 file: " + path);
  int32_t hdr[5];
  get(hdr, sizeof(hdr));
  MeshData m;
  m.dim = hdr[0];
  m.num_elem = hdr[1];
  m.P = hdr[2];
  m.num_nodes = hdr[3];
  m.num_vertices = hdr[4];
  require(m.dim >= 1 && m.dim <= 3 && m.num_elem > 0 && m.P >= 2 && m.num_nodes > 0 && m.num_vertices > 0,
          "mesh_read: bad header in " + path);
  const long long es = ipow_ll(m.P, m.dim), vs = ipow_ll(2, m.dim);
  require(m.num_elem * es < (1ll << 31) && static_cast<long long>(m.dim) * m.num_vertices < (1ll << 31),
          "mesh_read: sizes exceed int32 indexing");
  m.offsets.resize(static_cast<size_t>(m.num_elem * es));
  m.vertex_offsets.resize(static_cast<size_t>(m.num_elem * vs));
  m.coords.resize(static_cast<size_t>(m.dim) * m.num_vertices);
  get(m.offsets.data(), m.offsets.size() * sizeof(int32_t));
  get(m.vertex_offsets.data(), m.vertex_offsets.size() * sizeof(int32_t));
  get(m.coords.data(), m.coords.size() * sizeof(double));
  const uint64_t want = sum.h;
  uint64_t h = 0;
  in.read(reinterpret_cast<char*>(&h), sizeof(h));
  require(in.gcount() == sizeof(h) && h == want, "mesh_read: checksum mismatch in " + path);
  mesh_validate(m);
  return m;
}

}  // namespace femhip
