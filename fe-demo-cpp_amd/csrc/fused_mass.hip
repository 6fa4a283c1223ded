// Host side of the fused mass operator: registry of compiled shapes, qdata
// re-layout, block plans, algorithm choice and autotuning, dispatch, and the
// Jacobi diagonal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "block_plan.hpp"
#include "fused_box_kernel.hpp"
#include "fused_mass.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

// Reference backend-strided qdata[q*E + e] (q = qz*Q^2 + line, line = qy*Q
// + qx) -> private per-element tiles laid out by qdata_index: the fused
// kernels' z-line thread t reads its Q values in pairs, and the lanes of one
// load instruction read consecutive pairs.
template <class T>
__global__ __launch_bounds__(kThreads) void qdata_private_kernel(int E, int Q, const double* __restrict__ ref,
                                                                 T* __restrict__ priv) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  const size_t Q2 = static_cast<size_t>(Q) * Q, Q3 = Q2 * Q;
  if (idx >= Q3 * E) return;
  // idx walks the private layout (coalesced writes); reads stride E.
  const size_t e = idx / Q3, r = idx % Q3;
  size_t line, qz;  // the inverse of qdata_index
  if (Q % 2 == 0) {
    const size_t pr = r / (2 * Q2), rem = r % (2 * Q2);
    line = rem / 2;
    qz = 2 * pr + rem % 2;
  } else {
    line = r % Q2;
    qz = r / Q2;
  }
  priv[idx] = static_cast<T>(ref[(qz * Q2 + line) * E + e]);
}

template <class T>
__global__ __launch_bounds__(kThreads) void to_f32_kernel(size_t n, const double* __restrict__ a, T* __restrict__ b) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) b[i] = static_cast<T>(a[i]);
}

// Element diagonal of the mass matrix, one thread per (element, node):
// d[e][i] = sum_q qdata[e][q] * B2[qz][iz] * B2[qy][iy] * B2[qx][ix], written to
// Yd[t] (t = i + e*P^3); the transpose-map pass then sums each node's entries in
// ascending t, so the diagonal is the same bit for bit on every run.
__global__ __launch_bounds__(kThreads) void mass_diag_kernel(int E, int P, int Q, const double* __restrict__ qpriv,
                                                             const double* __restrict__ B2,
                                                             double* __restrict__ Yd) {
  const int P3 = P * P * P, Q2 = Q * Q;
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<size_t>(E) * P3) return;
  const size_t e = idx / P3;
  const int i = static_cast<int>(idx % P3);
  const int ix = i % P, iy = (i / P) % P, iz = i / (P * P);
  const double* qe = qpriv + e * Q2 * Q;  // laid out by qdata_index
  double sum = 0.0;
  for (int qy = 0; qy < Q; ++qy) {
    for (int qx = 0; qx < Q; ++qx) {
      const double bxy = B2[qy * P + iy] * B2[qx * P + ix];
      double s = 0.0;
      for (int qz = 0; qz < Q; ++qz) s += B2[qz * P + iz] * qe[qdata_index(Q, qy * Q + qx, qz)];
      sum += bxy * s;
    }
  }
  Yd[idx] = sum;
}

// Block-slot copy of the private qdata: dst[slot][q] = src[elem[slot]][q] (0 for padding slots).
template <class T>
__global__ __launch_bounds__(kThreads) void qdata_block_kernel(size_t n, int Q3, const int* __restrict__ elem,
                                                               const double* __restrict__ src, T* __restrict__ dst) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= n) return;
  const size_t slot = idx / Q3, q = idx % Q3;
  const int e = elem[slot];
  dst[idx] = e >= 0 ? static_cast<T>(src[static_cast<size_t>(e) * Q3 + q]) : T(0);
}

// Second passes of the deterministic scatters: for each row m (an L-node),
// v[node + c*cs] = sum over the row's entries k of vals[c*cl + idx[k]], added
// from 0 in ascending k. The transpose-map pass (entries t = i + e*P^3 in
// ascending t) reproduces the CPU scatter's order; the partial-sum pass adds a
// node's block partials in ascending block order. rows == nullptr means row m
// is node m. Each thread takes kRows rows
// (kThreads apart, so lanes stay coalesced) and issues their loads together.
constexpr int kRows = 4, kChunk = 4;
template <class T>
__global__ __launch_bounds__(kThreads) void row_sum_kernel(int nrows, int NC, int cs, const int* __restrict__ rows,
                                                           const int* __restrict__ ptr, const int* __restrict__ idx,
                                                           const T* __restrict__ vals, size_t cl,
                                                           T* __restrict__ v) {
  const int base = blockIdx.x * (kThreads * kRows) + threadIdx.x;
  int rb[kRows], re[kRows], node[kRows];
  int longest = 0;
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    const int m = base + j * kThreads;
    const bool ok = m < nrows;
    rb[j] = ok ? ptr[m] : 0;
    re[j] = ok ? ptr[m + 1] : 0;
    node[j] = ok ? (rows ? rows[m] : m) : -1;
    longest = max(longest, re[j] - rb[j]);
  }
  for (int c = 0; c < NC; ++c) {
    const T* vc = vals + static_cast<size_t>(c) * cl;
    T acc[kRows];
#pragma unroll
    for (int j = 0; j < kRows; ++j) acc[j] = T(0);
    for (int off = 0; off < longest; off += kChunk) {
      int id[kRows][kChunk];
      T val[kRows][kChunk];
#pragma unroll
      for (int j = 0; j < kRows; ++j)
#pragma unroll
        for (int q = 0; q < kChunk; ++q) {
          const int k = rb[j] + off + q;
          id[j][q] = k < re[j] ? idx[k] : -1;
        }
#pragma unroll
      for (int j = 0; j < kRows; ++j)
#pragma unroll
        for (int q = 0; q < kChunk; ++q) val[j][q] = id[j][q] >= 0 ? vc[id[j][q]] : T(0);
#pragma unroll
      for (int j = 0; j < kRows; ++j)
#pragma unroll
        for (int q = 0; q < kChunk; ++q)
          if (id[j][q] >= 0) acc[j] += val[j][q];
    }
#pragma unroll
    for (int j = 0; j < kRows; ++j)
      if (node[j] >= 0) v[static_cast<size_t>(node[j]) + static_cast<size_t>(c) * cs] = acc[j];
  }
}

template <class T>
void launch_row_sum(int nrows, int NC, int cs, const int* rows, const int* ptr, const int* idx, const T* vals,
                    size_t cl, T* v, hipStream_t s) {
  if (nrows <= 0) return;
  const size_t per = static_cast<size_t>(kThreads) * kRows;
  row_sum_kernel<T><<<static_cast<unsigned>((nrows + per - 1) / per), kThreads, 0, s>>>(nrows, NC, cs, rows, ptr,
                                                                                       idx, vals, cl, v);
  HIP_CHECK_LAUNCH();
}

std::map<FusedKey, int>& budget_registry() {
  static std::map<FusedKey, int> reg;
  return reg;
}

std::map<FusedKey, FusedOccupancy>& occupancy_registry() {
  static std::map<FusedKey, FusedOccupancy> reg;
  return reg;
}

FusedLauncher find(const FusedKey& k) {
  auto& reg = fused_registry();
  auto it = reg.find(k);
  return it == reg.end() ? nullptr : it->second;
}

}  // namespace

std::map<FusedKey, FusedLauncher>& fused_registry() {
  static std::map<FusedKey, FusedLauncher> reg;
  return reg;
}

void register_fused(const FusedKey& k, FusedLauncher f, int lds_node_budget) {
  fused_registry()[k] = f;
  budget_registry()[k] = lds_node_budget;
}

void register_fused_occupancy(const FusedKey& k, FusedOccupancy f) { occupancy_registry()[k] = f; }

int fused_wgs_per_cu(const FusedKey& k) {
  static std::mutex m;
  static std::map<FusedKey, int> cache;
  std::lock_guard<std::mutex> lk(m);
  auto c = cache.find(k);
  if (c != cache.end()) return c->second;
  auto it = occupancy_registry().find(k);
  const int n = it == occupancy_registry().end() ? 0 : std::max(0, it->second());
  cache[k] = n;
  return n;
}

int fused_node_budget(const FusedKey& k) {
  auto it = budget_registry().find(k);
  return it == budget_registry().end() ? 0 : it->second;
}

std::vector<int> fused_elem_variants(int P, int Q, int NC, int fp32, int alg, int exact) {
  std::vector<int> out;
  for (const auto& kv : fused_registry()) {
    const FusedKey& k = kv.first;
    if (k.P == P && k.Q == Q && k.NC == NC && k.fp32 == fp32 && k.alg == alg && k.exact == exact)
      out.push_back(k.elems);
  }
  return out;
}

bool fused_mass_supported(int P, int Q, int NC) {
  return !fused_elem_variants(P, Q, NC, 0, kAlgElemAtomic, 0).empty();
}

std::map<TuneKey, TuneRows>& tune_table() {
  static std::map<TuneKey, TuneRows> t;
  return t;
}

const TuneEntry* tune_lookup(const TuneKey& k, long long num_elem, double max_ratio) {
  auto it = tune_table().find(k);
  if (it == tune_table().end() || it->second.empty()) return nullptr;
  const TuneRows& rows = it->second;
  const double n = static_cast<double>(std::max(num_elem, 1LL));
  auto dist = [n](const TuneEntry& e) { return std::fabs(std::log(static_cast<double>(std::max(e.num_elem, 1LL)) / n)); };
  auto hi = rows.lower_bound(num_elem);  // first row measured on at least num_elem elements
  const TuneEntry* best = hi != rows.end() ? &hi->second : nullptr;
  if (hi != rows.begin()) {
    const TuneEntry& lo = std::prev(hi)->second;
    if (!best || dist(lo) < dist(*best)) best = &lo;
  }
  if (max_ratio > 0.0 && dist(*best) > std::log(max_ratio)) return nullptr;
  return best;
}

void tune_record(const TuneKey& k, const TuneEntry& e) { tune_table()[k][e.num_elem] = e; }

std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}

std::map<TuneAlgKey, TuneEntry>& tune_alg_table() {
  static std::map<TuneAlgKey, TuneEntry> t;
  return t;
}

void tune_table_save(const std::string& path) {
  std::ofstream out(path);
  require(static_cast<bool>(out), "tune_table_save: cannot write " + path);
  out.precision(6);
  std::lock_guard<std::mutex> lk(tune_mutex());
  for (const auto& kv : tune_table()) {
    const TuneKey& k = kv.first;
    for (const auto& row : kv.second) {
      const TuneEntry& e = row.second;
      out << "best " << k.P << ' ' << k.Q << ' ' << k.NC << ' ' << k.fp32 << ' ' << k.det << ' ' << k.box << ' '
          << e.alg << ' ' << e.elems << ' ' << e.ms << ' ' << e.num_elem << '\n';
    }
  }
  for (const auto& kv : tune_alg_table()) {
    const TuneAlgKey& k = kv.first;
    const TuneEntry& e = kv.second;
    out << "alg " << k.P << ' ' << k.Q << ' ' << k.NC << ' ' << k.fp32 << ' ' << k.alg << ' ' << e.elems << ' '
        << e.ms << ' ' << e.num_elem << '\n';
  }
  require(static_cast<bool>(out), "tune_table_save: write failed: " + path);
}

int tune_table_load(const std::string& path) {
  std::ifstream in(path);
  require(static_cast<bool>(in), "tune_table_load: cannot read " + path);
  std::string line;
  int taken = 0, lineno = 0;
  std::lock_guard<std::mutex> lk(tune_mutex());
  while (std::getline(in, line)) {
    ++lineno;
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string kind;
    ls >> kind;
    TuneEntry e;
    if (kind == "best") {
      std::vector<std::string> tok;
      for (std::string w; ls >> w;) tok.push_back(w);
      const bool has_box = tok.size() == 10;
      require(has_box || tok.size() == 9, "tune_table_load: bad row at line " + std::to_string(lineno));
      TuneKey k;
      try {
        size_t i = 0;
        auto next = [&]() { return std::stoi(tok[i++]); };
        k.P = next();
        k.Q = next();
        k.NC = next();
        k.fp32 = next();
        k.det = next();
        k.box = has_box ? next() : 0;
        e.alg = next();
        e.elems = next();
        e.ms = std::stod(tok[i++]);
        e.num_elem = std::stoll(tok[i++]);
      } catch (const std::exception&) {
        throw std::invalid_argument("tune_table_load: bad row at line " + std::to_string(lineno));
      }
      if (!find(FusedKey{k.P, k.Q, k.NC, k.fp32, e.alg, 0, e.elems})) continue;
      if (k.det && !alg_deterministic(e.alg)) continue;
      if (!k.box && alg_box(e.alg)) continue;
      tune_record(k, e);
      ++taken;
    } else if (kind == "alg") {
      TuneAlgKey k;
      ls >> k.P >> k.Q >> k.NC >> k.fp32 >> k.alg >> e.elems >> e.ms >> e.num_elem;
      require(!ls.fail(), "tune_table_load: bad row at line " + std::to_string(lineno));
      e.alg = k.alg;
      if (!find(FusedKey{k.P, k.Q, k.NC, k.fp32, k.alg, 0, e.elems})) continue;
      tune_alg_table()[k] = e;
      ++taken;
    } else {
      throw std::invalid_argument("tune_table_load: unknown row kind at line " + std::to_string(lineno));
    }
  }
  return taken;
}

namespace {

// Default packing of an algorithm: the variant keeping the most elements
// resident per CU when the CU holds 32 wavefronts (whole workgroups only);
// ties go to more elements per workgroup. LDS and registers are not modelled:
// the autotuner measures.
int default_elems(int P, int Q, int NC, int fp32, int alg, int exact) {
  auto v = fused_elem_variants(P, Q, NC, fp32, alg, exact);
  if (v.empty()) return 0;
  const int tpe = std::max({P * P, P * Q, Q * Q});
  int best = 0, best_res = -1;
  for (int ne : v) {
    const int waves = (ne * tpe + 63) / 64;
    const int resident = (32 / waves) * ne;
    if (resident > best_res || (resident == best_res && ne > best)) {
      best_res = resident;
      best = ne;
    }
  }
  return best;
}

bool valid_for(const FusedMass& f, int alg) {
  return (f.scatter != kScatterDeterministic || alg_deterministic(alg)) && (!alg_box(alg) || f.box);
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return cus;
}

// Segments per column of the key's box-column kernel on f's box: the
// operator's layer limit if set, else box_segments' choice for the kernel's
// resident workgroups (the occupancy query times the CUs; the component-split
// kernels run one workgroup per component of a column segment). Measured on
// MI355X, same box (apply ms, automatic -> this rule): 64^3, P=5 FP32 8x4
// columns 0.137 -> 0.115 (two resident 16-wave workgroups per CU, not one),
// 8x2 0.127 -> 0.118; 48^3, P=8, 3 components FP64 8x2 1.065 -> 1.049 (12
// layers, not 16); 64^3, P=5 FP64 unchanged (one 16-wave workgroup per CU).
int box_nsz_of(const FusedMass& f, const FusedKey& k) {
  int BX = 0, BY = 0;
  box_columns(k.elems, &BX, &BY);
  const long long ncols = static_cast<long long>(f.box_n[0] / BX) * (f.box_n[1] / BY);
  if (f.box_lz > 0) return box_segments(f.box_n[2], ncols, 1, f.P, f.box_lz);
  const int kind = k.elems / 100;  // 3, 5, 7: the component-split kernels
  const long long split = (kind == 3 || kind == 5 || kind == 7) ? f.NC : 1;
  const int wgs = fused_wgs_per_cu(k);
  const long long slots = wgs > 0 ? static_cast<long long>(device_cus()) * wgs
                                  : std::max<long long>(1, 32LL * device_cus() / (BX * BY));
  return box_segments(f.box_n[2], ncols * split, slots, f.P, 0);
}

// Whether the box-column variant with `elems` elements per layer tiles f's box
// (without asking the GPU: segments of 1 + 2/(P-1) layers at least exist).
bool box_fits(const FusedMass& f, int elems) {
  int BX = 0, BY = 0;
  if (!f.box || !f.basis_symmetric || !box_columns(elems, &BX, &BY)) return false;
  if (f.box_n[0] % BX || f.box_n[1] % BY || BX * (f.P - 1) < 2 || BY * (f.P - 1) < 2) return false;
  return box_segments(f.box_n[2], 1, 1, f.P, f.box_lz) > 0;
}

bool key_valid(const FusedMass& f, const FusedKey& k) {
  return valid_for(f, k.alg) && (!alg_box(k.alg) || box_fits(f, k.elems));
}

// The first compiled box-column variant that tiles the box (most elements per layer first).
int box_default_elems(const FusedMass& f, int fp32) {
  // Most elements per layer first; of one cross-section the layer-ahead qdata
  // kernel (elems < 100) first.
  std::vector<int> v = fused_elem_variants(f.P, f.Q, f.NC, fp32, kAlgBoxColumn, 0);
  std::sort(v.begin(), v.end(), [](int a, int b) { return a % 100 != b % 100 ? a % 100 > b % 100 : a < b; });
  for (int ne : v)
    if (box_fits(f, ne)) return ne;
  return 0;
}

DevPlan& ensure_plan(FusedMass& f, int NB, int budget, bool part, hipStream_t s) {
  auto& slot = f.plans[{NB, budget}];
  if (!slot) {
    ++f.gen;
    const int P3 = f.P * f.P * f.P, Q3 = f.Q * f.Q * f.Q;
    BlockPlan plan = build_block_plan(f.rstr->host_offsets().data(), f.num_elem, P3, f.l_size, NB,
                                      budget > 0 ? std::min(budget, NB * P3) : 0);
    auto d = std::make_unique<DevPlan>();
    d->NB = NB;
    d->budget = budget;
    d->nblk = plan.nblk;
    d->max_nodes = plan.max_nodes;
    d->node_ptr.upload(plan.node_ptr);
    d->nodes.upload(plan.nodes);
    d->ninterior.upload(plan.ninterior);
    d->nelem.upload(plan.nelem);
    d->loc.upload(plan.loc);
    d->color.upload(plan.color);
    d->ncolor.upload(plan.ncolor);
    d->part_ptr.upload(plan.part_ptr);
    d->npart = static_cast<size_t>(plan.part_ptr.back());
    DevBuf<int> elem;
    elem.upload(plan.elem);
    const size_t n = static_cast<size_t>(plan.nblk) * NB * Q3;
    d->q64.allocate(n);
    qdata_block_kernel<double><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, Q3, elem.data(), f.qdata64.data(),
                                                                         d->q64.data());
    HIP_CHECK_LAUNCH();
    HIP_CHECK(hipStreamSynchronize(s));  // `elem` is freed on return
    if (part) {
      PartialRows r = build_partial_rows(plan, f.l_size, f.NC, f.comp_stride);
      d->nrows = static_cast<int>(r.node.size());
      d->rows.upload(r.node);
      d->row_ptr.upload(r.ptr);
      d->row_part.upload(r.part);
      d->covers_all = r.covers_all;
    }
    slot = std::move(d);
  }
  DevPlan& d = *slot;
  if (part && d.row_ptr.empty()) {
    // The plan was made for the atomic kernel with the same budget: add the rows.
    BlockPlan plan = build_block_plan(f.rstr->host_offsets().data(), f.num_elem, f.P * f.P * f.P, f.l_size, NB,
                                      budget > 0 ? std::min(budget, NB * f.P * f.P * f.P) : 0);
    PartialRows r = build_partial_rows(plan, f.l_size, f.NC, f.comp_stride);
    d.nrows = static_cast<int>(r.node.size());
    d.rows.upload(r.node);
    d.row_ptr.upload(r.ptr);
    d.row_part.upload(r.part);
    d.covers_all = r.covers_all;
  }
  return d;
}

// Transpose map of the referenced nodes (host build, once): the element kernel
// writes every element output to Y, and this pass sums each node's outputs.
// Whether the scatter writes every L entry: every component of every
// referenced node, and those cover [0, l_size).
bool tmap_covers_all(const FusedMass& f, const std::vector<char>& referenced) {
  std::vector<char> hit(f.l_size, 0);
  for (int l = 0; l < f.l_size; ++l) {
    if (!referenced[l]) continue;
    for (int c = 0; c < f.NC; ++c) {
      const long long g = static_cast<long long>(l) + static_cast<long long>(c) * f.comp_stride;
      if (g >= f.l_size) return false;
      hit[g] = 1;
    }
  }
  for (int l = 0; l < f.l_size; ++l)
    if (!hit[l]) return false;
  return true;
}

void ensure_tmap(FusedMass& f) {
  if (f.tm_built) return;
  const Restriction& r = *f.rstr;
  if (r.type == 0 && r.scatter_mode == kScatterDeterministic && !r.t_offsets.empty()) {
    // The restriction's map is this map (rows: used nodes ascending; entries: t ascending).
    f.tm_nrows = r.num_nodes;
    // Rows are the distinct used offsets ascending: num_nodes of them in
    // [0, max_offset] are exactly 0..max_offset, and then row m is node m.
    f.tm_rows_p = r.num_nodes == r.max_offset + 1 ? nullptr : r.l_vec_indices.data();
    f.tm_ptr_p = r.t_offsets.data();
    f.tm_idx_p = r.t_indices.data();
    if (f.NC == 1) {
      f.tm_covers_all = r.identity_nodes;
    } else if (r.num_nodes == r.comp_stride && r.max_offset == r.comp_stride - 1 &&
               static_cast<long long>(f.NC) * r.comp_stride == f.l_size) {
      f.tm_covers_all = true;  // the nodes are exactly 0..comp_stride-1: the components tile [0, l_size)
    } else {
      std::vector<char> referenced(f.l_size, 0);
      for (int o : r.host_offsets()) referenced[o] = 1;
      f.tm_covers_all = tmap_covers_all(f, referenced);
    }
    f.tm_built = true;
    return;
  }
  const std::vector<int>& off = r.host_offsets();
  const size_t nt = off.size();
  std::vector<int> mult(static_cast<size_t>(f.l_size) + 1, 0);
  for (size_t t = 0; t < nt; ++t) ++mult[off[t]];
  std::vector<int> row_of(f.l_size, -1), rows, ptr{0};
  for (int l = 0; l < f.l_size; ++l) {
    if (mult[l] >= 1) {
      row_of[l] = static_cast<int>(rows.size());
      rows.push_back(l);
      ptr.push_back(ptr.back() + mult[l]);
    }
  }
  std::vector<int> idx(static_cast<size_t>(ptr.back())), fill(ptr.begin(), ptr.end() - 1);
  for (size_t t = 0; t < nt; ++t)  // ascending t within each row
    if (row_of[off[t]] >= 0) idx[fill[row_of[off[t]]]++] = static_cast<int>(t);
  std::vector<char> referenced(f.l_size, 0);
  for (int l = 0; l < f.l_size; ++l) referenced[l] = mult[l] > 0;
  const bool ok = tmap_covers_all(f, referenced);
  f.tm_nrows = static_cast<int>(rows.size());
  bool identity = true;
  for (size_t m = 0; m < rows.size() && identity; ++m) identity = rows[m] == static_cast<int>(m);
  if (identity) rows.clear();  // row m is node m: the pass reads no node ids
  f.tm_rows.upload(rows);
  f.tm_ptr.upload(ptr);
  f.tm_idx.upload(idx);
  f.tm_covers_all = ok;
  f.tm_rows_p = f.tm_rows.empty() ? nullptr : f.tm_rows.data();
  f.tm_ptr_p = f.tm_ptr.data();
  f.tm_idx_p = f.tm_idx.data();
  f.tm_built = true;
  ++f.gen;
}

template <class T>
DevBuf<T>& plan_qdata(FusedMass& f, DevPlan& d, hipStream_t s);
template <>
DevBuf<double>& plan_qdata<double>(FusedMass&, DevPlan& d, hipStream_t) {
  return d.q64;
}
template <>
DevBuf<float>& plan_qdata<float>(FusedMass& f, DevPlan& d, hipStream_t s) {
  if (d.q32.size() != d.q64.size()) {
    d.q32.allocate(d.q64.size());
    ++f.gen;
    to_f32_kernel<float><<<grid_for(d.q64.size(), kThreads), kThreads, 0, s>>>(d.q64.size(), d.q64.data(),
                                                                              d.q32.data());
    HIP_CHECK_LAUNCH();
  }
  return d.q32;
}

template <class T>
DevBuf<T>& plan_partials(DevPlan& d);
template <>
DevBuf<double>& plan_partials<double>(DevPlan& d) {
  return d.part64;
}
template <>
DevBuf<float>& plan_partials<float>(DevPlan& d) {
  return d.part32;
}

template <class T>
DevBuf<T>& elem_qdata(FusedMass& f, hipStream_t s);
template <>
DevBuf<double>& elem_qdata<double>(FusedMass& f, hipStream_t) {
  return f.qdata64;
}
template <>
DevBuf<float>& elem_qdata<float>(FusedMass& f, hipStream_t s) {
  if (f.qdata32.size() != f.qdata64.size()) {
    f.qdata32.allocate(f.qdata64.size());
    ++f.gen;
    to_f32_kernel<float><<<grid_for(f.qdata64.size(), kThreads), kThreads, 0, s>>>(f.qdata64.size(), f.qdata64.data(),
                                                                                  f.qdata32.data());
    HIP_CHECK_LAUNCH();
  }
  return f.qdata32;
}

template <class T>
DevBuf<T>& box_partials(FusedMass& f);
template <>
DevBuf<double>& box_partials<double>(FusedMass& f) {
  return f.box_part64;
}
template <>
DevBuf<float>& box_partials<float>(FusedMass& f) {
  return f.box_part32;
}

template <class T>
DevBuf<T>& box_eo(FusedMass& f);
template <>
DevBuf<double>& box_eo<double>(FusedMass& f) {
  return f.box_eo64;
}
template <>
DevBuf<float>& box_eo<float>(FusedMass& f) {
  return f.box_eo32;
}

// interp_1d B [Q][P] packed as the box kernels' even-odd coefficients
// (the layout and arithmetic of even_odd_arg in fused_box_kernel.hpp).
template <class T>
std::vector<T> even_odd_table(int P, int Q, const T* B) {
  const int HP = P / 2, HQ = (Q + 1) / 2, QF = Q / 2, PM = P & 1;
  const int BM = HQ * HP, BO = HQ * HP + PM * HQ;
  std::vector<T> m(static_cast<size_t>(BO + QF * HP));
  for (int q = 0; q < HQ; ++q)
    for (int j = 0; j < HP; ++j) m[q * HP + j] = (B[q * P + j] + B[q * P + P - 1 - j]) / T(2);
  if (PM)
    for (int q = 0; q < HQ; ++q) m[BM + q] = B[q * P + P / 2];
  for (int q = 0; q < QF; ++q)
    for (int j = 0; j < HP; ++j) m[BO + q * HP + j] = (B[q * P + j] - B[q * P + P - 1 - j]) / T(2);
  return m;
}

template <class T>
const T* box_eo_table(FusedMass& f) {
  DevBuf<T>& d = box_eo<T>(f);
  if (d.empty()) {
    const std::vector<T>& B = [&]() -> const std::vector<T>& {
      if constexpr (std::is_same<T, float>::value) {
        return f.B32;
      } else {
        return f.B64;
      }
    }();
    d.upload(even_odd_table<T>(f.P, f.Q, B.data()));
    ++f.gen;
  }
  return d.data();
}

// Per-component length of the box-column partial area for the key's variant.
size_t box_part_len(const FusedMass& f, const FusedKey& k) {
  int BX = 0, BY = 0;
  require(box_columns(k.elems, &BX, &BY), "fused mass: box-column variant not compiled");
  const long long ncols = static_cast<long long>(f.box_n[0] / BX) * (f.box_n[1] / BY);
  const int nsz = box_nsz_of(f, k);
  require(nsz > 0, "fused mass: no segment count fits the box");
  const int lz = (f.box_n[2] + nsz - 1) / nsz;  // the longest segment lays out every segment's slots
  const long long nseg = ncols * nsz;
  return static_cast<size_t>(nseg * box_face_size(BX * (f.P - 1) + 1, BY * (f.P - 1) + 1, lz * (f.P - 1) + 1));
}

template <class T>
DevBuf<T>& elem_outputs(FusedMass& f);
template <>
DevBuf<double>& elem_outputs<double>(FusedMass& f) {
  return f.Y64;
}
template <>
DevBuf<float>& elem_outputs<float>(FusedMass& f) {
  return f.Y32;
}

FusedMass::StageEvents* next_events(FusedMass& f) {
  if (!f.timing) return nullptr;
  if (f.events_used == f.events.size()) {
    FusedMass::StageEvents e;
    HIP_CHECK(hipEventCreate(&e.a));
    HIP_CHECK(hipEventCreate(&e.b));
    HIP_CHECK(hipEventCreate(&e.c));
    HIP_CHECK(hipEventCreate(&e.d));
    f.events.push_back(e);
  }
  return &f.events[f.events_used++];
}

// Makes every buffer the key's algorithm needs (so the apply itself allocates nothing).
template <class T>
void prepare_key(FusedMass& f, const FusedKey& k, hipStream_t s) {
  if (alg_box(k.alg)) {
    require(box_fits(f, k.elems), "fused mass: the box-column algorithm needs a structured box its columns tile");
    (void)elem_qdata<T>(f, s);
    (void)box_eo_table<T>(f);
    DevBuf<T>& pb = box_partials<T>(f);
    const size_t n = std::max<size_t>(box_part_len(f, k), 1) * f.NC;
    if (pb.size() < n) {
      pb.allocate(n);
      ++f.gen;
    }
    return;
  }
  if (alg_blocked(k.alg)) {
    const bool part = alg_partial(k.alg);
    DevPlan& d = ensure_plan(f, k.elems, fused_node_budget(k), part, s);
    (void)plan_qdata<T>(f, d, s);
    if (part) {
      DevBuf<T>& pb = plan_partials<T>(d);
      const size_t n = std::max<size_t>(d.npart, 1) * f.NC;
      if (pb.size() != n) {
        pb.allocate(n);
        ++f.gen;
      }
    }
  } else {
    (void)elem_qdata<T>(f, s);
    if (k.alg == kAlgElemAtomic && !k.exact) ensure_flagged_offsets(f, s);
    if (alg_tmap(k.alg)) {
      ensure_tmap(f);
      DevBuf<T>& Y = elem_outputs<T>(f);
      const size_t n = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P * f.NC;
      if (Y.size() != n) {
        Y.allocate(n);
        ++f.gen;
      }
    }
  }
}

// second = false (split applies): the element kernel only, without the
// zero fill (the split requires every entry covered) and the second pass.
template <class T>
void apply_key(FusedMass& f, const FusedKey& k, const T* u, T* v, hipStream_t s, bool second = true) {
  FusedLauncher L = find(k);
  require(L != nullptr, "fused mass: variant not compiled");
  prepare_key<T>(f, k, s);
  const bool f32 = std::is_same<T, float>::value;
  FusedArgs a;
  a.comp_stride = f.comp_stride;
  a.u = u;
  a.basis = f32 ? static_cast<const void*>(f.B32.data()) : static_cast<const void*>(f.B64.data());
  FusedMass::StageEvents* ev = next_events(f);
  if (ev) HIP_CHECK(hipEventRecord(ev->a, s));
  if (alg_box(k.alg)) {
    int BX = 0, BY = 0;
    box_columns(k.elems, &BX, &BY);
    a.qdata = elem_qdata<T>(f, s).data();
    a.box_n[0] = f.box_n[0];
    a.box_n[1] = f.box_n[1];
    a.box_n[2] = f.box_n[2];
    a.box_nsz = box_nsz_of(f, k);
    a.v = v;
    a.partials = box_partials<T>(f).data();
    a.part_comp_len = std::max<size_t>(box_part_len(f, k), 1);
    a.basis_eo = box_eo<T>(f).data();
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    a.box_pass = 0;
    L(a, s);
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
    a.box_pass = 1;
    L(a, s);
  } else if (alg_blocked(k.alg)) {
    DevPlan& d = ensure_plan(f, k.elems, fused_node_budget(k), alg_partial(k.alg), s);
    a.qdata = plan_qdata<T>(f, d, s).data();
    a.nblk = d.nblk;
    a.max_nodes = d.max_nodes;
    a.node_ptr = d.node_ptr.data();
    a.nodes = d.nodes.data();
    a.ninterior = d.ninterior.data();
    a.nelem = d.nelem.data();
    a.loc = d.loc.data();
    a.color = d.color.data();
    a.ncolor = d.ncolor.data();
    a.part_ptr = d.part_ptr.data();
    a.v = v;
    if (k.alg == kAlgBlockAtomic || !d.covers_all) zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(T), s);
    if (alg_partial(k.alg)) {
      a.partials = plan_partials<T>(d).data();
      a.part_comp_len = std::max<size_t>(d.npart, 1);
    }
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    L(a, s);
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
    if (alg_partial(k.alg) && second)
      launch_row_sum<T>(d.nrows, f.NC, f.comp_stride, d.rows.data(), d.row_ptr.data(), d.row_part.data(),
                        static_cast<const T*>(a.partials), a.part_comp_len, v, s);
  } else {
    a.qdata = elem_qdata<T>(f, s).data();
    a.num_elem = f.num_elem;
    a.offsets = f.rstr->offsets.data();
    // Only the atomic element kernel reads the single-contributor flags (the
    // others read plain offsets through the same argument).
    a.flagged_offsets = k.alg == kAlgElemAtomic && !k.exact ? f.flagged_offsets.data() : f.rstr->offsets.data();
    a.y_comp_len = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
    a.v = v;
    if (alg_tmap(k.alg)) {
      a.out = elem_outputs<T>(f).data();
      if (!f.tm_covers_all) zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(T), s);
      if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
      L(a, s);
      if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
      if (second)
        launch_row_sum<T>(f.tm_nrows, f.NC, f.comp_stride, f.tm_rows_p, f.tm_ptr_p, f.tm_idx_p,
                          elem_outputs<T>(f).data(), a.y_comp_len, v, s);
    } else {
      zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(T), s);
      a.out = v;
      if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
      L(a, s);
      if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
    }
  }
  if (ev) HIP_CHECK(hipEventRecord(ev->d, s));
}

}  // namespace

bool fused_box_operator(const FusedMass& f) {
  return f.box && f.basis_symmetric && !fused_elem_variants(f.P, f.Q, f.NC, 0, kAlgBoxColumn, 0).empty();
}

namespace {

// The only box the numbering can be, from offsets 0, P and P*P (element 0's
// nodes (0,0,0), (0,1,0) and (0,0,1)) and the sizes; false if none fits.
bool box_shape(const int* off, int num_elem, int P, int num_comp, int comp_stride, int l_size, long long* n3l) {
  if (P < 2 || num_elem <= 0 || off[0] != 0) return false;
  const long long Nx = off[P], Nxy = off[P * P];  // the nodes one step up in y and in z
  if (Nx < P || (Nx - 1) % (P - 1) != 0 || Nxy <= 0 || Nxy % Nx != 0) return false;
  const long long Ny = Nxy / Nx;
  if (Ny < P || (Ny - 1) % (P - 1) != 0) return false;
  const long long nx = (Nx - 1) / (P - 1), ny = (Ny - 1) / (P - 1);
  if (num_elem % (nx * ny) != 0) return false;
  const long long nz = num_elem / (nx * ny), Nz = nz * (P - 1) + 1;
  // One component: any comp_stride (unused); more: blocked components, one after another.
  if ((num_comp > 1 && Nx * Ny * Nz != comp_stride) || static_cast<long long>(num_comp) * Nx * Ny * Nz != l_size)
    return false;
  n3l[0] = nx;
  n3l[1] = ny;
  n3l[2] = nz;
  return true;
}

// mismatch[0] = 1 where an offset differs from the box formula (every writer
// stores the same value).
__global__ __launch_bounds__(kThreads) void box_check_kernel(size_t nt, int P, int nx, int ny, int Nx, long long Nxy,
                                                            const int* __restrict__ off, int* __restrict__ mismatch) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t >= nt) return;
  const int P3 = P * P * P;
  const long long e = static_cast<long long>(t / P3);
  const int i = static_cast<int>(t % P3), ix = i % P, iy = (i / P) % P, iz = i / (P * P);
  const long long ex = e % nx, ey = (e / nx) % ny, ez = e / (static_cast<long long>(nx) * ny);
  const long long want = ex * (P - 1) + ix + Nx * (ey * (P - 1) + iy) + Nxy * (ez * (P - 1) + iz);
  if (off[t] != want) mismatch[0] = 1;
}

// detect_box on the restriction's device copy of the offsets.
bool detect_box_device(const Restriction& r, int P, int num_comp, int* n3, hipStream_t s) {
  if (r.box_n[0] > 0 && r.type == 0 && r.elem_size == P * P * P && r.num_comp == num_comp) {
    // Made by the box generator: its offsets are the box formula (nothing to read back).
    for (int d = 0; d < 3; ++d) n3[d] = r.box_n[d];
    return true;
  }
  long long n3l[3];
  // The shape follows from offsets 0, P and P*P (box_shape); the device check
  // below tests all of them.
  std::vector<int> head(static_cast<size_t>(P) * P + 1);
  if (static_cast<size_t>(r.num_elem) * P * P * P < head.size()) return false;
  HIP_CHECK(hipMemcpyAsync(head.data(), r.offsets.data(), head.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (!box_shape(head.data(), r.num_elem, P, num_comp, r.comp_stride, r.l_size, n3l)) return false;
  const size_t nt = static_cast<size_t>(r.num_elem) * P * P * P;
  const int Nx = static_cast<int>(n3l[0] * (P - 1) + 1);
  const long long Nxy = static_cast<long long>(Nx) * (n3l[1] * (P - 1) + 1);
  StreamTmp<int> mismatch(1, s);
  zero_fill(mismatch.data(), sizeof(int), s);
  box_check_kernel<<<grid_for(nt, kThreads), kThreads, 0, s>>>(nt, P, static_cast<int>(n3l[0]),
                                                              static_cast<int>(n3l[1]), Nx, Nxy, r.offsets.data(),
                                                              mismatch.data());
  HIP_CHECK_LAUNCH();
  int bad = 1;
  HIP_CHECK(hipMemcpyAsync(&bad, mismatch.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (bad) return false;
  for (int d = 0; d < 3; ++d) n3[d] = static_cast<int>(n3l[d]);
  return true;
}

// Multiplicity of every node (atomics), then the offsets with bit 31 set on
// single-contributor nodes (the atomic element kernel stores those instead of adding).
__global__ __launch_bounds__(kThreads) void node_count_kernel(size_t nt, const int* __restrict__ off,
                                                             int* __restrict__ count) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t < nt) atomicAdd(&count[off[t]], 1);
}

__global__ __launch_bounds__(kThreads) void flag_single_kernel(size_t nt, const int* __restrict__ off,
                                                              const int* __restrict__ count,
                                                              int* __restrict__ flagged) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t >= nt) return;
  const int o = off[t];
  flagged[t] = count[o] == 1 ? static_cast<int>(static_cast<unsigned>(o) | 0x80000000u) : o;
}

}  // namespace

void ensure_flagged_offsets(FusedMass& f, hipStream_t s) {
  const size_t nt = f.rstr->offsets.size();
  if (f.flagged_offsets.size() == nt) return;
  StreamTmp<int> count(static_cast<size_t>(f.l_size), s);
  zero_fill(count.data(), static_cast<size_t>(f.l_size) * sizeof(int), s);
  f.flagged_offsets.allocate(nt);
  ++f.gen;
  node_count_kernel<<<grid_for(nt, kThreads), kThreads, 0, s>>>(nt, f.rstr->offsets.data(), count.data());
  flag_single_kernel<<<grid_for(nt, kThreads), kThreads, 0, s>>>(nt, f.rstr->offsets.data(), count.data(),
                                                                f.flagged_offsets.data());
  HIP_CHECK_LAUNCH();
}

bool detect_box(const int* off, int num_elem, int P, int num_comp, int comp_stride, int l_size, int* n3) {
  long long n3l[3];
  if (!box_shape(off, num_elem, P, num_comp, comp_stride, l_size, n3l)) return false;
  const long long nx = n3l[0], ny = n3l[1], nz = n3l[2];
  const long long Nx = nx * (P - 1) + 1, Ny = ny * (P - 1) + 1;
  long long t = 0;
  for (long long ez = 0; ez < nz; ++ez)
    for (long long ey = 0; ey < ny; ++ey)
      for (long long ex = 0; ex < nx; ++ex)
        for (int iz = 0; iz < P; ++iz)
          for (int iy = 0; iy < P; ++iy) {
            const long long row = ex * (P - 1) + Nx * ((ey * (P - 1) + iy) + Ny * (ez * (P - 1) + iz));
            for (int ix = 0; ix < P; ++ix, ++t)
              if (off[t] != row + ix) return false;
          }
  n3[0] = static_cast<int>(nx);
  n3[1] = static_cast<int>(ny);
  n3[2] = static_cast<int>(nz);
  return true;
}

int box_segments(int nz, long long ncols, long long slots, int P, int max_layers) {
  if (nz <= 0 || ncols <= 0 || slots <= 0) return 0;
  auto ok = [&](int nsz) { return (nz / nsz) * (P - 1) >= 2; };
  if (max_layers > 0) {
    const int nsz = (nz + max_layers - 1) / max_layers;
    return nsz <= nz && ok(nsz) ? nsz : 0;
  }
  auto steps = [&](int nsz) { return ((ncols * nsz + slots - 1) / slots) * ((nz + nsz - 1) / nsz); };
  long long least = 0;
  for (int nsz = 1; nsz <= nz && ok(nsz); ++nsz)
    if (!least || steps(nsz) < least) least = steps(nsz);
  for (int nsz = 1; nsz <= nz && ok(nsz); ++nsz)
    if (steps(nsz) * 20 <= least * 21) return nsz;
  return 0;
}

FusedKey fused_resolve(const FusedMass& f, int fp32) {
  FusedKey k{f.P, f.Q, f.NC, fp32, -1, 0, 0};
  const bool det = f.scatter == kScatterDeterministic;
  if (f.exact && !fp32) {
    // The reference's arithmetic: element kernels only (tmap reproduces the CPU sum order).
    k.exact = 1;
    k.alg = det ? kAlgElemTmap : kAlgElemAtomic;
    k.elems = f.choice.alg == k.alg && f.choice.elems > 0 ? f.choice.elems : default_elems(f.P, f.Q, f.NC, 0, k.alg, 1);
    require(k.elems > 0 && find(k), "fused mass: exact variant not compiled");
    return k;
  }
  if (f.choice.alg >= 0) {
    require(f.scatter != kScatterDeterministic || alg_deterministic(f.choice.alg),
            "fused mass: a deterministic scatter needs a deterministic algorithm");
    require(!alg_box(f.choice.alg) || f.box,
            "fused mass: the box-column algorithm needs a structured box mesh (lexicographic node numbering)");
    k.alg = f.choice.alg;
    k.elems = f.choice.elems;
    if (k.elems <= 0) {
      std::lock_guard<std::mutex> lk(tune_mutex());
      auto it = tune_alg_table().find(TuneAlgKey{f.P, f.Q, f.NC, fp32, k.alg});
      k.elems = it != tune_alg_table().end() ? it->second.elems : default_elems(f.P, f.Q, f.NC, fp32, k.alg, 0);
      if (alg_box(k.alg) && !box_fits(f, k.elems)) k.elems = box_default_elems(f, fp32);
    }
    require(k.elems > 0 && find(k), "fused mass: the chosen algorithm is not compiled for this shape");
    require(key_valid(f, k), "fused mass: the box-column variant does not tile this box");
    return k;
  }
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    for (int box = fused_box_operator(f) ? 1 : 0; box >= 0; --box) {
      const TuneEntry* e = tune_lookup(TuneKey{f.P, f.Q, f.NC, fp32, det ? 1 : 0, box}, f.num_elem);
      if (!e) continue;
      FusedKey t = k;
      t.alg = e->alg;
      t.elems = e->elems;
      if (t.alg >= 0 && key_valid(f, t) && find(t)) return t;
    }
  }
  // Default order: box columns (structured boxes of at least kBoxMinNodes
  // element nodes), partial-sum blocks, then (deterministic) transpose map or
  // (atomic) block atomics, then elements.
  const int order_det[] = {kAlgBoxColumn, kAlgBlockPart, kAlgElemTmap};
  const int order_any[] = {kAlgBoxColumn, kAlgBlockPartLds, kAlgBlockPart, kAlgBlockAtomic, kAlgElemAtomic,
                           kAlgElemTmap};
  const int* order = det ? order_det : order_any;
  const int n = det ? 3 : 6;
  const long long elem_nodes = static_cast<long long>(f.num_elem) * f.P * f.P * f.P * f.NC;
  for (int i = 0; i < n; ++i) {
    k.alg = order[i];
    if (!valid_for(f, k.alg) || (alg_box(k.alg) && elem_nodes < kBoxMinNodes)) continue;
    k.elems = alg_box(k.alg) ? box_default_elems(f, fp32) : default_elems(f.P, f.Q, f.NC, fp32, k.alg, 0);
    if (k.elems > 0 && find(k)) return k;
  }
  throw Unsupported("fused mass: no compiled variant for this shape");
}

std::unique_ptr<FusedMass> fused_mass_create(const Restriction& r, const HostBasis& b, const double* qdata_ref,
                                             hipStream_t s) {
  require(b.dim == 3, "fused mass: dim must be 3");
  require(fused_mass_supported(b.P, b.Q, b.num_comp), "fused mass: shape not compiled");
  auto f = std::make_unique<FusedMass>();
  f->P = b.P;
  f->Q = b.Q;
  f->NC = b.num_comp;
  f->num_elem = r.num_elem;
  f->comp_stride = r.comp_stride;
  f->l_size = r.l_size;
  f->rstr = &r;
  f->scatter = r.scatter_mode;
  f->B64 = b.interp;
  f->B32.assign(b.interp.begin(), b.interp.end());
  f->Bsq.resize(b.interp.size());
  for (size_t i = 0; i < b.interp.size(); ++i) f->Bsq[i] = b.interp[i] * b.interp[i];
  f->box = detect_box_device(r, b.P, b.num_comp, f->box_n, s);
  {
    double bmax = 0.0, asym = 0.0;
    for (int q = 0; q < b.Q; ++q)
      for (int p = 0; p < b.P; ++p) {
        bmax = std::max(bmax, std::fabs(b.interp[q * b.P + p]));
        asym = std::max(asym, std::fabs(b.interp[q * b.P + p] - b.interp[(b.Q - 1 - q) * b.P + (b.P - 1 - p)]));
      }
    f->basis_symmetric = asym <= 1e-14 * bmax;
  }
  const size_t nq = static_cast<size_t>(b.Q) * b.Q * b.Q * r.num_elem;
  f->qdata64.allocate(nq);
  qdata_private_kernel<double><<<grid_for(nq, kThreads), kThreads, 0, s>>>(r.num_elem, b.Q, qdata_ref,
                                                                          f->qdata64.data());
  HIP_CHECK_LAUNCH();
  return f;
}

FusedMass::~FusedMass() {
  for (auto& e : events) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
    (void)hipEventDestroy(e.c);
    (void)hipEventDestroy(e.d);
  }
}

void fused_timing_begin(FusedMass& f) {
  f.timing = true;
  f.events_used = 0;
}

void fused_timing_end(FusedMass& f, double* kernel_ms, double* aux_ms, int* count) {
  double k = 0.0, x = 0.0;
  for (size_t i = 0; i < f.events_used; ++i) {
    const auto& e = f.events[i];
    HIP_CHECK(hipEventSynchronize(e.d));
    float t1 = 0.f, t2 = 0.f, t3 = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t1, e.a, e.b));
    HIP_CHECK(hipEventElapsedTime(&t2, e.b, e.c));
    HIP_CHECK(hipEventElapsedTime(&t3, e.c, e.d));
    k += t2;
    x += t1 + t3;
  }
  if (kernel_ms) *kernel_ms = k;
  if (aux_ms) *aux_ms = x;
  if (count) *count = static_cast<int>(f.events_used);
  f.timing = false;
  f.events_used = 0;
}

namespace {

// Whether the element + transpose-map path can run with nothing made: the
// restriction carries its own map (deterministic restrictions build it at
// create).
bool tmap_without_alloc(const FusedMass& f) {
  const Restriction& r = *f.rstr;
  return f.tm_built || (r.type == 0 && r.scatter_mode == kScatterDeterministic && !r.t_offsets.empty());
}

}  // namespace

size_t fused_scratch_bytes64(const FusedMass& f) {
  const FusedKey k = fused_resolve(f, 0);
  if (alg_box(k.alg)) return std::max<size_t>(box_part_len(f, k), 1) * f.NC * sizeof(double);
  if (k.alg == kAlgElemTmap) return static_cast<size_t>(f.num_elem) * f.P * f.P * f.P * f.NC * sizeof(double);
  return 0;
}

void fused_mass_apply64_scratch(FusedMass& f, const double* u, double* v, void* scratch, size_t scratch_bytes,
                                hipStream_t s) {
  const FusedKey k = fused_resolve(f, 0);
  FusedLauncher L = find(k);
  require(L != nullptr, "fused mass: variant not compiled");
  if (!(alg_box(k.alg) || k.alg == kAlgElemTmap))
    throw Unsupported("apply_noalloc: the operator's algorithm is neither the box columns nor the element + "
                      "transpose-map scatter (the algorithms whose only per-apply buffer is the scratch)");
  const size_t need = fused_scratch_bytes64(f);
  if (need > scratch_bytes)
    throw std::invalid_argument("apply_noalloc: the apply needs " + std::to_string(need) +
                                " bytes of scratch, more than the " + std::to_string(scratch_bytes) + " available");
  require(scratch != nullptr && reinterpret_cast<uintptr_t>(scratch) % 16 == 0, "apply_noalloc: bad scratch pointer");
  FusedArgs a;
  a.comp_stride = f.comp_stride;
  a.u = u;
  a.basis = f.B64.data();
  a.qdata = f.qdata64.data();  // made at create
  if (alg_box(k.alg)) {
    require(box_fits(f, k.elems), "fused mass: the box-column algorithm needs a structured box its columns tile");
    a.box_n[0] = f.box_n[0];
    a.box_n[1] = f.box_n[1];
    a.box_n[2] = f.box_n[2];
    a.box_nsz = box_nsz_of(f, k);
    a.v = v;
    a.partials = scratch;  // every slot is written before the second pass reads it: no clearing
    a.part_comp_len = std::max<size_t>(box_part_len(f, k), 1);
    require(!f.box_eo64.empty(),
            "apply_noalloc: run one ordinary apply first (it puts the box kernels' coefficients on the device)");
    a.basis_eo = f.box_eo64.data();
    a.box_pass = 0;
    L(a, s);
    a.box_pass = 1;
    L(a, s);
    return;
  }
  if (!tmap_without_alloc(f))
    throw Unsupported("apply_noalloc: the transpose-map scatter needs a deterministic restriction (its map is made "
                      "at create)");
  ensure_tmap(f);  // points at the restriction's map: no device memory
  a.num_elem = f.num_elem;
  a.offsets = f.rstr->offsets.data();
  a.flagged_offsets = f.rstr->offsets.data();
  a.y_comp_len = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
  a.v = v;
  a.out = scratch;  // element outputs: every entry is written before the row sums read it
  if (!f.tm_covers_all) zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(double), s);  // a fill kernel
  L(a, s);
  launch_row_sum<double>(f.tm_nrows, f.NC, f.comp_stride, f.tm_rows_p, f.tm_ptr_p, f.tm_idx_p,
                         static_cast<const double*>(scratch), a.y_comp_len, v, s);
}

uint64_t fused_signature(const FusedMass& f, int fp32) {
  const FusedKey k = fused_resolve(f, fp32);
  const long long parts[] = {k.P, k.Q, k.NC, k.fp32, k.alg, k.exact, k.elems, alg_box(k.alg) ? box_nsz_of(f, k) : 0,
                             f.scatter, f.exact, static_cast<long long>(f.gen)};
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the parts
  for (long long v : parts) {
    for (int i = 0; i < 8; ++i) {
      h ^= static_cast<uint64_t>(v >> (8 * i)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

void fused_prepare(FusedMass& f, int fp32, hipStream_t s) {
  const FusedKey k = fused_resolve(f, fp32);
  if (fp32) {
    prepare_key<float>(f, k, s);
  } else {
    prepare_key<double>(f, k, s);
  }
}

namespace {

// Interior nodes (stored by the block kernel) are -1; then a row's node gets its row.
__global__ __launch_bounds__(kThreads) void node_row_kernel(size_t nodes, int* __restrict__ node_row) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < nodes) node_row[i] = -1;
}
__global__ __launch_bounds__(kThreads) void node_row_scatter_kernel(int nrows, const int* __restrict__ rows,
                                                                    int* __restrict__ node_row) {
  const int m = blockIdx.x * kThreads + threadIdx.x;
  if (m < nrows) node_row[rows[m]] = m;
}

template <class T>
bool split_impl(FusedMass& f, FusedSplit* o, hipStream_t s) {
  const int fp32 = std::is_same<T, float>::value ? 1 : 0;
  const FusedKey k = fused_resolve(f, fp32);
  // Entry i of a vector is node i % nodes of component i / nodes.
  const long long nodes = f.NC == 1 ? f.l_size : f.comp_stride;
  if (static_cast<long long>(f.NC) * nodes != f.l_size) return false;
  if (alg_blocked(k.alg) && alg_partial(k.alg)) {
    prepare_key<T>(f, k, s);
    DevPlan& d = ensure_plan(f, k.elems, fused_node_budget(k), true, s);
    if (!d.covers_all) return false;
    if (d.node_row.size() != static_cast<size_t>(nodes)) {
      d.node_row.allocate(static_cast<size_t>(nodes));
      ++f.gen;
      node_row_kernel<<<grid_for(static_cast<size_t>(nodes), kThreads), kThreads, 0, s>>>(
          static_cast<size_t>(nodes), d.node_row.data());
      if (d.nrows > 0)
        node_row_scatter_kernel<<<grid_for(static_cast<size_t>(d.nrows), kThreads), kThreads, 0, s>>>(
            d.nrows, d.rows.data(), d.node_row.data());
      HIP_CHECK_LAUNCH();
    }
    o->node_row = d.node_row.data();
    o->ptr = d.row_ptr.data();
    o->idx = d.row_part.data();
    o->vals = plan_partials<T>(d).data();
    o->cl = std::max<size_t>(d.npart, 1);
    o->cs = static_cast<int>(nodes);
    return true;
  }
  if (alg_tmap(k.alg)) {
    prepare_key<T>(f, k, s);
    if (!f.tm_covers_all || f.tm_rows_p != nullptr || f.tm_nrows != nodes) return false;
    o->node_row = nullptr;
    o->ptr = f.tm_ptr_p;
    o->idx = f.tm_idx_p;
    o->vals = elem_outputs<T>(f).data();
    o->cl = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
    o->cs = static_cast<int>(nodes);
    return true;
  }
  return false;
}

}  // namespace

bool fused_split(FusedMass& f, int fp32, FusedSplit* out, hipStream_t s) {
  return fp32 ? split_impl<float>(f, out, s) : split_impl<double>(f, out, s);
}

void fused_mass_apply_first64(FusedMass& f, const double* u, double* v, hipStream_t s) {
  apply_key<double>(f, fused_resolve(f, 0), u, v, s, false);
}

void fused_mass_apply_first32(FusedMass& f, const float* u, float* v, hipStream_t s) {
  apply_key<float>(f, fused_resolve(f, 1), u, v, s, false);
}

void fused_mass_apply64(FusedMass& f, const double* u, double* v, hipStream_t s) {
  apply_key<double>(f, fused_resolve(f, 0), u, v, s);
}

void fused_mass_apply32(FusedMass& f, const float* u, float* v, hipStream_t s) {
  apply_key<float>(f, fused_resolve(f, 1), u, v, s);
}

namespace {

template <class T>
TuneEntry autotune_impl(FusedMass& f, int reps, hipStream_t s, std::vector<TuneEntry>* all) {
  const int fp32 = std::is_same<T, float>::value ? 1 : 0;
  const size_t L = static_cast<size_t>(f.l_size);
  DevBuf<T> u(L), v(L);
  // Any input exercises the same memory traffic; ones keep the values finite.
  std::vector<T> ones(L, T(1));
  u.upload(ones);
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  TuneEntry best;
  best.ms = 1e300;
  std::vector<TuneEntry> local;
  const FusedChoice saved = f.choice;
  const bool saved_timing = f.timing;
  f.timing = false;
  try {
    for (int alg = 0; alg < kAlgCount; ++alg) {
      if (!valid_for(f, alg)) continue;
      for (int ne : fused_elem_variants(f.P, f.Q, f.NC, fp32, alg, 0)) {
        const FusedKey k{f.P, f.Q, f.NC, fp32, alg, 0, ne};
        if (!key_valid(f, k)) continue;
        apply_key<T>(f, k, u.data(), v.data(), s);  // warm-up, and builds the variant's state
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) apply_key<T>(f, k, u.data(), v.data(), s);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        TuneEntry t{alg, ne, ms / reps, f.num_elem};
        local.push_back(t);
        if (all) all->push_back(t);
        if (t.ms < best.ms) best = t;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f.choice = saved;
    f.timing = saved_timing;
    throw;
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  f.choice = saved;
  f.timing = saved_timing;
  require(best.alg >= 0, "autotune: no candidate");
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    tune_record(TuneKey{f.P, f.Q, f.NC, fp32, f.scatter == kScatterDeterministic ? 1 : 0, fused_box_operator(f) ? 1 : 0},
                best);
    for (const TuneEntry& t : local) {
      TuneEntry& slot = tune_alg_table()[TuneAlgKey{f.P, f.Q, f.NC, fp32, t.alg}];
      if (slot.alg < 0 || t.ms < slot.ms || slot.num_elem != t.num_elem) slot = t;
    }
  }
  // Drop block plans the winner does not use (each holds a slot-ordered qdata copy).
  const FusedKey wk{f.P, f.Q, f.NC, fp32, best.alg, 0, best.elems};
  for (auto it = f.plans.begin(); it != f.plans.end();) {
    const bool keep = alg_blocked(best.alg) && it->first == std::make_pair(best.elems, fused_node_budget(wk));
    it = keep ? std::next(it) : f.plans.erase(it);
  }
  if (!alg_tmap(best.alg)) {
    f.Y64.release();
    f.Y32.release();
  }
  ++f.gen;
  return best;
}

}  // namespace

TuneEntry fused_autotune(FusedMass& f, int fp32, int reps, hipStream_t s, std::vector<TuneEntry>* all) {
  require(reps > 0, "autotune: reps must be positive");
  return fp32 ? autotune_impl<float>(f, reps, s, all) : autotune_impl<double>(f, reps, s, all);
}

void fused_mass_diagonal(FusedMass& f, double* diag, hipStream_t s) {
  ensure_tmap(f);
  if (!f.tm_covers_all) zero_fill(diag, static_cast<size_t>(f.l_size) * sizeof(double), s);
  if (f.Bsq_dev.size() != f.Bsq.size()) {
    // From the operator's own host copy, which outlives the copy (nothing waits here).
    f.Bsq_dev.allocate(f.Bsq.size());
    HIP_CHECK(hipMemcpyAsync(f.Bsq_dev.data(), f.Bsq.data(), f.Bsq.size() * sizeof(double), hipMemcpyHostToDevice,
                             s));
  }
  const size_t n = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
  StreamTmp<double> Yd(n, s);
  mass_diag_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(f.num_elem, f.P, f.Q, f.qdata64.data(),
                                                               f.Bsq_dev.data(), Yd.data());
  HIP_CHECK_LAUNCH();
  // Every component gets the same diagonal (value stride 0 across components).
  launch_row_sum<double>(f.tm_nrows, f.NC, f.comp_stride, f.tm_rows_p, f.tm_ptr_p, f.tm_idx_p,
                         Yd.data(), 0, diag, s);
}

}  // namespace femhip
