// Element restriction kernels for gfx950 (wave64).
//
// Offset gather/scatter use a 64-element x 32-node tile per 256-thread workgroup:
// the tile's offsets are read row by row (32 consecutive ints of one element),
// staged in LDS, and the E-vector side is then accessed with 64 consecutive
// elements per node row, so both the index reads and the element-innermost
// E-vector accesses are coalesced. (The reference's one-thread-per-E-node
// mapping, src/cuda/cuda-elem-restriction.cu:37-67, coalesces only one side.)
//
// The deterministic scatter is one thread per used L-node summing its CSR row
// in ascending t, starting from the current L value, so the result equals the
// CPU loop ((v + c0) + c1) + ... of src/elem-restriction.cpp:73-81 bit for bit
// for any starting v (the reference kernel starts from 0 and adds the row sum
// to v once, src/cuda/cuda-elem-restriction.cu:85-92, which matches only for v = 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>
#include <numeric>

#include "restriction.hpp"
#include "mesh.hpp"

namespace femhip {

namespace {

constexpr int kTileE = 64;
constexpr int kTileI = 32;
constexpr int kThreads = 256;

template <class T>
__global__ __launch_bounds__(kThreads) void offset_gather_kernel(int ne, int es, int nc, int cs,
                                                                 const int* __restrict__ off,
                                                                 const T* __restrict__ L, T* __restrict__ E) {
  __shared__ int s_off[kTileE][kTileI + 1];
  const int e0 = blockIdx.x * kTileE, i0 = blockIdx.y * kTileI;
  for (int idx = threadIdx.x; idx < kTileE * kTileI; idx += kThreads) {
    const int il = idx % kTileI, el = idx / kTileI;
    const int e = e0 + el, i = i0 + il;
    s_off[el][il] = (e < ne && i < es) ? off[static_cast<size_t>(e) * es + i] : 0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kTileE * kTileI; idx += kThreads) {
    const int el = idx % kTileE, il = idx / kTileE;
    const int e = e0 + el, i = i0 + il;
    if (e < ne && i < es) {
      const size_t o = static_cast<size_t>(s_off[el][il]);
      for (int k = 0; k < nc; ++k)
        E[(static_cast<size_t>(k) * es + i) * ne + e] = L[o + static_cast<size_t>(k) * cs];
    }
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void offset_scatter_atomic_kernel(int ne, int es, int nc, int cs,
                                                                         const int* __restrict__ off,
                                                                         const T* __restrict__ E, T* __restrict__ L) {
  __shared__ int s_off[kTileE][kTileI + 1];
  const int e0 = blockIdx.x * kTileE, i0 = blockIdx.y * kTileI;
  for (int idx = threadIdx.x; idx < kTileE * kTileI; idx += kThreads) {
    const int il = idx % kTileI, el = idx / kTileI;
    const int e = e0 + el, i = i0 + il;
    s_off[el][il] = (e < ne && i < es) ? off[static_cast<size_t>(e) * es + i] : 0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kTileE * kTileI; idx += kThreads) {
    const int el = idx % kTileE, il = idx / kTileE;
    const int e = e0 + el, i = i0 + il;
    if (e < ne && i < es) {
      const size_t o = static_cast<size_t>(s_off[el][il]);
      for (int k = 0; k < nc; ++k)
        unsafeAtomicAdd(&L[o + static_cast<size_t>(k) * cs], E[(static_cast<size_t>(k) * es + i) * ne + e]);
    }
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void offset_scatter_det_kernel(int rows, int ne, int es, int nc, int cs,
                                                                      const int* __restrict__ lvec,
                                                                      const int* __restrict__ toff,
                                                                      const int* __restrict__ tind,
                                                                      const T* __restrict__ E, T* __restrict__ L) {
  const int n = blockIdx.x * kThreads + threadIdx.x;
  if (n >= rows) return;
  const size_t l = lvec ? static_cast<size_t>(lvec[n]) : static_cast<size_t>(n);
  const int t0 = toff[n], t1 = toff[n + 1];
  for (int k = 0; k < nc; ++k) {
    T acc = L[l + static_cast<size_t>(k) * cs];
    for (int t = t0; t < t1; ++t) {
      const int ti = tind[t];
      const int e = ti / es, i = ti - e * es;
      acc += E[(static_cast<size_t>(k) * es + i) * ne + e];
    }
    L[l + static_cast<size_t>(k) * cs] = acc;
  }
}

// Transpose-map sum of a [comp][t] element-output array (the fused operator's
// deterministic second pass). Rows ascending in L, t ascending within a row.
template <class T>
__global__ __launch_bounds__(kThreads) void csr_scatter_kernel(int rows, size_t comp_len, int nc, int cs,
                                                               const int* __restrict__ lvec,
                                                               const int* __restrict__ toff,
                                                               const int* __restrict__ tind, const T* __restrict__ Y,
                                                               T* __restrict__ L, int zero_start) {
  const int n = blockIdx.x * kThreads + threadIdx.x;
  if (n >= rows) return;
  const size_t l = lvec ? static_cast<size_t>(lvec[n]) : static_cast<size_t>(n);
  const int t0 = toff[n], t1 = toff[n + 1];
  // Rows of up to 8 entries (every node of a conforming hex mesh) issue all
  // index loads, then all value loads, before the in-order sum; longer rows
  // continue in chunks of 8.
  constexpr int kChunk = 8;
  for (int k = 0; k < nc; ++k) {
    const T* Yk = Y + k * comp_len;
    T acc = zero_start ? T(0) : L[l + static_cast<size_t>(k) * cs];
    for (int tb = t0; tb < t1; tb += kChunk) {
      int idx[kChunk];
      T val[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) idx[j] = (tb + j < t1) ? tind[tb + j] : -1;
#pragma unroll
      for (int j = 0; j < kChunk; ++j) val[j] = (idx[j] >= 0) ? Yk[idx[j]] : T(0);
#pragma unroll
      for (int j = 0; j < kChunk; ++j)
        if (idx[j] >= 0) acc += val[j];
    }
    L[l + static_cast<size_t>(k) * cs] = acc;
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void strided_gather_kernel(size_t e_size, int ne, int es, long long s0,
                                                                  long long s1, long long s2,
                                                                  const T* __restrict__ L, T* __restrict__ E) {
  const size_t n = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (n >= e_size) return;
  const size_t e = n % ne, r = n / ne;
  const size_t i = r % es, k = r / es;
  E[n] = L[i * s0 + k * s1 + e * s2];
}

template <class T>
__global__ __launch_bounds__(kThreads) void strided_scatter_kernel(size_t e_size, int ne, int es, long long s0,
                                                                   long long s1, long long s2,
                                                                   const T* __restrict__ E, T* __restrict__ L) {
  const size_t n = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (n >= e_size) return;
  const size_t e = n % ne, r = n / ne;
  const size_t i = r % es, k = r / es;
  L[i * s0 + k * s1 + e * s2] += E[n];
}

// Transpose map on the device, entry for entry the host build_transpose_map's:
// per-node counts (atomics), the used nodes and the row starts (exclusive
// scans), and the entries from a stable radix sort of the (offset, t) pairs,
// which keeps t ascending within each row for any row length.
__global__ __launch_bounds__(kThreads) void tmap_count_kernel(size_t nt, const int* __restrict__ off,
                                                             int* __restrict__ count) {
  const size_t t = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (t < nt) atomicAdd(&count[off[t]], 1);
}

__global__ __launch_bounds__(kThreads) void tmap_used_kernel(int l_size, const int* __restrict__ count,
                                                            int* __restrict__ used) {
  const int l = blockIdx.x * kThreads + threadIdx.x;
  if (l < l_size) used[l] = count[l] > 0 ? 1 : 0;
}

// Row m = row_of[l] of every used node l: its node id and its start.
__global__ __launch_bounds__(kThreads) void tmap_rows_kernel(int l_size, const int* __restrict__ count,
                                                            const int* __restrict__ row_of,
                                                            const int* __restrict__ start, int* __restrict__ l_vec,
                                                            int* __restrict__ t_offsets) {
  const int l = blockIdx.x * kThreads + threadIdx.x;
  if (l >= l_size || count[l] == 0) return;
  const int m = row_of[l];
  if (l_vec) l_vec[m] = l;
  t_offsets[m] = start[l];
}

__global__ __launch_bounds__(kThreads) void iota_kernel(size_t n, int* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) out[i] = static_cast<int>(i);
}

// Exclusive prefix sum of in[0..n) into out[0..n] (out[n] = the total).
void exclusive_scan(const int* in, int* out, int n, hipStream_t s) {
  size_t bytes = 0;
  HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out + 1, n, s));
  StreamTmp<unsigned char> tmp(bytes, s);
  zero_fill(out, sizeof(int), s);
  HIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp.data(), bytes, in, out + 1, n, s));
}

// Builds r's transpose map from r.offsets (on the device) on stream s.
// known_rows >= 0: the caller knows every L-node is used (rows = l_size), so
// the row count is not read back; with sync = false nothing waits for s.
void build_transpose_map_device(Restriction& r, hipStream_t s, int known_rows = -1, bool sync = true) {
  const size_t nt = static_cast<size_t>(r.num_elem) * r.elem_size;
  const int L = r.l_size;
  StreamTmp<int> count(static_cast<size_t>(L), s), used(static_cast<size_t>(L), s),
      row_of(static_cast<size_t>(L) + 1, s), start(static_cast<size_t>(L) + 1, s);
  zero_fill(count.data(), static_cast<size_t>(L) * sizeof(int), s);
  tmap_count_kernel<<<grid_for(nt, kThreads), kThreads, 0, s>>>(nt, r.offsets.data(), count.data());
  tmap_used_kernel<<<grid_for(L, kThreads), kThreads, 0, s>>>(L, count.data(), used.data());
  HIP_CHECK_LAUNCH();
  exclusive_scan(used.data(), row_of.data(), L, s);
  exclusive_scan(count.data(), start.data(), L, s);
  int rows = known_rows;
  if (rows < 0) {
    HIP_CHECK(hipMemcpyAsync(&rows, row_of.data() + L, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  r.num_nodes = rows;
  r.identity_nodes = (rows == L);
  if (!r.identity_nodes) r.l_vec_indices.allocate(static_cast<size_t>(rows));
  r.t_offsets.allocate(static_cast<size_t>(rows) + 1);
  r.t_indices.allocate(nt);
  tmap_rows_kernel<<<grid_for(L, kThreads), kThreads, 0, s>>>(
      L, count.data(), row_of.data(), start.data(), r.identity_nodes ? nullptr : r.l_vec_indices.data(),
      r.t_offsets.data());
  HIP_CHECK_LAUNCH();
  HIP_CHECK(hipMemcpyAsync(r.t_offsets.data() + rows, start.data() + L, sizeof(int), hipMemcpyDeviceToDevice, s));
  {
    // Entries: t sorted by offset, stably (t ascending within a row).
    StreamTmp<int> t_in(nt, s), keys_out(nt, s);
    iota_kernel<<<grid_for(nt, kThreads), kThreads, 0, s>>>(nt, t_in.data());
    HIP_CHECK_LAUNCH();
    int bits = 1;
    while (bits < 31 && (1LL << bits) <= static_cast<long long>(r.max_offset)) ++bits;
    size_t bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, r.offsets.data(), keys_out.data(), t_in.data(),
                                                 r.t_indices.data(), static_cast<int>(nt), 0, bits, s));
    StreamTmp<unsigned char> tmp(bytes, s);
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.data(), bytes, r.offsets.data(), keys_out.data(), t_in.data(),
                                                 r.t_indices.data(), static_cast<int>(nt), 0, bits, s));
  }
  if (sync) HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

bool Restriction::same_layout(const Restriction& o) const {
  return type == o.type && num_elem == o.num_elem && elem_size == o.elem_size && num_comp == o.num_comp &&
         comp_stride == o.comp_stride && l_size == o.l_size && strides == o.strides &&
         (type != 0 || host_offsets() == o.host_offsets());
}

const std::vector<int>& Restriction::host_offsets() const {
  std::lock_guard<std::mutex> lk(h_mu);
  if (!h_offsets_valid) {
    h_offsets = offsets.download();
    h_offsets_valid = true;
  }
  return h_offsets;
}

void build_transpose_map(int num_elem, int elem_size, int l_size, const int* offsets, TransposeMapHost& map) {
  const size_t nt = static_cast<size_t>(num_elem) * elem_size;
  // Count per L-node (a counting sort keyed on the offset; stable in t).
  std::vector<int> count(static_cast<size_t>(l_size) + 1, 0);
  for (size_t t = 0; t < nt; ++t) ++count[offsets[t]];
  std::vector<int> row(l_size, -1);
  map.l_vec_indices.clear();
  map.t_offsets.assign(1, 0);
  for (int l = 0; l < l_size; ++l) {
    if (count[l] == 0) continue;
    row[l] = static_cast<int>(map.l_vec_indices.size());
    map.l_vec_indices.push_back(l);
    map.t_offsets.push_back(map.t_offsets.back() + count[l]);
  }
  map.t_indices.assign(nt, 0);
  std::vector<int> cursor(map.t_offsets.begin(), map.t_offsets.end() - 1);
  for (size_t t = 0; t < nt; ++t) map.t_indices[cursor[row[offsets[t]]]++] = static_cast<int>(t);
}

Restriction* restriction_create_offset(int num_elem, int elem_size, int num_comp, int comp_stride, int l_size,
                                       const int* offsets, int scatter_mode, hipStream_t s) {
  // Validation of src/elem-restriction.cpp:102-138.
  require(num_elem > 0, "elem_restriction_create: num_elem must be positive");
  require(elem_size > 0, "elem_restriction_create: elem_size must be at least 1");
  require(num_comp > 0, "elem_restriction_create: num_comp must be at least 1");
  require(!(num_comp > 1 && comp_stride <= 0),
          "elem_restriction_create: comp_stride must be at least 1 when num_comp > 1");
  require(l_size > 0, "elem_restriction_create: l_size must be positive");
  require(offsets != nullptr, "elem_restriction_create: offsets size must be num_elem * elem_size");
  require(scatter_mode == kScatterAtomic || scatter_mode == kScatterDeterministic,
          "elem_restriction_create: unknown scatter mode");
  const size_t nt = static_cast<size_t>(num_elem) * elem_size;
  require(nt < (1ull << 31), "elem_restriction_create: num_elem * elem_size must fit in int32");
  const long long top_comp = static_cast<long long>(num_comp - 1) * comp_stride;
  int lo = 0, hi = 0;
  if (nt) lo = hi = offsets[0];
  for (size_t t = 0; t < nt; ++t) {
    lo = std::min(lo, offsets[t]);
    hi = std::max(hi, offsets[t]);
  }
  require(lo >= 0 && hi + top_comp < l_size,
          "elem_restriction_create: offsets (including component offsets) must be in the range [0, l_size - 1]");
  auto* r = new Restriction();
  try {
    r->type = 0;
    r->num_elem = num_elem;
    r->elem_size = elem_size;
    r->num_comp = num_comp;
    r->comp_stride = comp_stride;
    r->l_size = l_size;
    r->e_size = static_cast<size_t>(num_comp) * nt;
    r->scatter_mode = scatter_mode;
    r->max_offset = hi;
    r->h_offsets.assign(offsets, offsets + nt);
    r->h_offsets_valid = true;
    r->offsets.upload(r->h_offsets);
    if (scatter_mode == kScatterDeterministic) build_transpose_map_device(*r, s);
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}

Restriction* restriction_create_box(int nx, int ny, int nz, int P, int num_comp, int scatter_mode, hipStream_t s,
                                    bool sync) {
  require(nx > 0 && ny > 0 && nz > 0 && P >= 2 && num_comp > 0, "elem_restriction_create: bad box");
  require(scatter_mode == kScatterAtomic || scatter_mode == kScatterDeterministic,
          "elem_restriction_create: unknown scatter mode");
  const long long N = (static_cast<long long>(nx) * (P - 1) + 1) * (static_cast<long long>(ny) * (P - 1) + 1) *
                      (static_cast<long long>(nz) * (P - 1) + 1);
  const long long ne = static_cast<long long>(nx) * ny * nz, nt = ne * P * P * P;
  require(N * num_comp < (1ll << 31) && nt < (1ll << 31), "elem_restriction_create: box too large for int32");
  auto* r = new Restriction();
  try {
    r->type = 0;
    r->num_elem = static_cast<int>(ne);
    r->elem_size = P * P * P;
    r->num_comp = num_comp;
    r->comp_stride = static_cast<int>(N);
    r->l_size = static_cast<int>(N * num_comp);
    r->e_size = static_cast<size_t>(num_comp) * nt;
    r->scatter_mode = scatter_mode;
    r->max_offset = static_cast<int>(N - 1);
    r->box_n[0] = nx;
    r->box_n[1] = ny;
    r->box_n[2] = nz;
    r->offsets.allocate(static_cast<size_t>(nt));
    box_offsets_device(nx, ny, nz, P, 1, r->offsets.data(), s);
    if (scatter_mode == kScatterDeterministic) {
      build_transpose_map_device(*r, s, static_cast<int>(N), sync);  // every node of a box is used
    } else if (sync) {
      HIP_CHECK(hipStreamSynchronize(s));
    }
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}

Restriction* restriction_create_strided(int num_elem, int elem_size, int num_comp, int l_size, const int* strides) {
  require(num_elem > 0, "elem_restriction_create_strided: num_elem must be positive");
  require(elem_size > 0, "elem_restriction_create_strided: elem_size must be at least 1");
  require(num_comp > 0, "elem_restriction_create_strided: num_comp must be at least 1");
  require(l_size > 0, "elem_restriction_create_strided: l_size must be positive");
  std::array<int, 3> s;
  const bool backend = (strides == nullptr);
  if (backend) {
    require(static_cast<long long>(elem_size) * num_elem < (1ll << 31),
            "elem_restriction_create_strided: backend strides exceed int32");
    s = {num_elem, elem_size * num_elem, 1};
  } else {
    s = {strides[0], strides[1], strides[2]};
  }
  for (int v : s) require(v >= 0, "elem_restriction_create_strided: strides must be non-negative");
  const long long top = static_cast<long long>(elem_size - 1) * s[0] + static_cast<long long>(num_comp - 1) * s[1] +
                        static_cast<long long>(num_elem - 1) * s[2];
  require(top < l_size, "elem_restriction_create_strided: strided indices must be in the range [0, l_size - 1]");
  auto* r = new Restriction();
  r->type = 1;
  r->num_elem = num_elem;
  r->elem_size = elem_size;
  r->num_comp = num_comp;
  r->comp_stride = 0;
  r->l_size = l_size;
  r->e_size = static_cast<size_t>(num_comp) * elem_size * num_elem;
  r->strides = s;
  r->backend_strides = backend;
  return r;
}

Restriction* restriction_clone(const Restriction& src) {
  auto* r = new Restriction();
  try {
    r->type = src.type;
    r->num_elem = src.num_elem;
    r->elem_size = src.elem_size;
    r->num_comp = src.num_comp;
    r->comp_stride = src.comp_stride;
    r->l_size = src.l_size;
    r->e_size = src.e_size;
    r->strides = src.strides;
    r->backend_strides = src.backend_strides;
    r->scatter_mode = src.scatter_mode;
    r->h_offsets = src.host_offsets();
    r->h_offsets_valid = true;
    r->num_nodes = src.num_nodes;
    r->identity_nodes = src.identity_nodes;
    r->max_offset = src.max_offset;
    for (int d = 0; d < 3; ++d) r->box_n[d] = src.box_n[d];
    auto dup = [](const DevBuf<int>& from, DevBuf<int>& to) {
      to.allocate(from.size());
      SharedCapture lk;
      if (from.size()) HIP_CHECK(hipMemcpy(to.data(), from.data(), from.size() * sizeof(int), hipMemcpyDeviceToDevice));
      if (from.size()) HIP_CHECK(hipStreamSynchronize(nullptr));  // the copy is done, not only enqueued
    };
    dup(src.offsets, r->offsets);
    dup(src.l_vec_indices, r->l_vec_indices);
    dup(src.t_offsets, r->t_offsets);
    dup(src.t_indices, r->t_indices);
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}

RestrictionView restriction_view(const Restriction& r) {
  RestrictionView v;
  v.type = r.type;
  v.num_elem = r.num_elem;
  v.elem_size = r.elem_size;
  v.num_comp = r.num_comp;
  v.comp_stride = r.comp_stride;
  v.l_size = r.l_size;
  v.e_size = r.e_size;
  v.strides = r.strides;
  v.scatter_mode = r.scatter_mode;
  v.offsets = r.offsets.data();
  v.num_nodes = r.num_nodes;
  v.l_vec_indices = r.identity_nodes ? nullptr : r.l_vec_indices.data();
  v.t_offsets = r.t_offsets.data();
  v.t_indices = r.t_indices.data();
  return v;
}

template <class T>
void restriction_apply(const Restriction& r, int t_mode, const T* in, size_t in_len, T* out, size_t out_len,
                       hipStream_t s) {
  restriction_apply_view<T>(restriction_view(r), t_mode, in, in_len, out, out_len, s);
}

template <class T>
void restriction_apply_view(const RestrictionView& r, int t_mode, const T* in, size_t in_len, T* out, size_t out_len,
                            hipStream_t s) {
  require(t_mode == kNoTranspose || t_mode == kTranspose, "elem_restriction_apply: unknown transpose mode");
  const bool tr = (t_mode == kTranspose);
  require(in_len == (tr ? r.e_size : static_cast<size_t>(r.l_size)), "elem_restriction_apply: in has the wrong size");
  if (tr) {
    require(out_len == static_cast<size_t>(r.l_size),
            "elem_restriction_apply: Transpose adds into out, which must already have l_size entries");
  } else {
    require(out_len == r.e_size, "elem_restriction_apply: out must have e_size entries");
  }
  if (r.e_size == 0) return;
  require(in != nullptr && out != nullptr, "elem_restriction_apply: null device pointer");
  if (r.type == 1) {
    const auto& st = r.strides;
    if (!tr) {
      strided_gather_kernel<T><<<grid_for(r.e_size, kThreads), kThreads, 0, s>>>(r.e_size, r.num_elem, r.elem_size,
                                                                                 st[0], st[1], st[2], in, out);
    } else {
      strided_scatter_kernel<T><<<grid_for(r.e_size, kThreads), kThreads, 0, s>>>(r.e_size, r.num_elem, r.elem_size,
                                                                                  st[0], st[1], st[2], in, out);
    }
  } else if (!tr) {
    dim3 grid((r.num_elem + kTileE - 1) / kTileE, (r.elem_size + kTileI - 1) / kTileI);
    require(r.offsets != nullptr, "elem_restriction_apply: offsets are not on the device");
    offset_gather_kernel<T><<<grid, kThreads, 0, s>>>(r.num_elem, r.elem_size, r.num_comp, r.comp_stride, r.offsets,
                                                      in, out);
  } else if (r.scatter_mode == kScatterAtomic) {
    dim3 grid((r.num_elem + kTileE - 1) / kTileE, (r.elem_size + kTileI - 1) / kTileI);
    require(r.offsets != nullptr, "elem_restriction_apply: offsets are not on the device");
    offset_scatter_atomic_kernel<T><<<grid, kThreads, 0, s>>>(r.num_elem, r.elem_size, r.num_comp, r.comp_stride,
                                                              r.offsets, in, out);
  } else {
    require(r.t_offsets != nullptr && r.t_indices != nullptr,
            "elem_restriction_apply: the deterministic scatter needs the transpose map on the device");
    if (r.num_nodes == 0) return;
    offset_scatter_det_kernel<T><<<grid_for(r.num_nodes, kThreads), kThreads, 0, s>>>(
        r.num_nodes, r.num_elem, r.elem_size, r.num_comp, r.comp_stride, r.l_vec_indices, r.t_offsets, r.t_indices,
        in, out);
  }
  HIP_CHECK_LAUNCH();
}

template <class T>
void csr_scatter(const Restriction& r, const T* Y, T* L, bool zero_start, hipStream_t s) {
  require(r.type == 0 && r.scatter_mode == kScatterDeterministic, "csr_scatter: needs a deterministic offset restriction");
  if (zero_start && !r.identity_nodes) zero_fill(L, static_cast<size_t>(r.l_size) * sizeof(T), s);
  const size_t comp_len = static_cast<size_t>(r.num_elem) * r.elem_size;
  csr_scatter_kernel<T><<<grid_for(r.num_nodes, kThreads), kThreads, 0, s>>>(
      r.num_nodes, comp_len, r.num_comp, r.comp_stride, r.identity_nodes ? nullptr : r.l_vec_indices.data(),
      r.t_offsets.data(), r.t_indices.data(), Y, L, zero_start ? 1 : 0);
  HIP_CHECK_LAUNCH();
}

template void restriction_apply<double>(const Restriction&, int, const double*, size_t, double*, size_t, hipStream_t);
template void restriction_apply<float>(const Restriction&, int, const float*, size_t, float*, size_t, hipStream_t);
template void restriction_apply_view<double>(const RestrictionView&, int, const double*, size_t, double*, size_t,
                                             hipStream_t);
template void csr_scatter<double>(const Restriction&, const double*, double*, bool, hipStream_t);
template void csr_scatter<float>(const Restriction&, const float*, float*, bool, hipStream_t);

}  // namespace femhip
