// Host side of the fused mass operator: registry of compiled shapes, qdata
// re-layout, dispatch, and the Jacobi diagonal.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "block_plan.hpp"
#include "fused_mass.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

// Reference backend-strided qdata[q*E + e] -> private per-element tiles
// [e][q]: the fused kernel's z-line threads read qdata[e][qz*Q^2 + t], so lanes
// of one load instruction read consecutive words.
template <class T>
__global__ __launch_bounds__(kThreads) void qdata_private_kernel(int E, int Q, const double* __restrict__ ref,
                                                                 T* __restrict__ priv) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  const size_t Q3 = static_cast<size_t>(Q) * Q * Q;
  if (idx >= Q3 * E) return;
  // idx walks the private layout (coalesced writes); reads stride E.
  const size_t e = idx / Q3, q = idx % Q3;
  priv[idx] = static_cast<T>(ref[q * E + e]);
}

template <class T>
__global__ __launch_bounds__(kThreads) void to_f32_kernel(size_t n, const double* __restrict__ a, T* __restrict__ b) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) b[i] = static_cast<T>(a[i]);
}

// Element diagonal of the mass matrix, one thread per (element, node):
// d[e][i] = sum_q qdata[e][q] * B2[qz][iz] * B2[qy][iy] * B2[qx][ix], added into
// diag[offsets[e][i] + c*comp_stride] for every component.
__global__ __launch_bounds__(kThreads) void mass_diag_kernel(int E, int P, int Q, int NC, int cs,
                                                             const int* __restrict__ offsets,
                                                             const double* __restrict__ qpriv,
                                                             const double* __restrict__ B2,
                                                             double* __restrict__ diag) {
  const int P3 = P * P * P, Q2 = Q * Q;
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<size_t>(E) * P3) return;
  const size_t e = idx / P3;
  const int i = static_cast<int>(idx % P3);
  const int ix = i % P, iy = (i / P) % P, iz = i / (P * P);
  const double* qe = qpriv + e * Q2 * Q;  // [qz][qy][qx]
  double sum = 0.0;
  for (int qy = 0; qy < Q; ++qy) {
    for (int qx = 0; qx < Q; ++qx) {
      const double bxy = B2[qy * P + iy] * B2[qx * P + ix];
      double s = 0.0;
      for (int qz = 0; qz < Q; ++qz) s += B2[qz * P + iz] * qe[(qz * Q + qy) * Q + qx];
      sum += bxy * s;
    }
  }
  const size_t o = static_cast<size_t>(offsets[idx]);
  for (int c = 0; c < NC; ++c) unsafeAtomicAdd(&diag[o + static_cast<size_t>(c) * cs], sum);
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

void register_fused(const FusedKey& k, FusedLauncher f) { fused_registry()[k] = f; }

std::map<FusedKey, FusedBlockLauncher>& fused_block_registry() {
  static std::map<FusedKey, FusedBlockLauncher> reg;
  return reg;
}

void register_fused_block(const FusedKey& k, FusedBlockLauncher f) { fused_block_registry()[k] = f; }

// The block-accumulated launcher for this operator's shape and precision, or null.
static FusedBlockLauncher find_block(const FusedMass& f, int fp32, int* nb) {
  for (const auto& kv : fused_block_registry()) {
    const FusedKey& k = kv.first;
    if (k.P == f.P && k.Q == f.Q && k.NC == f.NC && k.fp32 == fp32) {
      *nb = k.elems;
      return kv.second;
    }
  }
  return nullptr;
}

// Builds the block plan on first use (host clustering + upload); false if the shape has no block kernel.
static bool ensure_block_plan(FusedMass& f, hipStream_t s) {
  if (f.block_tried) return f.block_nb > 0;
  f.block_tried = true;
  int nb = 0;
  if (!find_block(f, 0, &nb)) return false;
  int nb32 = 0;
  if (!find_block(f, 1, &nb32) || nb32 != nb) return false;
  const int P3 = f.P * f.P * f.P, Q3 = f.Q * f.Q * f.Q;
  BlockPlan plan = build_block_plan(f.rstr->h_offsets.data(), f.num_elem, P3, f.l_size, nb);
  f.b_node_ptr.upload(plan.node_ptr);
  f.b_nodes.upload(plan.nodes);
  f.b_ninterior.upload(plan.ninterior);
  f.b_nelem.upload(plan.nelem);
  f.b_loc.upload(plan.loc);
  DevBuf<int> elem;
  elem.upload(plan.elem);
  const size_t n = static_cast<size_t>(plan.nblk) * nb * Q3;
  f.b_q64.allocate(n);
  qdata_block_kernel<double><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, Q3, elem.data(), f.qdata64.data(),
                                                                       f.b_q64.data());
  HIP_CHECK_LAUNCH();
  f.b_q32.allocate(n);
  qdata_block_kernel<float><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, Q3, elem.data(), f.qdata64.data(),
                                                                      f.b_q32.data());
  HIP_CHECK_LAUNCH();
  HIP_CHECK(hipStreamSynchronize(s));  // `elem` is freed on return
  f.nblk = plan.nblk;
  f.block_nb = nb;
  return true;
}

template <class T>
static void launch_block(FusedMass& f, const T* u, T* v, hipStream_t s) {
  const int fp32 = std::is_same<T, float>::value ? 1 : 0;
  int nb = 0;
  FusedBlockLauncher L = find_block(f, fp32, &nb);
  FusedBlockArgs a{};
  a.nblk = f.nblk;
  a.comp_stride = f.comp_stride;
  a.node_ptr = f.b_node_ptr.data();
  a.nodes = f.b_nodes.data();
  a.ninterior = f.b_ninterior.data();
  a.nelem = f.b_nelem.data();
  a.loc = f.b_loc.data();
  a.qdata = fp32 ? static_cast<const void*>(f.b_q32.data()) : static_cast<const void*>(f.b_q64.data());
  a.u = u;
  a.v = v;
  a.basis = fp32 ? static_cast<const void*>(f.B32.data()) : static_cast<const void*>(f.B64.data());
  L(a, s);
}

void fused_prepare(FusedMass& f, hipStream_t s) {
  if (f.scatter == kScatterAtomic && f.elems == 0) ensure_block_plan(f, s);
}

std::vector<int> fused_elem_variants(int P, int Q, int NC, int fp32, int det, int exact) {
  std::vector<int> out;
  for (const auto& kv : fused_registry()) {
    const FusedKey& k = kv.first;
    if (k.P == P && k.Q == Q && k.NC == NC && k.fp32 == fp32 && k.det == det && k.exact == exact)
      out.push_back(k.elems);
  }
  return out;
}

int fused_default_elems(int P, int Q, int NC, int fp32) {
  // The first registered variant is the shape's default (lowest elems that was
  // compiled as default); the tuner may pick another.
  auto v = fused_elem_variants(P, Q, NC, fp32, 0, 0);
  if (v.empty()) return 0;
  // Prefer the variant registered as the default packing (largest multiple-of-64 fit).
  int best = v.front();
  double best_waste = 1.0;
  const int tpe = std::max({P * P, P * Q, Q * Q});
  for (int ne : v) {
    const int thr = ne * tpe, padded = (thr + 63) / 64 * 64;
    const double waste = 1.0 - static_cast<double>(thr) / padded;
    if (waste < best_waste - 1e-9 || (waste < best_waste + 1e-9 && ne < best)) {
      best_waste = waste;
      best = ne;
    }
  }
  return best;
}

bool fused_mass_supported(int P, int Q, int NC) { return !fused_elem_variants(P, Q, NC, 0, 0, 0).empty(); }

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
  {
    // Single-contributor nodes: the fast atomic kernel stores them instead of adding.
    const size_t nt = r.h_offsets.size();
    std::vector<int> mult(static_cast<size_t>(r.l_size), 0);
    for (size_t t = 0; t < nt; ++t) ++mult[r.h_offsets[t]];
    std::vector<int> flagged(r.h_offsets);
    for (size_t t = 0; t < nt; ++t)
      if (mult[flagged[t]] == 1) flagged[t] = static_cast<int>(static_cast<unsigned>(flagged[t]) | 0x80000000u);
    f->flagged_offsets.upload(flagged);
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

static FusedMass::StageEvents* next_events(FusedMass& f) {
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

static FusedKey pick(const FusedMass& f, int fp32) {
  FusedKey k{f.P, f.Q, f.NC, fp32, f.scatter == kScatterDeterministic ? 1 : 0, fp32 ? 0 : f.exact, 0};
  k.elems = f.elems > 0 ? f.elems : fused_default_elems(f.P, f.Q, f.NC, fp32);
  if (!find(k)) {
    auto v = fused_elem_variants(f.P, f.Q, f.NC, fp32, k.det, k.exact);
    require(!v.empty(), "fused mass: variant not compiled");
    k.elems = v.front();
  }
  return k;
}

void fused_mass_apply64(FusedMass& f, const double* u, double* v, hipStream_t s) {
  const FusedKey k = pick(f, 0);
  FusedArgs a{};
  a.num_elem = f.num_elem;
  a.comp_stride = f.comp_stride;
  a.offsets = f.rstr->offsets.data();
  a.flagged_offsets = f.flagged_offsets.data();
  a.qdata = f.qdata64.data();
  a.u = u;
  a.basis = f.B64.data();
  a.y_comp_len = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
  FusedMass::StageEvents* ev = next_events(f);
  if (ev) HIP_CHECK(hipEventRecord(ev->a, s));
  if (k.det) {
    if (f.Y64.size() != a.y_comp_len * f.NC) f.Y64.allocate(a.y_comp_len * f.NC);
    a.out = f.Y64.data();
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    find(k)(a, s);
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
    csr_scatter<double>(*f.rstr, f.Y64.data(), v, true, s);
  } else {
    zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(double), s);
    a.out = v;
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    if (f.elems == 0 && ensure_block_plan(f, s)) {
      launch_block<double>(f, u, v, s);
    } else {
      find(k)(a, s);
    }
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
  }
  if (ev) HIP_CHECK(hipEventRecord(ev->d, s));
}

void fused_mass_apply32(FusedMass& f, const float* u, float* v, hipStream_t s) {
  if (f.qdata32.empty()) {
    f.qdata32.allocate(f.qdata64.size());
    to_f32_kernel<float><<<grid_for(f.qdata64.size(), kThreads), kThreads, 0, s>>>(f.qdata64.size(), f.qdata64.data(),
                                                                                  f.qdata32.data());
    HIP_CHECK_LAUNCH();
  }
  const FusedKey k = pick(f, 1);
  FusedArgs a{};
  a.num_elem = f.num_elem;
  a.comp_stride = f.comp_stride;
  a.offsets = f.rstr->offsets.data();
  a.flagged_offsets = f.flagged_offsets.data();
  a.qdata = f.qdata32.data();
  a.u = u;
  a.basis = f.B32.data();
  a.y_comp_len = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
  FusedMass::StageEvents* ev = next_events(f);
  if (ev) HIP_CHECK(hipEventRecord(ev->a, s));
  if (k.det) {
    if (f.Y32.size() != a.y_comp_len * f.NC) f.Y32.allocate(a.y_comp_len * f.NC);
    a.out = f.Y32.data();
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    find(k)(a, s);
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
    csr_scatter<float>(*f.rstr, f.Y32.data(), v, true, s);
  } else {
    zero_fill(v, static_cast<size_t>(f.l_size) * sizeof(float), s);
    a.out = v;
    if (ev) HIP_CHECK(hipEventRecord(ev->b, s));
    if (f.elems == 0 && ensure_block_plan(f, s)) {
      launch_block<float>(f, u, v, s);
    } else {
      find(k)(a, s);
    }
    if (ev) HIP_CHECK(hipEventRecord(ev->c, s));
  }
  if (ev) HIP_CHECK(hipEventRecord(ev->d, s));
}

void fused_mass_diagonal(FusedMass& f, double* diag, hipStream_t s) {
  zero_fill(diag, static_cast<size_t>(f.l_size) * sizeof(double), s);
  StreamTmp<double> B2(f.Bsq.size(), s);
  HIP_CHECK(hipMemcpyAsync(B2.data(), f.Bsq.data(), f.Bsq.size() * sizeof(double), hipMemcpyHostToDevice, s));
  const size_t n = static_cast<size_t>(f.num_elem) * f.P * f.P * f.P;
  mass_diag_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(f.num_elem, f.P, f.Q, f.NC, f.comp_stride,
                                                               f.rstr->offsets.data(), f.qdata64.data(), B2.data(),
                                                               diag);
  HIP_CHECK_LAUNCH();
  // The host copy of Bsq must outlive the async copy.
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace femhip
