// fem::hip -- the C++ drop-in over libfemhip's C ABI (include/fem_hip.h).
//
// Mirrors the reference's device API (include/cuda/cuda-check.hpp,
// include/cuda/device-array.hpp, include/cuda/device-info.hpp,
// include/cuda/cuda-elem-restriction.hpp) with the same names, argument meaning
// and exceptions, and extends it to the rest of the operator path: device
// bases, QFunctions and operator_apply built from the reference's host types.
// Header-only; include it after the reference's own headers are on the include
// path (basis.hpp, elem-restriction.hpp, operator.hpp, qfunction.hpp,
// qfunctions/mass.hpp) and link libfemhip.so.
#pragma once

#include <array>
#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "basis.hpp"
#include "elem-restriction.hpp"
#include "fem_hip.h"
#include "operator.hpp"
#include "qfunction.hpp"
#include "qfunctions/mass.hpp"

namespace fem::hip {

// include/cuda/cuda-check.hpp:17-22: a failed call throws with "file:line: expr
// failed: ..."; invalid arguments keep the reference's std::invalid_argument.
inline void check(int rc, const char* expr, const char* file, int line) {
  if (rc == FEM_HIP_SUCCESS) return;
  const std::string msg = std::string(file) + ":" + std::to_string(line) + ": " + expr + " failed: " +
                          fem_hip_last_error();
  if (rc == FEM_HIP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

}  // namespace fem::hip

#define FEM_HIP_CHECK(expr) ::fem::hip::check((expr), #expr, __FILE__, __LINE__)

namespace fem::hip {

// include/cuda/device-array.hpp:18-95: move-only owner of one device allocation,
// explicit synchronous host <-> device copies.
template <class T>
class DeviceArray {
  static_assert(std::is_trivially_copyable_v<T>, "DeviceArray elements are copied as raw bytes");

 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { allocate(n); }
  explicit DeviceArray(const std::vector<T>& host) : DeviceArray(host.size()) { copy_from_host(host); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : data_(o.data_), size_(o.size_) {
    o.data_ = nullptr;
    o.size_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      data_ = o.data_;
      size_ = o.size_;
      o.data_ = nullptr;
      o.size_ = 0;
    }
    return *this;
  }
  T* data() { return data_; }
  const T* data() const { return data_; }
  std::size_t size() const { return size_; }
  std::size_t bytes() const { return size_ * sizeof(T); }
  bool empty() const { return size_ == 0; }
  void resize(std::size_t n) {
    if (n == size_) return;
    release();
    allocate(n);
  }
  void zero() {
    if (size_) FEM_HIP_CHECK(fem_hip_memset_zero(data_, bytes(), nullptr));
  }
  void copy_from_host(const std::vector<T>& host) {
    if (host.size() != size_) throw std::invalid_argument("DeviceArray::copy_from_host: size mismatch");
    if (size_) FEM_HIP_CHECK(fem_hip_memcpy_htod(data_, host.data(), bytes()));
  }
  void copy_to_host(std::vector<T>& host) const {
    host.resize(size_);
    if (size_) FEM_HIP_CHECK(fem_hip_memcpy_dtoh(host.data(), data_, bytes()));
  }
  std::vector<T> to_host() const {
    std::vector<T> h;
    copy_to_host(h);
    return h;
  }

 private:
  void allocate(std::size_t n) {
    if (n == 0) return;
    void* p = nullptr;
    FEM_HIP_CHECK(fem_hip_malloc(&p, n * sizeof(T)));
    data_ = static_cast<T*>(p);
    size_ = n;
  }
  void release() noexcept {
    if (data_) (void)fem_hip_free(data_);
    data_ = nullptr;
    size_ = 0;
  }
  T* data_ = nullptr;
  std::size_t size_ = 0;
};

// include/cuda/device-info.hpp:8-25 (wavefront size 64, LDS in place of shared memory).
struct DeviceInfo {
  int id = 0;
  std::string name, arch;
  int cu_count = 0, wavefront_size = 0, max_threads_per_block = 0, regs_per_block = 0;
  std::size_t lds_per_block = 0, lds_per_cu = 0, l2_bytes = 0, total_global_mem = 0, free_global_mem = 0;
};

inline int device_count() {
  int n = 0;
  FEM_HIP_CHECK(fem_hip_device_count(&n));
  return n;
}

inline DeviceInfo device_info(int device = 0) {
  fem_hip_device_info c;
  FEM_HIP_CHECK(fem_hip_device_info_get(device, &c));
  DeviceInfo d;
  d.id = c.id;
  d.name = c.name;
  d.arch = c.arch;
  d.cu_count = c.cu_count;
  d.wavefront_size = c.wavefront_size;
  d.max_threads_per_block = c.max_threads_per_block;
  d.regs_per_block = c.regs_per_block;
  d.lds_per_block = c.lds_per_block;
  d.lds_per_cu = c.lds_per_cu;
  d.l2_bytes = c.l2_bytes;
  d.total_global_mem = c.total_global_mem;
  d.free_global_mem = c.free_global_mem;
  return d;
}

// include/cuda/cuda-elem-restriction.hpp:21-40
enum class ScatterMode { Atomic, Deterministic };

struct TransposeMap {
  std::vector<int> l_vec_indices;
  std::vector<int> t_offsets;
  std::vector<int> t_indices;
};

inline void build_transpose_map(const fem::ElemRestriction& r, TransposeMap& map) {
  if (r.type != fem::RestrictionType::Offset) throw std::invalid_argument("build_transpose_map: only for Offset restrictions");
  std::vector<int> lv(r.l_size), to(static_cast<std::size_t>(r.l_size) + 1), ti(r.offsets.size());
  int rows = 0;
  FEM_HIP_CHECK(fem_hip_transpose_map_build(r.num_elem, r.elem_size, r.l_size, r.offsets.data(), &rows, lv.data(),
                                            to.data(), ti.data()));
  lv.resize(rows);
  to.resize(static_cast<std::size_t>(rows) + 1);
  map.l_vec_indices = std::move(lv);
  map.t_offsets = std::move(to);
  map.t_indices = std::move(ti);
}

// include/cuda/cuda-elem-restriction.hpp:46-58: the sizes, and the index arrays
// in device memory owned by the struct itself (move-only through DeviceArray).
struct ElemRestriction {
  RestrictionType type = RestrictionType::Offset;
  int num_elem = 0, elem_size = 0, num_comp = 0, comp_stride = 0;
  int l_size = 0, e_size = 0;
  std::array<int, 3> strides{};  // Strided only
  ScatterMode scatter_mode = ScatterMode::Atomic;

  DeviceArray<int> offsets;  // Offset only: copy of the host offsets, [num_elem][elem_size]

  // Offset + Deterministic only (empty otherwise): the TransposeMap on the device
  int num_nodes = 0;  // l_vec_indices.size()
  DeviceArray<int> l_vec_indices, t_offsets, t_indices;
};

namespace detail {
inline int cmode(ContractMode m) { return m == ContractMode::Transpose ? FEM_HIP_TRANSPOSE : FEM_HIP_NO_TRANSPOSE; }
inline int emode(EvalMode m) {
  switch (m) {
    case EvalMode::None: return FEM_HIP_EVAL_NONE;
    case EvalMode::Interp: return FEM_HIP_EVAL_INTERP;
    case EvalMode::Grad: return FEM_HIP_EVAL_GRAD;
    default: return FEM_HIP_EVAL_WEIGHT;
  }
}
inline int smode(ScatterMode m) {
  return m == ScatterMode::Deterministic ? FEM_HIP_SCATTER_DETERMINISTIC : FEM_HIP_SCATTER_ATOMIC;
}
// A library restriction handle (its own device copies) for the operator path.
inline std::shared_ptr<fem_hip_restriction_s> restriction_handle(const fem::ElemRestriction& host,
                                                                 ScatterMode scatter_mode) {
  fem_hip_restriction h = nullptr;
  if (host.type == RestrictionType::Offset) {
    if (host.offsets.size() != static_cast<std::size_t>(host.num_elem) * host.elem_size)
      throw std::invalid_argument("hip::elem_restriction_create: offsets size != num_elem * elem_size");
    FEM_HIP_CHECK(fem_hip_restriction_create(host.num_elem, host.elem_size, host.num_comp, host.comp_stride,
                                             host.l_size, host.offsets.data(), smode(scatter_mode), &h));
  } else {
    FEM_HIP_CHECK(fem_hip_restriction_create_strided(host.num_elem, host.elem_size, host.num_comp, host.l_size,
                                                     host.has_backend_strides ? nullptr : host.strides.data(), &h));
  }
  return std::shared_ptr<fem_hip_restriction_s>(h, [](fem_hip_restriction p) { (void)fem_hip_restriction_destroy(p); });
}
}  // namespace detail

// include/cuda/cuda-elem-restriction.hpp:60-66 (src/cuda/cuda-elem-restriction.cu:188-214):
// copies the host restriction's sizes and index arrays to the device once and,
// for Deterministic, builds the transpose map on the host and copies it;
// `restriction` is replaced only on success. The offsets are also checked to
// lie in [0, l_size) with every component, so a malformed host struct cannot
// make a kernel read out of bounds.
inline void elem_restriction_create(const fem::ElemRestriction& host, ScatterMode scatter_mode,
                                    ElemRestriction& restriction) {
  ElemRestriction r;
  r.type = host.type;
  r.num_elem = host.num_elem;
  r.elem_size = host.elem_size;
  r.num_comp = host.num_comp;
  r.comp_stride = host.comp_stride;
  r.l_size = host.l_size;
  r.e_size = host.e_size;
  r.strides = host.strides;
  r.scatter_mode = scatter_mode;
  if (host.type == RestrictionType::Offset) {
    if (host.offsets.size() != static_cast<std::size_t>(host.num_elem) * host.elem_size)
      throw std::invalid_argument("hip::elem_restriction_create: offsets size != num_elem * elem_size");
    const long long reach = static_cast<long long>(host.num_comp - 1) * host.comp_stride;
    for (int o : host.offsets)
      if (o < 0 || o + reach >= host.l_size)
        throw std::invalid_argument("hip::elem_restriction_create: an offset lies outside [0, l_size)");
    r.offsets = DeviceArray<int>(host.offsets);
    if (scatter_mode == ScatterMode::Deterministic) {
      TransposeMap map;
      build_transpose_map(host, map);
      r.num_nodes = static_cast<int>(map.l_vec_indices.size());
      r.l_vec_indices = DeviceArray<int>(map.l_vec_indices);
      r.t_offsets = DeviceArray<int>(map.t_offsets);
      r.t_indices = DeviceArray<int>(map.t_indices);
    }
  }
  restriction = std::move(r);
}

// include/cuda/cuda-elem-restriction.hpp:68-79: NoTranspose resizes out to e_size;
// Transpose adds into out, which must already have l_size entries. Asynchronous.
inline void elem_restriction_apply(const ElemRestriction& r, ContractMode t_mode, const DeviceArray<double>& in,
                                   DeviceArray<double>& out, void* stream = nullptr) {
  if (t_mode == ContractMode::NoTranspose) {
    if (in.size() != static_cast<std::size_t>(r.l_size))
      throw std::invalid_argument("hip::elem_restriction_apply: in has the wrong size");
    out.resize(r.e_size);
  }
  fem_hip_restriction_arrays a{};
  a.type = r.type == RestrictionType::Offset ? 0 : 1;
  a.num_elem = r.num_elem;
  a.elem_size = r.elem_size;
  a.num_comp = r.num_comp;
  a.comp_stride = r.comp_stride;
  a.l_size = r.l_size;
  a.e_size = r.e_size;
  for (int i = 0; i < 3; ++i) a.strides[i] = r.strides[i];
  a.scatter_mode = detail::smode(r.scatter_mode);
  a.offsets = r.offsets.data();
  a.num_nodes = r.num_nodes;
  a.l_vec_indices = r.l_vec_indices.data();
  a.t_offsets = r.t_offsets.data();
  a.t_indices = r.t_indices.data();
  FEM_HIP_CHECK(fem_hip_restriction_apply_arrays(&a, detail::cmode(t_mode), in.data(), in.size(), out.data(),
                                                 out.size(), stream));
}

// Device tensor basis built from the reference's host TensorBasis.
struct TensorBasis {
  int dim = 0, num_comp = 0, P_1d = 0, Q_1d = 0;
  std::shared_ptr<fem_hip_basis_s> handle;
};

inline void basis_create(const fem::TensorBasis& host, TensorBasis& basis) {
  fem_hip_basis h = nullptr;
  FEM_HIP_CHECK(fem_hip_basis_create_tensor(host.dim, host.num_comp, host.P_1d, host.Q_1d, host.q_ref_1d.data(),
                                            host.q_weight_1d.data(), host.interp_1d.data(), host.grad_1d.data(), &h));
  TensorBasis b;
  b.handle = std::shared_ptr<fem_hip_basis_s>(h, [](fem_hip_basis p) { (void)fem_hip_basis_destroy(p); });
  b.dim = host.dim;
  b.num_comp = host.num_comp;
  b.P_1d = host.P_1d;
  b.Q_1d = host.Q_1d;
  basis = std::move(b);
}

// tensor_basis_apply (include/tensor-contract.hpp:50-53); v must have the output size.
inline void tensor_basis_apply(const TensorBasis& b, int num_elem, ContractMode t_mode, EvalMode eval_mode,
                               const DeviceArray<double>& u, DeviceArray<double>& v, void* stream = nullptr) {
  FEM_HIP_CHECK(fem_hip_basis_apply(b.handle.get(), num_elem, detail::cmode(t_mode), detail::emode(eval_mode),
                                    u.data(), u.size(), v.data(), v.size(), stream));
}

// Device operator built from the reference's host Operator. The QFunction's
// host pointer selects the device implementation: fem::qfunctions::build_mass and
// fem::qfunctions::apply_mass (and vector mass operators built with
// operator_create_mass); anything else throws (no device implementation).
struct Operator {
  int active_in_l_size = 0, active_out_l_size = 0;
  std::shared_ptr<fem_hip_operator_s> handle;
};

inline void operator_create(const fem::Operator& host, Operator& op, ScatterMode scatter_mode = ScatterMode::Atomic) {
  int kind = 0;
  if (host.qf.user == &fem::qfunctions::build_mass) kind = FEM_HIP_QF_BUILD_MASS;
  if (host.qf.user == &fem::qfunctions::apply_mass) kind = FEM_HIP_QF_APPLY_MASS;
  if (!kind) throw std::runtime_error("hip::operator_create: the QFunction has no device implementation");
  fem_hip_qfunction qf = nullptr;
  FEM_HIP_CHECK(fem_hip_qfunction_create(kind, &qf));
  std::unique_ptr<fem_hip_qfunction_s, int (*)(fem_hip_qfunction)> qf_owner(qf, fem_hip_qfunction_destroy);
  for (const auto& f : host.qf.inputs) FEM_HIP_CHECK(fem_hip_qfunction_add_input(qf, f.name.c_str(), f.size, detail::emode(f.eval_mode)));
  for (const auto& f : host.qf.outputs)
    FEM_HIP_CHECK(fem_hip_qfunction_add_output(qf, f.name.c_str(), f.size, detail::emode(f.eval_mode)));
  if (!host.qf.ctx.empty()) FEM_HIP_CHECK(fem_hip_qfunction_set_context(qf, host.qf.ctx.data(), host.qf.ctx.size()));
  fem_hip_operator h = nullptr;
  FEM_HIP_CHECK(fem_hip_operator_create(qf, &h));
  Operator out;
  out.handle = std::shared_ptr<fem_hip_operator_s>(h, [](fem_hip_operator p) { (void)fem_hip_operator_destroy(p); });
  auto set = [&](const std::optional<OperatorField>& fo) {
    if (!fo) return;
    const OperatorField& f = *fo;
    std::shared_ptr<fem_hip_restriction_s> r;
    TensorBasis b;
    if (f.restriction) r = detail::restriction_handle(*f.restriction, scatter_mode);
    if (f.basis) basis_create(*f.basis, b);
    const int vk = f.vector.kind == VectorKind::Active    ? FEM_HIP_VECTOR_ACTIVE
                   : f.vector.kind == VectorKind::Passive ? FEM_HIP_VECTOR_PASSIVE
                                                          : FEM_HIP_VECTOR_NONE;
    FEM_HIP_CHECK(fem_hip_operator_set_field(h, f.name.c_str(), f.restriction ? r.get() : nullptr,
                                             f.basis ? b.handle.get() : nullptr, vk,
                                             f.vector.data.empty() ? nullptr : f.vector.data.data(),
                                             f.vector.data.size(), 0));
  };
  for (const auto& f : host.inputs) set(f);
  for (const auto& f : host.outputs) set(f);
  FEM_HIP_CHECK(fem_hip_operator_check_ready(h));
  long long info[8];
  FEM_HIP_CHECK(fem_hip_operator_get_info(h, info));
  out.active_in_l_size = static_cast<int>(info[2]);
  out.active_out_l_size = static_cast<int>(info[3]);
  op = std::move(out);
}

// operator_apply (include/operator.hpp:60-64): out resized to active_out_l_size,
// zeroed, then summed into. Asynchronous on `stream`.
inline void operator_apply(Operator& op, const DeviceArray<double>& in, DeviceArray<double>& out,
                           void* stream = nullptr) {
  out.resize(op.active_out_l_size);
  FEM_HIP_CHECK(fem_hip_operator_apply(op.handle.get(), in.data(), in.size(), out.data(), out.size(), stream));
}

inline void operator_apply(Operator& op, const DeviceArray<float>& in, DeviceArray<float>& out,
                           void* stream = nullptr) {
  out.resize(op.active_out_l_size);
  FEM_HIP_CHECK(fem_hip_operator_apply_f32(op.handle.get(), in.data(), in.size(), out.data(), out.size(), stream));
}

// Mass solve M x = b on a fused mass operator: mode FEM_HIP_SOLVE_FP64 (PCG),
// _MIXED (FP32 inner PCG, FP64 refinement) or _MIXED_RELIABLE (FP32 PCG with
// FP64 residual replacement).
// inner_rtol <= 0: 1e-5 per inner solve (MIXED), 1e-2 between replacements (MIXED_RELIABLE).
// cg_variant: FEM_HIP_CG_AUTO, _PLAIN, _PIPELINED or _PLAIN_WIDE.
inline fem_hip_solve_stats mass_solve(Operator& op, const DeviceArray<double>& b, DeviceArray<double>& x, int mode,
                                      double rtol = 1e-10, void* stream = nullptr, double inner_rtol = 0.0,
                                      int cg_variant = FEM_HIP_CG_AUTO) {
  if (inner_rtol <= 0.0) inner_rtol = 0.0;  // the library's default for the mode
  x.resize(b.size());
  fem_hip_solve_options o{mode, 2000, 30, rtol, inner_rtol, cg_variant};
  fem_hip_solve_stats st{};
  FEM_HIP_CHECK(fem_hip_mass_solve(op.handle.get(), b.data(), x.data(), &o, &st, stream));
  return st;
}
inline fem_hip_solve_stats mass_solve(Operator& op, const DeviceArray<double>& b, DeviceArray<double>& x,
                                      bool mixed, double rtol = 1e-10, void* stream = nullptr) {
  return mass_solve(op, b, x, mixed ? FEM_HIP_SOLVE_MIXED : FEM_HIP_SOLVE_FP64, rtol, stream);
}

// Jacobi diagonal of a fused mass operator (deterministic).
inline void mass_diagonal(Operator& op, DeviceArray<double>& diag, void* stream = nullptr) {
  diag.resize(op.active_out_l_size);
  FEM_HIP_CHECK(fem_hip_operator_mass_diagonal(op.handle.get(), diag.data(), diag.size(), stream));
}

// How a fused mass apply reaches the output (values of FEM_HIP_FUSED_*).
enum class FusedAlgorithm {
  Auto = FEM_HIP_FUSED_AUTO,
  ElemAtomic = FEM_HIP_FUSED_ELEM_ATOMIC,
  ElemTmap = FEM_HIP_FUSED_ELEM_TMAP,
  BlockAtomic = FEM_HIP_FUSED_BLOCK_ATOMIC,
  BlockPartial = FEM_HIP_FUSED_BLOCK_PARTIAL,
  BlockPartialLds = FEM_HIP_FUSED_BLOCK_PARTIAL_LDS,
  BoxColumn = FEM_HIP_FUSED_BOX_COLUMN,
};
struct FusedChoice {
  FusedAlgorithm alg = FusedAlgorithm::Auto;
  int elems = 0;
  double ms = 0.0;  // time per apply (autotune only)
};

// Fixes the algorithm (and elements per workgroup or block; 0 = its default) of
// the next fused applies; Auto returns to the tuning table's choice.
inline void operator_set_fused_algorithm(Operator& op, FusedAlgorithm alg, int elems = 0) {
  FEM_HIP_CHECK(fem_hip_operator_set_fused_algorithm(op.handle.get(), static_cast<int>(alg), elems));
}

inline FusedChoice operator_fused_choice(Operator& op, bool fp32 = false) {
  int alg = 0, elems = 0;
  FEM_HIP_CHECK(fem_hip_operator_get_fused_choice(op.handle.get(), fp32 ? 1 : 0, &alg, &elems));
  return {static_cast<FusedAlgorithm>(alg), elems, 0.0};
}

// Times every compiled algorithm and packing on the operator's own data and
// records the fastest in the process-wide tuning table. Synchronous.
inline FusedChoice operator_autotune(Operator& op, bool fp32 = false, int reps = 10, void* stream = nullptr) {
  int alg = 0, elems = 0, n = 0;
  double ms = 0.0;
  FEM_HIP_CHECK(fem_hip_operator_autotune(op.handle.get(), fp32 ? 1 : 0, reps, stream, &alg, &elems, &ms, nullptr, 0,
                                          &n));
  return {static_cast<FusedAlgorithm>(alg), elems, ms};
}

// Tuning tables as text; load returns the rows taken.
inline int tune_table_load(const std::string& path) {
  int taken = 0;
  FEM_HIP_CHECK(fem_hip_tune_table_load(path.c_str(), &taken));
  return taken;
}
inline void tune_table_save(const std::string& path) { FEM_HIP_CHECK(fem_hip_tune_table_save(path.c_str())); }

// A FEMHMSH1 mesh file: node restriction offsets [num_elem][P^dim], element
// vertices [num_elem][2^dim], coordinates [dim][num_vertices].
struct Mesh {
  fem_hip_mesh_info info{};
  std::vector<int> offsets, vertex_offsets;
  std::vector<double> coords;
};

inline Mesh mesh_read(const std::string& path) {
  Mesh m;
  FEM_HIP_CHECK(fem_hip_mesh_read(path.c_str(), &m.info, nullptr, nullptr, nullptr));
  size_t es = 1, vs = 1;
  for (int d = 0; d < m.info.dim; ++d) {
    es *= static_cast<size_t>(m.info.P);
    vs *= 2;
  }
  m.offsets.resize(es * m.info.num_elem);
  m.vertex_offsets.resize(vs * m.info.num_elem);
  m.coords.resize(static_cast<size_t>(m.info.dim) * m.info.num_vertices);
  FEM_HIP_CHECK(fem_hip_mesh_read(path.c_str(), &m.info, m.offsets.data(), m.vertex_offsets.data(), m.coords.data()));
  return m;
}

inline void mesh_write(const std::string& path, const Mesh& m) {
  FEM_HIP_CHECK(fem_hip_mesh_write(path.c_str(), &m.info, m.offsets.data(), m.vertex_offsets.data(), m.coords.data()));
}

// The node restriction of a mesh as the reference's host type, with num_comp
// blocked components (comp_stride = num_nodes).
inline fem::ElemRestriction mesh_restriction(const Mesh& m, int num_comp = 1) {
  fem::ElemRestriction r;
  const int es = static_cast<int>(m.offsets.size() / m.info.num_elem);
  fem::elem_restriction_create(m.info.num_elem, es, num_comp, m.info.num_nodes, num_comp * m.info.num_nodes,
                               m.offsets, r);
  return r;
}

// Continuous-batching mass-solve server (fem_hip_server_*): submit requests,
// wait for their results. Destruction finishes every submitted request.
class Server {
 public:
  explicit Server(const fem_hip_server_options& opt) {
    fem_hip_server s = nullptr;
    FEM_HIP_CHECK(fem_hip_server_create(&opt, &s));
    handle_.reset(s);
  }
  long long submit(const fem_hip_serve_request& req) {
    long long id = 0;
    FEM_HIP_CHECK(fem_hip_server_submit(handle_.get(), &req, &id));
    return id;
  }
  // Blocks until request `id` is done; throws if the request failed.
  fem_hip_serve_result wait(long long id) {
    fem_hip_serve_result res{};
    FEM_HIP_CHECK(fem_hip_server_wait(handle_.get(), id, -1.0, &res, nullptr, 0));
    return res;
  }
  // Same, and copies the solution (l_size = num_comp * (n*(P-1)+1)^3 entries;
  // the server must keep solutions).
  fem_hip_serve_result wait(long long id, long long l_size, std::vector<double>& x) {
    fem_hip_serve_result res{};
    x.resize(static_cast<size_t>(l_size));
    FEM_HIP_CHECK(fem_hip_server_wait(handle_.get(), id, -1.0, &res, x.data(), x.size()));
    return res;
  }

 private:
  struct Destroy {
    void operator()(fem_hip_server s) const { (void)fem_hip_server_destroy(s); }
  };
  std::unique_ptr<fem_hip_server_s, Destroy> handle_;
};

// Defaults: 16 requests at once, 8 setup workers, half the free HBM, at most
// 8 CG iterations per graph launch, no autotuning, solutions kept, the automatic
// algorithm where its setup is all on the device, small requests (up to 2^20
// element nodes) batched in shape cohorts.
inline fem_hip_server_options server_options() { return fem_hip_server_options{16, 8, 0.5, 8, 0, 1, 1, 1, 0}; }

inline void synchronize() { FEM_HIP_CHECK(fem_hip_device_synchronize()); }

}  // namespace fem::hip
