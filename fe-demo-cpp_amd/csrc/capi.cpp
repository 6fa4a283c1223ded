// extern "C" boundary of libfemhip (declared in include/fem_hip.h). Every entry
// point catches, maps the exception type to a status code and records the
// message for fem_hip_last_error() on the calling thread.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fem_hip.h"
#include "basis.hpp"
#include "basis_setup.hpp"
#include "block_plan.hpp"
#include "fused_box_kernel.hpp"
#include "mesh_io.hpp"
#include "server.hpp"
#include "batch_solve.hpp"
#include "common.hpp"
#include "mesh.hpp"
#include "operator.hpp"
#include "qfunction.hpp"
#include "restriction.hpp"
#include "solver.hpp"

struct fem_hip_restriction_s {
  std::shared_ptr<const femhip::Restriction> r;
};
struct fem_hip_basis_s {
  std::shared_ptr<const femhip::Basis> b;
};
struct fem_hip_qfunction_s {
  femhip::QFunctionDev q;
};
struct fem_hip_operator_s {
  std::unique_ptr<femhip::Operator> op;
  femhip::MassSolverState* solver = nullptr;
  ~fem_hip_operator_s() { femhip::mass_solver_state_destroy(solver); }
};

namespace femhip {
int probe_launch(int blocks, int threads);
size_t noalloc_scratch_capacity();
void* noalloc_scratch_device_ptr();
}

namespace {

thread_local std::string g_last_error;

template <class F>
int guard(F&& f) {
  try {
    f();
    return FEM_HIP_SUCCESS;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return FEM_HIP_ERR_INVALID_ARGUMENT;
  } catch (const femhip::OutOfMemory& e) {
    g_last_error = e.what();
    return FEM_HIP_ERR_OUT_OF_MEMORY;
  } catch (const femhip::DeviceError& e) {
    g_last_error = e.what();
    return FEM_HIP_ERR_DEVICE;
  } catch (const femhip::Unsupported& e) {
    g_last_error = e.what();
    return FEM_HIP_ERR_UNSUPPORTED;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return FEM_HIP_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return FEM_HIP_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return FEM_HIP_ERR_RUNTIME;
  }
}

hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }

void need(const void* p, const char* what) {
  if (!p) throw std::invalid_argument(std::string(what) + " must not be NULL");
}

}  // namespace

using namespace femhip;

extern "C" {

const char* fem_hip_last_error(void) { return g_last_error.c_str(); }
const char* fem_hip_version(void) { return "femhip 0.1 (gfx950)"; }

/* ---- device and memory ---- */
int fem_hip_device_count(int* count) {
  return guard([&] {
    need(count, "count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}

int fem_hip_device_info_get(int device, fem_hip_device_info* info) {
  return guard([&] {
    need(info, "info");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    require(device >= 0 && device < n, "device_info: no such HIP device");
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, device));
    std::memset(info, 0, sizeof(*info));
    info->id = device;
    std::strncpy(info->name, p.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
    info->cu_count = p.multiProcessorCount;
    info->wavefront_size = p.warpSize;
    info->max_threads_per_block = p.maxThreadsPerBlock;
    info->regs_per_block = p.regsPerBlock;
    info->lds_per_block = p.sharedMemPerBlock;
    info->lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
    info->l2_bytes = static_cast<size_t>(p.l2CacheSize);
    info->total_global_mem = p.totalGlobalMem;
    info->clock_khz = p.clockRate;
    info->memory_clock_khz = p.memoryClockRate;
    info->memory_bus_width = p.memoryBusWidth;
    int prev = 0;
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    HIP_CHECK(hipSetDevice(prev));
    info->free_global_mem = fr;
  });
}

int fem_hip_malloc(void** ptr, size_t bytes) {
  return guard([&] {
    need(ptr, "ptr");
    *ptr = nullptr;
    SharedCapture lk;
    if (bytes) HIP_CHECK(hipMalloc(ptr, bytes));
  });
}
int fem_hip_free(void* ptr) {
  return guard([&] {
    SharedCapture lk;
    if (ptr) HIP_CHECK(hipFree(ptr));
  });
}
int fem_hip_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    SharedCapture lk;
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}
int fem_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    SharedCapture lk;
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}
int fem_hip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* s) {
  return guard([&] {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st(s)));
  });
}
int fem_hip_stream_copy(void* dst, const void* src, size_t bytes, void* s) {
  return guard([&] { stream_copy(dst, src, bytes, st(s)); });
}
int fem_hip_memset_zero(void* dst, size_t bytes, void* s) {
  return guard([&] { zero_fill(dst, bytes, st(s)); });
}
int fem_hip_device_synchronize(void) {
  return guard([&] { HIP_CHECK(hipDeviceSynchronize()); });
}
int fem_hip_stream_create(void** s) {
  return guard([&] {
    need(s, "stream");
    hipStream_t x;
    HIP_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    *s = x;
  });
}
int fem_hip_stream_destroy(void* s) {
  return guard([&] {
    if (s) HIP_CHECK(hipStreamDestroy(st(s)));
  });
}
int fem_hip_stream_synchronize(void* s) {
  return guard([&] { HIP_CHECK(hipStreamSynchronize(st(s))); });
}
int fem_hip_event_create(void** e) {
  return guard([&] {
    need(e, "event");
    hipEvent_t x;
    HIP_CHECK(hipEventCreate(&x));
    *e = x;
  });
}
int fem_hip_event_destroy(void* e) {
  return guard([&] {
    if (e) HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(e)));
  });
}
int fem_hip_event_record(void* e, void* s) {
  return guard([&] { HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(e), st(s))); });
}
int fem_hip_event_synchronize(void* e) {
  return guard([&] { HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(e))); });
}
int fem_hip_event_elapsed_ms(void* a, void* b, float* ms) {
  return guard([&] {
    need(ms, "ms");
    HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
  });
}
int fem_hip_probe_launch(int blocks, int threads) {
  return guard([&] {
    const int e = probe_launch(blocks, threads);
    if (e != 0) hip_check(static_cast<hipError_t>(e), "probe_kernel<<<blocks, threads>>>", __FILE__, __LINE__);
  });
}

/* ---- 1D basis setup ---- */
int fem_hip_gauss_legendre(int Q, double* x, double* w) {
  return guard([&] {
    need(x, "q_ref_1d");
    need(w, "q_weight_1d");
    gauss_legendre(Q, x, w);
  });
}
int fem_hip_gauss_lobatto(int Q, double* x, double* w) {
  return guard([&] {
    need(x, "q_ref_1d");
    gauss_lobatto(Q, x, w);
  });
}
int fem_hip_lagrange_basis_matrix(int P, const double* nodes, int M, const double* pts, double* interp,
                                  double* grad) {
  return guard([&] {
    require(P > 0 && M >= 0, "lagrange_basis_matrix: bad sizes");
    need(nodes, "nodes");
    need(interp, "interp_1d");
    need(grad, "grad_1d");
    lagrange_matrices(P, nodes, M, pts, interp, grad);
  });
}

/* ---- restriction ---- */
int fem_hip_restriction_create(int ne, int es, int nc, int cs, int l_size, const int* offsets, int mode,
                               fem_hip_restriction* out) {
  return guard([&] {
    need(out, "out");
    std::shared_ptr<const Restriction> r(restriction_create_offset(ne, es, nc, cs, l_size, offsets, mode));
    *out = new fem_hip_restriction_s{std::move(r)};
  });
}
int fem_hip_restriction_create_strided(int ne, int es, int nc, int l_size, const int* strides,
                                       fem_hip_restriction* out) {
  return guard([&] {
    need(out, "out");
    std::shared_ptr<const Restriction> r(restriction_create_strided(ne, es, nc, l_size, strides));
    *out = new fem_hip_restriction_s{std::move(r)};
  });
}
int fem_hip_restriction_apply(fem_hip_restriction r, int t_mode, const double* in, size_t in_len, double* out,
                              size_t out_len, void* s) {
  return guard([&] {
    need(r, "restriction");
    restriction_apply<double>(*r->r, t_mode, in, in_len, out, out_len, st(s));
  });
}
int fem_hip_restriction_apply_f32(fem_hip_restriction r, int t_mode, const float* in, size_t in_len, float* out,
                                  size_t out_len, void* s) {
  return guard([&] {
    need(r, "restriction");
    restriction_apply<float>(*r->r, t_mode, in, in_len, out, out_len, st(s));
  });
}
int fem_hip_restriction_get_sizes(fem_hip_restriction h, long long* sizes) {
  return guard([&] {
    need(h, "restriction");
    need(sizes, "sizes");
    const Restriction& r = *h->r;
    const long long v[8] = {r.type, r.num_elem, r.elem_size, r.num_comp, r.comp_stride,
                            r.l_size, static_cast<long long>(r.e_size), r.num_nodes};
    std::memcpy(sizes, v, sizeof(v));
  });
}
int fem_hip_restriction_get_transpose_map(fem_hip_restriction h, int* lvec, int* toff, int* tind) {
  return guard([&] {
    need(h, "restriction");
    const Restriction& r = *h->r;
    require(r.type == 0 && r.scatter_mode == kScatterDeterministic,
            "get_transpose_map: only deterministic offset restrictions have one");
    if (lvec) {
      if (r.identity_nodes) {
        for (int i = 0; i < r.num_nodes; ++i) lvec[i] = i;
      } else {
        HIP_CHECK(hipMemcpy(lvec, r.l_vec_indices.data(), r.l_vec_indices.size() * sizeof(int), hipMemcpyDeviceToHost));
      }
    }
    if (toff) HIP_CHECK(hipMemcpy(toff, r.t_offsets.data(), r.t_offsets.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (tind) HIP_CHECK(hipMemcpy(tind, r.t_indices.data(), r.t_indices.size() * sizeof(int), hipMemcpyDeviceToHost));
  });
}
int fem_hip_restriction_destroy(fem_hip_restriction r) {
  return guard([&] { delete r; });
}
int fem_hip_restriction_apply_arrays(const fem_hip_restriction_arrays* a, int t_mode, const double* in,
                                     size_t in_len, double* out, size_t out_len, void* s) {
  return guard([&] {
    need(a, "restriction");
    require(a->type == 0 || a->type == 1, "elem_restriction_apply: unknown restriction type");
    require(a->num_elem >= 0 && a->elem_size >= 0 && a->num_comp >= 0 && a->l_size >= 0 && a->e_size >= 0,
            "elem_restriction_apply: negative size");
    require(a->e_size == static_cast<long long>(a->num_comp) * a->elem_size * a->num_elem,
            "elem_restriction_apply: e_size != num_comp * elem_size * num_elem");
    require(a->scatter_mode == kScatterAtomic || a->scatter_mode == kScatterDeterministic,
            "elem_restriction_apply: unknown scatter mode");
    RestrictionView v;
    v.type = a->type;
    v.num_elem = a->num_elem;
    v.elem_size = a->elem_size;
    v.num_comp = a->num_comp;
    v.comp_stride = a->comp_stride;
    v.l_size = a->l_size;
    v.e_size = static_cast<size_t>(a->e_size);
    v.strides = {a->strides[0], a->strides[1], a->strides[2]};
    v.scatter_mode = a->scatter_mode;
    v.offsets = a->offsets;
    v.num_nodes = a->num_nodes;
    v.l_vec_indices = a->l_vec_indices;
    v.t_offsets = a->t_offsets;
    v.t_indices = a->t_indices;
    if (v.type == 0 && v.scatter_mode == kScatterDeterministic && t_mode == kTranspose)
      require(v.num_nodes >= 0 && v.l_vec_indices != nullptr,
              "elem_restriction_apply: the deterministic scatter needs l_vec_indices on the device");
    restriction_apply_view<double>(v, t_mode, in, in_len, out, out_len, st(s));
  });
}
int fem_hip_transpose_map_build(int ne, int es, int l_size, const int* offsets, int* rows, int* lvec, int* toff,
                                int* tind) {
  return guard([&] {
    require(ne > 0 && es > 0 && l_size > 0, "build_transpose_map: bad sizes");
    need(offsets, "offsets");
    need(rows, "num_nodes");
    const size_t nt = static_cast<size_t>(ne) * es;
    for (size_t t = 0; t < nt; ++t)
      require(offsets[t] >= 0 && offsets[t] < l_size, "build_transpose_map: offsets must be in [0, l_size - 1]");
    TransposeMapHost m;
    build_transpose_map(ne, es, l_size, offsets, m);
    *rows = static_cast<int>(m.l_vec_indices.size());
    if (lvec) std::memcpy(lvec, m.l_vec_indices.data(), m.l_vec_indices.size() * sizeof(int));
    if (toff) std::memcpy(toff, m.t_offsets.data(), m.t_offsets.size() * sizeof(int));
    if (tind) std::memcpy(tind, m.t_indices.data(), m.t_indices.size() * sizeof(int));
  });
}

/* ---- basis ---- */
int fem_hip_basis_create_h1_lagrange(int dim, int nc, int P, int Q, int quad_mode, fem_hip_basis* out) {
  return guard([&] {
    need(out, "out");
    std::shared_ptr<const Basis> b(basis_create(make_h1_lagrange(dim, nc, P, Q, quad_mode)));
    *out = new fem_hip_basis_s{std::move(b)};
  });
}
int fem_hip_basis_create_tensor(int dim, int nc, int P, int Q, const double* qref, const double* qw,
                                const double* interp, const double* grad, fem_hip_basis* out) {
  return guard([&] {
    need(out, "out");
    require(dim >= 1 && dim <= 3 && nc >= 1 && P >= 1 && Q >= 1, "basis_create_tensor: bad sizes");
    need(interp, "interp_1d");
    need(grad, "grad_1d");
    need(qw, "q_weight_1d");
    HostBasis h;
    h.dim = dim;
    h.num_comp = nc;
    h.P = P;
    h.Q = Q;
    h.q_ref = qref ? std::vector<double>(qref, qref + Q) : std::vector<double>(Q, 0.0);
    h.q_weight.assign(qw, qw + Q);
    h.interp.assign(interp, interp + static_cast<size_t>(P) * Q);
    h.grad.assign(grad, grad + static_cast<size_t>(P) * Q);
    std::shared_ptr<const Basis> b(basis_create(h));
    *out = new fem_hip_basis_s{std::move(b)};
  });
}
int fem_hip_basis_get(fem_hip_basis h, int* dims, double* qref, double* qw, double* interp, double* grad) {
  return guard([&] {
    need(h, "basis");
    const HostBasis& b = h->b->h;
    if (dims) {
      dims[0] = b.dim;
      dims[1] = b.num_comp;
      dims[2] = b.P;
      dims[3] = b.Q;
    }
    if (qref) std::memcpy(qref, b.q_ref.data(), b.q_ref.size() * sizeof(double));
    if (qw) std::memcpy(qw, b.q_weight.data(), b.q_weight.size() * sizeof(double));
    if (interp) std::memcpy(interp, b.interp.data(), b.interp.size() * sizeof(double));
    if (grad) std::memcpy(grad, b.grad.data(), b.grad.size() * sizeof(double));
  });
}
int fem_hip_basis_apply(fem_hip_basis h, int ne, int t_mode, int eval_mode, const double* u, size_t u_len, double* v,
                        size_t v_len, void* s) {
  return guard([&] {
    need(h, "basis");
    basis_apply(*h->b, ne, t_mode, eval_mode, u, u_len, v, v_len, st(s));
  });
}
int fem_hip_basis_destroy(fem_hip_basis b) {
  return guard([&] { delete b; });
}
int fem_hip_contract(int A, int B, int C, int J, const double* t, int t_mode, int add, const double* u, double* v,
                     void* s) {
  return guard([&] {
    need(t, "t");
    need(u, "u");
    need(v, "v");
    require(t_mode == kNoTranspose || t_mode == kTranspose, "tensor_contract_apply: unknown transpose mode");
    contract(A, B, C, J, t, t_mode, add != 0, u, v, st(s));
  });
}

/* ---- qfunction ---- */
int fem_hip_qfunction_create(int kind, fem_hip_qfunction* out) {
  return guard([&] {
    need(out, "out");
    std::unique_ptr<QFunctionDev> q(qfunction_create(kind));
    *out = new fem_hip_qfunction_s{*q};
  });
}
int fem_hip_qfunction_add_input(fem_hip_qfunction qf, const char* name, int size, int mode) {
  return guard([&] {
    need(qf, "qfunction");
    need(name, "name");
    qfunction_add_input(qf->q, name, size, mode);
  });
}
int fem_hip_qfunction_add_output(fem_hip_qfunction qf, const char* name, int size, int mode) {
  return guard([&] {
    need(qf, "qfunction");
    need(name, "name");
    qfunction_add_output(qf->q, name, size, mode);
  });
}
int fem_hip_qfunction_set_context(fem_hip_qfunction qf, const void* data, size_t bytes) {
  return guard([&] {
    need(qf, "qfunction");
    const auto* p = static_cast<const unsigned char*>(data);
    qf->q.ctx.assign(p, p + (p ? bytes : 0));
  });
}
int fem_hip_qfunction_apply(fem_hip_qfunction qf, int Q, int num_in, const double* const* in, const size_t* in_lens,
                            int num_out, double* const* out, const size_t* out_lens, void* s) {
  return guard([&] {
    need(qf, "qfunction");
    require(num_in >= 0 && num_out >= 0, "qfunction_apply: negative field count");
    if (num_in) {
      need(in, "in");
      need(in_lens, "in_lens");
    }
    if (num_out) {
      need(out, "out");
      need(out_lens, "out_lens");
    }
    qfunction_apply(qf->q, Q, num_in, in, in_lens, num_out, out, out_lens, st(s));
  });
}
int fem_hip_qfunction_destroy(fem_hip_qfunction qf) {
  return guard([&] { delete qf; });
}

/* ---- operator ---- */
int fem_hip_operator_create(fem_hip_qfunction qf, fem_hip_operator* out) {
  return guard([&] {
    need(qf, "qfunction");
    need(out, "out");
    auto* h = new fem_hip_operator_s();
    h->op.reset(operator_create(qf->q));
    *out = h;
  });
}
int fem_hip_operator_set_field(fem_hip_operator op, const char* name, fem_hip_restriction r, fem_hip_basis b,
                               int vkind, const double* passive, size_t passive_len, int passive_on_device) {
  return guard([&] {
    need(op, "operator");
    need(name, "name");
    operator_set_field(*op->op, name, r ? r->r : nullptr, b ? b->b : nullptr, vkind, passive, passive_len,
                       passive_on_device != 0);
    mass_solver_state_destroy(op->solver);
    op->solver = nullptr;
  });
}
int fem_hip_operator_check_ready(fem_hip_operator op) {
  return guard([&] {
    need(op, "operator");
    operator_check_ready(*op->op);
  });
}
int fem_hip_operator_apply(fem_hip_operator op, const double* in, size_t in_len, double* out, size_t out_len,
                           void* s) {
  return guard([&] {
    need(op, "operator");
    operator_apply(*op->op, in, in_len, out, out_len, st(s));
  });
}
int fem_hip_operator_apply_f32(fem_hip_operator op, const float* in, size_t in_len, float* out, size_t out_len,
                               void* s) {
  return guard([&] {
    need(op, "operator");
    operator_apply_f32(*op->op, in, in_len, out, out_len, st(s));
  });
}
int fem_hip_operator_apply_noalloc(fem_hip_operator h, const double* in, size_t in_len, double* out, size_t out_len,
                                   void* s) {
  return guard([&] {
    need(h, "operator");
    Operator& op = *h->op;
    operator_check_ready(op);
    require(in_len == static_cast<size_t>(op.active_in_l_size),
            "operator_apply: input vector size does not match active_in_l_size");
    require(out_len == static_cast<size_t>(op.active_out_l_size),
            "operator_apply: output vector size does not match active_out_l_size");
    if (!op.fused_checked || op.fused_kind != 1 || !op.fused)
      throw Unsupported("apply_noalloc: needs a fused mass operator whose device state exists (apply it once with "
                        "fem_hip_operator_apply first)");
    FusedMass& f = operator_fused_mass(op, st(s));  // exists: takes the current options, allocates nothing
    fused_mass_apply64_scratch(f, in, out, noalloc_scratch_device_ptr(), noalloc_scratch_capacity(), st(s));
  });
}
size_t fem_hip_noalloc_scratch_bytes(void) { return femhip::noalloc_scratch_capacity(); }
int fem_hip_operator_get_info(fem_hip_operator h, long long* info) {
  return guard([&] {
    need(h, "operator");
    need(info, "info");
    Operator& op = *h->op;
    long long v[8] = {op.num_elem, op.num_qpts, op.active_in_l_size, op.active_out_l_size, 0, 0, 0, 0};
    bool ready = true;
    for (const auto* fs : {&op.inputs, &op.outputs})
      for (const auto& f : *fs) ready = ready && f.set;
    if (ready && op.num_elem) {
      try {
        FusedMass& f = operator_fused_mass(op, nullptr);
        v[4] = 1;
        v[5] = f.P;
        v[6] = f.Q;
        v[7] = f.NC;
      } catch (const Unsupported&) {
      }
    }
    std::memcpy(info, v, sizeof(v));
  });
}
// The solver's captured CG graphs hold the fused apply's kernels and buffers;
// a change of options or an autotune (which may free and remake buffers)
// drops them with the rest of the cached solver state (mass_solve also
// re-records graphs whose apply signature changed).
static void drop_solver(fem_hip_operator h) {
  femhip::mass_solver_state_destroy(h->solver);
  h->solver = nullptr;
}
int fem_hip_operator_set_fused_options(fem_hip_operator h, int scatter, int exact, int elems) {
  return guard([&] {
    need(h, "operator");
    require(scatter == -1 || scatter == kScatterAtomic || scatter == kScatterDeterministic,
            "set_fused_options: unknown scatter mode");
    require(elems >= 0, "set_fused_options: elems_per_block must be >= 0");
    h->op->opt_scatter = scatter;
    h->op->opt_exact = exact ? 1 : 0;
    h->op->opt_alg = elems > 0 ? -2 : -1;
    h->op->opt_elems = elems;
    drop_solver(h);
  });
}
int fem_hip_operator_set_fused_algorithm(fem_hip_operator h, int alg, int elems) {
  return guard([&] {
    need(h, "operator");
    require(alg >= -1 && alg < kAlgCount, "set_fused_algorithm: unknown algorithm");
    require(elems >= 0, "set_fused_algorithm: elems must be >= 0");
    h->op->opt_alg = alg;
    h->op->opt_elems = alg < 0 ? 0 : elems;
    drop_solver(h);
  });
}
int fem_hip_operator_set_box_layers(fem_hip_operator h, int layers) {
  return guard([&] {
    need(h, "operator");
    require(layers >= 0, "set_box_layers: layers must be >= 0");
    h->op->opt_box_lz = layers;
    drop_solver(h);
  });
}
int fem_hip_operator_get_fused_choice(fem_hip_operator h, int fp32, int* alg, int* elems) {
  return guard([&] {
    need(h, "operator");
    need(alg, "alg");
    need(elems, "elems");
    const FusedKey k = fused_resolve(operator_fused_mass(*h->op, nullptr), fp32 ? 1 : 0);
    *alg = k.alg;
    *elems = k.elems;
  });
}
int fem_hip_operator_autotune(fem_hip_operator h, int fp32, int reps, void* s, int* best_alg, int* best_elems,
                              double* best_ms, double* results, int max_results, int* n_results) {
  return guard([&] {
    need(h, "operator");
    require(reps > 0, "autotune: reps must be positive");
    require(max_results >= 0 && (results || max_results == 0), "autotune: bad results buffer");
    FusedMass& f = operator_fused_mass(*h->op, st(s));
    drop_solver(h);
    std::vector<TuneEntry> all;
    const TuneEntry best = fused_autotune(f, fp32 ? 1 : 0, reps, st(s), &all);
    if (best_alg) *best_alg = best.alg;
    if (best_elems) *best_elems = best.elems;
    if (best_ms) *best_ms = best.ms;
    for (int i = 0; i < max_results && i < static_cast<int>(all.size()); ++i) {
      results[3 * i] = all[i].alg;
      results[3 * i + 1] = all[i].elems;
      results[3 * i + 2] = all[i].ms;
    }
    if (n_results) *n_results = static_cast<int>(all.size());
  });
}
int fem_hip_tune_table_get(int* rows8, double* ms, long long* num_elem, int max_rows, int* n_rows) {
  return guard([&] {
    need(n_rows, "n_rows");
    require(max_rows >= 0 && (max_rows == 0 || rows8), "tune_table_get: bad rows buffer");
    int i = 0;
    std::lock_guard<std::mutex> lk(tune_mutex());
    for (const auto& kv : tune_table()) {
      for (const auto& row : kv.second) {
        if (i < max_rows) {
          const TuneKey& k = kv.first;
          const TuneEntry& e = row.second;
          const int r[8] = {k.P, k.Q, k.NC, k.fp32, k.det, k.box, e.alg, e.elems};
          std::memcpy(rows8 + 8 * i, r, sizeof(r));
          if (ms) ms[i] = e.ms;
          if (num_elem) num_elem[i] = e.num_elem;
        }
        ++i;
      }
    }
    *n_rows = i;
  });
}
int fem_hip_tune_table_set(const int* r, double ms, long long num_elem) {
  return guard([&] {
    need(r, "row");
    require(r[3] == 0 || r[3] == 1, "tune_table_set: fp32 must be 0 or 1");
    require(r[4] == 0 || r[4] == 1, "tune_table_set: deterministic must be 0 or 1");
    require(r[5] == 0 || r[5] == 1, "tune_table_set: box must be 0 or 1");
    require(r[6] >= 0 && r[6] < kAlgCount, "tune_table_set: unknown algorithm");
    require(!r[4] || alg_deterministic(r[6]), "tune_table_set: a deterministic row needs a deterministic algorithm");
    require(r[5] || !alg_box(r[6]), "tune_table_set: the box-column algorithm needs a box row");
    require(fused_registry().count(FusedKey{r[0], r[1], r[2], r[3], r[6], 0, r[7]}) != 0,
            "tune_table_set: algorithm/elems not compiled for this shape");
    TuneEntry e;
    e.alg = r[6];
    e.elems = r[7];
    e.ms = ms;
    e.num_elem = num_elem;
    require(num_elem > 0, "tune_table_set: num_elem (the mesh size the row was measured on) must be positive");
    std::lock_guard<std::mutex> lk(tune_mutex());
    tune_record(TuneKey{r[0], r[1], r[2], r[3], r[4], r[5]}, e);
  });
}
int fem_hip_tune_table_clear(void) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(tune_mutex());
    tune_table().clear();
    tune_alg_table().clear();
  });
}
int fem_hip_tune_table_save(const char* path) {
  return guard([&] {
    need(path, "path");
    tune_table_save(path);
  });
}
int fem_hip_tune_table_load(const char* path, int* rows_taken) {
  return guard([&] {
    need(path, "path");
    const int n = tune_table_load(path);
    if (rows_taken) *rows_taken = n;
  });
}
int fem_hip_operator_mass_diagonal(fem_hip_operator h, double* diag, size_t len, void* s) {
  return guard([&] {
    need(h, "operator");
    FusedMass& f = operator_fused_mass(*h->op, st(s));
    require(len == static_cast<size_t>(f.l_size), "mass_diagonal: diag must have l_size entries");
    fused_mass_diagonal(f, diag, st(s));
  });
}
int fem_hip_operator_timing_begin(fem_hip_operator h) {
  return guard([&] {
    need(h, "operator");
    fused_timing_begin(operator_fused_mass(*h->op, nullptr));
  });
}
int fem_hip_operator_timing_end(fem_hip_operator h, double* kernel_ms, double* aux_ms, int* count) {
  return guard([&] {
    need(h, "operator");
    fused_timing_end(operator_fused_mass(*h->op, nullptr), kernel_ms, aux_ms, count);
  });
}
int fem_hip_operator_destroy(fem_hip_operator op) {
  return guard([&] { delete op; });
}
int fem_hip_operator_create_mass(fem_hip_restriction r, fem_hip_basis b, const double* qdata, fem_hip_operator* out) {
  return guard([&] {
    need(r, "restriction");
    need(b, "basis");
    need(qdata, "qdata");
    need(out, "out");
    const Basis& bb = *b->b;
    const int nc = bb.h.num_comp;
    QFunctionDev qf;
    qf.kind = nc == 1 ? kQfApplyMass : kQfApplyMassVector;
    if (nc > 1) {
      qf.ctx.resize(sizeof(int));
      std::memcpy(qf.ctx.data(), &nc, sizeof(int));
    }
    qfunction_add_input(qf, "u", nc, kEvalInterp);
    qfunction_add_input(qf, "qdata", 1, kEvalNone);
    qfunction_add_output(qf, "v", nc, kEvalInterp);
    const int nq = ipow(bb.h.Q, bb.h.dim);
    const int ne = r->r->num_elem;
    std::shared_ptr<const Restriction> qr(restriction_create_strided(ne, nq, 1, ne * nq, nullptr));
    auto h = std::make_unique<fem_hip_operator_s>();
    h->op.reset(operator_create(qf));
    operator_set_field(*h->op, "u", r->r, b->b, kVectorActive, nullptr, 0, false);
    operator_set_field(*h->op, "qdata", qr, nullptr, kVectorPassive, qdata, static_cast<size_t>(ne) * nq, true);
    operator_set_field(*h->op, "v", r->r, b->b, kVectorActive, nullptr, 0, false);
    *out = h.release();
  });
}

/* ---- meshes ---- */
int fem_hip_rng_uniform(uint64_t seed, size_t n, double lo, double hi, double* out) {
  return guard([&] {
    if (n) need(out, "out");
    rng_uniform(seed, n, lo, hi, out);
  });
}
int fem_hip_rng_uniform_device(uint64_t seed, size_t n, double lo, double hi, double* out_device, void* stream) {
  return guard([&] {
    if (n) need(out_device, "out_device");
    rng_uniform_device(seed, n, lo, hi, out_device, static_cast<hipStream_t>(stream));
  });
}
int fem_hip_box_offsets(int nx, int ny, int nz, int P, int node_stride, int* offsets) {
  return guard([&] {
    need(offsets, "offsets");
    box_offsets(nx, ny, nz, P, node_stride, offsets);
  });
}
int fem_hip_box_offsets_device(int nx, int ny, int nz, int P, int node_stride, int* offsets, void* s) {
  return guard([&] { box_offsets_device(nx, ny, nz, P, node_stride, offsets, st(s)); });
}
int fem_hip_box_coords_device(int n, double jitter, uint64_t seed, double* coords, void* s) {
  return guard([&] { box_coords_device(n, jitter, seed, coords, st(s)); });
}
int fem_hip_box_coords(int n, double jitter, uint64_t seed, double* coords) {
  return guard([&] {
    need(coords, "coords");
    box_coords(n, jitter, seed, coords);
  });
}

int fem_hip_box_detect(int num_elem, int P, const int* offsets, int num_comp, int comp_stride, int l_size,
                       int* n3, int* is_box) {
  return guard([&] {
    need(offsets, "offsets");
    need(n3, "n3");
    need(is_box, "is_box");
    require(num_elem > 0 && P >= 2 && num_comp > 0 && l_size > 0, "box_detect: bad sizes");
    *is_box = detect_box(offsets, num_elem, P, num_comp, comp_stride, l_size, n3) ? 1 : 0;
  });
}

int fem_hip_box_column_check(int nx, int ny, int nz, int P, int elems, int layers, long long* stats) {
  return guard([&] {
    need(stats, "stats");
    int BX = 0, BY = 0;
    require(box_columns(elems, &BX, &BY), "box_column_check: no box-column variant with these elems");
    require(nx > 0 && ny > 0 && nz > 0 && P >= 2 && nx % BX == 0 && ny % BY == 0, "box_column_check: bad sizes");
    BoxGeom g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.Nx = nx * (P - 1) + 1;
    g.Ny = ny * (P - 1) + 1;
    g.Nz = nz * (P - 1) + 1;
    g.ncx = nx / BX;
    g.ncy = ny / BY;
    require(layers >= 0, "box_column_check: layers must be >= 0");
    // Automatic: a count box_segments picks for 32 elements per column layer per CU on 256 CUs (the
    // apply's own choice also counts its kernel's resident workgroups; the check holds for any count).
    g.nsz = box_segments(nz, static_cast<long long>(g.ncx) * g.ncy, 32 * 256 / (BX * BY), P, layers);
    require(g.nsz > 0, "box_column_check: no segment count fits");
    g.lz = (nz + g.nsz - 1) / g.nsz;
    g.cs = g.Nx * g.Ny * g.Nz;
    const int LXn = BX * (P - 1) + 1, LYn = BY * (P - 1) + 1, LZn = g.lz * (P - 1) + 1;
    require(LXn >= 3 && LYn >= 3 && LZn >= 3, "box_column_check: columns too thin");
    const long long nseg = static_cast<long long>(g.ncx) * g.ncy * g.nsz;
    const long long face = box_face_size(LXn, LYn, LZn);
    const long long N = static_cast<long long>(g.Nx) * g.Ny * g.Nz;
    require(N * 2 + nseg * face < (1ll << 31), "box_column_check: too large to check");
    auto fail = [](const char* m) { throw std::runtime_error(std::string("box columns: ") + m); };
    // What the column kernel writes: each region node once, to v or to a slot.
    std::vector<int> direct(N, 0);
    std::vector<char> slot(static_cast<size_t>(nseg * face), 0);
    long long slots_written = 0, direct_nodes = 0;
    for (long long w = 0; w < nseg; ++w) {
      const int cx = static_cast<int>(w % g.ncx), cy = static_cast<int>((w / g.ncx) % g.ncy),
                sz = static_cast<int>(w / (static_cast<long long>(g.ncx) * g.ncy));
      const int z0 = box_seg_z0(g, sz), LZs = (box_seg_z0(g, sz + 1) - z0) * (P - 1) + 1;  // this segment's planes
      for (int lz = 0; lz < LZs; ++lz)
        for (int ly = 0; ly < LYn; ++ly)
          for (int lx = 0; lx < LXn; ++lx) {
            const bool xs = (lx == 0 && cx > 0) || (lx == LXn - 1 && cx < g.ncx - 1);
            const bool ys = (ly == 0 && cy > 0) || (ly == LYn - 1 && cy < g.ncy - 1);
            const bool zs = (lz == 0 && sz > 0) || (lz == LZs - 1 && sz < g.nsz - 1);
            const long long node = (cx * (LXn - 1) + lx) +
                                   static_cast<long long>(g.Nx) * ((cy * (LYn - 1) + ly) +
                                                                   static_cast<long long>(g.Ny) * (z0 * (P - 1) + lz));
            if (!(xs || ys || zs)) {
              ++direct[node];
              ++direct_nodes;
            } else {
              const int sl = box_slot(lx, ly, lz, xs, ys, LXn, LYn, LZn);
              if (sl < 0 || sl >= face) fail("slot outside the segment's area");
              char& c = slot[static_cast<size_t>(w * face + sl)];
              if (c) fail("two nodes of a segment share a slot");
              c = 1;
              ++slots_written;
            }
          }
    }
    // What the shared-node pass reads, over its whole grid: each shared node
    // once, each written slot once.
    const long long count = box_shared_count(g, P, BX, BY);
    const dim3 grid = box_shared_grid(g);
    std::vector<char> shared(N, 0);
    int max_owners = 0;
    long long slots_read = 0, visited = 0;
    for (unsigned pl = 0; pl < grid.z; ++pl) {
      for (unsigned ln = 0; ln < grid.y * kSharedLines; ++ln) {
        for (unsigned pt = 0; pt < grid.x * 64; ++pt) {
          int gx, gy, gz, seg[8], sl[8];
          if (!box_shared_point(g, P, BX, BY, static_cast<int>(pl), static_cast<int>(ln), static_cast<int>(pt), &gx, &gy,
                                &gz)) {
            int seg2[8], sl2[8];
            size_t node2 = 0;
            if (box_shared_owners(g, P, BX, BY, static_cast<int>(pl), static_cast<int>(ln), static_cast<int>(pt), &node2,
                                  seg2, sl2) != 0)
              fail("one-step owners take a point that is not a shared node");
            continue;
          }
          ++visited;
          if (gx < 0 || gx >= g.Nx || gy < 0 || gy >= g.Ny || gz < 0 || gz >= g.Nz) fail("shared node out of the box");
          const long long node = gx + static_cast<long long>(g.Nx) * (gy + static_cast<long long>(g.Ny) * gz);
          if (shared[node]) fail("a shared node is enumerated twice");
          shared[node] = 1;
          const int n = box_owners(g, P, BX, BY, gx, gy, gz, seg, sl);
          if (n < 2) fail("a shared node has fewer than 2 segments");
          {  // the pass's one-step form gives the same node, owners and slots
            int seg2[8], sl2[8];
            size_t node2 = 0;
            const int n2 = box_shared_owners(g, P, BX, BY, static_cast<int>(pl), static_cast<int>(ln),
                                             static_cast<int>(pt), &node2, seg2, sl2);
            if (n2 != n || node2 != static_cast<size_t>(node)) fail("one-step owners disagree (count or node)");
            for (int i = 0; i < n; ++i)
              if (seg2[i] != seg[i] || sl2[i] != sl[i]) fail("one-step owners disagree (segment or slot)");
          }
          max_owners = std::max(max_owners, n);
          for (int i = 0; i < n; ++i) {
            if (i > 0 && seg[i] <= seg[i - 1]) fail("segments not ascending");
            if (seg[i] < 0 || seg[i] >= nseg || sl[i] < 0 || sl[i] >= face) fail("owner out of range");
            char& c = slot[static_cast<size_t>(static_cast<long long>(seg[i]) * face + sl[i])];
            if (c != 1) fail("a shared node reads a slot no segment wrote for it, or reads it twice");
            c = 2;
            ++slots_read;
          }
        }
      }
    }
    if (visited != count) fail("the shared-node grid misses or repeats nodes");
    for (long long node = 0; node < N; ++node)
      if (direct[node] + shared[node] != 1) fail("a node is not written exactly once");
    if (slots_read != slots_written) fail("a written slot is never read");
    stats[0] = nseg;
    stats[1] = g.lz;
    stats[2] = count;
    stats[3] = slots_written;
    stats[4] = direct_nodes;
    stats[5] = max_owners;
  });
}

int fem_hip_block_plan_stats(int ne, int es, int l_size, const int* offsets, int nb, int max_nodes, int num_comp,
                             int comp_stride, long long* stats) {
  return guard([&] {
    need(offsets, "offsets");
    need(stats, "stats");
    require(ne > 0 && es > 0 && l_size > 0 && nb > 0 && max_nodes >= 0 && num_comp > 0 && comp_stride >= 0,
            "block_plan_stats: bad sizes");
    const size_t nt = static_cast<size_t>(ne) * es;
    for (size_t t = 0; t < nt; ++t)
      require(offsets[t] >= 0 && offsets[t] < l_size, "block_plan_stats: offsets must be in [0, l_size - 1]");
    BlockPlan p = build_block_plan(offsets, ne, es, l_size, nb, max_nodes);
    auto fail = [](const char* m) { throw std::runtime_error(std::string("block plan: ") + m); };
    std::vector<int> seen(ne, 0), mult(l_size, 0), in_blk(l_size, 0), owner(l_size, -1);
    for (size_t t = 0; t < nt; ++t) ++mult[offsets[t]];
    long long uniq = 0, inter = 0;
    for (int b = 0; b < p.nblk; ++b) {
      const int n0 = p.node_ptr[b], nn = p.node_ptr[b + 1] - n0;
      if (max_nodes && nn > max_nodes) fail("a block exceeds max_nodes");
      if (p.part_ptr[b + 1] - p.part_ptr[b] != nn - p.ninterior[b]) fail("partial slots != boundary nodes");
      uniq += nn;
      inter += p.ninterior[b];
      for (int j = 0; j < nn; ++j) in_blk[p.nodes[n0 + j]] = 0;
      for (int j = 0; j < nb; ++j) {
        const size_t sl = static_cast<size_t>(b) * nb + j;
        const int e = p.elem[sl];
        if (j >= p.nelem[b]) {
          if (e != -1) fail("padding slot holds an element");
          continue;
        }
        if (e < 0 || e >= ne) fail("bad element id");
        if (p.color[sl] >= p.ncolor[b]) fail("colour out of range");
        ++seen[e];
        for (int i = 0; i < es; ++i) {
          const int l = p.loc[sl * es + i];
          if (l >= nn || p.nodes[n0 + l] != offsets[static_cast<size_t>(e) * es + i])
            fail("local index does not map back to the offset");
          ++in_blk[p.nodes[n0 + l]];
        }
      }
      // Slots of one colour share no node; interior nodes are referenced only here.
      for (int j = 0; j < nn; ++j) {
        const int node = p.nodes[n0 + j];
        if ((j < p.ninterior[b]) != (in_blk[node] == mult[node])) fail("interior/boundary split is wrong");
        if (j < p.ninterior[b]) {
          if (owner[node] >= 0) fail("node interior to two blocks");
          owner[node] = b;
        }
      }
      for (int j = 0; j < p.nelem[b]; ++j)
        for (int k = j + 1; k < p.nelem[b]; ++k) {
          const size_t sj = static_cast<size_t>(b) * nb + j, sk = static_cast<size_t>(b) * nb + k;
          if (p.color[sj] != p.color[sk]) continue;
          for (int i = 0; i < es; ++i)
            for (int i2 = 0; i2 < es; ++i2)
              if (p.loc[sj * es + i] == p.loc[sk * es + i2]) fail("two slots of one colour share a node");
        }
    }
    for (int e = 0; e < ne; ++e)
      if (seen[e] != 1) fail("element not in exactly one block");
    // Second-pass rows: each boundary (block, node) appears once, rows ascending,
    // partial slots ascending within a row and belonging to that node.
    PartialRows r = build_partial_rows(p, l_size, num_comp, comp_stride);
    std::vector<int> part_node(static_cast<size_t>(p.part_ptr.back()), -1);
    for (int b = 0; b < p.nblk; ++b)
      for (int j = p.ninterior[b]; j < p.node_ptr[b + 1] - p.node_ptr[b]; ++j)
        part_node[p.part_ptr[b] + j - p.ninterior[b]] = p.nodes[p.node_ptr[b] + j];
    std::vector<int> used(part_node.size(), 0);
    for (size_t m = 0; m < r.node.size(); ++m) {
      if (m && r.node[m] <= r.node[m - 1]) fail("rows not ascending");
      if (owner[r.node[m]] >= 0) fail("a row for an interior node");
      for (int k = r.ptr[m]; k < r.ptr[m + 1]; ++k) {
        if (k > r.ptr[m] && r.part[k] <= r.part[k - 1]) fail("row partials not ascending");
        if (part_node[r.part[k]] != r.node[m]) fail("partial of another node in a row");
        ++used[r.part[k]];
      }
    }
    for (int u : used)
      if (u != 1) fail("a partial not used exactly once");
    const long long v[9] = {p.nblk, p.max_nodes, uniq, inter, uniq - inter, static_cast<long long>(nt),
                            p.max_colors, static_cast<long long>(r.node.size()), r.covers_all ? 1 : 0};
    std::memcpy(stats, v, sizeof(v));
  });
}

/* ---- solve ---- */
int fem_hip_mass_solve(fem_hip_operator h, const double* b, double* x, const fem_hip_solve_options* opt,
                       fem_hip_solve_stats* stats, void* s) {
  return guard([&] {
    need(h, "operator");
    need(b, "b");
    need(x, "x");
    SolveOptions o;
    if (opt) {
      o.mode = opt->mode;
      o.max_iter = opt->max_iter;
      o.max_outer = opt->max_outer;
      o.rtol = opt->rtol;
      o.inner_rtol = opt->inner_rtol;
      o.cg_variant = opt->cg_variant;
      o.separate_second_pass = opt->separate_second_pass;
    }
    SolveStats out;
    mass_solve(*h->op, h->solver, b, x, o, out, st(s));
    if (stats) {
      stats->iterations = out.iterations;
      stats->outer = out.outer;
      stats->rel_residual = out.rel_residual;
      stats->ms = out.ms;
    }
  });
}

int fem_hip_mass_solve_batched(int count, const fem_hip_operator* ops, const double* const* b, double* const* x,
                               const fem_hip_solve_options* opts, int iters_per_launch, fem_hip_solve_stats* stats,
                               int* status) {
  return guard([&] {
    require(count >= 0, "mass_solve_batched: count must be >= 0");
    if (count == 0) return;
    need(ops, "ops");
    need(b, "b");
    need(x, "x");
    need(opts, "opts");
    for (int i = 0; i < count; ++i) {
      need(ops[i], "ops[i]");
      need(b[i], "b[i]");
      need(x[i], "x[i]");
      for (int k = 0; k < i; ++k)
        require(ops[k] != ops[i], "mass_solve_batched: an operator may appear once (its element outputs are per solve)");
    }
    HIP_CHECK(hipDeviceSynchronize());  // b and x written on any stream are complete
    std::vector<std::unique_ptr<femhip::BatchJob, void (*)(femhip::BatchJob*)>> jobs;
    for (int i = 0; i < count; ++i) {
      SolveOptions o;
      o.mode = opts[i].mode;
      o.max_iter = opts[i].max_iter;
      o.max_outer = opts[i].max_outer;
      o.rtol = opts[i].rtol;
      o.inner_rtol = opts[i].inner_rtol;
      jobs.emplace_back(femhip::batch_job_create(*ops[i]->op, b[i], x[i], o, nullptr), femhip::batch_job_destroy);
    }
    std::unique_ptr<femhip::BatchEngine, void (*)(femhip::BatchEngine*)> eng(
        femhip::batch_engine_create(iters_per_launch > 0 ? iters_per_launch : 8), femhip::batch_engine_destroy);
    for (auto& j : jobs) femhip::batch_engine_add(*eng, j.get());
    std::vector<femhip::BatchJob*> done;
    while (done.size() < jobs.size()) {
      if (!femhip::batch_engine_poll(*eng, done)) std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
    for (int i = 0; i < count; ++i) {
      const femhip::BatchJob& j = *jobs[i];
      const SolveStats& s = femhip::batch_job_stats(j);
      if (stats) {
        stats[i].iterations = s.iterations;
        stats[i].outer = s.outer;
        stats[i].rel_residual = s.rel_residual;
        stats[i].ms = s.ms;
      }
      if (status) status[i] = femhip::batch_job_status(j);
    }
    eng.reset();  // waits for the cohorts' streams
    for (int i = 0; i < count; ++i) {
      const femhip::BatchJob& j = *jobs[i];
      const std::string msg = "mass_solve_batched: solve " + std::to_string(i) + ": " + femhip::batch_job_error(j);
      switch (femhip::batch_job_status(j)) {
        case 0: continue;
        case 1: throw std::invalid_argument(msg);
        case 2: throw femhip::DeviceError(msg);
        case 4: throw femhip::Unsupported(msg);
        case 5: throw femhip::OutOfMemory(msg);
        default: throw std::runtime_error(msg);
      }
    }
  });
}

}  // extern "C"

/* ---- serving loop ---- */
struct fem_hip_server_s {
  femhip::MassServer* s = nullptr;
  ~fem_hip_server_s() { femhip::server_destroy(s); }
};

namespace {
femhip::ServeRequest to_req(const fem_hip_serve_request& r) {
  femhip::ServeRequest q;
  q.n = r.n;
  q.P = r.P;
  q.Q = r.Q;
  q.num_comp = r.num_comp;
  q.mode = r.mode;
  q.jitter = r.jitter;
  q.rtol = r.rtol;
  q.mesh_seed = r.mesh_seed;
  q.rhs_seed = r.rhs_seed;
  return q;
}
}  // namespace

int fem_hip_server_create(const fem_hip_server_options* o, fem_hip_server* out) {
  return guard([&] {
    need(out, "out");
    femhip::ServerOptions so;
    if (o) {
      so.max_active = o->max_active;
      so.workers = o->workers;
      so.hbm_fraction = o->hbm_fraction;
      so.batch_iters = o->batch_iters;
      so.autotune = o->autotune != 0;
      so.keep_x = o->keep_x != 0;
      so.cheap_setup = o->cheap_setup != 0;
      so.batched = o->batched != 0;
      if (o->batch_max_elem_nodes > 0) so.batch_max_nodes = o->batch_max_elem_nodes;
    }
    auto h = std::make_unique<fem_hip_server_s>();
    h->s = femhip::server_create(so);
    *out = h.release();
  });
}
int fem_hip_server_submit(fem_hip_server s, const fem_hip_serve_request* req, long long* id) {
  return guard([&] {
    need(s, "server");
    need(req, "request");
    need(id, "id");
    *id = femhip::server_submit(*s->s, to_req(*req));
  });
}
int fem_hip_server_wait(fem_hip_server s, long long id, double timeout_ms, fem_hip_serve_result* res, double* x,
                        size_t x_len) {
  int rc = 0;
  const int g = guard([&] {
    need(s, "server");
    need(res, "result");
    femhip::ServeResult r;
    if (!femhip::server_wait(*s->s, id, timeout_ms, &r)) {
      rc = FEM_HIP_ERR_NOT_READY;
      return;
    }
    res->status = r.status;
    res->iterations = r.iterations;
    res->outer = r.outer;
    res->rel_residual = r.rel_residual;
    res->queue_ms = r.queue_ms;
    res->setup_ms = r.setup_ms;
    res->solve_ms = r.solve_ms;
    res->latency_ms = r.latency_ms;
    res->l_size = r.l_size;
    res->est_bytes = r.est_bytes;
    res->build_ms = r.build_ms;
    res->graph_ms = r.graph_ms;
    if (r.status != 0) {
      rc = r.status;
      throw std::runtime_error(r.error);
    }
    if (x) {
      require(x_len == r.x.size(), "server_wait: x_len must equal the solution length (keep_x on)");
      std::memcpy(x, r.x.data(), x_len * sizeof(double));
    }
  });
  if (rc == FEM_HIP_ERR_NOT_READY) g_last_error = "server_wait: timed out";
  return rc ? rc : g;
}
int fem_hip_server_stats(fem_hip_server s, long long* stats) {
  return guard([&] {
    need(s, "server");
    need(stats, "stats");
    const femhip::ServerStats st = femhip::server_stats(*s->s);
    const long long v[9] = {st.submitted, st.completed, st.failed, st.active, st.queued, st.active_peak,
                            static_cast<long long>(st.budget), static_cast<long long>(st.used),
                            static_cast<long long>(st.used_peak)};
    std::memcpy(stats, v, sizeof(v));
  });
}
int fem_hip_server_destroy(fem_hip_server s) {
  return guard([&] { delete s; });
}
int fem_hip_serve_request_bytes(const fem_hip_serve_request* req, size_t* bytes) {
  return guard([&] {
    need(req, "request");
    need(bytes, "bytes");
    *bytes = femhip::serve_request_bytes(to_req(*req));
  });
}

/* ---- on-disk meshes ---- */
int fem_hip_mesh_read(const char* path, fem_hip_mesh_info* info, int* offsets, int* vertex_offsets, double* coords) {
  return guard([&] {
    need(path, "path");
    need(info, "info");
    const femhip::MeshData m = femhip::mesh_read(path);
    info->dim = m.dim;
    info->num_elem = m.num_elem;
    info->P = m.P;
    info->num_nodes = m.num_nodes;
    info->num_vertices = m.num_vertices;
    if (offsets) std::memcpy(offsets, m.offsets.data(), m.offsets.size() * sizeof(int));
    if (vertex_offsets) std::memcpy(vertex_offsets, m.vertex_offsets.data(), m.vertex_offsets.size() * sizeof(int));
    if (coords) std::memcpy(coords, m.coords.data(), m.coords.size() * sizeof(double));
  });
}
int fem_hip_mesh_write(const char* path, const fem_hip_mesh_info* info, const int* offsets,
                       const int* vertex_offsets, const double* coords) {
  return guard([&] {
    need(path, "path");
    need(info, "info");
    need(offsets, "offsets");
    need(vertex_offsets, "vertex_offsets");
    need(coords, "coords");
    require(info->dim >= 1 && info->dim <= 3 && info->num_elem > 0 && info->P >= 2 && info->num_vertices > 0,
            "mesh_write: bad sizes");
    femhip::MeshData m;
    m.dim = info->dim;
    m.num_elem = info->num_elem;
    m.P = info->P;
    m.num_nodes = info->num_nodes;
    m.num_vertices = info->num_vertices;
    long long es = 1, vs = 1;
    for (int d = 0; d < m.dim; ++d) {
      es *= m.P;
      vs *= 2;
    }
    m.offsets.assign(offsets, offsets + m.num_elem * es);
    m.vertex_offsets.assign(vertex_offsets, vertex_offsets + m.num_elem * vs);
    m.coords.assign(coords, coords + static_cast<size_t>(m.dim) * m.num_vertices);
    femhip::mesh_write(path, m);
  });
}
