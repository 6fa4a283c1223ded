// extern "C" boundary of libfemhip (declared in include/fem_hip.h). Every entry
// point catches, maps the exception type to a status code and records the
// message for fem_hip_last_error() on the calling thread.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/fem_hip.h"
#include "basis.hpp"
#include "basis_setup.hpp"
#include "block_plan.hpp"
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
    if (bytes) HIP_CHECK(hipMalloc(ptr, bytes));
  });
}
int fem_hip_free(void* ptr) {
  return guard([&] {
    if (ptr) HIP_CHECK(hipFree(ptr));
  });
}
int fem_hip_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}
int fem_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    if (bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}
int fem_hip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* s) {
  return guard([&] {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st(s)));
  });
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
int fem_hip_operator_set_fused_options(fem_hip_operator h, int scatter, int exact, int elems) {
  return guard([&] {
    need(h, "operator");
    require(scatter == -1 || scatter == kScatterAtomic || scatter == kScatterDeterministic,
            "set_fused_options: unknown scatter mode");
    require(elems >= 0, "set_fused_options: elems_per_block must be >= 0");
    h->op->opt_scatter = scatter;
    h->op->opt_exact = exact ? 1 : 0;
    h->op->opt_elems = elems;
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
int fem_hip_box_offsets(int nx, int ny, int nz, int P, int node_stride, int* offsets) {
  return guard([&] {
    need(offsets, "offsets");
    box_offsets(nx, ny, nz, P, node_stride, offsets);
  });
}
int fem_hip_box_coords(int n, double jitter, uint64_t seed, double* coords) {
  return guard([&] {
    need(coords, "coords");
    box_coords(n, jitter, seed, coords);
  });
}

int fem_hip_block_plan_stats(int ne, int es, int l_size, const int* offsets, int nb, long long* stats) {
  return guard([&] {
    need(offsets, "offsets");
    need(stats, "stats");
    require(ne > 0 && es > 0 && l_size > 0 && nb > 0, "block_plan_stats: bad sizes");
    const size_t nt = static_cast<size_t>(ne) * es;
    for (size_t t = 0; t < nt; ++t)
      require(offsets[t] >= 0 && offsets[t] < l_size, "block_plan_stats: offsets must be in [0, l_size - 1]");
    BlockPlan p = build_block_plan(offsets, ne, es, l_size, nb);
    std::vector<int> seen(ne, 0);
    long long uniq = 0, inter = 0;
    for (int b = 0; b < p.nblk; ++b) {
      const int n0 = p.node_ptr[b], nn = p.node_ptr[b + 1] - n0;
      uniq += nn;
      inter += p.ninterior[b];
      for (int j = 0; j < nb; ++j) {
        const int e = p.elem[static_cast<size_t>(b) * nb + j];
        if (j >= p.nelem[b]) {
          if (e != -1) throw std::runtime_error("block plan: padding slot holds an element");
          continue;
        }
        if (e < 0 || e >= ne) throw std::runtime_error("block plan: bad element id");
        ++seen[e];
        for (int i = 0; i < es; ++i) {
          const int l = p.loc[(static_cast<size_t>(b) * nb + j) * es + i];
          if (l >= nn || p.nodes[n0 + l] != offsets[static_cast<size_t>(e) * es + i])
            throw std::runtime_error("block plan: local index does not map back to the offset");
        }
      }
    }
    for (int e = 0; e < ne; ++e)
      if (seen[e] != 1) throw std::runtime_error("block plan: element not in exactly one block");
    const long long v[6] = {p.nblk, p.max_nodes, uniq, inter, uniq - inter, static_cast<long long>(nt)};
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

}  // extern "C"
