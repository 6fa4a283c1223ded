/*
 * fem_hip.h -- C ABI of the MI355X (gfx950) matrix-free mass-operator library
 * (libfemhip.so). Plain pointers and sizes only; every device pointer is a HIP
 * device allocation (from fem_hip_malloc or any other hipMalloc), every stream
 * a hipStream_t (NULL = the default stream).
 *
 * Each entry point returns FEM_HIP_SUCCESS (0) or an error code; the message of
 * the last failure on the calling thread is fem_hip_last_error(). The C++ layer
 * in fe-demo-cpp_amd/cpp/fem_hip.hpp turns the codes back into the reference's
 * exceptions (std::invalid_argument / std::runtime_error).
 *
 * Reference interfaces replaced (paths relative to the fe-demo-cpp tree):
 *   include/cuda/cuda-check.hpp:17-32          -> status codes + fem_hip_last_error
 *   include/cuda/device-array.hpp:18-95        -> fem_hip_malloc/free/memcpy/memset
 *   include/cuda/device-info.hpp:8-25          -> fem_hip_device_count/_info
 *   include/cuda/cuda-elem-restriction.hpp:21-79 -> fem_hip_restriction_*
 *   include/elem-restriction.hpp:43-57         -> fem_hip_restriction_* (same semantics)
 *   include/basis.hpp:7-40                     -> fem_hip_gauss_*, fem_hip_lagrange_*, fem_hip_basis_create_*
 *   include/tensor-contract.hpp:9-53           -> fem_hip_contract, fem_hip_basis_apply
 *   include/qfunction.hpp:20-61, qfunctions/mass.hpp:8-51 -> fem_hip_qfunction_*
 *   include/operator.hpp:15-64                 -> fem_hip_operator_*
 */
#ifndef FEM_HIP_H
#define FEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define FEM_HIP_SUCCESS 0
#define FEM_HIP_ERR_INVALID_ARGUMENT 1 /* reference: std::invalid_argument */
#define FEM_HIP_ERR_DEVICE 2           /* a HIP runtime call failed (reference: CUDA_CHECK runtime_error) */
#define FEM_HIP_ERR_RUNTIME 3          /* any other std::runtime_error */
#define FEM_HIP_ERR_UNSUPPORTED 4      /* e.g. a QFunction with no device implementation */
#define FEM_HIP_ERR_OUT_OF_MEMORY 5    /* device allocation failed, or admission refused */
#define FEM_HIP_ERR_NOT_READY 6        /* a wait timed out (the request is still in flight) */

/* ---- vocabulary (include/transpose-mode.hpp:13, include/eval-mode.hpp:11) --- */
#define FEM_HIP_NO_TRANSPOSE 0
#define FEM_HIP_TRANSPOSE 1
#define FEM_HIP_EVAL_NONE 0
#define FEM_HIP_EVAL_INTERP 1
#define FEM_HIP_EVAL_GRAD 2
#define FEM_HIP_EVAL_WEIGHT 3
#define FEM_HIP_QUAD_GAUSS 0 /* include/basis.hpp:23 QuadMode */
#define FEM_HIP_QUAD_GAUSS_LOBATTO 1
#define FEM_HIP_SCATTER_ATOMIC 0 /* include/cuda/cuda-elem-restriction.hpp:21 ScatterMode */
#define FEM_HIP_SCATTER_DETERMINISTIC 1
#define FEM_HIP_VECTOR_ACTIVE 0 /* include/operator.hpp:15 VectorKind */
#define FEM_HIP_VECTOR_PASSIVE 1
#define FEM_HIP_VECTOR_NONE 2
/* QFunctions with a device implementation (include/qfunctions/mass.hpp). */
#define FEM_HIP_QF_BUILD_MASS 1        /* ctx: int dim */
#define FEM_HIP_QF_APPLY_MASS 2        /* size-1 fields */
#define FEM_HIP_QF_APPLY_MASS_VECTOR 3 /* ctx: int num_comp; v[c][p] = qdata[p] u[c][p] */
/* Diffusion (Poisson), no reference counterpart; gradient fields [d][comp][Q]:
 * build: inputs dx (dim*dim, Grad), weights (1, Weight); output qdata
 *   (dim*(dim+1)/2, None) = w det(J) J^-1 J^-T, symmetric, stored
 *   (00),(11),(22),(12),(02),(01) for dim 3; (00),(11),(01) for dim 2. ctx: int dim.
 * apply: inputs du (dim*num_comp, Grad), qdata; output dv (dim*num_comp, Grad),
 *   dv[a][c] = sum_b qdata(a,b) du[b][c]. ctx: int dim, int num_comp. */
#define FEM_HIP_QF_BUILD_DIFFUSION 4
#define FEM_HIP_QF_APPLY_DIFFUSION 5
/* Arithmetic of the fused mass apply. */
#define FEM_HIP_FP64 0
#define FEM_HIP_FP32 1

const char* fem_hip_last_error(void);
const char* fem_hip_version(void);

/* ---- device and memory (include/cuda/device-info.hpp, device-array.hpp) ---- */
typedef struct {
  int id;
  char name[256];
  char arch[64];         /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int cu_count;          /* compute units */
  int wavefront_size;    /* 64 on CDNA */
  int max_threads_per_block;
  int regs_per_block;
  size_t lds_per_block;  /* bytes */
  size_t lds_per_cu;     /* bytes */
  size_t l2_bytes;
  size_t total_global_mem;
  size_t free_global_mem; /* at the time of the query (hipMemGetInfo) */
  int clock_khz;
  int memory_clock_khz;
  int memory_bus_width;
} fem_hip_device_info;

int fem_hip_device_count(int* count); /* 0 devices is not an error */
int fem_hip_device_info_get(int device, fem_hip_device_info* info);
int fem_hip_malloc(void** ptr, size_t bytes); /* bytes == 0 -> *ptr = NULL */
int fem_hip_free(void* ptr);
int fem_hip_memcpy_htod(void* dst, const void* src, size_t bytes);
int fem_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int fem_hip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int fem_hip_memset_zero(void* dst, size_t bytes, void* stream); /* a fill kernel, graph-capture safe */
/* dst = src by a streaming copy kernel (one 16-byte nontemporal load and store
 * per thread); pointers and bytes must be multiples of 16. No counterpart in
 * the reference: it measures the device's streaming bandwidth, the ceiling the
 * fused kernels are compared with. */
int fem_hip_stream_copy(void* dst, const void* src, size_t bytes, void* stream);
int fem_hip_device_synchronize(void);
int fem_hip_stream_create(void** stream);
int fem_hip_stream_destroy(void* stream);
int fem_hip_stream_synchronize(void* stream);
int fem_hip_event_create(void** event);
int fem_hip_event_destroy(void* event);
int fem_hip_event_record(void* event, void* stream);
int fem_hip_event_synchronize(void* event);
int fem_hip_event_elapsed_ms(void* start, void* stop, float* ms);
/* Error paths forced from the host (tests/cuda/test-device.cu:104-122): launch
 * a trivial kernel with `threads` threads per block; returns the launch status. */
int fem_hip_probe_launch(int blocks, int threads);

/* ---- 1D basis setup, host only (src/basis.cpp:16-228) --------------------- */
int fem_hip_gauss_legendre(int Q, double* q_ref_1d, double* q_weight_1d);
int fem_hip_gauss_lobatto(int Q, double* q_ref_1d, double* q_weight_1d /* nullable */);
int fem_hip_lagrange_basis_matrix(int P, const double* nodes, int M, const double* eval_points, double* interp_1d,
                                  double* grad_1d);

/* ---- element restriction (include/cuda/cuda-elem-restriction.hpp:46-79) ---- */
typedef struct fem_hip_restriction_s* fem_hip_restriction;

/* Offset restriction; offsets is a host array [num_elem][elem_size], validated
 * like src/elem-restriction.cpp:102-138 and copied to the device once. With
 * FEM_HIP_SCATTER_DETERMINISTIC the transpose map is built here too. */
int fem_hip_restriction_create(int num_elem, int elem_size, int num_comp, int comp_stride, int l_size,
                               const int* offsets, int scatter_mode, fem_hip_restriction* out);
/* Strided restriction; strides == NULL means backend strides {num_elem, elem_size*num_elem, 1}
 * (src/elem-restriction.cpp:155-204). */
int fem_hip_restriction_create_strided(int num_elem, int elem_size, int num_comp, int l_size, const int* strides,
                                       fem_hip_restriction* out);
/* NoTranspose: in = L-vector (in_len == l_size), out = E-vector overwritten (out_len == e_size).
 * Transpose:   in = E-vector (in_len == e_size), out = L-vector ADDED INTO (out_len == l_size).
 * Asynchronous on `stream`. */
int fem_hip_restriction_apply(fem_hip_restriction r, int t_mode, const double* in, size_t in_len, double* out,
                              size_t out_len, void* stream);
int fem_hip_restriction_apply_f32(fem_hip_restriction r, int t_mode, const float* in, size_t in_len, float* out,
                                  size_t out_len, void* stream);
/* sizes[8] = {type(0 offset,1 strided), num_elem, elem_size, num_comp, comp_stride, l_size, e_size, num_nodes} */
int fem_hip_restriction_get_sizes(fem_hip_restriction r, long long* sizes);
/* Copies the device transpose map back (deterministic offset restrictions only):
 * l_vec_indices[num_nodes], t_offsets[num_nodes+1], t_indices[num_elem*elem_size]. */
int fem_hip_restriction_get_transpose_map(fem_hip_restriction r, int* l_vec_indices, int* t_offsets, int* t_indices);
int fem_hip_restriction_destroy(fem_hip_restriction r);
/* A restriction whose device index arrays the caller owns: the shape of the
 * reference's fem::cuda::ElemRestriction (include/cuda/cuda-elem-restriction.hpp:46-58),
 * which holds its sizes and DeviceArray<int> offsets, l_vec_indices, t_offsets,
 * t_indices itself. Every pointer is device memory the caller keeps alive
 * during the apply; offsets (offset restrictions) hold num_elem*elem_size
 * entries in [0, l_size) less the component offsets, and the transpose map
 * (deterministic offset restrictions) holds num_nodes rows as
 * fem_hip_transpose_map_build makes them. */
typedef struct {
  int type; /* 0 offset, 1 strided (include/elem-restriction.hpp:8 RestrictionType) */
  int num_elem, elem_size, num_comp, comp_stride, l_size;
  long long e_size;
  int strides[3];   /* strided: (node, component, element) strides */
  int scatter_mode; /* offset: FEM_HIP_SCATTER_* */
  const int* offsets;
  int num_nodes;
  const int* l_vec_indices;
  const int* t_offsets;
  const int* t_indices;
} fem_hip_restriction_arrays;
/* elem_restriction_apply on such a restriction (src/cuda/cuda-elem-restriction.cu:228-273):
 * the sizes and semantics of fem_hip_restriction_apply, nothing allocated, asynchronous. */
int fem_hip_restriction_apply_arrays(const fem_hip_restriction_arrays* r, int t_mode, const double* in,
                                     size_t in_len, double* out, size_t out_len, void* stream);
/* Host-only transpose map (src/cuda/cuda-elem-restriction.cu:149-177). Arrays must
 * hold l_size, l_size + 1 and num_elem*elem_size ints; *num_nodes receives the row count. */
int fem_hip_transpose_map_build(int num_elem, int elem_size, int l_size, const int* offsets, int* num_nodes,
                                int* l_vec_indices, int* t_offsets, int* t_indices);

/* ---- tensor basis (include/basis.hpp:28-40, include/tensor-contract.hpp) ---- */
typedef struct fem_hip_basis_s* fem_hip_basis;

int fem_hip_basis_create_h1_lagrange(int dim, int num_comp, int P_1d, int Q_1d, int quad_mode, fem_hip_basis* out);
/* From an existing host TensorBasis: interp_1d/grad_1d are [Q_1d][P_1d] row-major. */
int fem_hip_basis_create_tensor(int dim, int num_comp, int P_1d, int Q_1d, const double* q_ref_1d,
                                const double* q_weight_1d, const double* interp_1d, const double* grad_1d,
                                fem_hip_basis* out);
/* dims[4] = {dim, num_comp, P_1d, Q_1d}; any of the arrays may be NULL. */
int fem_hip_basis_get(fem_hip_basis b, int* dims, double* q_ref_1d, double* q_weight_1d, double* interp_1d,
                      double* grad_1d);
/* tensor_basis_apply (src/tensor-contract.cpp:207-222) on device E/Q-vectors in the
 * reference layout [comp][node][elem] / [dim][comp][qpt][elem]. v_len must be the
 * exact output size. Grad with Transpose overwrites v with the sum of the dim
 * blocks (the reference accumulates into whatever v held). */
int fem_hip_basis_apply(fem_hip_basis b, int num_elem, int t_mode, int eval_mode, const double* u, size_t u_len,
                        double* v, size_t v_len, void* stream);
int fem_hip_basis_destroy(fem_hip_basis b);
/* tensor_contract_apply (src/tensor-contract.cpp:35-59) on device arrays; t is a
 * device [J][B] (NoTranspose) or [B][J] (Transpose) matrix. */
int fem_hip_contract(int A, int B, int C, int J, const double* t, int t_mode, int add, const double* u, double* v,
                     void* stream);

/* ---- QFunction (include/qfunction.hpp:20-61) -------------------------------- */
typedef struct fem_hip_qfunction_s* fem_hip_qfunction;

int fem_hip_qfunction_create(int qf_kind, fem_hip_qfunction* out);
int fem_hip_qfunction_add_input(fem_hip_qfunction qf, const char* name, int size, int eval_mode);
int fem_hip_qfunction_add_output(fem_hip_qfunction qf, const char* name, int size, int eval_mode);
int fem_hip_qfunction_set_context(fem_hip_qfunction qf, const void* data, size_t bytes);
/* in[f] / out[f] are device arrays of inputs[f].size*Q / outputs[f].size*Q doubles. */
int fem_hip_qfunction_apply(fem_hip_qfunction qf, int Q, int num_in, const double* const* in, const size_t* in_lens,
                            int num_out, double* const* out, const size_t* out_lens, void* stream);
int fem_hip_qfunction_destroy(fem_hip_qfunction qf);

/* ---- operator (include/operator.hpp:34-64) ------------------------------------ */
typedef struct fem_hip_operator_s* fem_hip_operator;

/* The QFunction, restrictions and bases are copied (by value, as the reference
 * stores them); the handles may be destroyed afterwards. */
int fem_hip_operator_create(fem_hip_qfunction qf, fem_hip_operator* out);
/* restriction / basis may be NULL ("std::nullopt"). For FEM_HIP_VECTOR_PASSIVE,
 * passive_data holds passive_len doubles (host memory if passive_on_device == 0). */
int fem_hip_operator_set_field(fem_hip_operator op, const char* name, fem_hip_restriction restriction,
                               fem_hip_basis basis, int vector_kind, const double* passive_data, size_t passive_len,
                               int passive_on_device);
int fem_hip_operator_check_ready(fem_hip_operator op);
/* out = A(in): out (out_len == active_out_l_size) is zeroed, then summed into.
 * Recognised mass operators run the fused per-element kernel. */
int fem_hip_operator_apply(fem_hip_operator op, const double* in, size_t in_len, double* out, size_t out_len,
                           void* stream);
/* info[8] = {num_elem, num_qpts, active_in_l_size, active_out_l_size, fused(0 none,1 mass,2 build),
 *            P_1d, Q_1d, num_comp} */
int fem_hip_operator_get_info(fem_hip_operator op, long long* info);
/* Fused mass apply options: scatter (FEM_HIP_SCATTER_*; -1 = the restriction's
 * mode), exact (1 = no fused multiply-add, element sums in the reference order;
 * 0 = FMA), elems_per_block (0 = automatic algorithm choice; > 0 = the element
 * kernels of the scatter mode with that many elements per workgroup). Takes
 * effect on the next apply. Deterministic results are the same bit for bit on
 * every run; with exact = 1 they equal the reference's CPU operator_apply. */
int fem_hip_operator_set_fused_options(fem_hip_operator op, int scatter_mode, int exact, int elems_per_block);

/* How a fused mass apply reaches v (no counterpart in the reference; its CUDA
 * restriction offers the atomic and transpose-map scatters,
 * include/cuda/cuda-elem-restriction.hpp:21). */
#define FEM_HIP_FUSED_AUTO -1         /* tuned table, else the default order */
#define FEM_HIP_FUSED_ELEM_ATOMIC 0   /* per element, one atomic add per shared node */
#define FEM_HIP_FUSED_ELEM_TMAP 1     /* per element into element outputs, then the transpose-map sum (deterministic) */
#define FEM_HIP_FUSED_BLOCK_ATOMIC 2  /* element blocks summed in LDS, boundary nodes one atomic add each */
#define FEM_HIP_FUSED_BLOCK_PARTIAL 3 /* element blocks summed in LDS, boundary partials summed by a second pass
                                         (deterministic) */
/* 4 and 6 are retired (persistent software-pipelined forms of 1 and 5 that
 * never measured faster); selecting them reports "not compiled". */
#define FEM_HIP_FUSED_BLOCK_PARTIAL_LDS 5 /* BLOCK_PARTIAL summing in LDS with atomics (any order) */
#define FEM_HIP_FUSED_BOX_COLUMN 7 /* structured boxes (fem_hip_box_detect): element columns marching in z with node
                                      planes in LDS, shared column faces summed by a second pass (deterministic;
                                      no index arrays read). elems = the column cross-section (16: 4x4 elements,
                                      8: 4x2, 4: 2x2; for the wavefront kernels also 18: 8x2, 32: 8x4, 34: 16x2)
                                      plus the kernel: + 0 qdata loaded a layer ahead, + 100 qdata within its own
                                      layer, + 200 (FP64) u through an LDS ring of node planes, + 300 one component
                                      per workgroup (several components), + 400 / + 600 one / two elements per
                                      wavefront (+ 500 / + 700 with one component per workgroup). Which are
                                      compiled depends on the shape: fem_hip_operator_autotune lists them. */
/* Explicit algorithm and its elements per workgroup (element kernels) or per
 * block (block kernels); 0 = that algorithm's default. A deterministic scatter
 * mode accepts only the deterministic algorithms. exact = 1 overrides. */
int fem_hip_operator_set_fused_algorithm(fem_hip_operator op, int alg, int elems);
/* Most element layers per workgroup of FEM_HIP_FUSED_BOX_COLUMN: each column
 * of the box is cut into ceil(nz / layers) segments of floor or ceil of
 * nz / segments layers (0 = automatic: the segment count that fills every
 * workgroup slot the kernel has on the device, by its occupancy, in the fewest
 * layer steps). The box-column algorithm is unavailable when a segment would
 * hold fewer than 2 node planes beyond its first. */
int fem_hip_operator_set_box_layers(fem_hip_operator op, int layers);
/* The algorithm and elems the next apply in this precision (fp32 = 0/1) will run. */
int fem_hip_operator_get_fused_choice(fem_hip_operator op, int fp32, int* alg, int* elems);
/* Autotuner: times every compiled algorithm/packing valid for the operator's
 * scatter mode in one precision on the operator's own data (reps applies each,
 * after a warm-up; the operator's output is not touched), records the fastest
 * for (P, Q, num_comp, fp32, deterministic, box) in the process-wide tuning table,
 * which automatic choice then follows, and returns it. results (may be NULL)
 * receives up to max_results rows {alg, elems, ms_per_apply} of every candidate;
 * *n_results is set to the number of candidates. Synchronous. */
int fem_hip_operator_autotune(fem_hip_operator op, int fp32, int reps, void* stream, int* best_alg, int* best_elems,
                              double* best_ms, double* results, int max_results, int* n_results);
/* Tuning table rows {P, Q, num_comp, fp32, deterministic, box, alg, elems} with
 * the time per apply and the element count it was measured on; box = 1 rows
 * serve operators on structured boxes (fem_hip_box_detect), box = 0 rows any
 * operator (a box operator without a box row uses the box = 0 row). A key may
 * hold rows for several mesh sizes; an operator follows the row measured on
 * the size nearest its own element count (by ratio). get: copies up to
 * max_rows rows and sets *n_rows to the table size; set: adds the key's row for
 * num_elem (> 0), replacing one of that size (alg must be compiled for the
 * shape); clear: empties it. */
int fem_hip_tune_table_get(int* rows8, double* ms, long long* num_elem, int max_rows, int* n_rows);
int fem_hip_tune_table_set(const int* row8, double ms, long long num_elem);
int fem_hip_tune_table_clear(void);
/* Save / load the tuning tables as text (the best algorithm per shape,
 * determinism and mesh size, and the fastest packing per algorithm). Load
 * replaces rows with the same key and size, skips rows for kernels this library
 * lacks, and reports the rows taken. */
int fem_hip_tune_table_save(const char* path);
int fem_hip_tune_table_load(const char* path, int* rows_taken);
/* FP32 fused mass apply on FP32 L-vectors (qdata and basis rounded to FP32 once). */
int fem_hip_operator_apply_f32(fem_hip_operator op, const float* in, size_t in_len, float* out, size_t out_len,
                               void* stream);
/* fem_hip_operator_apply with no device allocation. No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). A fused mass operator
 * whose device state exists (made by its first fem_hip_operator_apply) runs
 * its FP64 apply with the per-apply scratch (box-column partials, or the
 * element outputs of the transpose-map scatter) in a library-owned device
 * array of fem_hip_noalloc_scratch_bytes() per GPU (64 MiB, held until the
 * process ends); results equal fem_hip_operator_apply's bit for bit. Rejects
 * an apply needing more scratch (FEM_HIP_ERR_INVALID_ARGUMENT) and any other
 * algorithm, or a missing device state (FEM_HIP_ERR_UNSUPPORTED). Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race and must be ordered by the caller. */
int fem_hip_operator_apply_noalloc(fem_hip_operator op, const double* in, size_t in_len, double* out, size_t out_len,
                                   void* stream);
size_t fem_hip_noalloc_scratch_bytes(void);
/* Jacobi diagonal of a fused mass operator (FP64 L-vector of active_out_l_size). */
int fem_hip_operator_mass_diagonal(fem_hip_operator op, double* diag, size_t len, void* stream);
/* Kernel-level timing of the fused path with HIP events on the apply's stream:
 * between begin and end every fused apply records events around its stages;
 * end synchronizes and returns the summed fused-kernel time, the summed time of
 * the other stages (output zeroing or the transpose-map pass) and the count. */
int fem_hip_operator_timing_begin(fem_hip_operator op);
int fem_hip_operator_timing_end(fem_hip_operator op, double* kernel_ms, double* aux_ms, int* count);
int fem_hip_operator_destroy(fem_hip_operator op);

/* Convenience: the reference mass operator (tests/test-operator.cpp:17-24,75-82)
 * from a node restriction, a basis and device qdata in the backend-strided layout
 * qdata[q*num_elem + e]. */
int fem_hip_operator_create_mass(fem_hip_restriction r, fem_hip_basis b, const double* qdata_device,
                                 fem_hip_operator* out);

/* ---- synthetic structured meshes (host only) ---------------------------------- */
/* splitmix64 stream: out[i] = lo + (hi-lo) * ((x_i >> 11) * 2^-53). */
int fem_hip_rng_uniform(uint64_t seed, size_t n, double lo, double hi, double* out);
/* The same stream, bit for bit, written to device memory out_device[0..n) by a
 * kernel enqueued on stream (element i depends only on seed and i). No
 * counterpart in the reference, which reads its vectors from the host. */
int fem_hip_rng_uniform_device(uint64_t seed, size_t n, double lo, double hi, double* out_device, void* stream);
/* Lexicographic offsets of an nx*ny*nz box of P^3-node hexes on the (nx(P-1)+1)x.. grid,
 * offsets (gx + Nx*(gy + Ny*gz)) * node_stride (tests/cuda/test-cuda-elem-restriction.cu:24-42). */
int fem_hip_box_offsets(int nx, int ny, int nz, int P, int node_stride, int* offsets);
/* The same offsets, written to device memory by a kernel enqueued on stream (the
 * serving loop makes its meshes this way). No counterpart in the reference. */
int fem_hip_box_offsets_device(int nx, int ny, int nz, int P, int node_stride, int* offsets_device, void* stream);
/* Vertex coordinates of an n^3 unit-cube mesh, blocked [x...,y...,z...] over the (n+1)^3
 * vertices, interior vertices moved by jitter*h*U[-1,1) per axis, drawn x,y,z per vertex
 * in lexicographic order from splitmix64(seed). */
int fem_hip_box_coords(int n, double jitter, uint64_t seed, double* coords);
/* The same coordinates, bit for bit, written to device memory (3*(n+1)^3
 * doubles) by a kernel enqueued on stream (the splitmix64 stream in counter
 * form). No counterpart in the reference. */
int fem_hip_box_coords_device(int n, double jitter, uint64_t seed, double* coords_device, void* stream);

/* Host-only statistics and invariant checks of the element-block plan used by
 * the block algorithms: blocks of at most nb elements and max_nodes unique nodes
 * (0 = no limit), greedy clustering on shared nodes. stats[9] = {blocks, max
 * unique nodes in a block, total unique nodes over blocks, total block-interior
 * nodes (plain stores), total block-boundary nodes (atomic adds or partials),
 * element nodes (num_elem*elem_size), most colours in a block, second-pass rows,
 * 1 if interior nodes and rows write every L entry}. Checks: every element in
 * exactly one block, local indices map back to the offsets, slots of one colour
 * share no node, interior nodes belong to one block, each partial is summed once
 * by the row of its node, in ascending order. */
int fem_hip_block_plan_stats(int num_elem, int elem_size, int l_size, const int* offsets, int nb, int max_nodes,
                             int num_comp, int comp_stride, long long* stats);

/* Structured boxes (no counterpart in the reference). is_box = 1 when offsets
 * [num_elem][P^3] are an nx*ny*nz box of P^3-node hexes with lexicographic
 * node numbering (tests/cuda/test-cuda-elem-restriction.cu:24-42 with node
 * stride 1), l_size = num_comp * node count and, for num_comp > 1, comp_stride =
 * node count; n3 = {nx, ny, nz}. Fused mass operators on such restrictions can
 * run FEM_HIP_FUSED_BOX_COLUMN. */
int fem_hip_box_detect(int num_elem, int P, const int* offsets, int num_comp, int comp_stride, int l_size, int* n3,
                       int* is_box);
/* Host-only check of the box-column algorithm's index arithmetic for an
 * nx*ny*nz box with `elems` elements per column layer and `layers` element
 * layers per workgroup (0 = automatic): every node is written
 * once (to v by its segment, or by the shared-node pass), every partial slot is
 * written once and read once, owners ascend. stats[6] = {segments, layers per
 * segment, shared nodes, partial slots, nodes written directly, most owners of a node}. */
int fem_hip_box_column_check(int nx, int ny, int nz, int P, int elems, int layers, long long* stats);

/* ---- mass solve ------------------------------------------------------------ */
#define FEM_HIP_SOLVE_FP64 0  /* Jacobi-PCG, FP64 throughout */
#define FEM_HIP_SOLVE_MIXED 1 /* FP32 inner Jacobi-PCG, FP64 iterative refinement */
#define FEM_HIP_SOLVE_MIXED_RELIABLE 2 /* one FP32 Jacobi-PCG recursion whose residual is replaced by the FP64
                                          b - M x whenever it has dropped by inner_rtol (x accumulated in FP64) */
#define FEM_HIP_CG_AUTO 0      /* fem_hip_solve_options.cg_variant */
#define FEM_HIP_CG_PLAIN 1
#define FEM_HIP_CG_PIPELINED 2
#define FEM_HIP_CG_PLAIN_WIDE 3
typedef struct {
  int mode;          /* FEM_HIP_SOLVE_* */
  int max_iter;      /* CG iterations (FP64) / per inner solve (mixed) */
  int max_outer;     /* refinement steps (mixed) / residual replacements (mixed reliable) */
  double rtol;       /* target ||b - M x|| / ||b|| (FP64 residual) */
  double inner_rtol; /* inner FP32 solve reduction per refinement step / residual drop between
                        replacements; <= 0: 1e-5 (MIXED: two steps to 1e-10) / 1e-2 (MIXED_RELIABLE:
                        keeps the FP64 iteration count). A refinement step never goes below half
                        of what the target still needs. */
  int cg_variant;    /* FEM_HIP_CG_*. 0: automatic (FP64: pipelined CG up to 2^17 unknowns, plain CG above; the
                        FP32 inner solves and MIXED_RELIABLE: plain; the plain form wide on vectors of 96 MiB
                        and more: FP64 from 12.6M unknowns, FP32 from 25.2M);
                        1: plain Jacobi-PCG (apply, dot, update, direction: five dependent launches an
                        iteration); 2: pipelined Jacobi-PCG (Ghysels & Vanroose 2014: apply and one fused
                        kernel, three launches); 3: plain Jacobi-PCG with one 16-byte pack per thread and
                        alpha, beta formed by one-block kernels between the streaming ones (the per-entry
                        arithmetic of 1, another fixed summation order; for long vectors) */
  int separate_second_pass; /* pipelined form: 0 (default) folds the apply's second pass (block partial sums,
                        transpose map) into the step kernel where the apply splits, one launch fewer an
                        iteration with the same arithmetic; 1 keeps it as its own kernel */
} fem_hip_solve_options;
typedef struct {
  int iterations;    /* total CG iterations (inner ones for mixed) */
  int outer;         /* refinement steps (mixed), 0 for FP64 */
  double rel_residual; /* final ||b - M x|| / ||b|| computed in FP64 */
  double ms;         /* wall time of the solve */
} fem_hip_solve_stats;
int fem_hip_mass_solve(fem_hip_operator op, const double* b, double* x, const fem_hip_solve_options* opt,
                       fem_hip_solve_stats* stats, void* stream);
/* No reference counterpart (the reference solves nothing): `count` solves of
 * M_i x_i = b_i on distinct fused mass operators, run together. Solves whose
 * operators share (P, Q, num_comp, basis) and precision (FEM_HIP_SOLVE_FP64, or
 * FEM_HIP_SOLVE_MIXED's FP32 inner solves) form a cohort: each launch of its
 * captured graph advances every one of them by iters_per_launch (<= 0: 8)
 * pipelined Jacobi-PCG iterations, the element work of all of them in one
 * persistent element kernel (transpose-map algorithm, deterministic) and their
 * vector updates and reductions in one persistent step kernel. A solve's result
 * is the same bits whichever other solves share its launches. Modes FP64 and
 * MIXED only; the operators are set to the transpose-map algorithm. Synchronous:
 * the x_i are complete on return. stats (optional) and status (optional) have
 * count entries; status[i] is 0 or the FEM_HIP_ERR_* of solve i (the call then
 * returns the first such code). */
int fem_hip_mass_solve_batched(int count, const fem_hip_operator* ops, const double* const* b, double* const* x,
                               const fem_hip_solve_options* opts, int iters_per_launch, fem_hip_solve_stats* stats,
                               int* status);

/* ---- on-disk meshes ------------------------------------------------------------
 * No reference counterpart (its meshes are built in code: tests/test-operator.cpp:90-121,
 * tests/cuda/test-cuda-elem-restriction.cu:24-42). A FEMHMSH1 file holds a hex
 * (quad, line) mesh: the node restriction offsets [num_elem][P^dim] in any
 * numbering, the element vertices [num_elem][2^dim] and the vertex coordinates
 * [dim][num_vertices], with an FNV-1a checksum. read fills info and, where the
 * pointers are not NULL, the arrays (call once with NULL arrays for the sizes);
 * it rejects bad headers, sizes, out-of-range indices and checksum mismatches. */
typedef struct {
  int dim, num_elem, P, num_nodes, num_vertices;
} fem_hip_mesh_info;
int fem_hip_mesh_read(const char* path, fem_hip_mesh_info* info, int* offsets, int* vertex_offsets, double* coords);
int fem_hip_mesh_write(const char* path, const fem_hip_mesh_info* info, const int* offsets,
                       const int* vertex_offsets, const double* coords);

/* ---- serving loop ------------------------------------------------------------
 * No counterpart in the reference (it solves nothing): continuous batching of
 * mass-solve requests on one GPU. Requests are admitted while a slot is free and
 * their estimated device bytes fit the budget (hbm_fraction of the HBM free when
 * the server starts); they are set up on worker threads and their CG iterations
 * (graph-captured batches, one stream per request) overlap on the device; a
 * request arriving meanwhile starts as soon as a slot and memory free up. Every
 * solve uses the deterministic scatter, so a request's solution is the same bit
 * for bit whether it runs alone or in a mix (as long as the tuning table, which
 * picks the algorithm, is the same). */
typedef struct fem_hip_server_s* fem_hip_server;
typedef struct {
  int max_active;      /* requests set up or solving at once */
  int workers;         /* host threads for setup and teardown */
  double hbm_fraction; /* (0, 0.95] */
  int batch_iters;     /* most CG iterations per graph launch (a launch takes fewer when the measured
                          convergence rate says the runs in it need fewer) */
  int autotune;        /* 1: tune each new (P, Q, num_comp, precision) on first sight */
  int keep_x;          /* 1: keep solutions for fem_hip_server_wait */
  int cheap_setup;     /* 1: the automatic (tuned) algorithm where its setup is all on the device (box
                          columns, element + transpose map), else the transpose map; 0: the tuning
                          table's fastest algorithm where it has a row for the size (block algorithms
                          cluster elements on the host), else the transpose map */
  int batched;         /* 1: FP64 and MIXED requests whose shape has a batched kernel and that have at
                          most batch_max_elem_nodes element nodes run in cohorts (see
                          fem_hip_mass_solve_batched): the element blocks of concurrent requests of one
                          shape share each persistent launch of one captured graph; larger requests (and
                          all with 0): one stream and captured graph per request, tuned algorithm */
  long long batch_max_elem_nodes; /* num_elem * P^3 * num_comp; <= 0: 2^20 (requests too small to fill
                          the GPU alone; the box columns, which cohorts do not use, win above it) */
} fem_hip_server_options;
typedef struct {
  int n, P, Q, num_comp; /* n^3 hexes of the unit cube, order P, Q Gauss points */
  int mode;              /* FEM_HIP_SOLVE_FP64, _MIXED or _MIXED_RELIABLE */
  double jitter, rtol;   /* vertex jitter (fraction of h), relative residual target */
  uint64_t mesh_seed, rhs_seed; /* b ~ U[-1,1] from rhs_seed (splitmix64) */
} fem_hip_serve_request;
typedef struct {
  int status; /* 0 or the FEM_HIP_ERR_* code of the failure */
  int iterations, outer;
  double rel_residual;
  double queue_ms, setup_ms, solve_ms, latency_ms; /* submit->admit, admit->ready, solve, submit->done */
  long long l_size;
  size_t est_bytes;
  /* Parts of setup_ms: building the problem (mesh, maps, qdata, solver state)
   * and recording the CG graphs (with the wait for other threads' recordings). */
  double build_ms, graph_ms;
} fem_hip_serve_result;
int fem_hip_server_create(const fem_hip_server_options* opt, fem_hip_server* out);
int fem_hip_server_submit(fem_hip_server s, const fem_hip_serve_request* req, long long* id);
/* Waits up to timeout_ms (< 0: forever); FEM_HIP_ERR_NOT_READY on timeout. On
 * completion fills res and, if x is not NULL, copies the solution (x_len must
 * equal res->l_size). Returns the request's own status; a request is reported once. */
int fem_hip_server_wait(fem_hip_server s, long long id, double timeout_ms, fem_hip_serve_result* res, double* x,
                        size_t x_len);
/* stats[9] = {submitted, completed, failed, active, queued, active_peak, budget bytes, used bytes, used peak} */
int fem_hip_server_stats(fem_hip_server s, long long* stats);
/* Finishes every submitted request, then stops. */
int fem_hip_server_destroy(fem_hip_server s);
int fem_hip_serve_request_bytes(const fem_hip_serve_request* req, size_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* FEM_HIP_H */
