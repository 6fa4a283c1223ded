// Fused per-element mass operator: gather -> interp (x,y,z) -> qdata * u ->
// interp^T (x,y,z) -> scatter, in one kernel per apply plus, for some scatter
// algorithms, a second pass.
#pragma once

#include <cstdint>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "basis.hpp"
#include "common.hpp"
#include "restriction.hpp"

namespace femhip {

// How the element outputs reach v.
enum FusedAlg : int {
  kAlgElemAtomic = 0,  // per element; one atomic add per shared node, plain store for single-contributor nodes
  kAlgElemTmap = 1,    // per element into Y[E*P^3], then the transpose-map pass (deterministic; CPU order)
  kAlgBlockAtomic = 2, // per element block; sums in LDS, boundary nodes one atomic each
  kAlgBlockPart = 3,   // per element block; sums in LDS, boundary nodes one partial each, then a partial-sum pass
  // 4 and 6: retired (persistent software-pipelined forms of 1 and 5, which
  // measured the same as or slower than them on MI355X); no kernels.
  kAlgBlockPartLds = 5,  // kAlgBlockPart with LDS atomic sums in place of the colour rounds (any order)
  kAlgBoxColumn = 7,  // structured boxes only: element columns marching in z, node planes in LDS,
                      // shared column faces summed by a second pass (deterministic)
  kAlgCount = 8
};
inline bool alg_deterministic(int alg) {
  return alg == kAlgElemTmap || alg == kAlgBlockPart || alg == kAlgBoxColumn;
}
inline bool alg_box(int alg) { return alg == kAlgBoxColumn; }
// Element nodes (num_elem * P^3 * NC) from which an untuned box operator
// defaults to the box columns. Below it there are too few column segments of
// enough layers to fill the GPU: on MI355X the box columns lost to the element
// and block algorithms at 0.2-0.4M element nodes and won from about 1M
// (P = 3..6, one component).
constexpr long long kBoxMinNodes = 1LL << 20;
inline bool alg_tmap(int alg) { return alg == kAlgElemTmap; }
inline bool alg_blocked(int alg) {
  return alg == kAlgBlockAtomic || alg == kAlgBlockPart || alg == kAlgBlockPartLds;
}
inline bool alg_partial(int alg) {
  return alg == kAlgBlockPart || alg == kAlgBlockPartLds;
}

// One compiled kernel.
struct FusedKey {
  int P, Q, NC;
  int fp32;   // 0 FP64, 1 FP32
  int alg;    // FusedAlg
  int exact;  // 1: no FMA, reference association (FP64, element algorithms only)
  int elems;  // elements per workgroup (element algorithms) or per block (block algorithms)
  bool operator<(const FusedKey& o) const {
    return std::tie(P, Q, NC, fp32, alg, exact, elems) < std::tie(o.P, o.Q, o.NC, o.fp32, o.alg, o.exact, o.elems);
  }
};

// Launch parameters shared by every instantiation (each kernel reads its part).
struct FusedArgs {
  int comp_stride = 0;
  const void* qdata = nullptr;  // element algorithms: [E][Q^3]; block algorithms: [nblk*NB][Q^3] (T)
  const void* u = nullptr;      // L-vector, T
  const void* basis = nullptr;  // host pointer to Q*P values of T (interp_1d), copied into the kernel arguments
  // element algorithms
  int num_elem = 0;
  const int* offsets = nullptr;          // [E][P^3]
  const int* flagged_offsets = nullptr;  // the same with bit 31 set on single-contributor nodes
  void* out = nullptr;                   // atomic: v (pre-zeroed); tmap: Y [NC][E*P^3]
  size_t y_comp_len = 0;                 // E*P^3
  // block algorithms
  int nblk = 0, max_nodes = 0;
  const int* node_ptr = nullptr;   // [nblk + 1]
  const int* nodes = nullptr;      // per block: interior then boundary nodes (comp-0 L indices)
  const int* ninterior = nullptr;  // [nblk]
  const int* nelem = nullptr;      // [nblk]
  const uint16_t* loc = nullptr;   // [nblk*NB][P^3]
  const uint8_t* color = nullptr;  // [nblk*NB]
  const int* ncolor = nullptr;     // [nblk]
  const int* part_ptr = nullptr;   // [nblk + 1]
  void* v = nullptr;               // block atomic: pre-zeroed v; block part: v (interior nodes stored)
  void* partials = nullptr;        // block part: [NC][npart]; box column: [NC][segments * face slots]
  const void* basis_eo = nullptr;  // box column: the even-odd coefficients of interp_1d on the device (T)
  size_t part_comp_len = 0;
  // box column: elements per direction, segments per column, pass (0 the column kernel, 1 the shared-node sums)
  int box_n[3] = {0, 0, 0};
  int box_nsz = 0;
  int box_pass = 0;
};

using FusedLauncher = void (*)(const FusedArgs&, hipStream_t);
// Registry of compiled instantiations (filled by the fused_inst_*.hip units).
std::map<FusedKey, FusedLauncher>& fused_registry();
void register_fused(const FusedKey& k, FusedLauncher f, int lds_node_budget);
// Box-column variants: workgroups of the variant's column kernel resident on
// one CU (the HIP occupancy query on the kernel; needs the current device).
using FusedOccupancy = int (*)();
void register_fused_occupancy(const FusedKey& k, FusedOccupancy f);
// The registered query's answer (cached per key), 0 when none is registered.
int fused_wgs_per_cu(const FusedKey& k);
// For block kernels: the most unique nodes per block the kernel's LDS is sized for.
int fused_node_budget(const FusedKey& k);
// Every compiled elems variant of one algorithm for a shape.
std::vector<int> fused_elem_variants(int P, int Q, int NC, int fp32, int alg, int exact);
bool fused_mass_supported(int P, int Q, int NC);

// Device copy of a block plan (one per (NB, node budget) in use).
struct DevPlan {
  int NB = 0, budget = 0, nblk = 0, max_nodes = 0, nrows = 0;
  size_t npart = 0;
  bool covers_all = false;
  DevBuf<int> node_ptr, nodes, ninterior, nelem, ncolor, part_ptr, rows, row_ptr, row_part;
  DevBuf<uint16_t> loc;
  DevBuf<uint8_t> color;
  DevBuf<double> q64;  // qdata in slot order [slot][Q^3]
  DevBuf<float> q32;   // lazily made
  DevBuf<double> part64;
  DevBuf<float> part32;
  DevBuf<int> node_row;  // split applies: a node's row, or -1 for a node the block kernel stores (lazy)
};

// The selected way to apply: algorithm and its elems (0 = the shape's default).
struct FusedChoice {
  int alg = -1;  // -1: automatic (tuned table, else the default order)
  int elems = 0;
};

// The device state of a fused mass operator.
struct FusedMass {
  int P = 0, Q = 0, NC = 0, num_elem = 0, comp_stride = 0, l_size = 0;
  const Restriction* rstr = nullptr;  // owned by the operator (node restriction)
  DevBuf<int> flagged_offsets;        // offsets with bit 31 set where the node has one contributor (lazy)
  DevBuf<double> qdata64;             // private layout [E][Q^3]
  DevBuf<float> qdata32;              // lazily made from qdata64
  std::vector<double> B64;            // interp_1d [Q][P]
  std::vector<float> B32;
  std::vector<double> Bsq;            // interp_1d squared entry-wise (Jacobi diagonal)
  DevBuf<double> Bsq_dev;             // its device copy (made by the first diagonal; after Bsq, so freed first)
  std::map<std::pair<int, int>, std::unique_ptr<DevPlan>> plans;  // by (NB, node budget)
  DevBuf<double> Y64;                 // element outputs of the transpose-map algorithm (lazy)
  DevBuf<float> Y32;
  // Transpose map over the referenced nodes: rows ascending, entries
  // t = i + e*P^3 ascending (the CPU scatter's order). The restriction's own
  // map when it has one (deterministic restrictions), else built here; the
  // pointers below are the ones in use (rows null: row m is node m).
  bool tm_built = false, tm_covers_all = false;
  int tm_nrows = 0;
  DevBuf<int> tm_rows, tm_ptr, tm_idx;
  const int *tm_rows_p = nullptr, *tm_ptr_p = nullptr, *tm_idx_p = nullptr;
  // Structured box (every offset is the lexicographic box formula with node
  // stride 1, comp_stride = node count, l_size = NC * node count): elements per direction.
  bool box = false;
  bool basis_symmetric = false;  // B(Q-1-q, P-1-p) = B(q, p) to rounding (the box kernel's even-odd form)
  int box_n[3] = {0, 0, 0};
  int box_lz = 0;             // most element layers per box-column segment (0: box_segments chooses)
  DevBuf<double> box_eo64;    // box columns: interp_1d's even-odd coefficients on the device (lazy)
  DevBuf<float> box_eo32;
  DevBuf<double> box_part64;  // box-column partials (lazy)
  DevBuf<float> box_part32;
  int scatter = kScatterAtomic;  // requested: atomic (any order) or deterministic (reproducible)
  int exact = 0;                 // requested: the reference's arithmetic (bit-equal with deterministic)
  FusedChoice choice;            // explicit algorithm, or automatic
  // Bumped whenever a device buffer an apply passes to its kernels is
  // allocated, reallocated or released (plans, FP32 copies, partials, element
  // outputs, maps, flags): work captured into a graph stays valid only while
  // it and the resolved kernel are unchanged (fused_signature).
  uint64_t gen = 0;

  // Optional per-apply stage timing with HIP events on the apply's stream:
  // a..b = pre-pass (output zeroing), b..c = the fused kernel, c..d = post-pass
  // (transpose-map or partial sums). Events are pooled and read back by timing_end.
  struct StageEvents {
    hipEvent_t a, b, c, d;
  };
  bool timing = false;
  std::vector<StageEvents> events;
  size_t events_used = 0;
  ~FusedMass();
};

void fused_timing_begin(FusedMass& f);
// Sums of the stage times (ms) over the applies since timing_begin.
void fused_timing_end(FusedMass& f, double* kernel_ms, double* aux_ms, int* count);

// qdata_ref is the backend-strided layout qdata[q*E + e] (device).
std::unique_ptr<FusedMass> fused_mass_create(const Restriction& r, const HostBasis& b, const double* qdata_ref,
                                             hipStream_t s);
void fused_mass_apply64(FusedMass& f, const double* u, double* v, hipStream_t s);
void fused_mass_apply32(FusedMass& f, const float* u, float* v, hipStream_t s);
// The kernel an apply in this precision would launch now (after automatic choice).
FusedKey fused_resolve(const FusedMass& f, int fp32);
// Identifies what an apply in this precision would launch now: the resolved
// kernel, its box layer count, the scatter and exact settings and f.gen. Two
// equal signatures mean a graph captured at the first still launches the same
// kernels on live buffers at the second.
uint64_t fused_signature(const FusedMass& f, int fp32);
// FP64 apply that allocates nothing: the resolved algorithm's per-apply
// scratch (the box-column partials, or the transpose-map algorithm's element
// outputs) lives at `scratch` (scratch_bytes, device memory the caller owns);
// everything else the kernels read must exist already (made at create, or
// from the restriction's own transpose map). Throws Unsupported for any other
// algorithm, and invalid_argument when the scratch is too small.
// fused_scratch_bytes64 gives the bytes an apply would need.
size_t fused_scratch_bytes64(const FusedMass& f);
void fused_mass_apply64_scratch(FusedMass& f, const double* u, double* v, void* scratch, size_t scratch_bytes,
                                hipStream_t s);
// An apply split for a solver that folds the apply's second pass into its own
// vector kernel. The first stage (fused_mass_apply_first*) runs only the
// element kernel, which writes v at the nodes it completes and partial sums
// or element outputs for the others; then for entry i (component
// c = i / cs, node i - c cs)
//   row = node_row ? node_row[node] : node,
//   (A u)_i = row < 0 ? v_i : sum over k in [ptr[row], ptr[row+1]) of vals[c cl + idx[k]]
// (added into 0 in ascending k), which is what the second pass computes, in
// its order.
struct FusedSplit {
  const int* node_row = nullptr;  // null: row m is node m
  const int* ptr = nullptr;
  const int* idx = nullptr;
  const void* vals = nullptr;  // partial sums or element outputs, in the apply's precision
  size_t cl = 0;               // vals per component
  int cs = 0;                  // vector entries per component
};
// Whether the apply in this precision splits so (block partial sums whose rows
// and interior nodes cover every entry, or the transpose-map algorithm whose
// rows are all the nodes); fills `out` (buffers and the node-to-row map made now).
bool fused_split(FusedMass& f, int fp32, FusedSplit* out, hipStream_t s);
void fused_mass_apply_first64(FusedMass& f, const double* u, double* v, hipStream_t s);
void fused_mass_apply_first32(FusedMass& f, const float* u, float* v, hipStream_t s);
// Builds lazily made state (block plans, FP32 copies) for the current choice,
// so later applies allocate nothing (graph capture).
void fused_prepare(FusedMass& f, int fp32, hipStream_t s);
// Whether offsets [E][P^3] are a structured nx*ny*nz box of P^3-node hexes
// with lexicographic node numbering (node stride 1), l_size = NC * node count
// and (NC > 1) comp_stride = node count; on success n3 = {nx, ny, nz}.
bool detect_box(const int* offsets, int num_elem, int P, int num_comp, int comp_stride, int l_size, int* n3);
// Makes f.flagged_offsets (on the device; the atomic element kernel's input).
void ensure_flagged_offsets(FusedMass& f, hipStream_t s);
// Whether the operator is a structured box with box-column kernels compiled
// for its shape (its tuning rows are the box = 1 ones).
bool fused_box_operator(const FusedMass& f);
// Segments per column of the box-column algorithm for nz element layers,
// ncols columns (times the workgroups per column segment) and `slots`
// workgroups running at once: the fewest segments whose layer steps of the
// longest-running CU, ceil(ncols*nsz / slots) * ceil(nz / nsz), are within
// 5% of the least (the pass over shared nodes grows with every segment end,
// and a last partial round of workgroups costs less than a full one, since
// the few CUs still running then share the memory); segments hold at least
// 2 node planes beyond their first ((nz / nsz) * (P-1) >= 2).
// With max_layers > 0 (the operator's setting) the fewest segments of at most
// max_layers layers. 0 if none qualifies.
int box_segments(int nz, long long ncols, long long slots, int P, int max_layers);
// Jacobi diagonal: diag = E^T [ (B o B)^T x3 qdata ] (FP64).
void fused_mass_diagonal(FusedMass& f, double* diag, hipStream_t s);

// Autotuning table: the fastest (alg, elems) per (P, Q, NC, fp32,
// deterministic, box) and per mesh size (elements), measured by
// fused_autotune and consulted by automatic choice; box = 1 rows are for
// structured boxes (which can also run the box-column algorithm), box = 0 rows
// for any node numbering. A box operator without a box row uses the box = 0
// row. A key may hold rows measured on several mesh sizes (the winner changes
// with size: the box columns need many elements per column segment); an
// operator uses the row whose size is nearest its own (by ratio).
struct TuneKey {
  int P, Q, NC, fp32, det, box;
  bool operator<(const TuneKey& o) const {
    return std::tie(P, Q, NC, fp32, det, box) < std::tie(o.P, o.Q, o.NC, o.fp32, o.det, o.box);
  }
};
struct TuneEntry {
  int alg = -1, elems = 0;
  double ms = 0.0;  // measured time per apply
  long long num_elem = 0;  // size it was measured on
};
using TuneRows = std::map<long long, TuneEntry>;  // by TuneEntry::num_elem
std::map<TuneKey, TuneRows>& tune_table();  // guard every access with tune_mutex()
std::mutex& tune_mutex();
// The row of `k` measured on the mesh size nearest num_elem (ratio of sizes;
// ties go to the larger mesh) within a factor max_ratio (0: any), or null.
// The caller holds tune_mutex().
const TuneEntry* tune_lookup(const TuneKey& k, long long num_elem, double max_ratio = 0.0);
// Adds e as k's row for e.num_elem, replacing a row of that size. The caller
// holds tune_mutex().
void tune_record(const TuneKey& k, const TuneEntry& e);
// The fastest packing of each algorithm: (P, Q, NC, fp32, alg) -> (elems, ms).
// An explicit algorithm with elems 0 uses it (else the default packing).
struct TuneAlgKey {
  int P, Q, NC, fp32, alg;
  bool operator<(const TuneAlgKey& o) const {
    return std::tie(P, Q, NC, fp32, alg) < std::tie(o.P, o.Q, o.NC, o.fp32, o.alg);
  }
};
std::map<TuneAlgKey, TuneEntry>& tune_alg_table();  // guarded by tune_mutex()
// Text form of both tables: one row per line,
//   best P Q NC fp32 det box alg elems ms num_elem
//   alg  P Q NC fp32 alg elems ms num_elem
// (a best row without the box field reads as box = 0)
// load replaces best rows with the same key and num_elem (alg rows: the same
// key), skips rows not compiled in this library, and returns the rows taken.
void tune_table_save(const std::string& path);
int tune_table_load(const std::string& path);
// Times every compiled candidate valid for the operator's requested scatter and
// precision on its own data (reps applies each, after one warm-up), records the
// fastest in the table and returns it. Restores the operator's choice (automatic).
TuneEntry fused_autotune(FusedMass& f, int fp32, int reps, hipStream_t s, std::vector<TuneEntry>* all = nullptr);

}  // namespace femhip
