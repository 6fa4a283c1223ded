// Box-column algorithm of the fused mass apply, for node restrictions that are
// a structured nx x ny x nz box of P^3-node hexes with lexicographic node
// numbering (offset of local node (ix,iy,iz) of element (ex,ey,ez) is
// gx + Nx*(gy + Ny*gz), g = e*(P-1) + i; e = ex + nx*(ey + ny*ez)). Nothing
// about the mesh is read from memory: only qdata, u and v move.
//
// The elements are cut into columns of BX x BY elements in x and y, and each
// column into nsz segments in z, of floor(nz/nsz) or ceil(nz/nsz) element
// layers (segment s starts at layer s*nz/nsz, rounded down); one workgroup owns
// one segment and marches up its layers. Per layer:
//   * each of the BX*BY elements runs the fast kernel's five in-place stages
//     in its own LDS tile (u gathered straight from its x-rows), each 1D
//     product in the even-odd form of the symmetric basis (eo_forward);
//   * one thread per node position of the column cross-section then sums,
//     plane by plane, the (up to 4) element tiles holding the node, in
//     ascending element order;
//   * the layer's lower P-1 planes are then complete (the top one is kept in
//     LDS for the next layer, whose bottom plane it is): their nodes go
//     straight to v, except nodes on a face the segment shares with another
//     segment (a column face inside the box, or a segment end inside the box),
//     which go to the segment's partial area.
// A second pass gives each shared node the sum of its 2, 4 or 8 partials in
// ascending segment order. Every node of v is written exactly once, the
// result is the same bit for bit on every run, and no atomics are used.
//
// Partial area of segment w: face_size = 2*LYn*LZn + 2*LXn*LZn + 2*LXn*LYn
// slots (LXn = BX*(P-1)+1, LYn = BY*(P-1)+1, LZn = lz*(P-1)+1 nodes with lz
// the longest segment's layer count), slot by the first shared direction of
// the node (x, else y, else z): see box_slot.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "fused_mass_kernel.hpp"

namespace femhip {

// Geometry of one launch (all sizes in elements unless noted).
struct BoxGeom {
  int nx, ny, nz;     // elements per direction
  int Nx, Ny, Nz;     // nodes per direction
  int ncx, ncy, nsz;  // columns in x, y; segments per column
  int lz;             // element layers of the longest segment (ceil(nz / nsz))
  int cs;             // component stride of u and v (= Nx*Ny*Nz)
};

// First element layer of segment s (s = nsz gives nz).
__host__ __device__ inline int box_seg_z0(const BoxGeom& g, int s) {
  return static_cast<int>(static_cast<long long>(s) * g.nz / g.nsz);
}
// The segment holding element layer l (0 <= l < nz).
__host__ __device__ inline int box_seg_of_layer(const BoxGeom& g, int l) {
  return static_cast<int>((static_cast<long long>(l + 1) * g.nsz - 1) / g.nz);
}

__host__ __device__ inline long long box_face_size(int LXn, int LYn, int LZn) {
  return 2LL * LYn * LZn + 2LL * LXn * LZn + 2LL * LXn * LYn;
}

// Slot of local node (lx, ly, lz) in its segment's partial area; xs / ys say
// whether the node's x / y plane is a column face inside the box (else the
// node lies on a shared segment end).
__host__ __device__ inline int box_slot(int lx, int ly, int lz, bool xs, bool ys, int LXn, int LYn, int LZn) {
  if (xs) return (lx == 0 ? 0 : 1) * LYn * LZn + lz * LYn + ly;
  int base = 2 * LYn * LZn;
  if (ys) return base + (ly == 0 ? 0 : 1) * LXn * LZn + lz * LXn + lx;
  base += 2 * LXn * LZn;
  return base + (lz == 0 ? 0 : 1) * LXn * LYn + ly * LXn + lx;
}

// The q-th node coordinate, in ascending order, that is not an interior cut
// (cuts at multiples of S strictly inside [0, N-1], nb of them).
__host__ __device__ inline int box_uncut(int q, int S, int nb) {
  if (q == 0) return 0;
  const int b = (q - 1) / (S - 1);
  return q + (b < nb ? b : nb);
}

// Shared nodes: interior column faces in x (all y, z), then in y (x not a
// cut), then interior segment ends (x and y not cuts).
__host__ __device__ inline long long box_shared_count(const BoxGeom& g, int P, int BX, int BY) {
  const long long nbx = g.ncx - 1, nby = g.ncy - 1, nbz = g.nsz - 1;
  return nbx * g.Ny * g.Nz + nby * (g.Nx - nbx) * g.Nz + nbz * (g.Nx - nbx) * (g.Ny - nby);
}

// The shared-node pass runs on a grid of (face planes) x (lines) x (points
// along a line): plane f < ncx-1 is the x-face gx = (f+1)*SX with lines gz and
// points gy; the next ncy-1 planes are y-faces gy = (f'+1)*SY, lines gz,
// points gx; the last nsz-1 are segment ends gz = (P-1)*box_seg_z0(f''+1), lines gy, points
// gx. A point on a cut of an earlier kind (x before y before z) belongs to
// that plane and is skipped here. Returns whether (plane, line, point) is a
// shared node of this plane, and its coordinates.
__host__ __device__ inline bool box_shared_point(const BoxGeom& g, int P, int BX, int BY, int plane, int line,
                                                 int point, int* gx, int* gy, int* gz) {
  const int SX = BX * (P - 1), SY = BY * (P - 1);
  const int nbx = g.ncx - 1, nby = g.ncy - 1;
  auto xcut = [&](int x) { return x % SX == 0 && x > 0 && x < g.Nx - 1; };
  auto ycut = [&](int y) { return y % SY == 0 && y > 0 && y < g.Ny - 1; };
  if (plane < nbx) {
    if (line >= g.Nz || point >= g.Ny) return false;
    *gx = (plane + 1) * SX;
    *gy = point;
    *gz = line;
    return true;
  }
  plane -= nbx;
  if (plane < nby) {
    if (line >= g.Nz || point >= g.Nx || xcut(point)) return false;
    *gx = point;
    *gy = (plane + 1) * SY;
    *gz = line;
    return true;
  }
  plane -= nby;
  if (line >= g.Ny || point >= g.Nx || ycut(line) || xcut(point)) return false;
  *gx = point;
  *gy = line;
  *gz = box_seg_z0(g, plane + 1) * (P - 1);
  return true;
}

// The segments (ascending) whose node regions hold node (gx, gy, gz), with the
// node's slot in each; returns how many (1, 2, 4 or 8).
__host__ __device__ inline int box_owners(const BoxGeom& g, int P, int BX, int BY, int gx, int gy, int gz, int* seg,
                                          int* slot) {
  const int SX = BX * (P - 1), SY = BY * (P - 1);
  const int LXn = SX + 1, LYn = SY + 1, LZn = g.lz * (P - 1) + 1;
  const bool xs = gx % SX == 0 && gx > 0 && gx < g.Nx - 1;
  const bool ys = gy % SY == 0 && gy > 0 && gy < g.Ny - 1;
  int lyr = gz / (P - 1);
  if (lyr > g.nz - 1) lyr = g.nz - 1;
  const int sz1 = box_seg_of_layer(g, lyr);
  const bool zs = gz % (P - 1) == 0 && sz1 > 0 && gz == box_seg_z0(g, sz1) * (P - 1);
  int cx1 = gx / SX, cy1 = gy / SY;
  if (cx1 > g.ncx - 1) cx1 = g.ncx - 1;
  if (cy1 > g.ncy - 1) cy1 = g.ncy - 1;
  const int cx0 = xs ? cx1 - 1 : cx1, cy0 = ys ? cy1 - 1 : cy1, sz0 = zs ? sz1 - 1 : sz1;
  int n = 0;
  for (int s = sz0; s <= sz1; ++s)
    for (int cy = cy0; cy <= cy1; ++cy)
      for (int cx = cx0; cx <= cx1; ++cx) {
        seg[n] = (s * g.ncy + cy) * g.ncx + cx;
        slot[n] = box_slot(gx - cx * SX, gy - cy * SY, gz - box_seg_z0(g, s) * (P - 1), xs, ys, LXn, LYn, LZn);
        ++n;
      }
  return n;
}

// The shared-node pass's point (plane, line, point) in one step: whether it is
// a shared node of that plane (as box_shared_point), its node index, and its
// owners and slots as box_owners gives them (same order), each passed to
// f(i, segment, slot) with i = 0, 1, ... Each plane kind fixes one coordinate
// and one pair of owner columns or segments, so only the other two are
// derived; divisions by P-1, SX and SY become constants when the kernel
// inlines this with its template arguments, and the segment of a node plane
// takes 32-bit divisions (nz * nsz < 2^31, checked at launch). The owners are
// at most 2 x 2 x 2 (segments x column rows x columns), visited with fixed
// loop bounds so the pass compiles to straight-line code. Returns how many
// owners (0: not a shared node of this plane).
template <class F>
__host__ __device__ __forceinline__ int box_shared_visit(const BoxGeom& g, int P, int BX, int BY, int plane, int line,
                                                         int point, size_t* node, F&& f) {
  const int SX = BX * (P - 1), SY = BY * (P - 1);
  const int LXn = SX + 1, LYn = SY + 1, LZn = g.lz * (P - 1) + 1;
  const int nbx = g.ncx - 1, nby = g.ncy - 1;
  const unsigned unz = static_cast<unsigned>(g.nz), unsz = static_cast<unsigned>(g.nsz);
  auto seg_z0 = [&](int sg) { return static_cast<int>(static_cast<unsigned>(sg) * unz / unsz); };
  int gx, gy, gz, cx0, cx1, cy0, cy1, sz0, sz1;
  bool xs = false, ys = false;
  // The segments holding node plane gz (a shared segment end belongs to two).
  auto z_owners = [&]() {
    int lyr = gz / (P - 1);
    if (lyr > g.nz - 1) lyr = g.nz - 1;
    sz1 = static_cast<int>((static_cast<unsigned>(lyr + 1) * unsz - 1u) / unz);
    const bool zs = gz % (P - 1) == 0 && sz1 > 0 && gz == seg_z0(sz1) * (P - 1);
    sz0 = zs ? sz1 - 1 : sz1;
  };
  if (plane < nbx) {  // x-face gx = (plane+1) SX: lines gz, points gy
    if (line >= g.Nz || point >= g.Ny) return 0;
    gx = (plane + 1) * SX;
    gy = point;
    gz = line;
    xs = true;
    cx0 = plane;
    cx1 = plane + 1;
    const int qy = gy / SY;
    ys = gy == qy * SY && gy > 0 && gy < g.Ny - 1;
    cy1 = qy < g.ncy - 1 ? qy : g.ncy - 1;
    cy0 = ys ? cy1 - 1 : cy1;
    z_owners();
  } else if (plane < nbx + nby) {  // y-face gy = (f+1) SY: lines gz, points gx not on an x cut
    const int f = plane - nbx;
    if (line >= g.Nz || point >= g.Nx) return 0;
    const int qx = point / SX;
    if (point == qx * SX && point > 0 && point < g.Nx - 1) return 0;
    gx = point;
    gy = (f + 1) * SY;
    gz = line;
    ys = true;
    cy0 = f;
    cy1 = f + 1;
    cx0 = cx1 = qx < g.ncx - 1 ? qx : g.ncx - 1;
    z_owners();
  } else {  // segment end gz = (P-1) z0(f+1): lines gy, points gx, neither on a cut
    const int f = plane - nbx - nby;
    if (line >= g.Ny || point >= g.Nx) return 0;
    const int qx = point / SX, qy = line / SY;
    if ((point == qx * SX && point > 0 && point < g.Nx - 1) || (line == qy * SY && line > 0 && line < g.Ny - 1))
      return 0;
    gx = point;
    gy = line;
    gz = seg_z0(f + 1) * (P - 1);
    cx0 = cx1 = qx < g.ncx - 1 ? qx : g.ncx - 1;
    cy0 = cy1 = qy < g.ncy - 1 ? qy : g.ncy - 1;
    sz0 = f;
    sz1 = f + 1;
  }
  *node = static_cast<size_t>(gx) + static_cast<size_t>(g.Nx) * (gy + static_cast<size_t>(g.Ny) * gz);
  int n = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    if (a > sz1 - sz0) continue;
    const int sg = sz0 + a;
    const int lz = gz - seg_z0(sg) * (P - 1);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b > cy1 - cy0) continue;
      const int cy = cy0 + b;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (c > cx1 - cx0) continue;
        const int cx = cx0 + c;
        f(n, (sg * g.ncy + cy) * g.ncx + cx, box_slot(gx - cx * SX, gy - cy * SY, lz, xs, ys, LXn, LYn, LZn));
        ++n;
      }
    }
  }
  return n;
}

// box_shared_visit's owners collected into arrays (the host check's form).
__host__ __device__ inline int box_shared_owners(const BoxGeom& g, int P, int BX, int BY, int plane, int line,
                                                 int point, size_t* node, int* seg, int* slot) {
  return box_shared_visit(g, P, BX, BY, plane, line, point, node, [&](int i, int sg, int sl) {
    seg[i] = sg;
    slot[i] = sl;
  });
}

// qdata is read exactly once per apply. FP64 loads it as streaming
// (nontemporal) data, which leaves the caches to the u rows that neighbouring
// elements and columns read again: 64^3, P=5 FP64 box kernel 0.192 -> 0.177 ms
// on MI355X. FP32 measured the same either way and keeps plain loads.
template <int Q, class T, bool NT = sizeof(T) == 8>
__device__ __forceinline__ void load_qdata(const T* tile, int t, T (&q)[Q]) {
  load_qline<Q, T, NT>(tile, t, q);
}

// One x-row of P consecutive values of u, read as 2-vectors of T (16-byte
// loads for FP64) and a single value for odd P: fewer memory instructions per
// row than P single loads. The row starts at any multiple of sizeof(T), so the
// vectors are declared with T's alignment (the hardware needs dword alignment
// for these loads).
template <int P, class T>
__device__ __forceinline__ void load_row(const T* __restrict__ r, T (&x)[P]) {
  typedef T t2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
#pragma unroll
  for (int i = 0; i < P / 2; ++i) {
    const t2 v = *reinterpret_cast<const t2*>(r + 2 * i);
    x[2 * i] = v[0];
    x[2 * i + 1] = v[1];
  }
  if constexpr (P % 2 == 1) x[P - 1] = r[P - 1];
}

// Even-odd form of a symmetric 1D interpolation matrix B [Q][P]
// (B(Q-1-q, P-1-p) = B(q, p), true for Lagrange bases on symmetric nodes at
// symmetric points): with x_e = x_j + x_(P-1-j), x_o = x_j - x_(P-1-j),
//   out_q, out_(Q-1-q) = sum_j Be(q,j) x_e[j] (+ Bm(q) x_mid) +- sum_j Bo(q,j) x_o[j],
//   Be(q,j) = (B(q,j) + B(q,P-1-j))/2, Bo(q,j) = (B(q,j) - B(q,P-1-j))/2, Bm(q) = B(q,(P-1)/2),
// and the same values give the transposed product. Half the basis values and
// about half the multiply-adds of the plain product (Kronbichler & Kormann,
// "Fast matrix-free evaluation of discontinuous Galerkin finite element
// operators", ACM TOMS 45(3), 2019, sec. 4.2).
template <int P, int Q>
struct EvenOdd {
  static constexpr int HP = P / 2, HQ = (Q + 1) / 2, QF = Q / 2, PM = P & 1, QM = Q & 1;
  static constexpr int BE = 0, BM = HQ * HP, BO = HQ * HP + PM * HQ, N = BO + QF * HP;
};

// Packs Be, Bm, Bo (EvenOdd layout) from B [Q][P].
template <int P, int Q, class T>
MatArg<T, EvenOdd<P, Q>::N> even_odd_arg(const void* basis) {
  using E = EvenOdd<P, Q>;
  const T* B = static_cast<const T*>(basis);
  MatArg<T, E::N> M;
  for (int q = 0; q < E::HQ; ++q)
    for (int j = 0; j < E::HP; ++j) M.m[E::BE + q * E::HP + j] = (B[q * P + j] + B[q * P + P - 1 - j]) / T(2);
  if (E::PM)
    for (int q = 0; q < E::HQ; ++q) M.m[E::BM + q] = B[q * P + P / 2];
  for (int q = 0; q < E::QF; ++q)
    for (int j = 0; j < E::HP; ++j) M.m[E::BO + q * E::HP + j] = (B[q * P + j] - B[q * P + P - 1 - j]) / T(2);
  return M;
}

// Even-odd coefficients read from memory through scalar loads (a uniform
// address in the constant address space): for bases with many coefficients,
// which would not fit the SGPRs as kernel arguments (P = 8, Q = 10: 40
// doubles), each stage reloads them instead of holding them for the kernel.
template <class T>
struct EoLoads {
  const __attribute__((address_space(4))) T* m;
};

// out[Q] = B x[P]
template <int P, int Q, class T, class CM = MatArg<T, EvenOdd<P, Q>::N>>
__device__ __forceinline__ void eo_forward(const T (&x)[P], T (&out)[Q], const CM& M) {
  using E = EvenOdd<P, Q>;
  T xe[E::HP > 0 ? E::HP : 1], xo[E::HP > 0 ? E::HP : 1];
#pragma unroll
  for (int j = 0; j < E::HP; ++j) {
    xe[j] = x[j] + x[P - 1 - j];
    xo[j] = x[j] - x[P - 1 - j];
  }
#pragma unroll
  for (int q = 0; q < E::HQ; ++q) {
    T e = E::PM ? M.m[E::BM + q] * x[P / 2] : T(0);
#pragma unroll
    for (int j = 0; j < E::HP; ++j) e = fma_t(M.m[E::BE + q * E::HP + j], xe[j], e);
    if (q < E::QF) {
      T o = T(0);
#pragma unroll
      for (int j = 0; j < E::HP; ++j) o = fma_t(M.m[E::BO + q * E::HP + j], xo[j], o);
      out[q] = e + o;
      out[Q - 1 - q] = e - o;
    } else {
      out[q] = e;  // the middle point (Q odd)
    }
  }
}

// out[P] = B^T y[Q]
template <int P, int Q, class T, class CM = MatArg<T, EvenOdd<P, Q>::N>>
__device__ __forceinline__ void eo_transpose(const T (&y)[Q], T (&out)[P], const CM& M) {
  using E = EvenOdd<P, Q>;
  T ye[E::QF > 0 ? E::QF : 1], yo[E::QF > 0 ? E::QF : 1];
#pragma unroll
  for (int k = 0; k < E::QF; ++k) {
    ye[k] = y[k] + y[Q - 1 - k];
    yo[k] = y[k] - y[Q - 1 - k];
  }
#pragma unroll
  for (int p = 0; p < E::HP; ++p) {
    T e = E::QM ? M.m[E::BE + E::QF * E::HP + p] * y[E::QF] : T(0);
#pragma unroll
    for (int k = 0; k < E::QF; ++k) e = fma_t(M.m[E::BE + k * E::HP + p], ye[k], e);
    T o = T(0);
#pragma unroll
    for (int k = 0; k < E::QF; ++k) o = fma_t(M.m[E::BO + k * E::HP + p], yo[k], o);
    out[p] = e + o;
    out[P - 1 - p] = e - o;
  }
  if (E::PM) {
    T m = E::QM ? M.m[E::BM + E::QF] * y[E::QF] : T(0);
#pragma unroll
    for (int k = 0; k < E::QF; ++k) m = fma_t(M.m[E::BM + k], ye[k], m);
    out[P / 2] = m;
  }
}

// Element tile strides of the box kernel: tile [iz][y][x], y and x below
// max(P, Q), with y stride RY, iz stride ZS, and FBUF entries per element.
template <int P, int Q>
struct BoxTile {
  static constexpr int RY = Shape<P, Q>::RY, ZS = Shape<P, Q>::ZS, FBUF = Shape<P, Q>::FBUF;
};
// P=5, Q=6: strides from a search over a model of LDS bank conflicts of every
// stage's 64-bit accesses (32 and 64 banks); on MI355X at 64^3 the FP32 apply
// took 0.127 instead of 0.130 ms, FP64 0.1997 instead of 0.2008 ms.
template <>
struct BoxTile<5, 6> {
  static constexpr int RY = 11, ZS = 71, FBUF = 357;
};

// Element tile strides of the wavefront kernels, per value type: a model of
// LDS bank conflicts (32 banks of 4 bytes; 16 lanes a cycle for 64-bit
// accesses, 32 for 32-bit ones) over every stage's accesses and the node sums
// agrees with the measured conflict cycles of BoxTile's strides (P=5: 28% /
// 43% modelled against 26% / 41% measured in FP64 / FP32; P=8: 48% / 44%
// against 42% / 41%). These strides are the model's best; the rest keep
// BoxTile's.
template <int P, int Q, class T>
struct WaveTile : BoxTile<P, Q> {};
template <>
struct WaveTile<5, 6, float> {
  static constexpr int RY = 9, ZS = 51, FBUF = 279;
};
template <>
struct WaveTile<8, 10, double> {
  static constexpr int RY = 10, ZS = 106, FBUF = 855;
};
template <>
struct WaveTile<8, 10, float> {
  static constexpr int RY = 12, ZS = 138, FBUF = 1095;
};

// UM: how u reaches the x-rows. 0: each row thread loads its row when the
// step starts; 1: the row of the next step is loaded one step ahead; 2: the
// node planes go through an LDS ring, the next layer's loaded one layer ahead.
// QA: 1: a layer's qdata is loaded during the layer before it (held in
// registers for a whole layer); 0: it is loaded in its own layer, right after
// the first step's u rows, two stages before its use (12 fewer FP64 VGPRs at
// Q=6, so more workgroups fit on a CU).
// SPLIT > 1 (with NC = 1): the components of a blocked SPLIT-component field
// are independent scalar masses; workgroup w takes component w % SPLIT of
// segment w / SPLIT (after the XCD remap, so a segment's components run
// together on one XCD and share its qdata through that XCD's L2, loaded
// cached instead of streaming).
template <int P, int Q, int NC, class T, int BX, int BY, int UM, int QA, int SPLIT = 1>
__global__ __launch_bounds__(round64(BX* BY* Shape<P, Q>::TPE)) void fused_box_kernel(
    BoxGeom g, const T* __restrict__ qdata, const T* __restrict__ u, T* __restrict__ v, T* __restrict__ partials,
    size_t part_comp_len, MatArg<T, EvenOdd<P, Q>::N> Em) {
  using S = Shape<P, Q>;
  constexpr int NE = BX * BY;
  constexpr int P2 = S::P2, Q2 = S::Q2, Q3 = S::Q3, TPE = S::TPE;
  constexpr int RY = BoxTile<P, Q>::RY, ZS = BoxTile<P, Q>::ZS, FB = BoxTile<P, Q>::FBUF;
  constexpr int M = P > Q ? P : Q;
  static_assert(RY >= M && ZS >= (M - 1) * RY + M && FB >= (P - 1) * ZS + (M - 1) * RY + M,
                "element tile strides hold every (iz, y, x) entry");
  constexpr int LXn = BX * (P - 1) + 1, LYn = BY * (P - 1) + 1, PLANE = LXn * LYn;
  __shared__ T s_b[NE + 1][FB];  // element tiles, then one tile of zeros (absent contributions)
  __shared__ T s_top[NC][PLANE];  // top node plane of the last layer, completed by the next one
  __shared__ T s_u[UM == 2 ? NC : 1][UM == 2 ? P : 1][UM == 2 ? PLANE : 1];  // u ring (UM 2), slot gz % P
  static_assert(PLANE <= NE * S::TPE, "one node position per thread");

  // Segment of this workgroup; XCD x (which receives workgroups h with
  // h % 8 == x) gets the contiguous range [x*per, (x+1)*per) when 8 divides the grid.
  const int nwg = gridDim.x;
  int w = blockIdx.x;
  if (nwg % 8 == 0) w = (blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8;
  static_assert(SPLIT == 1 || NC == 1, "a component-split kernel runs one component per workgroup");
  if constexpr (SPLIT > 1) {
    const int comp = w % SPLIT;
    w /= SPLIT;
    u += static_cast<size_t>(comp) * g.cs;
    v += static_cast<size_t>(comp) * g.cs;
    partials += static_cast<size_t>(comp) * part_comp_len;
  }
  constexpr bool kQdataNT = sizeof(T) == 8 && SPLIT == 1;
  const int cx = w % g.ncx, cy = (w / g.ncx) % g.ncy, sz = w / (g.ncx * g.ncy);
  const int ez0 = box_seg_z0(g, sz), nl = box_seg_z0(g, sz + 1) - ez0;  // this segment's layers
  const int LZn = g.lz * (P - 1) + 1;  // slot layout: the longest segment's node planes
  const int LZs = nl * (P - 1) + 1;    // this segment's node planes
  const int gx0 = cx * BX * (P - 1), gy0 = cy * BY * (P - 1), gz0 = ez0 * (P - 1);
  const size_t Nxy = static_cast<size_t>(g.Nx) * g.Ny;
  T* my_part = partials + static_cast<size_t>(w) * box_face_size(LXn, LYn, LZn);

  const int el = threadIdx.x / TPE;
  const int t = threadIdx.x - el * TPE;
  const bool live = el < NE;  // the workgroup is rounded up to whole wavefronts
  const int lex = live ? el % BX : 0, ley = live ? el / BX : 0;
  const int ex = cx * BX + lex, ey = cy * BY + ley;
  T* b = live ? s_b[el] : nullptr;
  for (int i = threadIdx.x; i < FB; i += blockDim.x) s_b[NE][i] = T(0);  // read after the first barriers

  // Thread `pos` owns node position (plx, ply) of the column cross-section in
  // every plane: the layer's elements holding it (1, 2 or 4, ascending) and
  // the offsets of (ix, iy) in their tiles, its v address (less the plane)
  // and whether it lies on a column face shared with another column.
  const int pos = threadIdx.x;
  const bool pos_thread = pos < PLANE;
  const int ply = pos_thread ? pos / LXn : 0, plx = pos_thread ? pos - (pos / LXn) * LXn : 0;
  // The four candidates, in ascending element order: (lower y, lower x),
  // (lower y, upper x), (upper y, lower x), (upper y, upper x), where "lower"
  // is the element below a cut holding the node at index P-1 and "upper" the
  // element holding it at index r (0 on a cut). Absent contributions (a node
  // held by fewer than 4 elements) read the zero tile, whose +0 leaves every
  // sum unchanged. Plain selects, no per-thread arrays: a dynamically indexed
  // array would be placed in LDS or scratch memory by the compiler.
  constexpr int ZT = NE * FB;
  int toff0 = ZT, toff1 = ZT, toff2 = ZT, toff3 = ZT;
  if (pos_thread) {
    const int qx = plx / (P - 1), rx = plx - qx * (P - 1), qy = ply / (P - 1), ry = ply - qy * (P - 1);
    const bool xlo = rx == 0 && qx > 0, xup = rx != 0 || qx < BX;
    const bool ylo = ry == 0 && qy > 0, yup = ry != 0 || qy < BY;
    const int oxl = (qx - 1) * FB + (P - 1), oxu = qx * FB + rx;
    const int oyl = (qy - 1) * BX * FB + (P - 1) * RY, oyu = qy * BX * FB + ry * RY;
    if (ylo && xlo) toff0 = oyl + oxl;
    if (ylo && xup) toff1 = oyl + oxu;
    if (yup && xlo) toff2 = oyu + oxl;
    if (yup && xup) toff3 = oyu + oxu;
  }
  const bool pxs = (plx == 0 && cx > 0) || (plx == LXn - 1 && cx < g.ncx - 1);
  const bool pys = (ply == 0 && cy > 0) || (ply == LYn - 1 && cy < g.ncy - 1);
  T* const pv = v + static_cast<size_t>(gx0 + plx) + static_cast<size_t>(g.Nx) * (gy0 + ply);

  // Software pipeline over the (layer, component) steps (see UM); qdata of
  // the next layer is loaded one layer ahead. With the LDS ring, plane gz of
  // the column's node region sits in slot gz % P and is loaded as whole rows;
  // the P-1 new planes of the next layer are stored into the slots the
  // current layer's x-rows have left (the shared plane stays).
  constexpr int NT = round64(NE * S::TPE);
  constexpr int NEW = (P - 1) * PLANE;       // new u values per layer and component
  constexpr int UPT = UM == 2 ? (NEW + NT - 1) / NT : 1;
  const bool row_thread = live && t < P2, line_thread = live && t < Q2;
  const int riz = t / P, riy = t - (t / P) * P;
  auto node_addr = [&](int gz, int p) -> size_t {
    const int ly = p / LXn, lx = p - ly * LXn;
    return static_cast<size_t>(gx0 + lx) + static_cast<size_t>(g.Nx) * (gy0 + ly) + Nxy * gz;
  };
  auto row_ptr = [&](int ez, int c) -> const T* {
    return u + static_cast<size_t>(c) * g.cs + static_cast<size_t>(ex * (P - 1)) +
           static_cast<size_t>(g.Nx) * (ey * (P - 1) + riy) + Nxy * (ez * (P - 1) + riz);
  };
  auto qd_ptr = [&](int ez) -> const T* {
    const size_t e = static_cast<size_t>(ex) + static_cast<size_t>(g.nx) * (ey + static_cast<size_t>(g.ny) * ez);
    return qdata + e * Q3;
  };
  if constexpr (UM == 2) {
    const int gzb = gz0;  // all P planes of the first layer
    for (int idx = threadIdx.x; idx < NC * P * PLANE; idx += NT) {
      const int c = idx / (P * PLANE), r = idx - c * (P * PLANE), j = r / PLANE, p = r - j * PLANE;
      s_u[c][(gzb + j) % P][p] = u[static_cast<size_t>(c) * g.cs + node_addr(gzb + j, p)];
    }
  }
  T xc[UM == 1 ? P : 1] = {};
  if constexpr (UM == 1) {
    if (row_thread) {
      load_row<P, T>(row_ptr(ez0, 0), xc);
    }
  }
  T qc[Q] = {};
  if (QA && line_thread) load_qdata<Q, T, kQdataNT>(qd_ptr(ez0), t, qc);
  if constexpr (UM == 2) __syncthreads();
  // Nothing loaded above is still in flight when the layer loop starts. (Those
  // loads target registers that later stages reuse; left in flight, the
  // compiler's wait analysis would carry them around the loop and wait for
  // every outstanding load before those stages, the qdata prefetch included.)
  __builtin_amdgcn_s_waitcnt(0);

#pragma unroll 1
  for (int k = 0; k < nl; ++k) {
    const int ez = ez0 + k;
    T qn[Q] = {};
    T qd[Q];
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) qd[qz] = qc[qz];
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      const bool more = k + 1 < nl;
      const int gzn = (ez + 1) * (P - 1) + 1;  // first new plane of the next layer
      T un[UPT] = {};
      T xn[UM == 1 ? P : 1] = {};
      if constexpr (UM == 2) {
#pragma unroll
        for (int r = 0; r < UPT; ++r) {
          const int idx = threadIdx.x + r * NT;
          if (more && idx < NEW) {
            const int j = idx / PLANE, p = idx - j * PLANE;
            un[r] = u[static_cast<size_t>(c) * g.cs + node_addr(gzn + j, p)];
          }
        }
      } else if constexpr (UM == 1) {
        const int cn = c + 1 < NC ? c + 1 : 0, kn = c + 1 < NC ? k : k + 1;
        if (row_thread && kn < nl) {
          load_row<P, T>(row_ptr(ez0 + kn, cn), xn);
        }
      }
      // S1: x-rows (iz,iy): contract x.
      T x[P];
      if (row_thread) {
        const int iz = riz, iy = riy;
        if constexpr (UM == 2) {
          const T* ur = &s_u[c][(ez * (P - 1) + iz) % P][(ley * (P - 1) + iy) * LXn + lex * (P - 1)];
#pragma unroll
          for (int ix = 0; ix < P; ++ix) x[ix] = ur[ix];
        } else if constexpr (UM == 1) {
#pragma unroll
          for (int ix = 0; ix < P; ++ix) x[ix] = xc[ix];
        } else {
          load_row<P, T>(row_ptr(ez, c), x);
        }
      }
      // The next layer's qdata (QA = 1) or this layer's (QA = 0), issued
      // after this step's u loads: outstanding loads complete in issue order,
      // so S1's wait for u leaves it in flight. Every lane loads (lanes past the
      // lines and the last layer read a valid row they do not use), so the
      // number of loads in flight is known where S1 waits.
      if (NC == 1 || c == 0) {
        const int tq = t < Q2 ? t : Q2 - 1, ezq = QA && k + 1 < nl ? ez + 1 : ez;
        const size_t e = static_cast<size_t>(ex) + static_cast<size_t>(g.nx) * (ey + static_cast<size_t>(g.ny) * ezq);
        const T* r = qdata + e * Q3;
        if constexpr (QA) {
          load_qdata<Q, T, kQdataNT>(r, tq, qn);
        } else {
          load_qdata<Q, T, kQdataNT>(r, tq, qd);
        }
      }
      if (row_thread) {
        const int iz = riz, iy = riy;
        T y[Q];
        eo_forward<P, Q, T>(x, y, Em);
        const int row = iz * ZS + iy * RY;
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) b[row + qx] = y[qx];
      }
      __syncthreads();
      // S2: y-columns (iz,qx).
      if (live && t < P * Q) {
        const int col = (t / Q) * ZS + t % Q;
        T x[P], y[Q];
#pragma unroll
        for (int iy = 0; iy < P; ++iy) x[iy] = b[col + iy * RY];
        eo_forward<P, Q, T>(x, y, Em);
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) b[col + qy * RY] = y[qy];
      }
      __syncthreads();
      // S3: z-lines (qy,qx): contract z, times qdata, contract z back.
      if (live && t < Q2) {
        const int line = (t / Q) * RY + t % Q;
        T x[P], d[Q];
#pragma unroll
        for (int iz = 0; iz < P; ++iz) x[iz] = b[iz * ZS + line];
        eo_forward<P, Q, T>(x, d, Em);
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) d[qz] *= qd[qz];
        eo_transpose<P, Q, T>(d, x, Em);
#pragma unroll
        for (int iz = 0; iz < P; ++iz) b[iz * ZS + line] = x[iz];
      }
      __syncthreads();
      // S4: y-columns back.
      if (live && t < P * Q) {
        const int col = (t / Q) * ZS + t % Q;
        T y[Q], x[P];
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) y[qy] = b[col + qy * RY];
        eo_transpose<P, Q, T>(y, x, Em);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) b[col + iy * RY] = x[iy];
      }
      __syncthreads();
      // S5: x-rows back.
      if (live && t < P2) {
        const int row = (t / P) * ZS + (t % P) * RY;
        T y[Q], x[P];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) y[qx] = b[row + qx];
        eo_transpose<P, Q, T>(y, x, Em);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) b[row + ix] = x[ix];
      }
      __syncthreads();
      // Node sums: each position adds its elements' outputs in ascending
      // element order (the bottom plane after what the layer below left in
      // s_top), then the plane goes to v, to the partial area (shared faces),
      // or, for the top plane below the segment's last layer, to s_top.
      if (pos_thread) {
        const T* tb = &s_b[0][0];
        T* vc = pv + static_cast<size_t>(c) * g.cs;
        T* pc = my_part + static_cast<size_t>(c) * part_comp_len;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          // All four reads issue together; absent ones add the zero tile.
          const T t0 = tb[toff0 + j * ZS], t1 = tb[toff1 + j * ZS], t2 = tb[toff2 + j * ZS],
                  t3 = tb[toff3 + j * ZS];
          T sum = (j == 0 && k > 0) ? s_top[c][pos] + t0 : t0;
          sum = ((sum + t1) + t2) + t3;
          if (j == P - 1 && k + 1 < nl) {
            s_top[c][pos] = sum;
          } else {
            const int gz = ez * (P - 1) + j, lzl = gz - gz0;
            const bool zs = (lzl == 0 && sz > 0) || (lzl == LZs - 1 && sz < g.nsz - 1);
            if (!(pxs || pys || zs)) {
              vc[Nxy * gz] = sum;
            } else {
              pc[box_slot(plx, ply, lzl, pxs, pys, LXn, LYn, LZn)] = sum;
            }
          }
        }
      }
      if constexpr (UM == 2) {
        // The next layer's new planes into the slots this layer's S1 has left.
#pragma unroll
        for (int r = 0; r < UPT; ++r) {
          const int idx = threadIdx.x + r * NT;
          if (more && idx < NEW) {
            const int j = idx / PLANE, p = idx - j * PLANE;
            s_u[c][(gzn + j) % P][p] = un[r];
          }
        }
      } else if constexpr (UM == 1) {
#pragma unroll
        for (int ix = 0; ix < P; ++ix) xc[ix] = xn[ix];
      }
      (void)more;
      (void)gzn;
      __syncthreads();  // the next step's S1 overwrites the tiles (and reads the ring)
    }
    if constexpr (QA) {
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) qc[qz] = qn[qz];
    }
    (void)qn;
  }
}

// Orders a wavefront's own LDS accesses: LDS executes one wavefront's
// instructions in order, so a lane's read after this sees what any lane of
// the same wavefront wrote before it; the fences keep the compiler from
// moving memory operations across (the wave_barrier idiom of rocPRIM).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The box-column kernel with whole wavefronts per element: the five stages of
// an element run on the lanes of its own wavefront (EPW = 1: all 64; EPW = 2:
// two elements per wavefront, 32 lanes each) — x-rows one per lane, y lines
// (P*Q) and z-lines (Q^2) taken LPE at a time — so they are ordered by
// wave_sync alone, and the workgroup waits twice per layer (before the node
// sums, which read every element's tile, and before the next layer's first
// stage overwrites the tiles) instead of six times. Segments, node sums, the
// partial area and the shared-node pass are those of fused_box_kernel; u rows
// are loaded straight from u and qdata within its own layer (UM = 0, QA = 0).
// WPE > 0 asks the compiler to fit WPE wavefronts per SIMD (a register budget
// of 512 / WPE VGPRs).
template <int P, int Q, int NC, class T, int BX, int BY, int SPLIT = 1, int EPW = 1, int WPE = 0>
__global__ __launch_bounds__(64 * BX * BY / EPW) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void fused_box_wave_kernel(
    BoxGeom g, const T* __restrict__ qdata, const T* __restrict__ u, T* __restrict__ v, T* __restrict__ partials,
    size_t part_comp_len, MatArg<T, EvenOdd<P, Q>::N> Em, const T* __restrict__ eo) {
  using S = Shape<P, Q>;
  constexpr int NE = BX * BY;
  constexpr int LPE = 64 / EPW;  // lanes per element
  // More than 24 coefficients (P = 8): read each stage's from eo (the same
  // values as Em) with scalar loads; the asm hides the pointer's origin so the
  // loads stay in their stage.
  constexpr bool kEoLoads = EvenOdd<P, Q>::N > 24;
  auto coef = [&]() {
    if constexpr (kEoLoads) {
      auto p = (const __attribute__((address_space(4))) T*)eo;
      asm volatile("" : "+s"(p));
      return EoLoads<T>{p};
    } else {
      return Em;
    }
  };
  static_assert((EPW == 1 || EPW == 2) && NE % EPW == 0, "one or two elements per wavefront");
  constexpr int P2 = S::P2, Q2 = S::Q2, Q3 = S::Q3, PQ = P * Q;
  constexpr int RY = WaveTile<P, Q, T>::RY, ZS = WaveTile<P, Q, T>::ZS, FB = WaveTile<P, Q, T>::FBUF;
  constexpr int M = P > Q ? P : Q;
  static_assert(RY >= M && ZS >= (M - 1) * RY + M && FB >= (P - 1) * ZS + (M - 1) * RY + M,
                "element tile strides hold every (iz, y, x) entry");
  static_assert(P2 <= LPE, "one x-row per lane");
  constexpr int QL = (Q2 + LPE - 1) / LPE;  // z-lines per lane
  constexpr int LXn = BX * (P - 1) + 1, LYn = BY * (P - 1) + 1, PLANE = LXn * LYn;
  static_assert(PLANE <= NE * LPE, "one node position per thread");
  __shared__ T s_b[NE + 1][FB];   // element tiles, then one tile of zeros (absent contributions)
  __shared__ T s_top[NC][PLANE];  // top node plane of the last layer, completed by the next one

  const int nwg = gridDim.x;
  int w = blockIdx.x;
  if (nwg % 8 == 0) w = (blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8;
  static_assert(SPLIT == 1 || NC == 1, "a component-split kernel runs one component per workgroup");
  if constexpr (SPLIT > 1) {
    const int comp = w % SPLIT;
    w /= SPLIT;
    u += static_cast<size_t>(comp) * g.cs;
    v += static_cast<size_t>(comp) * g.cs;
    partials += static_cast<size_t>(comp) * part_comp_len;
  }
  constexpr bool kQdataNT = sizeof(T) == 8 && SPLIT == 1;
  const int cx = w % g.ncx, cy = (w / g.ncx) % g.ncy, sz = w / (g.ncx * g.ncy);
  const int ez0 = box_seg_z0(g, sz), nl = box_seg_z0(g, sz + 1) - ez0;
  const int LZn = g.lz * (P - 1) + 1;
  const int LZs = nl * (P - 1) + 1;
  const int gx0 = cx * BX * (P - 1), gy0 = cy * BY * (P - 1), gz0 = ez0 * (P - 1);
  const size_t Nxy = static_cast<size_t>(g.Nx) * g.Ny;
  T* my_part = partials + static_cast<size_t>(w) * box_face_size(LXn, LYn, LZn);

  const int el = threadIdx.x / LPE, lane = threadIdx.x % LPE;
  const int lex = el % BX, ley = el / BX;
  const int ex = cx * BX + lex, ey = cy * BY + ley;
  T* b = s_b[el];
  for (int i = threadIdx.x; i < FB; i += blockDim.x) s_b[NE][i] = T(0);

  // Node positions of the column cross-section (as in fused_box_kernel).
  const int pos = threadIdx.x;
  const bool pos_thread = pos < PLANE;
  const int ply = pos_thread ? pos / LXn : 0, plx = pos_thread ? pos - (pos / LXn) * LXn : 0;
  constexpr int ZT = NE * FB;
  int toff0 = ZT, toff1 = ZT, toff2 = ZT, toff3 = ZT;
  if (pos_thread) {
    const int qx = plx / (P - 1), rx = plx - qx * (P - 1), qy = ply / (P - 1), ry = ply - qy * (P - 1);
    const bool xlo = rx == 0 && qx > 0, xup = rx != 0 || qx < BX;
    const bool ylo = ry == 0 && qy > 0, yup = ry != 0 || qy < BY;
    const int oxl = (qx - 1) * FB + (P - 1), oxu = qx * FB + rx;
    const int oyl = (qy - 1) * BX * FB + (P - 1) * RY, oyu = qy * BX * FB + ry * RY;
    if (ylo && xlo) toff0 = oyl + oxl;
    if (ylo && xup) toff1 = oyl + oxu;
    if (yup && xlo) toff2 = oyu + oxl;
    if (yup && xup) toff3 = oyu + oxu;
  }
  const bool pxs = (plx == 0 && cx > 0) || (plx == LXn - 1 && cx < g.ncx - 1);
  const bool pys = (ply == 0 && cy > 0) || (ply == LYn - 1 && cy < g.ncy - 1);
  T* const pv = v + static_cast<size_t>(gx0 + plx) + static_cast<size_t>(g.Nx) * (gy0 + ply);

  const bool row_lane = lane < P2;
  const int riz = lane / P, riy = lane - (lane / P) * P;
  auto row_ptr = [&](int ez, int c) -> const T* {
    return u + static_cast<size_t>(c) * g.cs + static_cast<size_t>(ex * (P - 1)) +
           static_cast<size_t>(g.Nx) * (ey * (P - 1) + riy) + Nxy * (ez * (P - 1) + riz);
  };
  __syncthreads();  // the zero tile

  T qd[QL][Q];
#pragma unroll 1
  for (int k = 0; k < nl; ++k) {
    const int ez = ez0 + k;
#pragma unroll 1
    for (int c = 0; c < NC; ++c) {
      // S1: x-rows (iz, iy): the row of u, then this layer's qdata (the
      // first component's step loads it for all), both before S1 waits.
      T x[P];
      if (row_lane) load_row<P, T>(row_ptr(ez, c), x);
      if (NC == 1 || c == 0) {
        const size_t e = static_cast<size_t>(ex) + static_cast<size_t>(g.nx) * (ey + static_cast<size_t>(g.ny) * ez);
#pragma unroll
        for (int j = 0; j < QL; ++j) {
          const int line = lane + LPE * j;
          load_qdata<Q, T, kQdataNT>(qdata + e * Q3, line < Q2 ? line : Q2 - 1, qd[j]);
        }
      }
      if (row_lane) {
        T y[Q];
        eo_forward<P, Q, T>(x, y, coef());
        const int row = riz * ZS + riy * RY;
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) b[row + qx] = y[qx];
      }
      wave_sync();
      // S2: y-columns (iz, qx).
#pragma unroll
      for (int t0 = 0; t0 < PQ; t0 += LPE) {
        const int t = t0 + lane;
        if (t < PQ) {
          const int col = (t / Q) * ZS + t % Q;
          T xx[P], y[Q];
#pragma unroll
          for (int iy = 0; iy < P; ++iy) xx[iy] = b[col + iy * RY];
          eo_forward<P, Q, T>(xx, y, coef());
#pragma unroll
          for (int qy = 0; qy < Q; ++qy) b[col + qy * RY] = y[qy];
        }
      }
      wave_sync();
      // S3: z-lines (qy, qx): contract z, times qdata, contract z back.
#pragma unroll
      for (int j = 0; j < QL; ++j) {
        const int t = lane + LPE * j;
        if (t < Q2) {
          const int line = (t / Q) * RY + t % Q;
          T xx[P], d[Q];
#pragma unroll
          for (int iz = 0; iz < P; ++iz) xx[iz] = b[iz * ZS + line];
          const auto cz = coef();
          eo_forward<P, Q, T>(xx, d, cz);
#pragma unroll
          for (int qz = 0; qz < Q; ++qz) d[qz] *= qd[j][qz];
          eo_transpose<P, Q, T>(d, xx, cz);
#pragma unroll
          for (int iz = 0; iz < P; ++iz) b[iz * ZS + line] = xx[iz];
        }
      }
      wave_sync();
      // S4: y-columns back.
#pragma unroll
      for (int t0 = 0; t0 < PQ; t0 += LPE) {
        const int t = t0 + lane;
        if (t < PQ) {
          const int col = (t / Q) * ZS + t % Q;
          T y[Q], xx[P];
#pragma unroll
          for (int qy = 0; qy < Q; ++qy) y[qy] = b[col + qy * RY];
          eo_transpose<P, Q, T>(y, xx, coef());
#pragma unroll
          for (int iy = 0; iy < P; ++iy) b[col + iy * RY] = xx[iy];
        }
      }
      wave_sync();
      // S5: x-rows back.
      if (row_lane) {
        const int row = riz * ZS + riy * RY;
        T y[Q], xx[P];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) y[qx] = b[row + qx];
        eo_transpose<P, Q, T>(y, xx, coef());
#pragma unroll
        for (int ix = 0; ix < P; ++ix) b[row + ix] = xx[ix];
      }
      __syncthreads();
      // Node sums (as in fused_box_kernel).
      if (pos_thread) {
        const T* tb = &s_b[0][0];
        T* vc = pv + static_cast<size_t>(c) * g.cs;
        T* pc = my_part + static_cast<size_t>(c) * part_comp_len;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const T t0 = tb[toff0 + j * ZS], t1 = tb[toff1 + j * ZS], t2 = tb[toff2 + j * ZS],
                  t3 = tb[toff3 + j * ZS];
          T sum = (j == 0 && k > 0) ? s_top[c][pos] + t0 : t0;
          sum = ((sum + t1) + t2) + t3;
          if (j == P - 1 && k + 1 < nl) {
            s_top[c][pos] = sum;
          } else {
            const int gz = ez * (P - 1) + j, lzl = gz - gz0;
            const bool zs = (lzl == 0 && sz > 0) || (lzl == LZs - 1 && sz < g.nsz - 1);
            if (!(pxs || pys || zs)) {
              vc[Nxy * gz] = sum;
            } else {
              pc[box_slot(plx, ply, lzl, pxs, pys, LXn, LYn, LZn)] = sum;
            }
          }
        }
      }
      __syncthreads();  // the next step's S1 overwrites the tiles
    }
  }
}

// v at every shared node = its partials summed in ascending segment order
// (grid: see box_shared_point; a block is kSharedLines lines of 64 points).
// The NC components' loads of an owner are issued together (each thread's
// work is a short chain of dependent loads: several chains in flight hide
// their latency).
constexpr int kSharedLines = 4;
template <class T, int P, int BX, int BY, int NC>
__global__ __launch_bounds__(64 * kSharedLines) void box_shared_sum_kernel(BoxGeom g, const T* __restrict__ partials,
                                                                           size_t part_comp_len, T* __restrict__ v) {
  const int LZn = g.lz * (P - 1) + 1;
  const size_t face = static_cast<size_t>(box_face_size(BX * (P - 1) + 1, BY * (P - 1) + 1, LZn));
  // A wavefront is one line (64 points of one threadIdx.y): the line, and
  // everything derived from it (the segments of a node plane and their first
  // layers, runtime divisions), is wavefront-uniform and taken on the scalar unit.
  const int line = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.y * kSharedLines + threadIdx.y));
  T s[NC];
  size_t node;
  const int n = box_shared_visit(g, P, BX, BY, blockIdx.z, line, blockIdx.x * 64 + threadIdx.x, &node,
                                 [&](int i, int sg, int sl) {
                                   const T* pc = partials + static_cast<size_t>(sg) * face + sl;
#pragma unroll
                                   for (int c = 0; c < NC; ++c) {
                                     const T t = pc[static_cast<size_t>(c) * part_comp_len];
                                     s[c] = i == 0 ? t : s[c] + t;
                                   }
                                 });
  if (n == 0) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) v[static_cast<size_t>(c) * g.cs + node] = s[c];
}

// Grid of the shared-node pass.
inline dim3 box_shared_grid(const BoxGeom& g) {
  const int planes = (g.ncx - 1) + (g.ncy - 1) + (g.nsz - 1);
  const int lines = g.Nz > g.Ny ? g.Nz : g.Ny;
  const int points = g.Nx > g.Ny ? g.Nx : g.Ny;
  return dim3(static_cast<unsigned>((points + 63) / 64), static_cast<unsigned>((lines + kSharedLines - 1) / kSharedLines),
              static_cast<unsigned>(planes));
}

inline BoxGeom box_geom(const FusedArgs& a, int P, int BX, int BY) {
  BoxGeom g;
  g.nx = a.box_n[0];
  g.ny = a.box_n[1];
  g.nz = a.box_n[2];
  g.Nx = g.nx * (P - 1) + 1;
  g.Ny = g.ny * (P - 1) + 1;
  g.Nz = g.nz * (P - 1) + 1;
  g.ncx = g.nx / BX;
  g.ncy = g.ny / BY;
  g.nsz = a.box_nsz;
  g.lz = g.nsz > 0 ? (g.nz + g.nsz - 1) / g.nsz : 0;
  g.cs = g.Nx * g.Ny * g.Nz;  // blocked components (checked when the box was detected)
  return g;
}

template <int P, int Q, int NC, class T, int BX, int BY, int UM, int QA, bool SPLIT = false, int WAVE = 0,
          int WPE = 0>
void launch_box(const FusedArgs& a, hipStream_t s) {
  using S = Shape<P, Q>;
  const BoxGeom g = box_geom(a, P, BX, BY);
  require(g.nx % BX == 0 && g.ny % BY == 0 && g.nsz > 0 && g.nsz <= g.nz && BX * (P - 1) >= 2 &&
              BY * (P - 1) >= 2 && (g.nz / g.nsz) * (P - 1) >= 2,
          "fused mass: box does not divide into the kernel's columns");
  const long long nseg = static_cast<long long>(g.ncx) * g.ncy * g.nsz;
  require(nseg < (1ll << 31) && static_cast<long long>(g.Nx) * g.Ny * g.Nz < (1ll << 31), "fused mass: box too large");
  require(a.part_comp_len >= static_cast<size_t>(nseg * box_face_size(BX * (P - 1) + 1, BY * (P - 1) + 1,
                                                                      g.lz * (P - 1) + 1)),
          "fused mass: box partial buffer too small");
  if (a.box_pass == 0) {
    constexpr int threads = round64(BX * BY * S::TPE);
    if constexpr (WAVE > 0) {  // WAVE: elements per wavefront
      require(a.basis_eo != nullptr, "fused mass: the box kernel's even-odd coefficients are not on the device");
      constexpr int NCK = SPLIT ? 1 : NC, SP = SPLIT ? NC : 1;
      require(nseg * SP < (1ll << 31), "fused mass: box too large");
      fused_box_wave_kernel<P, Q, NCK, T, BX, BY, SP, WAVE, WPE>
          <<<static_cast<unsigned>(nseg * SP), 64 * BX * BY / WAVE, 0, s>>>(
          g, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u), static_cast<T*>(a.v),
          static_cast<T*>(a.partials), a.part_comp_len, even_odd_arg<P, Q, T>(a.basis),
          static_cast<const T*>(a.basis_eo));
    } else if constexpr (SPLIT) {
      require(nseg * NC < (1ll << 31), "fused mass: box too large");
      fused_box_kernel<P, Q, 1, T, BX, BY, UM, QA, NC><<<static_cast<unsigned>(nseg * NC), threads, 0, s>>>(
          g, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u), static_cast<T*>(a.v),
          static_cast<T*>(a.partials), a.part_comp_len, even_odd_arg<P, Q, T>(a.basis));
    } else {
      fused_box_kernel<P, Q, NC, T, BX, BY, UM, QA><<<static_cast<unsigned>(nseg), threads, 0, s>>>(
          g, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u), static_cast<T*>(a.v),
          static_cast<T*>(a.partials), a.part_comp_len, even_odd_arg<P, Q, T>(a.basis));
    }
  } else {
    const dim3 grid = box_shared_grid(g);
    if (grid.z == 0) return;
    require(grid.y < 65536 && grid.z < 65536 && static_cast<long long>(g.nz) * g.nsz < (1ll << 31),
            "fused mass: box too large for the shared-node pass");
    box_shared_sum_kernel<T, P, BX, BY, NC><<<grid, dim3(64, kSharedLines), 0, s>>>(
        g, static_cast<const T*>(a.partials), a.part_comp_len, static_cast<T*>(a.v));
  }
  HIP_CHECK_LAUNCH();
}

// elems % 100 of a column cross-section: its element count, except 8 x 2
// (18; 16 is 4 x 4) and 16 x 2 (34; 32 is 8 x 4).
constexpr int box_code(int BX, int BY) {
  return BX == 8 && BY == 2 ? 18 : BX == 16 && BY == 2 ? 34 : BX * BY;
}

// Workgroups of a box-column variant's column kernel resident on one CU.
template <int P, int Q, int NC, class T, int BX, int BY, int UM, int QA, bool SPLIT = false, int WAVE = 0,
          int WPE = 0>
int box_occupancy() {
  int n = 0;
  if constexpr (WAVE > 0) {
    constexpr int NCK = SPLIT ? 1 : NC, SP = SPLIT ? NC : 1;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, fused_box_wave_kernel<P, Q, NCK, T, BX, BY, SP, WAVE, WPE>, 64 * BX * BY / WAVE, 0));
  } else if constexpr (SPLIT) {
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_box_kernel<P, Q, 1, T, BX, BY, UM, QA, NC>,
                                                           round64(BX * BY * Shape<P, Q>::TPE), 0));
  } else {
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_box_kernel<P, Q, NC, T, BX, BY, UM, QA>,
                                                           round64(BX * BY * Shape<P, Q>::TPE), 0));
  }
  return n;
}

// Key elems of a box-column variant: box_code(BX, BY), plus 100 for
// the kernel that loads qdata within its own layer (QA = 0), plus 200 for FP64
// with u through the LDS ring, plus 300 (several components) for the
// component-split kernel (one component per workgroup, qdata within its own
// layer). By default FP64 loads its x-rows straight from u, FP32 goes through
// the LDS ring.
template <int P, int Q, int NC, int BX, int BY>
struct FusedBoxReg {
  FusedBoxReg() {
    reg<double, 0, 1, false>(0);
    reg<double, 0, 0, false>(100);
    reg<double, 2, 1, false>(200);
    reg<float, 2, 1, false>(0);
    reg<float, 2, 0, false>(100);
    if constexpr (NC > 1) {
      reg<double, 0, 0, true>(300);
      reg<float, 2, 0, true>(300);
    }
  }
  template <class T, int UM, int QA, bool SPLIT>
  static void reg(int code) {
    const FusedKey k{P, Q, NC, std::is_same<T, float>::value ? 1 : 0, kAlgBoxColumn, 0, box_code(BX, BY) + code};
    register_fused(k, &launch_box<P, Q, NC, T, BX, BY, UM, QA, SPLIT>, 0);
    register_fused_occupancy(k, &box_occupancy<P, Q, NC, T, BX, BY, UM, QA, SPLIT>);
  }
};

// Only the component-split kernels (code 300) of a multi-component shape.
template <int P, int Q, int NC, int BX, int BY>
struct FusedBoxSplitReg {
  static_assert(NC > 1, "the split kernels run the components of a multi-component field");
  FusedBoxSplitReg() {
    FusedBoxReg<P, Q, NC, BX, BY>::template reg<double, 0, 0, true>(300);
    FusedBoxReg<P, Q, NC, BX, BY>::template reg<float, 2, 0, true>(300);
  }
};

// The kernels with whole wavefronts per element (fused_box_wave_kernel), EPW
// elements per wavefront: code 400 (EPW 1) or 600 (EPW 2), plus 100 for the
// component split of a multi-component field.
template <int P, int Q, int NC, int BX, int BY, int EPW = 1>
struct FusedBoxWaveReg {
  FusedBoxWaveReg() {
    constexpr int base = EPW == 1 ? 400 : 600;
    if constexpr (NC == 1) {
      reg<double, false>(base);
      reg<float, false>(base);
    } else {
      reg<double, true>(base + 100);
      reg<float, true>(base + 100);
    }
  }
  template <class T, bool SPLIT>
  static void reg(int code) {
    const FusedKey k{P, Q, NC, std::is_same<T, float>::value ? 1 : 0, kAlgBoxColumn, 0, box_code(BX, BY) + code};
    // FP32, P = 8: held to 64 VGPRs (it needs 65-68 otherwise), so that two
    // 16-wave workgroups fit a CU instead of one.
    constexpr int W = std::is_same<T, float>::value && P == 8 ? 8 : 0;
    register_fused(k, &launch_box<P, Q, NC, T, BX, BY, 0, 0, SPLIT, EPW, W>, 0);
    register_fused_occupancy(k, &box_occupancy<P, Q, NC, T, BX, BY, 0, 0, SPLIT, EPW, W>);
  }
};

// Column cross-section (BX, BY) of each compiled elems value (elems % 100).
inline bool box_columns(int elems, int* BX, int* BY) {
  switch (elems % 100) {
    case 16: *BX = 4; *BY = 4; return elems >= 0;
    case 8: *BX = 4; *BY = 2; return elems >= 0;
    case 4: *BX = 2; *BY = 2; return elems >= 0;
    case 18: *BX = 8; *BY = 2; return elems >= 0;
    case 32: *BX = 8; *BY = 4; return elems >= 0;
    case 34: *BX = 16; *BY = 2; return elems >= 0;
    default: return false;
  }
}

}  // namespace femhip
