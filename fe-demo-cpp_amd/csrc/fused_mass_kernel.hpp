// The fused per-element mass kernel template (gfx950, wave64).
//
// Workgroup = NE elements x TPE threads (TPE = max(P^2, PQ, Q^2), padded to a
// whole number of wavefronts). Each 1D contraction is done by one thread per 1D
// line of the element tile, holding the line in registers (P or Q loads feed
// P*Q fused multiply-adds); lines change orientation between stages through LDS.
//
// Fast path (FMA, fused_mass_fast_kernel), 5 barriers per element and component:
//   S1  x-rows (iz,iy):  gather u[offsets] straight into registers, contract x  -> A [iz][iy][qx]
//   S2  y-lines (iz,qx): contract y                                             -> Bf[iz][qy][qx]
//   S3  z-lines (qy,qx): contract z, multiply by qdata, contract z back (registers)
//   S4  y-lines (iz,qx): contract y back                                        -> A'[iz][iy][qx]
//   S5  x-rows (iz,iy):  contract x back; then a coalesced scatter of the element outputs
// Exact path (no FMA, fused_mass_exact_kernel): the reference's axis order in
// both directions, x -> y -> z then x -> y -> z (src/tensor-contract.cpp:74-91),
// every sum from 0 in ascending order and added into 0, so each element output
// equals the reference's bit for bit; 6 barriers.
//
// qdata is read in a private per-element layout [e][q] (q = (qz*Q + qy)*Q + qx),
// so lanes of one load instruction read consecutive words. LDS rows that lanes
// stride across use an odd row length (no 2/4-way bank conflicts for 64-bit
// words). Workgroup ids are remapped so each of the 8 XCDs (which receive
// workgroups round-robin) walks one contiguous range of elements.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "fused_mass.hpp"

namespace femhip {

template <class T, int N>
struct MatArg {
  T m[N];
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int round64(int n) { return (n + 63) / 64 * 64; }
constexpr int odd(int n) { return n | 1; }

// Stride of an iz-slab so that lanes (iz, qx) of a y-line stage hit distinct
// 64-bit LDS banks: the slab stride is congruent to Q modulo 32 words-of-8-bytes.
constexpr int slab(int min_len, int Q) {
  return Q * Q < 16 ? min_len : min_len + (((Q - min_len) % 32) + 32) % 32;
}

template <int P, int Q>
struct Shape {
  static constexpr int P2 = P * P, P3 = P * P * P, Q2 = Q * Q, Q3 = Q * Q * Q;
  static constexpr int TPE = cmax(cmax(P2, P * Q), Q2);
  static constexpr int RQ = odd(Q);  // padded row length of x-lines with Q entries
  static constexpr int RP = odd(P);  // padded row length of x-lines with P entries
  static constexpr int ZS0 = slab(Q2, Q);      // Bf[iz] slab stride in buffer 0
  static constexpr int ZS1 = slab(P * RQ, Q);  // A[iz] slab stride in buffer 1
  // buffer 0: Bf [P][ZS0] (both paths) / D [Q][Q][RQ], A2 [Q][P][P] (exact)
  static constexpr int BUF0F = P * ZS0;
  static constexpr int BUF0 = cmax(BUF0F, cmax(Q2 * RQ, Q * P2));
  // buffer 1: A [P][ZS1] (both paths) / A1 [Q][Q][RP] (exact)
  static constexpr int BUF1F = P * ZS1;
  static constexpr int BUF1 = cmax(BUF1F, Q2 * RP);
  // Fast path: one buffer per element, [iz][y][x] with y, x < M = max(P, Q).
  // Every stage owns whole lines of it (rows (iz,y) or columns (iz,x) or z-lines
  // (y,x)), so each contraction reads its line and writes its result in place.
  static constexpr int M = cmax(P, Q);
  static constexpr int RY = odd(M);   // y stride
  static constexpr int ZS = M * RY;   // iz stride
  static constexpr int FBUF = P * ZS;
};

// Fused multiply-add in T's own precision (__builtin_fma is the double one and
// would promote float operands).
template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (std::is_same_v<T, float>) {
    return __builtin_fmaf(a, b, c);
  } else {
    return __builtin_fma(a, b, c);
  }
}

// Private qdata layout of one element tile (made by the operator at create):
// for even Q, [qz/2][line][qz%2] with line = qy*Q + qx, so that z-line thread t
// reads its values as Q/2 pairs and the lanes of one load instruction read
// consecutive pairs; for odd Q, [qz][line] (consecutive lanes, single values).
__host__ __device__ constexpr int qdata_index(int Q, int line, int qz) {
  return Q % 2 == 0 ? (qz / 2) * 2 * Q * Q + line * 2 + (qz % 2) : qz * Q * Q + line;
}

// The Q qdata values of z-line t of the element tile at p (p = qdata +
// e*Q^3): 16-byte pairs for FP64 and 8-byte pairs for FP32 when Q is even
// (an even Q makes every tile and pair offset even, so the pairs are
// aligned), single values otherwise. NT: as streaming (read-once) data.
template <int Q, class T, bool NT = false>
__device__ __forceinline__ void load_qline(const T* __restrict__ p, int t, T (&q)[Q]) {
  constexpr int Q2 = Q * Q;
  if constexpr (Q % 2 == 0) {
    typedef T t2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int i = 0; i < Q / 2; ++i) {
      const t2* v = reinterpret_cast<const t2*>(p + i * 2 * Q2 + 2 * t);
      t2 x;
      if constexpr (NT) {
        x = __builtin_nontemporal_load(v);
      } else {
        x = *v;
      }
      q[2 * i] = x[0];
      q[2 * i + 1] = x[1];
    }
  } else {
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) {
      if constexpr (NT) {
        q[qz] = __builtin_nontemporal_load(p + qz * Q2 + t);
      } else {
        q[qz] = p[qz * Q2 + t];
      }
    }
  }
}

template <bool EXACT, class T>
__device__ __forceinline__ T mac(T a, T b, T acc) {
  if constexpr (EXACT) {
    return acc + a * b;  // separate rounding (the library is built with -ffp-contract=off)
  } else {
    return fma_t(a, b, acc);
  }
}

template <bool EXACT, class T>
__device__ __forceinline__ T finish(T acc) {
  if constexpr (EXACT) {
    return T(0) + acc;  // the reference adds the sum into a zeroed output
  } else {
    return acc;
  }
}

template <int P, int Q, int NC, class T, int NE, bool DET>
__global__ __launch_bounds__(round64(NE* Shape<P, Q>::TPE)) void fused_mass_exact_kernel(
    int num_elem, int comp_stride, const int* __restrict__ offsets, const T* __restrict__ qdata,
    const T* __restrict__ u, T* __restrict__ out, size_t y_comp_len, MatArg<T, P * Q> Bm) {
  using S = Shape<P, Q>;
  constexpr bool EXACT = true;
  constexpr int P2 = S::P2, P3 = S::P3, Q2 = S::Q2, Q3 = S::Q3, TPE = S::TPE, RQ = S::RQ, RP = S::RP;
  constexpr int ZS0 = S::ZS0, ZS1 = S::ZS1;
  __shared__ T s_b0[NE][EXACT ? S::BUF0 : S::BUF0F];
  __shared__ T s_b1[NE][EXACT ? S::BUF1 : S::BUF1F];
  __shared__ int s_off[EXACT ? NE * P3 : 1];

  // XCD-aware block order: hardware id h runs on XCD h % 8; give XCD x the
  // contiguous logical range [x*per, (x+1)*per). The grid is a multiple of 8.
  const int per = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long e0 = static_cast<long long>(blk) * NE;
  if (e0 >= num_elem) return;  // the whole workgroup leaves together: no barrier is skipped
  const int n_here = static_cast<int>(min(static_cast<long long>(NE), num_elem - e0));

  const int el = threadIdx.x / TPE;
  const int t = threadIdx.x - el * TPE;
  const bool live = el < n_here;
  const long long e = e0 + el;
  T* b0 = live ? s_b0[el] : nullptr;
  T* b1 = live ? s_b1[el] : nullptr;
  auto B = [&](int q, int p) -> T { return Bm.m[q * P + p]; };

  // qdata of the thread's z-line (qy,qx) = t, issued first to overlap the gather.
  T qd[Q];
  if (live && t < Q2) {
    load_qline<Q, T>(qdata + static_cast<size_t>(e) * Q3, t, qd);
  }
  // Offsets of the thread's x-row (iz,iy) = t, kept in registers for the scatter.
  int o[P];
  if (live && t < P2) {
    const int* op = offsets + e * P3 + t * P;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) o[ix] = op[ix];
  }
  if constexpr (EXACT) {
    for (int idx = threadIdx.x; idx < n_here * P3; idx += blockDim.x) s_off[idx] = offsets[e0 * P3 + idx];
  }

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const size_t coff = static_cast<size_t>(c) * comp_stride;
    // S1: x-rows (iz,iy) = t: gather, contract x -> b1[iz*ZS1 + iy*RQ + qx]
    if (live && t < P2) {
      const int r1 = (t / P) * ZS1 + (t % P) * RQ;
      T x[P];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) x[ix] = u[static_cast<size_t>(o[ix]) + coff];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T acc = T(0);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) acc = mac<EXACT>(B(qx, ix), x[ix], acc);
        b1[r1 + qx] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // S2: y-lines (iz,qx): b1[iz*ZS1 + iy*RQ + qx] -> b0[iz*ZS0 + qy*Q + qx]
    if (live && t < P * Q) {
      const int iz = t / Q, qx = t - iz * Q;
      T x[P];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) x[iy] = b1[iz * ZS1 + iy * RQ + qx];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        T acc = T(0);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) acc = mac<EXACT>(B(qy, iy), x[iy], acc);
        b0[iz * ZS0 + qy * Q + qx] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    {
      // S3 (exact): z-lines (qy,qx) = t: D = qdata * (z contraction) -> b0 as [qz][qy][RQ rows]
      const bool zl = live && t < Q2;
      const int qy = t / Q, qx = t - (t / Q) * Q;
      T d[Q];
      if (zl) {
        T x[P];
#pragma unroll
        for (int iz = 0; iz < P; ++iz) x[iz] = b0[iz * ZS0 + t];
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
          T acc = T(0);
#pragma unroll
          for (int iz = 0; iz < P; ++iz) acc = mac<true>(B(qz, iz), x[iz], acc);
          d[qz] = qd[qz] * finish<true>(acc);
        }
      }
      __syncthreads();  // every z-line has been read before D overwrites b0 with padded rows
      if (zl) {
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) b0[(qz * Q + qy) * RQ + qx] = d[qz];
      }
      __syncthreads();
      // T1: x-rows (qz,qy) = t: b0[t*RQ + qx] -> b1[t*RP + ix]
      if (live && t < Q2) {
        T x[Q];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) x[qx] = b0[t * RQ + qx];
#pragma unroll
        for (int ix = 0; ix < P; ++ix) {
          T acc = T(0);
#pragma unroll
          for (int qx = 0; qx < Q; ++qx) acc = mac<true>(B(qx, ix), x[qx], acc);
          b1[t * RP + ix] = finish<true>(acc);
        }
      }
      __syncthreads();
      // T2: y-lines (qz,ix): b1[(qz*Q + qy)*RP + ix] -> b0[(qz*P + iy)*P + ix]
      if (live && t < Q * P) {
        const int qz = t / P, ix = t - qz * P;
        T x[Q];
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) x[qy] = b1[(qz * Q + qy) * RP + ix];
#pragma unroll
        for (int iy = 0; iy < P; ++iy) {
          T acc = T(0);
#pragma unroll
          for (int qy = 0; qy < Q; ++qy) acc = mac<true>(B(qy, iy), x[qy], acc);
          b0[(qz * P + iy) * P + ix] = finish<true>(acc);
        }
      }
      __syncthreads();
      // T3: z-lines (iy,ix) = t: b0[qz*P2 + t] -> V[iz] -> scatter
      if (live && t < P2) {
        T x[Q];
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) x[qz] = b0[qz * P2 + t];
        const int* off = s_off + el * P3;
#pragma unroll
        for (int iz = 0; iz < P; ++iz) {
          T acc = T(0);
#pragma unroll
          for (int qz = 0; qz < Q; ++qz) acc = mac<true>(B(qz, iz), x[qz], acc);
          const T val = finish<true>(acc);
          const int i = iz * P2 + t;
          if constexpr (DET) {
            out[static_cast<size_t>(c) * y_comp_len + static_cast<size_t>(e) * P3 + i] = val;
          } else {
            unsafeAtomicAdd(&out[static_cast<size_t>(off[i]) + coff], val);
          }
        }
      }
    }
    if (c + 1 < NC) __syncthreads();
  }
}


// Fast path (FMA): one LDS buffer per element (layout [iz][y][x], Shape::ZS/RY),
// every stage contracting its own lines in place; the element outputs are then
// scattered with lanes walking consecutive element nodes (coalesced), from an
// LDS copy of the element block's offsets. In the atomic mode, nodes with a single
// contributor (bit 31 of their offset set at create time) are plain stores.
// One workgroup's group of NE elements (elements blk*NE ...), given its LDS:
// the body of fused_mass_fast_kernel, also run per work item by the batched
// element kernel (batch_kernel.hpp).
template <int P, int Q, int NC, class T, int NE, bool DET>
__device__ __forceinline__ void fast_elem_group(int blk, int num_elem, int comp_stride,
                                                const int* __restrict__ offsets, const T* __restrict__ qdata,
                                                const T* __restrict__ u, T* __restrict__ out, size_t y_comp_len,
                                                const MatArg<T, P * Q>& Bm, T (*s_b)[Shape<P, Q>::FBUF],
                                                int* s_off) {
  using S = Shape<P, Q>;
  constexpr int P2 = S::P2, P3 = S::P3, Q2 = S::Q2, Q3 = S::Q3, TPE = S::TPE, RY = S::RY, ZS = S::ZS;
  const long long e0 = static_cast<long long>(blk) * NE;
  if (e0 >= num_elem) return;  // the whole workgroup leaves together
  const int n_here = static_cast<int>(min(static_cast<long long>(NE), num_elem - e0));
  const int el = threadIdx.x / TPE;
  const int t = threadIdx.x - el * TPE;
  const bool live = el < n_here;
  const long long e = e0 + el;
  T* b = live ? s_b[el] : nullptr;
  auto B = [&](int q, int p) -> T { return Bm.m[q * P + p]; };

  T qd[Q];
  if (live && t < Q2) {
    load_qline<Q, T>(qdata + static_cast<size_t>(e) * Q3, t, qd);
  }
  int o[P];
  if (live && t < P2) {
    const int* op = offsets + e * P3 + t * P;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) o[ix] = op[ix] & 0x7fffffff;
  }
  if constexpr (!DET) {
    for (int idx = threadIdx.x; idx < n_here * P3; idx += blockDim.x) s_off[idx] = offsets[e0 * P3 + idx];
  }

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const size_t coff = static_cast<size_t>(c) * comp_stride;
    // S1: x-rows (iz,iy): gather and contract x.
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[P];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) x[ix] = u[static_cast<size_t>(o[ix]) + coff];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T acc = T(0);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) acc = fma_t(B(qx, ix), x[ix], acc);
        b[row + qx] = acc;
      }
    }
    __syncthreads();
    // S2: y-columns (iz,qx), in place.
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[P];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) x[iy] = b[col + iy * RY];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        T acc = T(0);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) acc = fma_t(B(qy, iy), x[iy], acc);
        b[col + qy * RY] = acc;
      }
    }
    __syncthreads();
    // S3: z-lines (qy,qx): contract z, times qdata, contract z back, in place.
    if (live && t < Q2) {
      const int line = (t / Q) * RY + t % Q;
      T x[P];
#pragma unroll
      for (int iz = 0; iz < P; ++iz) x[iz] = b[iz * ZS + line];
      T d[Q];
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        T acc = T(0);
#pragma unroll
        for (int iz = 0; iz < P; ++iz) acc = fma_t(B(qz, iz), x[iz], acc);
        d[qz] = qd[qz] * acc;
      }
#pragma unroll
      for (int iz = 0; iz < P; ++iz) {
        T acc = T(0);
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) acc = fma_t(B(qz, iz), d[qz], acc);
        b[iz * ZS + line] = acc;
      }
    }
    __syncthreads();
    // S4: y-columns back (iz,qx), in place.
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[Q];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) x[qy] = b[col + qy * RY];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) {
        T acc = T(0);
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) acc = fma_t(B(qy, iy), x[qy], acc);
        b[col + iy * RY] = acc;
      }
    }
    __syncthreads();
    // S5: x-rows back (iz,iy), in place.
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[Q];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) x[qx] = b[row + qx];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) {
        T acc = T(0);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) acc = fma_t(B(qx, ix), x[qx], acc);
        b[row + ix] = acc;
      }
    }
    __syncthreads();
    // Scatter, lanes over consecutive element nodes i = (iz*P + iy)*P + ix.
    if (live) {
      for (int i = t; i < P3; i += TPE) {
        const int iz = i / P2, iy = (i / P) % P, ix = i % P;
        const T val = b[iz * ZS + iy * RY + ix];
        if constexpr (DET) {
          out[static_cast<size_t>(c) * y_comp_len + static_cast<size_t>(e) * P3 + i] = val;
        } else {
          const int raw = s_off[el * P3 + i];
          T* dst = &out[static_cast<size_t>(raw & 0x7fffffff) + coff];
          if (raw < 0) {
            *dst = val;  // the only contribution to this node
          } else {
            unsafeAtomicAdd(dst, val);
          }
        }
      }
    }
    if (c + 1 < NC) __syncthreads();
  }
}

template <int P, int Q, int NC, class T, int NE, bool DET>
__global__ __launch_bounds__(round64(NE* Shape<P, Q>::TPE)) void fused_mass_fast_kernel(
    int num_elem, int comp_stride, const int* __restrict__ offsets, const T* __restrict__ qdata,
    const T* __restrict__ u, T* __restrict__ out, size_t y_comp_len, MatArg<T, P * Q> Bm) {
  using S = Shape<P, Q>;
  __shared__ T s_b[NE][S::FBUF];
  __shared__ int s_off[DET ? 1 : NE * S::P3];
  const int per = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
  fast_elem_group<P, Q, NC, T, NE, DET>(blk, num_elem, comp_stride, offsets, qdata, u, out, y_comp_len, Bm, s_b, s_off);
}

// Block-accumulated variant of the fast path (atomic scatter mode). A workgroup
// handles one block of NB elements chosen at create time (a compact patch of
// the mesh, see block_plan.cpp). S5 adds each element output into an LDS
// accumulator per unique block node; at the end, block-interior nodes (no
// reference outside the block) are written with plain stores and the block's
// boundary nodes with one device atomic each, lanes walking the sorted node
// list. Compared with one atomic per element node, the atomic count drops to the
// block boundary (about a third at P=5 for 16-element blocks).
template <int P, int Q, int NC, class T, int NB>
__global__ __launch_bounds__(round64(NB* Shape<P, Q>::TPE)) void fused_mass_block_kernel(
    int nblk, int comp_stride, const int* __restrict__ node_ptr, const int* __restrict__ nodes,
    const int* __restrict__ ninterior, const int* __restrict__ nelem, const uint16_t* __restrict__ loc,
    const T* __restrict__ qdata, const T* __restrict__ u, T* __restrict__ v, MatArg<T, P * Q> Bm) {
  using S = Shape<P, Q>;
  constexpr int P2 = S::P2, P3 = S::P3, Q2 = S::Q2, Q3 = S::Q3, TPE = S::TPE, RY = S::RY, ZS = S::ZS;
  constexpr int MAXN = NB * P3;
  __shared__ T s_b[NB][S::FBUF];
  __shared__ int s_nodes[MAXN];
  __shared__ T s_acc[NC][MAXN];

  const int per = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (blk >= nblk) return;  // whole workgroup
  const int n0 = node_ptr[blk], nn = node_ptr[blk + 1] - n0, nint = ninterior[blk], ne_here = nelem[blk];

  for (int j = threadIdx.x; j < nn; j += blockDim.x) {
    s_nodes[j] = nodes[n0 + j];
#pragma unroll
    for (int c = 0; c < NC; ++c) s_acc[c][j] = T(0);
  }
  const int el = threadIdx.x / TPE;
  const int t = threadIdx.x - el * TPE;
  const bool live = el < ne_here;
  const size_t slot = static_cast<size_t>(blk) * NB + el;
  T* b = live ? s_b[el] : nullptr;
  auto B = [&](int q, int p) -> T { return Bm.m[q * P + p]; };

  T qd[Q];
  if (live && t < Q2) {
    load_qline<Q, T>(qdata + slot * Q3, t, qd);
  }
  int lo[P];
  if (live && t < P2) {
    const uint16_t* lp = loc + slot * P3 + t * P;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) lo[ix] = lp[ix];
  }
  __syncthreads();

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const size_t coff = static_cast<size_t>(c) * comp_stride;
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[P];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) x[ix] = u[static_cast<size_t>(s_nodes[lo[ix]]) + coff];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T acc = T(0);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) acc = fma_t(B(qx, ix), x[ix], acc);
        b[row + qx] = acc;
      }
    }
    __syncthreads();
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[P];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) x[iy] = b[col + iy * RY];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        T acc = T(0);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) acc = fma_t(B(qy, iy), x[iy], acc);
        b[col + qy * RY] = acc;
      }
    }
    __syncthreads();
    if (live && t < Q2) {
      const int line = (t / Q) * RY + t % Q;
      T x[P];
#pragma unroll
      for (int iz = 0; iz < P; ++iz) x[iz] = b[iz * ZS + line];
      T d[Q];
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        T acc = T(0);
#pragma unroll
        for (int iz = 0; iz < P; ++iz) acc = fma_t(B(qz, iz), x[iz], acc);
        d[qz] = qd[qz] * acc;
      }
#pragma unroll
      for (int iz = 0; iz < P; ++iz) {
        T acc = T(0);
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) acc = fma_t(B(qz, iz), d[qz], acc);
        b[iz * ZS + line] = acc;
      }
    }
    __syncthreads();
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[Q];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) x[qy] = b[col + qy * RY];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) {
        T acc = T(0);
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) acc = fma_t(B(qy, iy), x[qy], acc);
        b[col + iy * RY] = acc;
      }
    }
    __syncthreads();
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[Q];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) x[qx] = b[row + qx];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) {
        T acc = T(0);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) acc = fma_t(B(qx, ix), x[qx], acc);
        unsafeAtomicAdd(&s_acc[c][lo[ix]], acc);  // LDS add
      }
    }
    if (c + 1 < NC) __syncthreads();
  }
  __syncthreads();
  // Flush: interior nodes are complete here; boundary nodes need the other blocks' parts.
#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const size_t coff = static_cast<size_t>(c) * comp_stride;
    for (int j = threadIdx.x; j < nn; j += blockDim.x) {
      const size_t g = static_cast<size_t>(s_nodes[j]) + coff;
      const T val = s_acc[c][j];
      if (j < nint) {
        v[g] = val;
      } else {
        unsafeAtomicAdd(&v[g], val);
      }
    }
  }
}

// Block-accumulated variant with a deterministic scatter: like the kernel
// above, but the block's u values are staged in LDS first (one load per unique
// node instead of one per element node), element outputs are added into the
// node sums colour by colour (slots of one colour share no node, so the adds
// need no atomics and their order is fixed), interior nodes are stored and each
// boundary node's block sum goes to its own partial-sum slot; a second pass
// (the row-sum pass) adds a node's partials in ascending block order. The
// result is the same bit for bit on every run. With LDS_ATOMIC the element
// outputs are added with LDS atomics instead (one barrier in place of the
// colour rounds; the order of the adds, and so the rounding, varies by run).
template <int P, int Q, int NC, class T, int NB, bool LDS_ATOMIC>
__global__ __launch_bounds__(round64(NB* Shape<P, Q>::TPE)) void fused_mass_part_kernel(
    int nblk, int comp_stride, const int* __restrict__ node_ptr, const int* __restrict__ nodes,
    const int* __restrict__ ninterior, const int* __restrict__ nelem, const uint8_t* __restrict__ color,
    const int* __restrict__ ncolor, const int* __restrict__ part_ptr, const uint16_t* __restrict__ loc,
    const T* __restrict__ qdata, const T* __restrict__ u, T* __restrict__ v, T* __restrict__ partials,
    size_t part_comp_len, int ldn, MatArg<T, P * Q> Bm) {
  using S = Shape<P, Q>;
  constexpr int P2 = S::P2, P3 = S::P3, Q2 = S::Q2, Q3 = S::Q3, TPE = S::TPE, RY = S::RY, ZS = S::ZS;
  constexpr int NT = round64(NB * TPE);
  constexpr int KN = (NB * P3 + NT - 1) / NT;  // node-list entries per thread (a block has <= NB*P3 nodes)
  __shared__ T s_b[NB][S::FBUF];
  extern __shared__ __align__(16) unsigned char s_dyn[];
  T* s_n = reinterpret_cast<T*>(s_dyn);  // [NC][ldn]: the block's u values, then its node sums

  const int per = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (blk >= nblk) return;  // whole workgroup
  const int n0 = node_ptr[blk], nn = node_ptr[blk + 1] - n0, nint = ninterior[blk], ne_here = nelem[blk];
  const int ncol = ncolor[blk];
  const int tid = threadIdx.x;

  int gn[KN];
#pragma unroll
  for (int k = 0; k < KN; ++k) {
    const int j = tid + k * NT;
    gn[k] = j < nn ? nodes[n0 + j] : 0;
  }
  const int el = tid / TPE;
  const int t = tid - el * TPE;
  const bool live = el < ne_here;
  const size_t slot = static_cast<size_t>(blk) * NB + el;
  T* b = live ? s_b[el] : nullptr;
  auto B = [&](int q, int p) -> T { return Bm.m[q * P + p]; };

  T qd[Q];
  if (live && t < Q2) {
    load_qline<Q, T>(qdata + slot * Q3, t, qd);
  }
  int lo[P];
  if (live && t < P2) {
    const uint16_t* lp = loc + slot * P3 + t * P;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) lo[ix] = lp[ix];
  }
  const int my_color = live ? color[slot] : -1;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      const int j = tid + k * NT;
      if (j < nn) s_n[c * ldn + j] = u[static_cast<size_t>(gn[k]) + static_cast<size_t>(c) * comp_stride];
    }
  }
  __syncthreads();

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    T* sn = s_n + c * ldn;
    // S1: x-rows (iz,iy): gather from the staged node values, contract x.
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[P];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) x[ix] = sn[lo[ix]];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T acc = T(0);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) acc = fma_t(B(qx, ix), x[ix], acc);
        b[row + qx] = acc;
      }
    }
    __syncthreads();
    // S2: y-columns (iz,qx), in place; the component's u values are consumed, so
    // its slice of s_n is cleared for the sums.
    for (int j = tid; j < nn; j += NT) sn[j] = T(0);
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[P];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) x[iy] = b[col + iy * RY];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        T acc = T(0);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) acc = fma_t(B(qy, iy), x[iy], acc);
        b[col + qy * RY] = acc;
      }
    }
    __syncthreads();
    // S3: z-lines (qy,qx): contract z, times qdata, contract z back.
    if (live && t < Q2) {
      const int line = (t / Q) * RY + t % Q;
      T x[P];
#pragma unroll
      for (int iz = 0; iz < P; ++iz) x[iz] = b[iz * ZS + line];
      T d[Q];
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        T acc = T(0);
#pragma unroll
        for (int iz = 0; iz < P; ++iz) acc = fma_t(B(qz, iz), x[iz], acc);
        d[qz] = qd[qz] * acc;
      }
#pragma unroll
      for (int iz = 0; iz < P; ++iz) {
        T acc = T(0);
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) acc = fma_t(B(qz, iz), d[qz], acc);
        b[iz * ZS + line] = acc;
      }
    }
    __syncthreads();
    // S4: y-columns back (iz,qx), in place.
    if (live && t < P * Q) {
      const int col = (t / Q) * ZS + t % Q;
      T x[Q];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) x[qy] = b[col + qy * RY];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) {
        T acc = T(0);
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) acc = fma_t(B(qy, iy), x[qy], acc);
        b[col + iy * RY] = acc;
      }
    }
    __syncthreads();
    // S5: x-rows back (iz,iy) into registers, then the colour rounds.
    T out[P];
    if (live && t < P2) {
      const int row = (t / P) * ZS + (t % P) * RY;
      T x[Q];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) x[qx] = b[row + qx];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) {
        T acc = T(0);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) acc = fma_t(B(qx, ix), x[qx], acc);
        out[ix] = acc;
      }
    }
    if constexpr (LDS_ATOMIC) {
      // Any order (atomic scatter mode): LDS atomic adds, one barrier.
      if (live && t < P2) {
#pragma unroll
        for (int ix = 0; ix < P; ++ix) unsafeAtomicAdd(&sn[lo[ix]], out[ix]);
      }
      __syncthreads();
    } else {
#pragma unroll 1
      for (int k = 0; k < ncol; ++k) {
        if (my_color == k && t < P2) {
#pragma unroll
          for (int ix = 0; ix < P; ++ix) sn[lo[ix]] += out[ix];
        }
        __syncthreads();
      }
    }
  }
  // Flush: interior nodes are complete; each boundary node's block sum is a partial.
  const size_t pb = static_cast<size_t>(part_ptr[blk]);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      const int j = tid + k * NT;
      if (j < nn) {
        const T val = s_n[c * ldn + j];
        if (j < nint) {
          v[static_cast<size_t>(gn[k]) + static_cast<size_t>(c) * comp_stride] = val;
        } else {
          partials[static_cast<size_t>(c) * part_comp_len + pb + (j - nint)] = val;
        }
      }
    }
  }
}

// Node-sum values a block kernel can keep in LDS (0 = the kernel sizes its own).
template <int P, int Q, int NC, class T, int NB>
constexpr int part_node_budget() {
  using S = Shape<P, Q>;
  constexpr int fixed = NB * S::FBUF * static_cast<int>(sizeof(T));
  // Aim for four workgroups per CU (40 KiB each); large elements get 64 KiB.
  constexpr int target = (40 * 1024 - fixed) / (NC * static_cast<int>(sizeof(T))) >= 2 * S::P3 ? 40 * 1024 : 64 * 1024;
  return (target - fixed) / (NC * static_cast<int>(sizeof(T)));
}

template <int P, int Q, int NC, class T>
MatArg<T, P * Q> basis_arg(const void* basis) {
  MatArg<T, P * Q> Bm;
  const T* src = static_cast<const T*>(basis);
  for (int i = 0; i < P * Q; ++i) Bm.m[i] = src[i];
  return Bm;
}

template <int P, int Q, int NC, class T, int NB>
void launch_block_atomic(const FusedArgs& a, hipStream_t s) {
  using S = Shape<P, Q>;
  const long long grid = (static_cast<long long>(a.nblk) + 7) / 8 * 8;
  require(grid < (1ll << 31), "fused mass: too many blocks for one launch");
  require(a.max_nodes <= NB * S::P3, "fused mass: block plan exceeds the kernel's node capacity");
  constexpr int threads = round64(NB * S::TPE);
  fused_mass_block_kernel<P, Q, NC, T, NB><<<static_cast<unsigned>(grid), threads, 0, s>>>(
      a.nblk, a.comp_stride, a.node_ptr, a.nodes, a.ninterior, a.nelem, a.loc, static_cast<const T*>(a.qdata),
      static_cast<const T*>(a.u), static_cast<T*>(a.v), basis_arg<P, Q, NC, T>(a.basis));
  HIP_CHECK_LAUNCH();
}

template <int P, int Q, int NC, class T, int NB, bool LDS_ATOMIC>
void launch_block_part(const FusedArgs& a, hipStream_t s) {
  using S = Shape<P, Q>;
  const long long grid = (static_cast<long long>(a.nblk) + 7) / 8 * 8;
  require(grid < (1ll << 31), "fused mass: too many blocks for one launch");
  require(a.max_nodes > 0 && a.max_nodes <= part_node_budget<P, Q, NC, T, NB>() && a.max_nodes <= NB * S::P3,
          "fused mass: block plan exceeds the kernel's LDS node budget");
  constexpr int threads = round64(NB * S::TPE);
  const size_t lds = static_cast<size_t>(NC) * a.max_nodes * sizeof(T);
  fused_mass_part_kernel<P, Q, NC, T, NB, LDS_ATOMIC><<<static_cast<unsigned>(grid), threads, lds, s>>>(
      a.nblk, a.comp_stride, a.node_ptr, a.nodes, a.ninterior, a.nelem, a.color, a.ncolor, a.part_ptr, a.loc,
      static_cast<const T*>(a.qdata), static_cast<const T*>(a.u), static_cast<T*>(a.v), static_cast<T*>(a.partials),
      a.part_comp_len, a.max_nodes, basis_arg<P, Q, NC, T>(a.basis));
  HIP_CHECK_LAUNCH();
}

template <int P, int Q, int NC, class T, int NE, bool EXACT, bool DET>
void launch_fused(const FusedArgs& a, hipStream_t s) {
  using S = Shape<P, Q>;
  const auto Bm = basis_arg<P, Q, NC, T>(a.basis);
  const long long blocks = (static_cast<long long>(a.num_elem) + NE - 1) / NE;
  const long long grid = (blocks + 7) / 8 * 8;
  require(grid < (1ll << 31), "fused mass: too many elements for one launch");
  require(static_cast<long long>(a.num_elem) * S::P3 < (1ll << 31), "fused mass: offsets exceed int32 indexing");
  constexpr int threads = round64(NE * S::TPE);
  if constexpr (EXACT) {
    fused_mass_exact_kernel<P, Q, NC, T, NE, DET><<<static_cast<unsigned>(grid), threads, 0, s>>>(
        a.num_elem, a.comp_stride, a.offsets, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u),
        static_cast<T*>(a.out), a.y_comp_len, Bm);
  } else {
    fused_mass_fast_kernel<P, Q, NC, T, NE, DET><<<static_cast<unsigned>(grid), threads, 0, s>>>(
        a.num_elem, a.comp_stride, a.flagged_offsets, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u),
        static_cast<T*>(a.out), a.y_comp_len, Bm);
  }
  HIP_CHECK_LAUNCH();
}

template <int P, int Q, int NC, class T, int NE, bool EXACT, bool DET>
struct FusedReg {
  FusedReg() {
    register_fused(FusedKey{P, Q, NC, std::is_same<T, float>::value ? 1 : 0, DET ? kAlgElemTmap : kAlgElemAtomic,
                            EXACT ? 1 : 0, NE},
                   &launch_fused<P, Q, NC, T, NE, EXACT, DET>, 0);
  }
};

// All element-algorithm variants of one shape: FP64 {atomic, tmap} x {fast, exact}, FP32 {atomic, tmap}.
template <int P, int Q, int NC, int NE64, int NE32>
struct FusedShapeReg {
  FusedReg<P, Q, NC, double, NE64, false, false> a;
  FusedReg<P, Q, NC, double, NE64, false, true> b;
  FusedReg<P, Q, NC, double, NE64, true, true> c;
  FusedReg<P, Q, NC, double, NE64, true, false> d;
  FusedReg<P, Q, NC, float, NE32, false, false> e;
  FusedReg<P, Q, NC, float, NE32, false, true> f;
};

template <int P, int Q, int NC, class T, int NB>
struct FusedBlockReg {
  FusedBlockReg() {
    const int fp32 = std::is_same<T, float>::value ? 1 : 0;
    register_fused(FusedKey{P, Q, NC, fp32, kAlgBlockAtomic, 0, NB}, &launch_block_atomic<P, Q, NC, T, NB>, 0);
    register_fused(FusedKey{P, Q, NC, fp32, kAlgBlockPart, 0, NB}, &launch_block_part<P, Q, NC, T, NB, false>,
                   part_node_budget<P, Q, NC, T, NB>());
    register_fused(FusedKey{P, Q, NC, fp32, kAlgBlockPartLds, 0, NB}, &launch_block_part<P, Q, NC, T, NB, true>,
                   part_node_budget<P, Q, NC, T, NB>());
  }
};

// Block-algorithm variants (atomic and partial-sum; FP64 and FP32).
template <int P, int Q, int NC, int NB64, int NB32>
struct FusedBlockShapeReg {
  FusedBlockReg<P, Q, NC, double, NB64> a;
  FusedBlockReg<P, Q, NC, float, NB32> b;
};

}  // namespace femhip
