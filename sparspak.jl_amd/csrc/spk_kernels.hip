// Kernels of the numeric phase. See spk_kernels.h for the storage model.
#include "spk_kernels.h"

namespace spk {

// Index of the calling thread's wave in its workgroup, as a wave-uniform
// (scalar) value so that per-wave index arithmetic stays on the scalar unit.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / 64); }

// Orders one wave's LDS accesses: every access before it is complete and
// visible to the wave's other lanes before any access after it. The
// wavefront-scope fences emit no instructions (one wave's LDS operations
// complete in issue order) but keep the compiler from moving LDS loads or
// stores across the barrier, which the barrier intrinsic alone does not.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int64_t ld_i64(const int64_t* p, int64_t i) { return p[i]; }

// Column range of piece p (0-based supernode): first column (0-based) and
// width.
__device__ inline void piece_cols(const DevSym& s, int64_t p, int64_t& c0,
                                  int64_t& w) {
  c0 = s.xsuper[p] - 1;
  w = s.xsuper[p + 1] - s.xsuper[p];
}

// First (front-local) column of the piece holding own column c.
__device__ inline int piece_start_of_col(const DevSym& s, int64_t fc0, int c) {
  return (int)(s.xsuper[s.snode[fc0 + c] - 1] - 1 - fc0);
}

// Unit-stride element of the reference layout for front entry (r, j).
// Returns the lnz (is_u = false) or unz (is_u = true) offset, or -1 for the
// update-matrix block (r >= W and j >= W), which has no slot.
__device__ inline int64_t slot_of(const DevSym& s, int64_t fc0, int64_t W,
                                  int64_t r, int64_t j, bool& is_u) {
  if (j < W) {
    int64_t g = fc0 + j;
    int64_t p = s.snode[g] - 1, pc0, pw;
    piece_cols(s, p, pc0, pw);
    int64_t cq = pc0 - fc0;
    if (r >= cq) {
      is_u = false;
      return s.xlnz[g] - 1 + (r - cq);
    }
  } else if (r >= W) {
    return -1;
  }
  // U entry: row r lies in an earlier piece (or j is an update column).
  int64_t gr = fc0 + r;
  int64_t pr = s.snode[gr] - 1, rc0, rw;
  piece_cols(s, pr, rc0, rw);
  int64_t cr = rc0 - fc0;
  is_u = true;
  return s.xunz[gr] - 1 + (j - cr - rw);
}

// ---------------------------------------------------------------------------
// Assembly: zero lnz/unz, scatter A's values through the precomputed map.
// amap[k] >= 0: lnz slot; amap[k] <= -2: unz slot (-amap-2); -1: dropped
// (the reference's search runs off the end of the index list and ignores the
// entry, SpkSparseBase.jl:328-341).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T* __restrict__ a, int64_t len, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < len; i += stride) a[i] = v;
}

// A's values added into their slots of the zeroed lnz / unz (amap[k] >= 0:
// lnz slot, <= -2: unz slot -(amap[k] + 2), -1: not stored; distinct entries
// of a CSC pattern map to distinct slots). Value set blockIdx.y: its nnz
// values, its lnz / unz at vs_l / vs_u apart.
template <typename T>
__global__ void scatter_a_kernel(const double* __restrict__ nzval,
                                 const int64_t* __restrict__ amap,
                                 int64_t nnz, T* __restrict__ lnz,
                                 T* __restrict__ unz, int64_t vs_l, int64_t vs_u) {
  const int64_t vs = blockIdx.y;
  nzval += vs * nnz;
  lnz += vs * vs_l;
  unz += vs * vs_u;
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int64_t d = amap[k];
  T v = (T)nzval[k];
  if (d >= 0)
    lnz[d] += v;
  else if (d <= -2)
    unz[-d - 2] += v;
}

// ---------------------------------------------------------------------------
// Partial LU of one front held in LDS (F, leading dimension L): pieces are
// eliminated column by column with row pivoting confined to each piece's
// diagonal block, exactly as _getrf! (SpkLUFactor.jl:230) followed by the two
// TRSMs and the U12 row swap (:235-240): a row interchange spans the piece's
// own columns and everything to their right, never earlier pieces' columns;
// multipliers are formed with the reciprocal of the pivot as ggetrf!/gtrsm!
// do (GenericBlasLapackFragments.jl:80-84, 376-395).
// ---------------------------------------------------------------------------
template <typename T, int G>
__device__ void lds_partial_lu(T* F, int L, int cq, int nq, int pivot,
                               int* piv_out, int* zero_out, int gt, int* s_kp_, T* s_inv_) {
  // G threads (a workgroup of 256 or one wave of 64), gt = thread index in
  // the group; s_kp_ / s_inv_ one LDS slot of the group
  auto gsync = [] {
    if constexpr (G == kWG)
      __syncthreads();
    else
      wave_lds_sync();
  };
  for (int k = 0; k < nq; ++k) {
    const int col = cq + k;
    if (gt == 0) {
      int kp = col;
      if (pivot) {
        T amax = fabs(F[col + col * L]);
        for (int i = col + 1; i < cq + nq; ++i) {
          T a = fabs(F[i + col * L]);
          if (a > amax) {
            amax = a;
            kp = i;
          }
        }
      }
      *s_kp_ = kp;
      T pv = F[kp + col * L];
      if (pv == T(0)) *zero_out = 1;
      *s_inv_ = T(1) / pv;
      piv_out[k] = kp - cq;
    }
    gsync();
    const int kp = *s_kp_;
    const bool nonzero = (F[kp + col * L] != T(0));
    if (kp != col && nonzero) {
      for (int c = cq + gt; c < L; c += G) {
        T t = F[col + c * L];
        F[col + c * L] = F[kp + c * L];
        F[kp + c * L] = t;
      }
    }
    gsync();
    const T inv = *s_inv_;
    // Rows inside the diagonal block are scaled only for a nonzero pivot
    // (ggetrf!); rows below it always (gtrsm! multiplies by 1/U(j,j)).
    for (int r = col + 1 + gt; r < L; r += G)
      if (r >= cq + nq || nonzero) F[r + col * L] *= inv;
    gsync();
    const int m = L - col - 1;
    const int cnt = m * m;
    for (int idx = gt; idx < cnt; idx += G) {
      int r = col + 1 + idx % m;
      int c = col + 1 + idx / m;
      F[r + c * L] -= F[r + col * L] * F[col + c * L];
    }
    gsync();
  }
}

// One small front per group of G threads (G = 256: a workgroup; G = 64: a
// wave, fpw fronts per workgroup, for levels of many tiny fronts): gather
// the pieces' panels and the children's update matrices into LDS (L x L per
// front), factor, write the panels back to lnz/unz/ipvt and the update
// matrix to the arena.
template <typename T, int G>
__global__ void __launch_bounds__(kWG)
small_front_kernel(DevSym s, const int32_t* __restrict__ list, int cnt, int fpw, int maxL,
                   T* __restrict__ lnz, T* __restrict__ unz,
                   int64_t* __restrict__ ipvt, T* __restrict__ arena,
                   unsigned long long* __restrict__ first_fail) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  unz += vs * s.vs_unz;
  ipvt += vs * s.n;
  arena += vs * s.vs_arena;
  extern __shared__ unsigned char smem_raw[];
  __shared__ int s_piv_all[kWG / G][kSmallLmax];  // a piece of a small front has <= L columns
  __shared__ int s_zero_all[kWG / G];
  __shared__ int s_kp_all[kWG / G];
  __shared__ T s_inv_all[kWG / G];
  const int grp = (G == kWG) ? 0 : wave_id();
  const int tid = (G == kWG) ? (int)threadIdx.x : (int)threadIdx.x % 64;
  const int item = blockIdx.x * fpw + grp;
  if (grp >= fpw || item >= cnt) return;
  auto gsync = [] {
    if constexpr (G == kWG)
      __syncthreads();
    else
      wave_lds_sync();
  };
  T* F = reinterpret_cast<T*>(smem_raw) + (int64_t)grp * maxL * maxL;
  int* s_piv = s_piv_all[grp];
  int& s_zero = s_zero_all[grp];
  const int f = list[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  for (int i = tid; i < L * L; i += G) F[i] = T(0);
  gsync();
  // pieces' panels
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    const int jl = L - cq;
    const T* lp = lnz + (s.xlnz[pc0] - 1);
    for (int idx = tid; idx < jl * nq; idx += G) {
      int r = idx % jl, c = idx / jl;
      if (!s.sym || r >= c) F[(cq + r) + (cq + c) * L] = lp[idx];
    }
    if (!s.sym) {
      const int ul = jl - (int)nq;
      const T* up = unz + (s.xunz[pc0] - 1);
      for (int idx = tid; idx < ul * nq; idx += G) {
        int r = idx % ul, c = idx / ul;
        F[(cq + c) + (cq + (int)nq + r) * L] = up[idx];
      }
    }
  }
  gsync();
  if (s.sym) {  // upper rows of the own columns from the lower triangle
    for (int idx = tid; idx < W * L; idx += G) {
      int r = idx % W, j = idx / W;
      if (j > r) F[r + j * L] = F[j + r * L];
    }
    gsync();
  }
  // children's update matrices (one child at a time: rows of different
  // children may coincide)
  for (int ci = s.fchp[f]; ci < s.fchp[f + 1]; ++ci) {
    const int c = s.fch[ci];
    const int Rc = s.fL[c] - s.fW[c];
    const int32_t* rel = s.relmap + s.frel[c];
    const T* cb = arena + s.fcb[c];
    const int ld = s.fcbld[c];
    for (int idx = tid; idx < Rc * Rc; idx += G) {
      int tp = idx % Rc, t = idx / Rc;
      F[rel[tp] + rel[t] * L] += cb[tp + (int64_t)t * ld];
    }
    gsync();
  }
  // factor piece by piece
  if (tid == 0) s_zero = 0;
  gsync();
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    lds_partial_lu<T, G>(F, L, cq, (int)nq, s.pivot, s_piv, &s_zero, tid, &s_kp_all[grp],
                         &s_inv_all[grp]);
    for (int k = tid; k < nq; k += G) ipvt[pc0 + k] = s_piv[k] + 1;
    if (tid == 0 && s_zero) {
      atomicMin(first_fail, (unsigned long long)(p + 1));
      s_zero = 0;
    }
    gsync();
  }
  // write back panels
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    const int jl = L - cq;
    T* lp = lnz + (s.xlnz[pc0] - 1);
    for (int idx = tid; idx < jl * nq; idx += G) {
      int r = idx % jl, c = idx / jl;
      lp[idx] = F[(cq + r) + (cq + c) * L];
    }
    const int ul = jl - (int)nq;
    T* up = unz + (s.xunz[pc0] - 1);
    for (int idx = tid; idx < ul * nq; idx += G) {
      int r = idx % ul, c = idx / ul;
      up[idx] = F[(cq + c) + (cq + (int)nq + r) * L];
    }
  }
  // update matrix
  const int R = L - W;
  T* cb = arena + s.fcb[f];
  for (int idx = tid; idx < R * R; idx += G) {
    int tp = idx % R, t = idx / R;
    cb[idx] = F[(W + tp) + (W + t) * L];
  }
}

// ---------------------------------------------------------------------------
// Large fronts: the L x L work matrix lives in the arena.
// ---------------------------------------------------------------------------

// U part of a large front (rows of a piece, columns right of it) moves
// between unz (U12^T panels: a piece's U rows are its panel's columns) and
// the work matrix through 64-row x nq tiles transposed in LDS, so that both
// sides are read/written along their contiguous dimension. For SPD input the
// U entries are the transposed L entries (lnz panel rows below the block).
template <typename T, bool STORE>
__global__ void __launch_bounds__(kWG)
large_utile_kernel(DevSym s, const int32_t* __restrict__ it_front,
                   const int32_t* __restrict__ it_piece, const int32_t* __restrict__ it_r0,
                   T* __restrict__ lnz, T* __restrict__ unz, T* __restrict__ arena) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  unz += vs * s.vs_unz;
  arena += vs * s.vs_arena;
  __shared__ T tile[64 * 97];
  const int f = it_front[blockIdx.x];
  const int p = it_piece[blockIdx.x];
  const int r0 = it_r0[blockIdx.x];
  const int L = s.fL[f];
  const int64_t fc0 = s.fc0[f];
  int64_t pc0, nq64;
  piece_cols(s, p, pc0, nq64);
  const int nq = (int)nq64;
  const int cq = (int)(pc0 - fc0);
  const int ul = L - cq - nq;
  const int nr = min(64, ul - r0);
  const int ld = nq + 1;
  T* F = arena + s.foff[f];
  const bool from_l = s.sym && !STORE;
  T* src = from_l ? lnz + (s.xlnz[pc0] - 1) + nq : unz + (s.xunz[pc0] - 1);
  const int64_t sld = from_l ? (int64_t)(L - cq) : (int64_t)ul;
  if (!STORE) {
    for (int idx = threadIdx.x; idx < 64 * nq; idx += kWG) {
      const int r = idx % 64, c = idx / 64;
      if (r < nr) tile[r * ld + c] = src[(r0 + r) + c * sld];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * nq; idx += kWG) {
      const int c = idx % nq, r = idx / nq;
      if (r < nr) F[(cq + c) + (int64_t)(cq + nq + r0 + r) * L] = tile[r * ld + c];
    }
  } else {
    for (int idx = threadIdx.x; idx < 64 * nq; idx += kWG) {
      const int c = idx % nq, r = idx / nq;
      if (r < nr) tile[r * ld + c] = F[(cq + c) + (int64_t)(cq + nq + r0 + r) * L];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * nq; idx += kWG) {
      const int r = idx % 64, c = idx / 64;
      if (r < nr) unz[(s.xunz[pc0] - 1) + (r0 + r) + (int64_t)c * ul] = tile[r * ld + c];
    }
  }
}

// One workgroup per (front, column j): the column's L part from lnz (rows
// from its piece's first row), zeros in the update block, then the
// children's update-matrix columns that map to j, one child after the other
// (rows of different children may coincide). Runs after large_utile_kernel
// has placed the U part. With staged set the column is assembled in LDS
// (dynamic, L elements): read once (the U part from the front, the rest from
// lnz or zero), the children's columns added there, written once; otherwise
// the children's entries are added in the front in place. The item and each
// child's entry come resolved (InitItem, EaItem); the next child's entry is
// read while the current child's column is added.
template <typename T>
__global__ void __launch_bounds__(kWG)
large_init_kernel(DevSym s, const InitItem* __restrict__ items,
                  const int64_t* __restrict__ it_ea, const EaItem* __restrict__ ea,
                  const T* __restrict__ lnz, T* __restrict__ arena, int staged) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  arena += vs * s.vs_arena;
  extern __shared__ unsigned char smem_raw[];
  const int item = blockIdx.x;
  const InitItem m = items[item];
  const int L = m.L, W = m.W, j = m.j;
  T* F = arena + m.fcol;
  const int64_t e0 = it_ea[item], e1 = it_ea[item + 1];
  EaItem en = ea[e0 < e1 ? e0 : 0];
  if (staged && !s.sym) {
    T* Fs = reinterpret_cast<T*>(smem_raw);
    // rows below r1 come from lnz (own column) or are zero (update column);
    // rows above it are the U part already in the front
    const int r1 = m.r1;
    const T* lc = (m.lcb >= 0) ? lnz + m.lcb : nullptr;
    for (int r0 = threadIdx.x; r0 < L; r0 += 16 * kWG) {
      T v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int r = r0 + u * kWG;
        v[u] = (r >= L) ? T(0) : (r < r1) ? F[r] : (lc ? lc[r] : T(0));
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (r0 + u * kWG < L) Fs[r0 + u * kWG] = v[u];
    }
    __syncthreads();
    for (int64_t e = e0; e < e1; ++e) {
      const EaItem ec = en;
      if (e + 1 < e1) en = ea[e + 1];
      const int Rc = ec.Rc;
      const int32_t* rel = s.relmap + ec.rel;
      const T* col = arena + ec.col;
      for (int tp0 = threadIdx.x; tp0 < Rc; tp0 += 8 * kWG) {
        int ri[8];
        T cv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int tp = tp0 + u * kWG;
          ri[u] = (tp < Rc) ? rel[tp] : 0;
          cv[u] = (tp < Rc) ? col[tp] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (tp0 + u * kWG < Rc) Fs[ri[u]] += cv[u];
      }
      __syncthreads();
    }
    for (int r0 = threadIdx.x; r0 < L; r0 += 8 * kWG) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (r0 + u * kWG < L) ? Fs[r0 + u * kWG] : T(0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (r0 + u * kWG < L) F[r0 + u * kWG] = v[u];
    }
    return;
  }
  if (j < W) {
    const int cp = m.r1;
    const T* lc = lnz + m.lcb;
    if (s.sym) {  // lower triangle only: the diagonal block's upper part is U
      for (int r = cp + threadIdx.x; r < L; r += kWG)
        if (r >= j) F[r] = lc[r];
        else F[r] = lnz[s.lcol[m.fc0 + r] + j];
    } else {
      // eight loads in flight per thread
      for (int r0 = cp + threadIdx.x; r0 < L; r0 += 8 * kWG) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = (r0 + u * kWG < L) ? lc[r0 + u * kWG] : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (r0 + u * kWG < L) F[r0 + u * kWG] = v[u];
      }
    }
  } else {
    for (int r = W + threadIdx.x; r < L; r += kWG) F[r] = T(0);
  }
  __syncthreads();
  for (int64_t e = e0; e < e1; ++e) {
    const EaItem ec = en;
    if (e + 1 < e1) en = ea[e + 1];
    const int Rc = ec.Rc;
    const int32_t* rel = s.relmap + ec.rel;
    const T* col = arena + ec.col;
    for (int tp0 = threadIdx.x; tp0 < Rc; tp0 += 4 * kWG) {
      int ri[4];
      T cv[4], fv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int tp = tp0 + u * kWG;
        ri[u] = (tp < Rc) ? rel[tp] : 0;
        cv[u] = (tp < Rc) ? col[tp] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) fv[u] = F[ri[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (tp0 + u * kWG < Rc) F[ri[u]] = fv[u] + cv[u];
    }
    __syncthreads();
  }
}

// One wave per front column, four per workgroup: the staged (LDS) assembly
// of large_init_kernel for LU levels whose longest column fits a quarter of
// the LDS budget, so that the many short columns of the lower levels take a
// quarter of the workgroups. A child's update rows are distinct, and one
// wave's LDS operations complete in issue order, so the children are added
// one after another without barriers, in the same order as there.
template <typename T>
__global__ void __launch_bounds__(kWG)
large_init_wave_kernel(DevSym s, const InitItem* __restrict__ items,
                       const int64_t* __restrict__ it_ea, const EaItem* __restrict__ ea,
                       const T* __restrict__ lnz, T* __restrict__ arena, int nitems, int maxL) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  arena += vs * s.vs_arena;
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x % 64, wave = wave_id();
  const int item = blockIdx.x * 4 + wave;
  if (item >= nitems) return;  // no workgroup barrier below
  T* Fs = reinterpret_cast<T*>(smem_raw) + (size_t)wave * maxL;
  const InitItem m = items[item];
  const int L = m.L;
  T* F = arena + m.fcol;
  const int64_t e0 = it_ea[item], e1 = it_ea[item + 1];
  EaItem en = ea[e0 < e1 ? e0 : 0];
  const int r1 = m.r1;
  const T* lc = (m.lcb >= 0) ? lnz + m.lcb : nullptr;
  for (int r0 = lane; r0 < L; r0 += 8 * 64) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = r0 + u * 64;
      v[u] = (r >= L) ? T(0) : (r < r1) ? F[r] : (lc ? lc[r] : T(0));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (r0 + u * 64 < L) Fs[r0 + u * 64] = v[u];
  }
  wave_lds_sync();
  for (int64_t e = e0; e < e1; ++e) {
    const EaItem ec = en;
    if (e + 1 < e1) en = ea[e + 1];
    const int Rc = ec.Rc;
    const int32_t* rel = s.relmap + ec.rel;
    const T* col = arena + ec.col;
    for (int tp0 = lane; tp0 < Rc; tp0 += 8 * 64) {
      int ri[8];
      T cv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int tp = tp0 + u * 64;
        ri[u] = (tp < Rc) ? rel[tp] : 0;
        cv[u] = (tp < Rc) ? col[tp] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (tp0 + u * 64 < Rc) Fs[ri[u]] += cv[u];
    }
    wave_lds_sync();
  }
  for (int r0 = lane; r0 < L; r0 += 8 * 64) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (r0 + u * 64 < L) ? Fs[r0 + u * 64] : T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (r0 + u * 64 < L) F[r0 + u * 64] = v[u];
  }
}

// L part of the own columns back to lnz (contiguous per column).
template <typename T>
__global__ void __launch_bounds__(kWG)
large_writeback_kernel(DevSym s, const int32_t* __restrict__ it_front,
                       const int32_t* __restrict__ it_col,
                       T* __restrict__ lnz, T* __restrict__ unz,
                       const T* __restrict__ arena) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  unz += vs * s.vs_unz;
  arena += vs * s.vs_arena;
  const int item = blockIdx.x;
  const int f = it_front[item];
  const int j = it_col[item];
  const int L = s.fL[f], W = s.fW[f];
  if (j >= W) return;
  const int64_t fc0 = s.fc0[f];
  const T* F = arena + s.foff[f] + (int64_t)j * L;
  const int cp = piece_start_of_col(s, fc0, j);
  T* lc = lnz + s.lcol[fc0 + j];
  for (int r0 = cp + threadIdx.x; r0 < L; r0 += 8 * kWG) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (r0 + u * kWG < L) ? F[r0 + u * kWG] : T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (r0 + u * kWG < L) lc[r0 + u * kWG] = v[u];
  }
}

// Descriptor of workgroup b: the last one whose wg_begin <= b.
template <typename Desc>
__device__ inline int find_blk(const Desc* d, int nd, int b) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (d[mid].wg_begin <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Large fronts, blocked. The pieces of a front are taken in groups (blocks)
// of consecutive pieces spanning columns [c0, c1) (at most 128 columns).
// Each block takes three launches:
//
//  blk_diag_kernel     one workgroup per block factors the bw x bw diagonal
//                      region [c0, c1)^2 in LDS, piece by piece, exactly as
//                      the reference eliminates a supernode
//                      (SpkLUFactor.jl:230-240): LU with row interchanges
//                      inside the piece's nq x nq diagonal block, the
//                      interchanges applied to the piece's own columns and
//                      those right of it but never to earlier pieces' L
//                      columns, L21 = A21 U11^-1 for the region rows below,
//                      U12 = L11^-1 P A12 for the region columns right of
//                      it, then the update of the rest of the region.
//  blk_offdiag_kernel  rows below the block: X = A U_blk^-1 (U_blk = the
//                      upper triangle of the factored region, one piece at a
//                      time), and columns right of the block: per piece the
//                      earlier pieces' contribution, the piece's interchanges
//                      and its unit-lower solve — the left-looking order of
//                      the reference, so every entry is the reference's.
//  blk_gemm_kernel     trailing update A22 -= L21 U12. Blocks are taken in
//                      groups (8 by default): after each block but the last
//                      of a group only the border of the group's remaining
//                      blocks (their columns, and their rows right of them)
//                      is updated; after the last, the rest of the trailing
//                      matrix receives the whole group at once, with inner
//                      dimension up to 8 x 128 — an eighth of the
//                      read-modify-write traffic of one update per block.
// ---------------------------------------------------------------------------

// Copy of an nr x nc tile (nr <= 128) from a column-major HBM array (leading
// dimension lds) into LDS at dst[r * DR + c * DC]: lane = row (two halves),
// the four waves take interleaved columns, eight loads in flight per thread.
template <typename T, int DR, int DC>
__device__ inline void load_tile(const T* __restrict__ src, int64_t lds, int nr, int nc,
                                 T* __restrict__ dst) {
  const int w = threadIdx.x / 64;
  for (int r = threadIdx.x % 64; r < nr; r += 64) {
    for (int cb = w; cb < nc; cb += 32) {
      const T* sp = src + r + (int64_t)cb * lds;
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (cb + 4 * u < nc) ? sp[(int64_t)(4 * u) * lds] : T(0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (cb + 4 * u < nc) dst[r * DR + (cb + 4 * u) * DC] = v[u];
    }
  }
}

// The reverse copy: dst[r + c * ldd] = src[r * DR + c * DC].
template <typename T, int DR, int DC>
__device__ inline void store_tile(T* __restrict__ dst, int64_t ldd, int nr, int nc,
                                  const T* __restrict__ src) {
  const int w = threadIdx.x / 64;
  for (int r = threadIdx.x % 64; r < nr; r += 64)
    for (int c = w; c < nc; c += 4) dst[r + (int64_t)c * ldd] = src[r * DR + c * DC];
}

// Cross-lane moves inside rows of 16 lanes (DPP): quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror.
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ inline double rdlane(double v, int l) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}
__device__ inline float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Wave-uniform maximum of v over the 64 lanes (all lanes active).
template <typename T>
__device__ inline T wave_max(T v) {
  v = fmax(v, dpp_mov<0xB1>(v));   // lanes ^1
  v = fmax(v, dpp_mov<0x4E>(v));   // lanes ^2
  v = fmax(v, dpp_mov<0x141>(v));  // mirror inside 8
  v = fmax(v, dpp_mov<0x140>(v));  // mirror inside 16
  return fmax(fmax(rdlane(v, 0), rdlane(v, 16)), fmax(rdlane(v, 32), rdlane(v, 48)));
}

// Wave-uniform minimum of v over the 64 lanes (all lanes active).
__device__ inline int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Leading dimension of the diagonal region in LDS (blocks are at most 128
// columns wide); a compile-time stride turns every column offset into an
// immediate LDS offset.
constexpr int kRLD = 129;
// Row stride of a block's diagonal region (at most 64 columns) held row by
// row in LDS: lanes storing consecutive rows fall in different banks.
constexpr int kRW = 65;
// Most columns of one piece (the SPD limit 90, rounded up).
constexpr int kMaxNq = 96;

// One thread's triangular solve against a piece's diagonal block, in place
// in LDS: X[j XS] (j < nq) becomes
//   x_j = (a_j - sum_{k<j} x_k M(k, j)) * r_j,   M(k, j) = M[k MK + j MJ],
// r_j = 1 / M(j, j) (or rd[j] when given; 1 when UNIT). Columns go in
// blocks of eight: the earlier columns' contributions in a runtime loop over
// k with eight independent sums, then the block's own triangle in
// registers. Each x_j receives the same operations in the same order as the
// plain dot form (k ascending, then the scaling), so the result is the
// same, while the code stays a few hundred bytes (a fully unrolled solve is
// fetched from L2 once per piece and runs at instruction-fetch speed).
template <typename T, int XS, int MK, int MJ, bool UNIT>
__device__ inline void blk_trsv(T* X, const T* M, const T* rd, int nq) {
  for (int jb = 0; jb < nq; jb += 8) {
    int jc[8];  // columns of the block, clamped inside the piece
#pragma unroll
    for (int c = 0; c < 8; ++c) jc[c] = min(jb + c, nq - 1);
    T xb[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) xb[c] = X[jc[c] * XS];
#pragma unroll 2
    for (int kk = 0; kk < jb; ++kk) {
      const T xk = X[kk * XS];
#pragma unroll
      for (int c = 0; c < 8; ++c) xb[c] -= xk * M[kk * MK + jc[c] * MJ];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
#pragma unroll
      for (int cp = 0; cp < c; ++cp) xb[c] -= xb[cp] * M[jc[cp] * MK + jc[c] * MJ];
      if (!UNIT) xb[c] = xb[c] * (rd ? rd[jc[c]] : T(1) / M[jc[c] * MK + jc[c] * MJ]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (jb + c < nq) X[(jb + c) * XS] = xb[c];
  }
}

// Pivoted LU of the nq x nq block D (LDS, leading dimension LD, nq <= 64) by
// one wave, with rows that never move during the elimination: lane i owns
// LDS row i of the block and tracks the position pos its row holds after the
// interchanges so far. Step k takes the row of smallest position among those
// holding the maximum |a| over positions >= k (the first maximum of ggetrf!,
// GenericBlasLapackFragments.jl:56-100; a zero maximum keeps row k and skips
// the scaling), reads the pivot row where it lies and updates the rows of
// position > k, eight columns per batch of loads. At the end every row moves
// to its position, together with the columns [nq, ncols) of D. piv[k] is the
// position swapped with k at step k, perm[p] the row now at position p,
// dinv[k] = 1 / U(k, k); *zero is set by a zero pivot.
template <typename T, int LD>
__device__ void pos_getrf(T* D, int nq, int ncols, int* piv, int* perm, T* dinv, int* zero) {
  const int lane = threadIdx.x % 64;
  const bool mine = lane < nq;
  int pos = lane;
  for (int k = 0; k < nq; ++k) {
    const bool cand = mine && pos >= k;
    const T v = mine ? D[lane + k * LD] : T(0);
    const T a = cand ? fabs(v) : T(-1);
    const T m = wave_max(a);
    const bool tie = cand && a == m;
    const unsigned long long msk = __ballot(tie);
    int blane;
    if (msk == 0ull) {  // no comparable candidate (NaN column): keep row k
      blane = __ffsll((long long)__ballot(mine && pos == k)) - 1;
    } else if (__popcll(msk) == 1) {
      blane = __ffsll((long long)msk) - 1;
    } else {
      const int mp = wave_min(tie ? pos : (1 << 20));
      blane = __ffsll((long long)__ballot(tie && pos == mp)) - 1;
    }
    blane = __builtin_amdgcn_readfirstlane(blane);
    const int kp = __builtin_amdgcn_readlane(pos, blane);
    const T pv = rdlane(v, blane);
    const bool nonzero = pv != T(0);
    if (nonzero && kp != k) {
      if (mine && pos == k)
        pos = kp;
      else if (lane == blane)
        pos = k;
    }
    const T inv = T(1) / pv;
    if (lane == 0) {
      piv[k] = kp;
      dinv[k] = inv;
      if (!nonzero) *zero = 1;
    }
    const bool below = mine && pos > k;
    const T l = nonzero ? v * inv : v;
    if (below && nonzero) D[lane + k * LD] = l;
    int j0 = k + 1;
    for (; j0 + 8 <= nq; j0 += 8) {
      const T* pr = D + blane + j0 * LD;
      T* xr = D + lane + j0 * LD;
      T u[8], x[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        u[jj] = pr[jj * LD];
        x[jj] = xr[jj * LD];
      }
      if (below) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) xr[jj * LD] = x[jj] - l * u[jj];
      }
    }
    for (; j0 < nq; ++j0)
      if (below) D[lane + j0 * LD] -= l * D[blane + j0 * LD];
    wave_lds_sync();
  }
  if (mine) perm[pos] = lane;
  wave_lds_sync();
  // rows to their positions: position p takes LDS row perm[p]
  const T* srow = D + (mine ? perm[lane] : 0);
  T* drow = D + lane;
  int c0 = 0;
  for (; c0 + 8 <= ncols; c0 += 8) {
    T x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = srow[(c0 + u) * LD];
    wave_lds_sync();
    if (mine) {
#pragma unroll
      for (int u = 0; u < 8; ++u) drow[(c0 + u) * LD] = x[u];
    }
    wave_lds_sync();
  }
  for (; c0 < ncols; ++c0) {
    const T x = srow[c0 * LD];
    wave_lds_sync();
    if (mine) drow[c0 * LD] = x;
    wave_lds_sync();
  }
}

// Pivot of column k among the rows whose position is k or later (v: this
// lane's entry of the column, pos: its row's position): largest |v|, ties
// to the smallest position; a column without a comparable candidate (NaN)
// keeps the row at position k. Returns the lane of the pivot row.
template <typename T>
__device__ inline int pivot_lane(T v, int pos, bool mine, int k) {
  const bool cand = mine && pos >= k;
  const T a = cand ? fabs(v) : T(-1);
  const T m = wave_max(a);
  const bool tie = cand && a == m;
  const unsigned long long msk = __ballot(tie);
  int blane;
  if (msk == 0ull) {
    blane = __ffsll((long long)__ballot(mine && pos == k)) - 1;
  } else if (__popcll(msk) == 1) {
    blane = __ffsll((long long)msk) - 1;
  } else {
    const int mp = wave_min(tie ? pos : (1 << 20));
    blane = __ffsll((long long)__ballot(tie && pos == mp)) - 1;
  }
  return __builtin_amdgcn_readfirstlane(blane);
}

// pos_getrf with the whole workgroup, one barrier per step. Wave 0 updates
// the next pivot column first and runs the next step's pivot search on it
// (pivot_lane), publishing the pivot row and value through LDS (two slots,
// by step parity); waves 1-3 update the columns j = k + 2 + (wave - 1) + 3 i
// meanwhile. The multipliers of column k are written by wave 0 one step
// later, when no wave reads that column any more. Must be called by all
// threads of the workgroup.
template <typename T, int LD>
__device__ void wg_getrf(T* D, int nq, int ncols, int* piv, int* perm, T* dinv, int* zero) {
  __shared__ int s_blane[2];
  __shared__ T s_pv[2];
  const int lane = threadIdx.x % 64, wave = wave_id();
  const bool mine = lane < nq;
  int pos = lane;
  T lprev = T(0);
  bool wprev = false;
  if (wave == 0) {
    const T v = mine ? D[lane] : T(0);
    const int b = pivot_lane(v, pos, mine, 0);
    const T pv = rdlane(v, b);
    if (lane == 0) {
      s_blane[0] = b;
      s_pv[0] = pv;
    }
  }
  __syncthreads();
  for (int k = 0; k < nq; ++k) {
    const int par = k & 1;
    const int blane = s_blane[par];
    const T pv = s_pv[par];
    const T v = mine ? D[lane + k * LD] : T(0);
    if (wave == 0 && wprev) D[lane + (k - 1) * LD] = lprev;
    // the pivot row's position before this step's interchange
    const int kp = __builtin_amdgcn_readlane(pos, blane);
    const bool nonzero = pv != T(0);
    if (nonzero && kp != k) {
      if (mine && pos == k)
        pos = kp;
      else if (lane == blane)
        pos = k;
    }
    const T inv = T(1) / pv;
    if (threadIdx.x == 0) {
      piv[k] = kp;
      dinv[k] = inv;
      if (!nonzero) *zero = 1;
    }
    const bool below = mine && pos > k;
    const T l = nonzero ? v * inv : v;
    lprev = l;
    wprev = below && nonzero;
    if (wave == 0) {
      if (k + 1 < nq) {
        const int j = k + 1;
        const T u = D[blane + j * LD];
        T x = D[lane + j * LD];
        if (below) {
          x -= l * u;
          D[lane + j * LD] = x;
        }
        const int b = pivot_lane(x, pos, mine, j);
        const T pn = rdlane(x, b);
        if (lane == 0) {
          s_blane[par ^ 1] = b;
          s_pv[par ^ 1] = pn;
        }
      }
    } else {
      // columns j = k + 2 + (wave - 1) + 3 i in batches of eight: loads of a
      // batch are issued together (indices clamped inside the block), stores
      // guarded
      for (int jb = k + 1 + wave; jb < nq; jb += 24) {
        T u[8], x[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int j = min(jb + 3 * jj, nq - 1);
          u[jj] = D[blane + j * LD];
          x[jj] = D[lane + j * LD];
        }
        if (below) {
#pragma unroll
          for (int jj = 0; jj < 8; ++jj)
            if (jb + 3 * jj < nq) D[lane + (jb + 3 * jj) * LD] = x[jj] - l * u[jj];
        }
      }
    }
    __syncthreads();
  }
  if (wave == 0 && wprev) D[lane + (nq - 1) * LD] = lprev;
  if (wave == 0 && mine) perm[pos] = lane;
  __syncthreads();
  // rows to their positions (position p takes LDS row perm[p]), the waves
  // taking interleaved columns; all reads of a batch precede its writes
  const T* srow = D + (mine ? perm[lane] : 0);
  T* drow = D + lane;
  for (int c0 = 0; c0 < ncols; c0 += 32) {
    T x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = c0 + wave + 4 * u;
      x[u] = (c < ncols) ? srow[c * LD] : T(0);
    }
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = c0 + wave + 4 * u;
        if (c < ncols) drow[c * LD] = x[u];
      }
    }
    __syncthreads();
  }
}

// The same factorisation without pivoting for SPD pieces of up to kMaxNq
// columns (lane = row, rows lane and lane + 64): a zero pivot sets *zero and
// skips the scaling; piv[k] = k.
template <typename T, int LD>
__device__ void nopiv_getrf(T* D, int nq, int* piv, T* dinv, int* zero) {
  const int lane = threadIdx.x % 64;
  for (int k = 0; k < nq; ++k) {
    const T pv = D[k + k * LD];
    const bool nonzero = pv != T(0);
    const T inv = T(1) / pv;
    if (lane == 0) {
      piv[k] = k;
      dinv[k] = inv;
      if (!nonzero) *zero = 1;
    }
    for (int i = k + 1 + lane; i < nq; i += 64) {
      T l = D[i + k * LD];
      if (nonzero) {
        l *= inv;
        D[i + k * LD] = l;
      }
      for (int j0 = k + 1; j0 < nq; j0 += 8) {
        T u[8], x[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int j = min(j0 + jj, nq - 1);
          u[jj] = D[k + j * LD];
          x[jj] = D[i + j * LD];
        }
#pragma unroll
        for (int jj = 0; jj < 8; ++jj)
          if (j0 + jj < nq) D[i + (j0 + jj) * LD] = x[jj] - l * u[jj];
      }
    }
    wave_lds_sync();
  }
}

// v_mfma_*_16x16x4: both forms take A(i, k) and B(k, j) from lane i + 16 k
// (j + 16 k); accumulator element v of lane l is D(i, l % 16) with i = 4
// (l / 16) + v for the f32 form and i = 4 v + l / 16 for the f64 form.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lgrp, int v) { return 4 * v + lgrp; }
};
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lgrp, int v) { return 4 * lgrp + v; }
};

// Diagonal region of one block per workgroup (see the section comment);
// NQ bounds the piece width. Dynamic LDS: 128 * kRLD elements.
template <typename T, int NQ>
__global__ void __launch_bounds__(kWG)
blk_diag_kernel(DevSym s, const BlkDesc* __restrict__ desc, T* __restrict__ arena,
                int64_t* __restrict__ ipvt, unsigned long long* __restrict__ first_fail) {
  const int64_t vs = blockIdx.y;  // value set
  arena += vs * s.vs_arena;
  ipvt += vs * s.n;
  constexpr int LD = kRLD;
  extern __shared__ unsigned char smem_raw[];
  __shared__ int piv[NQ];
  __shared__ int perm[NQ];
  __shared__ T dinv[NQ];
  __shared__ int s_zero;
  T* R = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x, wave = wave_id();
  const BlkDesc d = desc[blockIdx.x];
  const int f = d.front;
  const int L = s.fL[f];
  const int64_t fc0 = s.fc0[f];
  const int c0 = d.c0, bw = d.c1 - d.c0;
  T* Fr = arena + s.foff[f] + c0 + (int64_t)c0 * L;
  load_tile<T, 1, LD>(Fr, L, bw, bw, R);
  __syncthreads();
  for (int pi = 0; pi < d.np; ++pi) {
    int64_t pc0, nq64;
    piece_cols(s, d.piece + pi, pc0, nq64);
    const int nq = (int)nq64;
    const int o = (int)(pc0 - fc0) - c0;  // piece offset inside the region
    const int rest = bw - o - nq;         // region rows below / columns right
    T* D = R + o + o * LD;
    if (tid == 0) s_zero = 0;
    __syncthreads();
    // interchanges reach the region columns right of the piece too
    if (s.pivot && nq <= 64) {
      wg_getrf<T, LD>(D, nq, nq + rest, piv, perm, dinv, &s_zero);
    } else if (wave == 0) {
      nopiv_getrf<T, LD>(D, nq, piv, dinv, &s_zero);
    }
    __syncthreads();
    if (tid < nq) ipvt[pc0 + tid] = piv[tid] + 1;
    if (tid == 0 && s_zero) atomicMin(first_fail, (unsigned long long)(d.piece + pi + 1));
    if (tid >= 128) {
      // U12 of the region (rows already interchanged): unit-lower solve,
      // u_i = a_i - sum_k<i L(i, k) u_k, one column per thread
      for (int c = tid - 128; c < rest; c += 128)
        blk_trsv<T, 1, LD, 1, true>(D + (nq + c) * LD, D, nullptr, nq);
    } else {
      // L21 of the region: x_j = (a_j - sum_k<j x_k U(k, j)) / U(j, j), one
      // row per thread
      for (int r = tid; r < rest; r += 128)
        blk_trsv<T, LD, 1, LD, false>(D + nq + r, D, dinv, nq);
    }
    __syncthreads();
    if (rest > 0) {
      // rest x rest update on the matrix cores: 16 x 16 tiles, every fourth
      // to a wave; per tile all operand reads are issued before the MFMAs
      // (k in steps of four, cut at nq); products summed from zero and
      // subtracted once, as in blk_gemm_kernel
      typedef typename Mfma16<T>::v4 v4;
      constexpr int KS = (NQ + 3) / 4;
      const int lane = tid % 64, l16 = lane % 16, lk = lane / 16;
      const int nt = (rest + 15) / 16;
      const T* Lb = D + nq;       // L21(i, k) = Lb[i + k LD]
      const T* Ub = D + nq * LD;  // U12(k, j) = Ub[k + j LD]
      T* C = D + nq + nq * LD;
      for (int t = wave; t < nt * nt; t += 4) {
        const int tr = (t % nt) * 16, tc = (t / nt) * 16;
        const int ri = tr + l16, cj = tc + l16;
        const bool rin = ri < rest, cin = cj < rest;
        T av[KS], bv[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int kk = 4 * ks + lk;
          const bool kin = kk < nq;
          av[ks] = (kin && cin) ? Ub[kk + cj * LD] : T(0);
          bv[ks] = (kin && rin) ? Lb[ri + kk * LD] : T(0);
        }
        v4 acc = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
          if (4 * ks < nq) acc = Mfma16<T>::op(av[ks], bv[ks], acc);
        if (rin) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int j = tc + Mfma16<T>::row(lk, v);
            if (j < rest) C[ri + j * LD] -= acc[v];
          }
        }
      }
    }
    __syncthreads();
  }
  store_tile<T, 1, LD>(Fr, L, bw, bw, R);
}

// Diagonal region of one block of at most 64 columns per workgroup, held
// in registers: wave w keeps columns [16 w, 16 w + 16) of the region, one
// row per lane. The region is factored right-looking, one column at a time,
// with the reference's pivoting confined to each piece's rows (LU; an SPD
// front takes the row at position k, no interchanges)
// (SpkLUFactor.jl:230-240, GenericBlasLapackFragments.jl:56-100). Step k,
// a column of piece p, takes as pivot the row of p with the largest |a|
// among those whose position is k or later (ties: the smallest position; no
// comparable candidate: the row at position k), scales the rows below it by
// the reciprocal of the pivot (not when it is zero) and subtracts their
// multiple of the pivot row from the columns right of k. The rows below are
// p's rows at later positions and every row of the later pieces: for p's
// rows this is the piece's pivoted LU and, right of the piece, its
// unit-lower solve U12 = L11^-1 P A12; for the later pieces' rows it is
// L21 = A21 U11^-1 and the update of their part of the region. The
// triangular parts receive their terms one at a time in ascending k, in the
// order of the left-looking forms; the later pieces' part right of the
// piece sums the piece's products from zero and subtracts the sum once,
// as the reference's update of one supernode by another does (the sum is
// kept in the entry itself, the entry's value aside until the piece's last
// column). Rows never move while factoring: each lane tracks the position
// its row holds. The wave owning column k publishes the step (multipliers,
// pivot lane, pivot position) through LDS; the other waves apply published
// steps to their columns as they come, without barriers (a wave only ever
// waits for steps of columns left of its own, which the waves to its left
// publish without waiting on it). At the end each row is stored at its
// final position in the columns of its piece and right of it, and at its
// own row in the columns of earlier pieces, which the reference never
// interchanges.
// When d.fk1 > d.fk0 the region first takes the product of the previous
// block (columns [fk0, fk1)) with its own columns, summed from zero over
// those columns in order and subtracted once (the trailing update's form),
// so that the rest of that block's update can run beside this kernel.
// lds: kSlabLds<T> bytes (the three 64 x 64 arrays below).
constexpr int kMaxBlkPieces = 8;  // most pieces in one block
template <typename T>
constexpr int kSlabLds = 3 * 64 * 64 * (int)sizeof(T);
template <typename T>
__device__ __attribute__((always_inline)) void blk_diag_slab_body(
    const DevSym& s, const BlkDesc& d, T* __restrict__ arena, int64_t* __restrict__ ipvt,
    unsigned long long* __restrict__ first_fail, unsigned char* lds) {
  constexpr int SW = 16;                  // columns per wave
  // step k: the multiplier of each row; finished column c, rows in lane
  // order; entries set aside while a piece's sum forms
  T (*sl)[64] = reinterpret_cast<T (*)[64]>(lds);
  T (*sout)[64] = reinterpret_cast<T (*)[64]>(lds + 64 * 64 * sizeof(T));
  T (*ssv)[64] = reinterpret_cast<T (*)[64]>(lds + 2 * 64 * 64 * sizeof(T));
  __shared__ int smeta[64];               // step k: pivot lane | position << 8 | nonzero << 16
  __shared__ int spo[kMaxBlkPieces + 1];  // piece offsets in the region; spo[np] = bw
  __shared__ int scp[64];                 // piece of each region column
  __shared__ int spos[64];                // final position of each row
  __shared__ int spiv[64];                // step k: pivot position, 1-based in its piece
  __shared__ int sdone;                   // steps published
  const int lane = threadIdx.x % 64, w = wave_id();
  const int L = d.L;
  const int64_t fc0 = d.fc0;
  const int c0 = d.c0, bw = d.c1 - d.c0, np = d.np;
  const T* const Fb = arena + d.foff;
  T* R = arena + d.foff + c0 + (int64_t)c0 * L;
  const int cb = SW * w;
  // the front and the piece widths come from the descriptor, so the region,
  // the previous block's panels and the widths are read in one batch
  // a[i] holds column c + i, c = cb while the earlier waves' steps are
  // applied, then the column of the current step: the slab shifts left by
  // one column after each of this wave's steps, so that the pivot column is
  // always a[0] and the loop body indexes registers statically
  T a[SW];
#pragma unroll
  for (int i = 0; i < SW; ++i)
    a[i] = (lane < bw && cb + i < bw) ? R[lane + (int64_t)(cb + i) * L] : T(0);
  const int fkw = d.fk1 - d.fk0;
  if (fkw > 0) {
    // L(c0 + r, fk0 + k) at sout[k][r], U(fk0 + k, c0 + c) at sl[k][c]
    for (int idx = threadIdx.x; idx < 64 * fkw; idx += kWG) {
      const int r = idx % 64, k = idx / 64;
      sout[k][r] = (r < bw) ? Fb[c0 + r + (int64_t)(d.fk0 + k) * L] : T(0);
      sl[k][r] = (r < bw) ? Fb[d.fk0 + k + (int64_t)(c0 + r) * L] : T(0);
    }
  }
  if ((int)threadIdx.x <= np) {
    int o = 0;  // spo[np] = bw: the pieces fill the block
    for (int q = 0; q < (int)threadIdx.x; ++q) o += d.pw[q];
    spo[threadIdx.x] = o;
  }
  if (threadIdx.x == 0) sdone = 0;
  __syncthreads();
  if (threadIdx.x < 64) {
    int p = 0;
    while (p + 1 < np && (int)threadIdx.x >= spo[p + 1]) ++p;
    scp[threadIdx.x] = p;
  }
  __syncthreads();
  // this lane's row: its piece pr (np for the rows past the region)
  const int pr = lane < bw ? scp[lane] : np;
  const int rlo = pr < np ? spo[pr] : 0;
  int pos = lane;
  if (fkw > 0) {
    T acc[SW];
#pragma unroll
    for (int i = 0; i < SW; ++i) acc[i] = T(0);
    for (int k = 0; k < fkw; ++k) {
      const T lv = sout[k][lane];
#pragma unroll
      for (int i = 0; i < SW; ++i) acc[i] += lv * sl[k][cb + i];
    }
#pragma unroll
    for (int i = 0; i < SW; ++i)
      if (lane < bw && cb + i < bw) a[i] = a[i] - acc[i];
    __syncthreads();
  }
  // first step of piece pk (current column c): the later pieces' entries
  // right of the piece are set aside and the piece's products summed from
  // zero in their place; last step: the entries take the sum off. Only this
  // wave's own columns [cb, cb + SW): past them a[] holds placeholders, and
  // the set-aside slots there belong to the wave owning those columns, which
  // may still be restoring an earlier piece from them. Called under
  // wave-uniform branches that hold a compiler barrier, so that they stay
  // branches instead of selects executed at every step.
  const int cown = cb + SW;
  auto piece_begin = [&](int pk, int c) {
    __asm__ volatile("" ::: "memory");
    const int cn = __builtin_amdgcn_readfirstlane(spo[pk + 1]);
    const bool later = pr > pk && pr < np;
#pragma unroll
    for (int i = 0; i < SW; ++i)
      if (c + i >= cn && c + i < cown) {
        ssv[c + i][lane] = a[i];
        if (later) a[i] = T(0);
      }
  };
  auto piece_end = [&](int pk, int c) {
    __asm__ volatile("" ::: "memory");
    const int cn = __builtin_amdgcn_readfirstlane(spo[pk + 1]);
    const bool later = pr > pk && pr < np;
#pragma unroll
    for (int i = 0; i < SW; ++i)
      if (c + i >= cn && c + i < cown && later) a[i] = ssv[c + i][lane] + a[i];
  };
  // the piece of the current step, [plo, phi): advanced at its boundaries
  int pk = 0;
  int plo = 0;
  int phi = __builtin_amdgcn_readfirstlane(spo[1]);
  // steps of the columns left of this wave's slab, as they are published
  const int nprev = min(cb, bw);
  for (int k = 0; k < nprev; ++k) {
    if (k == phi) {
      ++pk;
      plo = phi;
      phi = __builtin_amdgcn_readfirstlane(spo[pk + 1]);
    }
    for (int it = 0; it < (1 << 24) &&
         __hip_atomic_load(&sdone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) <= k;
         ++it)
      __builtin_amdgcn_s_sleep(1);
    const int meta = __builtin_amdgcn_readfirstlane(smeta[k]);
    const int bl = meta & 0xff;
    const int kp = (meta >> 8) & 0xff;
    const bool nz = (meta >> 16) & 1;
    if (nz && kp != k) {
      if (pos == k)
        pos = kp;
      else if (lane == bl)
        pos = k;
    }
    const bool below = (pr == pk && pos > k) || (pr > pk && pr < np);
    // rows not below subtract 0 u, which leaves a finite entry unchanged
    const T lb = below ? sl[k][lane] : T(0);
    if (k == plo) piece_begin(pk, cb);
    T u[SW];
#pragma unroll
    for (int i = 0; i < SW; ++i) u[i] = rdlane(a[i], bl);
#pragma unroll
    for (int i = 0; i < SW; ++i) a[i] = a[i] - lb * u[i];
    if (k == phi - 1) piece_end(pk, cb);
  }
  // this wave's own columns; step k is published (sdone = k + 1) at the
  // start of step k + 1, when its LDS writes have long completed, so that
  // the release never waits
  const int kend = min(cb + SW, bw);
  for (int k = cb; k < kend; ++k) {
    if (k == phi) {
      ++pk;
      plo = phi;
      phi = __builtin_amdgcn_readfirstlane(spo[pk + 1]);
    }
    if (k > cb)
      __hip_atomic_store(&sdone, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == plo) piece_begin(pk, k);
    const T v = a[0];
    // without pivoting (SPD) the pivot row is the one at position k: lane k
    const int bl = s.pivot ? pivot_lane(v, pos, pr == pk, k) : k;
    const int kp = __builtin_amdgcn_readlane(pos, bl);
    const T pv = rdlane(v, bl);
    T u[SW];
#pragma unroll
    for (int i = 1; i < SW; ++i) u[i] = rdlane(a[i], bl);
    const bool nz = pv != T(0);
    if (nz && kp != k) {
      if (pr == pk && pos == k)
        pos = kp;
      else if (lane == bl)
        pos = k;
    }
    const T inv = T(1) / pv;
    const bool below = (pr == pk && pos > k) || (pr > pk && pr < np);
    const T l = nz ? v * inv : v;
    // (no global memory access until the end: a release would wait for it)
    sl[k][lane] = l;
    if (lane == 0) {
      smeta[k] = bl | (kp << 8) | ((int)nz << 16);
      spiv[k] = kp - plo + 1;
    }
    sout[k][lane] = (below && nz) ? l : v;
    const T lb = below ? l : T(0);
#pragma unroll
    for (int i = 1; i < SW; ++i) a[i] = a[i] - lb * u[i];
    if (k == phi - 1) piece_end(pk, k);
#pragma unroll
    for (int i = 0; i + 1 < SW; ++i) a[i] = a[i + 1];
    a[SW - 1] = T(0);
  }
  if (kend > cb)
    __hip_atomic_store(&sdone, kend, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  // the wave of the last column has seen every interchange
  if (w == (bw - 1) / SW) spos[lane] = pos;
  __syncthreads();
  const int fpos = spos[lane];
  if ((int)threadIdx.x < bw) {
    ipvt[fc0 + c0 + threadIdx.x] = spiv[threadIdx.x];
    // a zero pivot: the first piece that met one
    if (!((smeta[threadIdx.x] >> 16) & 1))
      atomicMin(first_fail, (unsigned long long)(d.piece + scp[threadIdx.x] + 1));
  }
  // column c of this wave: row r at its position in the columns of its
  // piece and right of it, at its own row in earlier pieces' columns
  if (lane < bw) {
    for (int i = 0; i < SW; ++i) {
      const int c = cb + i;
      if (c < bw) R[(c >= rlo ? fpos : lane) + (int64_t)c * L] = sout[c][lane];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kWG)
blk_diag_slab_kernel(DevSym s, const BlkDesc* __restrict__ desc, T* __restrict__ arena,
                     int64_t* __restrict__ ipvt, unsigned long long* __restrict__ first_fail) {
  const int64_t vs = blockIdx.y;  // value set
  arena += vs * s.vs_arena;
  ipvt += vs * s.n;
  __shared__ __attribute__((aligned(16))) unsigned char lds[kSlabLds<T>];
  blk_diag_slab_body<T>(s, desc[blockIdx.x], arena, ipvt, first_fail, lds);
}

// Off-diagonal parts of a block of at most 64 columns. The tile, the
// block's whole diagonal region and (columns) the block's interchanges are
// read in one batch of loads, with the front and the piece widths (pw) taken
// from the descriptor; the pieces then follow each other in LDS with the
// operations, in the order, of the general path below (same results bit for
// bit). Dynamic LDS: 2 * 64 * kRW elements.
template <typename T>
__device__ __attribute__((always_inline)) inline void
offdiag64_body(const BlkDesc& d, const int32_t* __restrict__ pw, int b, T* __restrict__ arena,
               const int64_t* __restrict__ ipvt, T* __restrict__ S, int* __restrict__ spiv) {
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int L = d.L;
  const int c0 = d.c0, bw = d.c1 - d.c0;
  T* F = arena + d.foff;
  const T* Fd = F + c0 + (int64_t)c0 * L;  // the diagonal region
  if (b < d.nl) {
    // rows [r0, r0 + nr): X(i, k) at X[i + 64 k]; U(k, j) of the region at
    // R[k kRW + j]
    const int r0 = d.c1 + b * 64, nr = min(64, L - r0);
    T* X = S;
    T* R = S + 64 * kRW;
    T* Fx = F + r0 + (int64_t)c0 * L;
    T xv[16], rv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = wave + 4 * u;
      xv[u] = (lane < nr && c < bw) ? Fx[lane + (int64_t)c * L] : T(0);
      rv[u] = (lane < bw && c < bw) ? Fd[lane + (int64_t)c * L] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = wave + 4 * u;
      X[lane + c * 64] = xv[u];
      R[lane * kRW + c] = rv[u];
    }
    __syncthreads();
    int o = 0;
    for (int pi = 0; pi < d.np; ++pi) {
      const int nq = pw[pi];
      if (o > 0) {  // earlier pieces of the block: X_q -= X_<q U_<q,q
        for (int cb = wave * 8; cb < nq; cb += 32) {
          T acc[8];
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) acc[jj] = T(0);
          const int cs = min(cb, nq - 8 >= 0 ? nq - 8 : 0);  // shifted start
          const T* xr = X + lane;
          const T* ur = R + o + cs;
          for (int k = 0; k < o; ++k) {
            const T xk = xr[k * 64];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) acc[jj] += xk * ur[k * kRW + jj];
          }
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) {
            const int j = cs + jj;
            if (j >= cb && j < nq) X[lane + (o + j) * 64] -= acc[jj];
          }
        }
        __syncthreads();
      }
      if (wave == 0)  // X_q U_qq^-1, one row per lane
        blk_trsv<T, 64, kRW, 1, false>(X + lane + o * 64, R + o * kRW + o, nullptr, nq);
      __syncthreads();
      o += nq;
    }
    store_tile<T, 1, 64>(Fx, L, nr, bw, X);
  } else {
    // columns [j0, j0 + nc): A(r, c) at A[r + c kRW]; L(r, k) of the region
    // at R[r + k kRW]
    const int j0 = d.c1 + (b - d.nl) * 64, nc = min(64, L - j0);
    T* A = S;
    T* R = S + 64 * kRW;
    T* Fa = F + c0 + (int64_t)j0 * L;
    const int pv = (tid < bw) ? (int)ipvt[d.fc0 + c0 + tid] - 1 : 0;
    T av[16], rv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = wave + 4 * u;
      av[u] = (lane < bw && c < nc) ? Fa[lane + (int64_t)c * L] : T(0);
      rv[u] = (lane < bw && c < bw) ? Fd[lane + (int64_t)c * L] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = wave + 4 * u;
      A[lane + c * kRW] = av[u];
      R[lane + c * kRW] = rv[u];
    }
    if (tid < bw) spiv[tid] = pv;  // piece-local positions
    __syncthreads();
    int o = 0;
    for (int pi = 0; pi < d.np; ++pi) {
      const int nq = pw[pi];
      if (o > 0) {  // earlier pieces: A_q -= L_q,<q A_<q (rows not yet swapped)
        T* ac = A + lane * kRW;
        for (int ib = wave * 8; ib < nq; ib += 32) {
          const int is = min(ib, nq - 8 >= 0 ? nq - 8 : 0);
          T acc[8];
#pragma unroll
          for (int ii = 0; ii < 8; ++ii) acc[ii] = T(0);
          for (int k = 0; k < o; ++k) {
            const T a = ac[k];
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) acc[ii] += R[o + is + ii + k * kRW] * a;
          }
#pragma unroll
          for (int ii = 0; ii < 8; ++ii) {
            const int i = is + ii;
            if (i >= ib && i < nq) ac[o + i] -= acc[ii];
          }
        }
        __syncthreads();
      }
      if (wave == 0) {  // interchanges, then L_qq^-1, one column per lane
        T* a = A + lane * kRW + o;
        for (int k = 0; k < nq; ++k) {
          const int kp = spiv[o + k];
          if (kp != k) {
            const T t = a[k];
            a[k] = a[kp];
            a[kp] = t;
          }
        }
        blk_trsv<T, 1, kRW, 1, true>(a, R + o + o * kRW, nullptr, nq);
      }
      __syncthreads();
      o += nq;
    }
    store_tile<T, 1, kRW>(Fa, L, bw, nc, A);
  }
}

// Off-diagonal parts of one block after blk_diag_kernel (see the section
// comment); NQ bounds the piece width. Workgroups b < d.nl take 64 rows
// below the block, the others 64 columns right of it. Dynamic LDS:
// 64 * kRLD + 128 * NQ elements.
template <typename T, int NQ>
__global__ void __launch_bounds__(kWG)
blk_offdiag_kernel(DevSym s, const BlkDesc* __restrict__ desc, int ndesc,
                   T* __restrict__ arena, const int64_t* __restrict__ ipvt) {
  const int64_t vs = blockIdx.y;  // value set
  arena += vs * s.vs_arena;
  ipvt += vs * s.n;
  extern __shared__ unsigned char smem_raw[];
  __shared__ int spiv[NQ > 64 ? NQ : 64];
  T* S = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int di = find_blk(desc, ndesc, blockIdx.x);
  const BlkDesc d = desc[di];
  const int b = blockIdx.x - d.wg_begin;
  if (d.c1 - d.c0 <= 64) {
    offdiag64_body<T>(d, desc[di].pw, b, arena, ipvt, S, spiv);
    return;
  }
  const int f = d.front;
  const int L = s.fL[f];
  const int64_t fc0 = s.fc0[f];
  const int c0 = d.c0, c1 = d.c1, bw = c1 - c0;
  T* F = arena + s.foff[f];
  if (b < d.nl) {
    // rows [r0, r0 + nr): X(i, k) at X[i + 64 k]
    const int r0 = c1 + b * 64, nr = min(64, L - r0);
    T* X = S;
    T* Us = S + 64 * kRLD;  // strip of the piece's columns: Us[k * NQ + j]
    T* Fx = F + r0 + (int64_t)c0 * L;
    for (int idx = tid; idx < 64 * bw; idx += kWG) X[idx] = T(0);
    __syncthreads();
    load_tile<T, 1, 64>(Fx, L, nr, bw, X);
    for (int pi = 0; pi < d.np; ++pi) {
      int64_t pc0, nq64;
      piece_cols(s, d.piece + pi, pc0, nq64);
      const int nq = (int)nq64;
      const int o = (int)(pc0 - fc0) - c0;
      load_tile<T, NQ, 1>(F + c0 + (int64_t)(c0 + o) * L, L, o + nq, nq, Us);
      __syncthreads();
      if (o > 0) {  // earlier pieces of the block: X_q -= X_<q U_<q,q
        for (int cb = wave * 8; cb < nq; cb += 32) {
          T acc[8];
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) acc[jj] = T(0);
          const T* xr = X + lane;
          const T* ur = Us + min(cb, nq - 8 >= 0 ? nq - 8 : 0);
          for (int k = 0; k < o; ++k) {
            const T xv = xr[k * 64];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) acc[jj] += xv * ur[k * NQ + jj];
          }
          const int cs = min(cb, nq - 8 >= 0 ? nq - 8 : 0);  // shifted start
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) {
            const int j = cs + jj;
            if (j >= cb && j < nq) X[lane + (o + j) * 64] -= acc[jj];
          }
        }
        __syncthreads();
      }
      if (wave == 0)  // X_q U_qq^-1, one row per lane
        blk_trsv<T, 64, NQ, 1, false>(X + lane + o * 64, Us + o * NQ, nullptr, nq);
      __syncthreads();
    }
    store_tile<T, 1, 64>(Fx, L, nr, bw, X);
  } else {
    // columns [j0, j0 + nc): A(r, c) at A[r + c * kRLD]
    const int j0 = c1 + (b - d.nl) * 64, nc = min(64, L - j0);
    T* A = S;
    T* Ls = S + 64 * kRLD;  // the piece's rows of L: Ls[i + k * NQ], k < o + nq
    T* Fa = F + c0 + (int64_t)j0 * L;
    for (int idx = tid; idx < 64 * kRLD; idx += kWG) A[idx] = T(0);
    __syncthreads();
    load_tile<T, 1, kRLD>(Fa, L, bw, nc, A);
    for (int pi = 0; pi < d.np; ++pi) {
      int64_t pc0, nq64;
      piece_cols(s, d.piece + pi, pc0, nq64);
      const int nq = (int)nq64;
      const int o = (int)(pc0 - fc0) - c0;
      load_tile<T, 1, NQ>(F + c0 + o + (int64_t)c0 * L, L, nq, o + nq, Ls);
      if (tid < nq) spiv[tid] = (int)ipvt[pc0 + tid] - 1;
      __syncthreads();
      if (o > 0) {  // earlier pieces: A_q -= L_q,<q A_<q (rows not yet swapped)
        T* ac = A + lane * kRLD;
        for (int ib = wave * 8; ib < nq; ib += 32) {
          const int is = min(ib, nq - 8 >= 0 ? nq - 8 : 0);
          T acc[8];
#pragma unroll
          for (int ii = 0; ii < 8; ++ii) acc[ii] = T(0);
          for (int k = 0; k < o; ++k) {
            const T av = ac[k];
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) acc[ii] += Ls[is + ii + k * NQ] * av;
          }
#pragma unroll
          for (int ii = 0; ii < 8; ++ii) {
            const int i = is + ii;
            if (i >= ib && i < nq) ac[o + i] -= acc[ii];
          }
        }
        __syncthreads();
      }
      if (wave == 0) {  // interchanges, then L_qq^-1, one column per lane
        T* a = A + lane * kRLD + o;
        for (int k = 0; k < nq; ++k) {
          const int kp = spiv[k];
          if (kp != k) {
            const T t = a[k];
            a[k] = a[kp];
            a[kp] = t;
          }
        }
        blk_trsv<T, 1, NQ, 1, true>(a, Ls + o * NQ, nullptr, nq);
      }
      __syncthreads();
    }
    store_tile<T, 1, kRLD>(Fa, L, bw, nc, A);
  }
}

// The trailing update on the matrix cores: 128 x 128 tiles, each of the
// four waves owns a 64 x 64 quadrant as 4 x 4 MFMA 16x16x4 tiles. The
// operands are passed as (U^T, L^T) so that the accumulator holds the tile
// transposed: lane l keeps row l % 16 of four columns, and each group of 16
// lanes writes 16 contiguous rows of the column-major front.

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

// The 16-bit trailing update of one 128 x 128 tile on the gfx950 bf16
// matrix cores (v_mfma_f32_16x16x32_bf16), K-chunks of 32, double-buffered
// in LDS as rows of kB16 halves: A(i, k) = L21 rows at As[i kB16 + k],
// B(k, j) = U12 columns at Bs[j kB16 + k]. Global reads are four-float
// vectors (rows of an L21 column, k of a U12 column); each thread converts
// them and stores four bf16 per row or column (8 bytes); each lane reads its
// MFMA operand (eight consecutive k) with one 16-byte LDS read. Split mode:
// x = hi + lo, both bf16, products lo_b hi_a + hi_b lo_a + hi_b hi_a.
constexpr int kB16 = 40;                       // halves per LDS row (80 bytes)
constexpr int kB16Part = 128 * kB16;           // one operand array, halves
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

// p[0 .. 3], element e read only when e < nvalid (0 otherwise)
__device__ inline f4u load4(const float* p, int nvalid) {
  if (nvalid >= 4) return *reinterpret_cast<const f4u*>(p);
  f4u v = {0.f, 0.f, 0.f, 0.f};
  if (nvalid > 0) v.x = p[0];
  if (nvalid > 1) v.y = p[1];
  if (nvalid > 2) v.z = p[2];
  return v;
}

// four floats to bf16 (round to nearest even), two per v_cvt_pk_bf16_f32
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint2 pack_bf16(float a, float b, float c, float d) {
  const bf16x2 p = {(__bf16)a, (__bf16)b}, q = {(__bf16)c, (__bf16)d};
  return uint2{__builtin_bit_cast(unsigned, p), __builtin_bit_cast(unsigned, q)};
}
// the part of x below its bf16 rounding (exact: x and its rounding are floats
// within a factor of two of each other)
__device__ inline float bf_lo(float x) { return x - (float)(__bf16)x; }

__device__ void gemm_bf16_chunks(const float* __restrict__ F, int L, int k0, int bw, int r0,
                                 int rend, int j0, int cend, bool split,
                                 unsigned char* lds_raw, f32x4 (&acc)[4][4]) {
  constexpr int KC = 32;
  unsigned short* base = reinterpret_cast<unsigned short*>(lds_raw);
  // buffer b: [A hi, B hi, A lo, B lo], each kB16Part halves
  auto part = [&](int buf, int which) { return base + (buf * 4 + which) * kB16Part; };
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int wr = (wave % 2) * 64, wc = (wave / 2) * 64;
  const int l16 = lane % 16, lk = lane / 16;
  // A: rows 4 ag .. 4 ag + 3 of the tile, k = 4 aq + u; B: columns
  // bc + 32 u, k = 4 bq .. 4 bq + 3. The 32 lanes of a half-wave store rows
  // 4 apart (A: ag = tid / 8; B: bc = 4 ((tid / 8) % 4) + ...), which at 80
  // bytes per LDS row start 16 banks apart: their 8-byte stores hit 64
  // distinct banks (rows 1 apart would collide every fourth row, rows 16
  // apart every row)
  const int ag = tid / 8, aq = tid % 8;
  const int bq = tid % 8, bc = 4 * ((tid / 8) % 4) + (tid / 32) % 4 + 16 * (tid / 128);
  const int arow = r0 + 4 * ag;
  const int anv = rend - arow;  // valid rows of this thread's A groups
  f4u ra[4], rb[4];
  auto gload = [&](int kc) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = kc + 4 * aq + u;
      ra[u] = load4(F + arow + (int64_t)(k0 + k) * L, k < bw ? anv : 0);
      const int c = j0 + bc + 32 * u;
      const int kb = kc + 4 * bq;
      rb[u] = load4(F + (k0 + kb) + (int64_t)c * L, c < cend ? bw - kb : 0);
    }
  };
  auto lstore = [&](int buf) {
    unsigned short* Ah = part(buf, 0);
    unsigned short* Bh = part(buf, 1);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      *reinterpret_cast<uint2*>(Ah + (4 * ag + e) * kB16 + 4 * aq) =
          pack_bf16(ra[0][e], ra[1][e], ra[2][e], ra[3][e]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      *reinterpret_cast<uint2*>(Bh + (bc + 32 * u) * kB16 + 4 * bq) =
          pack_bf16(rb[u][0], rb[u][1], rb[u][2], rb[u][3]);
    if (split) {
      unsigned short* Al = part(buf, 2);
      unsigned short* Bl = part(buf, 3);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        *reinterpret_cast<uint2*>(Al + (4 * ag + e) * kB16 + 4 * aq) =
            pack_bf16(bf_lo(ra[0][e]), bf_lo(ra[1][e]), bf_lo(ra[2][e]), bf_lo(ra[3][e]));
#pragma unroll
      for (int u = 0; u < 4; ++u)
        *reinterpret_cast<uint2*>(Bl + (bc + 32 * u) * kB16 + 4 * bq) =
            pack_bf16(bf_lo(rb[u][0]), bf_lo(rb[u][1]), bf_lo(rb[u][2]), bf_lo(rb[u][3]));
    }
  };
  const int nch = (bw + KC - 1) / KC;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int cur = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * KC);
    bf16x8 ah[4], bh[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      ah[t] = *reinterpret_cast<const bf16x8*>(part(cur, 0) + (wr + t * 16 + l16) * kB16 + 8 * lk);
      bh[t] = *reinterpret_cast<const bf16x8*>(part(cur, 1) + (wc + t * 16 + l16) * kB16 + 8 * lk);
    }
    if (split) {
      bf16x8 al[4], bl[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        al[t] = *reinterpret_cast<const bf16x8*>(part(cur, 2) + (wr + t * 16 + l16) * kB16 + 8 * lk);
        bl[t] = *reinterpret_cast<const bf16x8*>(part(cur, 3) + (wc + t * 16 + l16) * kB16 + 8 * lk);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[c], ah[a], acc[a][c], 0, 0, 0);
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[c], al[a], acc[a][c], 0, 0, 0);
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[c], ah[a], acc[a][c], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[c], ah[a], acc[a][c], 0, 0, 0);
    }
    if (ch + 1 < nch) lstore(cur ^ 1);
    __syncthreads();
  }
}

// With lowp set, a float update of a front wider than 256 columns takes its
// L21 / U12 operands in 16 bits and accumulates in float: lowp 1 rounds them
// to bf16 for v_mfma_f32_16x16x16bf16_1k, lowp 2 scales column k of L21 up
// and row k of U12 down by the power of two of |U(k, k)| and rounds to fp16
// for v_mfma_f32_16x16x16f16 (operand k = 4 (l / 16) .. + 3 per lane, the
// accumulator layout of the f32 form); refinement in double recovers the
// accuracy.
// double-buffered: As[buf][i * KP + k] = L(r0 + i, k0 + k),
//                  Bs[buf][j * KP + k] = U(k0 + k, j0 + j)
// (the split-bf16 form stages four bf16 arrays, hi and lo of A and B, in
// the same bytes: 2 buffers x 4 x 128 rows of 20 halves)
template <typename T>
constexpr int kGemmLds = (std::is_same<T, float>::value && 2 * 4 * kB16Part * 2 >
                                                              2 * 2 * 128 * 17 * (int)sizeof(T))
                             ? 2 * 4 * kB16Part * 2
                             : 2 * 2 * 128 * 17 * (int)sizeof(T);

// Workgroup bid of a trailing-update launch over desc[0, ndesc); lds_raw:
// kGemmLds<T> bytes.
template <typename T>
__device__ __attribute__((always_inline)) void blk_gemm_body(
    const DevSym& s, const GemmDesc* __restrict__ desc, int ndesc, T* __restrict__ arena,
    int lowp, int bid, unsigned char* lds_raw) {
  constexpr int TM = 128, KC = 16, KP = KC + 1, PER = TM * KC / kWG;  // 8 loads each
  typedef typename Mfma16<T>::v4 v4;
  T (*As)[TM * KP] = reinterpret_cast<T (*)[TM * KP]>(lds_raw);
  T (*Bs)[TM * KP] = reinterpret_cast<T (*)[TM * KP]>(lds_raw + 2 * TM * KP * sizeof(T));
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int di = find_blk(desc, ndesc, bid);
  const GemmDesc d = desc[di];
  const int b = bid - d.wg_begin;
  const int L = s.fL[d.front];
  T* F = arena + s.foff[d.front];
  const int bw = d.k1 - d.k0;
  const int rend = d.r0 + d.nr, cend = d.j0 + d.nc;
  const int r0 = d.r0 + (b % d.ntr) * TM, j0 = d.j0 + (b / d.ntr) * TM;
  // symmetric fronts: an entry whose row lies past the tile's last column
  // also stands for its transpose, which no tile computes (a tile column's
  // own square, its first tile, is computed whole)
  const int mend = d.mirror ? min(j0 + TM, cend) : (1 << 30);
  const int wr = (wave % 2) * 64, wc = (wave / 2) * 64;
  const int l16 = lane % 16, lk = lane / 16;
  // this thread's share of a chunk: A element (ai, ak + 2 t), B element
  // (bk, bj + 16 t) for t < 8
  const int ai = tid % TM, ak = tid / TM;       // ak in {0, 1}
  const int bk = tid % KC, bj = tid / KC;       // bj in 0..15
  const bool arow = r0 + ai < rend;
  const T* Ap = F + (r0 + ai) + (int64_t)d.k0 * L;
  const T* Bp = F + d.k0 + bk;
  T ra[PER], rb[PER];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int kk = ak + 2 * t;
      ra[t] = (arow && k0 + kk < bw) ? Ap[(int64_t)(k0 + kk) * L] : T(0);
      const int c = j0 + bj + 16 * t;
      rb[t] = (k0 + bk < bw && c < cend) ? Bp[k0 + (int64_t)c * L] : T(0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      As[buf][ai * KP + ak + 2 * t] = ra[t];
      Bs[buf][(bj + 16 * t) * KP + bk] = rb[t];
    }
  };
  // the products are summed from zero and subtracted from A22 once at the
  // end: starting the accumulators from A22 instead loses accuracy when
  // |A22| >> |L21 U12| (measured 8x larger backward error on the 3D
  // convection-diffusion operator)
  v4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = v4{T(0), T(0), T(0), T(0)};
  const int nch = (bw + KC - 1) / KC;
  bool done = false;
  if constexpr (std::is_same<T, float>::value) {
    if (lowp == 2 && d.big) {
      // fp16 operands under diagonal scaling: A(i, k) times and B(k, j)
      // divided by 2^e_k, 2^e_k the power of two nearest below |U(k, k)|
      // (the product is unchanged; both stay in fp16 range when A is
      // diagonally scaled); rows of KB = 20 halves as in the bf16 form
      constexpr int KB = 20;
      _Float16* Ah = reinterpret_cast<_Float16*>(&As[0][0]);
      _Float16* Bh = reinterpret_cast<_Float16*>(&Bs[0][0]);
      const int akw = __builtin_amdgcn_readfirstlane(ak);
      float sa[PER], sb = 1.0f;
      // 2^e and 2^-e of the diagonal entry of front column k (1 when the
      // entry is zero, subnormal or not finite)
      auto pow2 = [&](int k, bool inv) {
        const float dg = F[(int64_t)k * (L + 1)];
        const unsigned be = (__float_as_uint(dg) >> 23) & 0xffu;
        if (be == 0u || be >= 254u) return 1.0f;
        return __uint_as_float((inv ? 254u - be : be) << 23);
      };
      auto scales = [&](int k0) {
#pragma unroll
        for (int t = 0; t < PER; ++t) {
          const int kk = akw + 2 * t;
          sa[t] = (k0 + kk < bw) ? pow2(d.k0 + k0 + kk, false) : 1.0f;
        }
        sb = (k0 + bk < bw) ? pow2(d.k0 + k0 + bk, true) : 1.0f;
      };
      auto hstore = [&](int buf) {
#pragma unroll
        for (int t = 0; t < PER; ++t) {
          Ah[buf * TM * KB + ai * KB + ak + 2 * t] = (_Float16)(ra[t] * sa[t]);
          Bh[buf * TM * KB + (bj + 16 * t) * KB + bk] = (_Float16)(rb[t] * sb);
        }
      };
      gload(0);
      scales(0);
      hstore(0);
      __syncthreads();
      for (int ch = 0; ch < nch; ++ch) {
        const int cur = ch & 1;
        if (ch + 1 < nch) {
          gload((ch + 1) * KC);
          scales((ch + 1) * KC);
        }
        f16x4 av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          av[t] = *reinterpret_cast<const f16x4*>(Ah + cur * TM * KB + (wr + t * 16 + l16) * KB + 4 * lk);
          bv[t] = *reinterpret_cast<const f16x4*>(Bh + cur * TM * KB + (wc + t * 16 + l16) * KB + 4 * lk);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x16f16(bv[c], av[a], acc[a][c], 0, 0, 0);
        if (ch + 1 < nch) hstore(cur ^ 1);
        __syncthreads();
      }
      done = true;
    }
    if ((lowp == 1 || lowp == 3) && d.big) {
      gemm_bf16_chunks(F, L, d.k0, bw, r0, rend, j0, cend, lowp == 3, lds_raw, acc);
      done = true;
    }
  }
  if (!done) {
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int cur = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * KC);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kk = ks * 4 + lk;
      T av[4], bv[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        av[t] = As[cur][(wr + t * 16 + l16) * KP + kk];
        bv[t] = Bs[cur][(wc + t * 16 + l16) * KP + kk];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = Mfma16<T>::op(bv[c], av[a], acc[a][c]);
    }
    if (ch + 1 < nch) lstore(cur ^ 1);
    __syncthreads();
  }
  }
  // A22 -= the products, AG x 16 entries at a time: all reads of a group
  // are issued before its writes (one memory round trip per group instead
  // of one per entry; all 64 floats at once, doubles in two groups of 32 to
  // stay within the register budget of two workgroups per CU)
  constexpr int AG = std::is_same<T, float>::value ? 4 : 2;
#pragma unroll
  for (int a0 = 0; a0 < 4; a0 += AG) {
    T cv[AG][4][4];
#pragma unroll
    for (int ag = 0; ag < AG; ++ag) {
      const int r = r0 + wr + (a0 + ag) * 16 + l16;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int col = j0 + wc + c * 16 + Mfma16<T>::row(lk, v);
          cv[ag][c][v] = (r < rend && col < cend) ? F[r + (int64_t)col * L] : T(0);
        }
    }
#pragma unroll
    for (int ag = 0; ag < AG; ++ag) {
      const int r = r0 + wr + (a0 + ag) * 16 + l16;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int col = j0 + wc + c * 16 + Mfma16<T>::row(lk, v);
          if (r < rend && col < cend) {
            const T nv = cv[ag][c][v] - acc[a0 + ag][c][v];
            F[r + (int64_t)col * L] = nv;
            if (r >= mend) F[col + (int64_t)r * L] = nv;
          }
        }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kWG, 2)
blk_gemm_kernel(DevSym s, const GemmDesc* __restrict__ desc, int ndesc,
                T* __restrict__ arena, int lowp) {
  const int64_t vs = blockIdx.y;  // value set
  arena += vs * s.vs_arena;
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[kGemmLds<T>];
  blk_gemm_body<T>(s, desc, ndesc, arena, lowp, blockIdx.x, lds_raw);
}

// The diagonal regions of one block step (workgroups b < ndiag, with the
// previous block's product folded in) beside the rest of the previous
// block's update within its block group (the others): the two touch
// disjoint parts of each front, so one launch holds both and the update
// leaves the step's critical path.
template <typename T>
__global__ void __launch_bounds__(kWG)
blk_diag_border_kernel(DevSym s, const BlkDesc* __restrict__ ddesc, int ndiag,
                       const GemmDesc* __restrict__ gdesc, int ngemm, T* __restrict__ arena,
                       int64_t* __restrict__ ipvt, unsigned long long* __restrict__ first_fail,
                       int lowp) {
  const int64_t vs = blockIdx.y;  // value set
  arena += vs * s.vs_arena;
  ipvt += vs * s.n;
  constexpr int kL = kSlabLds<T> > kGemmLds<T> ? kSlabLds<T> : kGemmLds<T>;
  __shared__ __attribute__((aligned(16))) unsigned char lds[kL];
  if ((int)blockIdx.x < ndiag)
    blk_diag_slab_body<T>(s, ddesc[blockIdx.x], arena, ipvt, first_fail, lds);
  else
    blk_gemm_body<T>(s, gdesc, ngemm, arena, lowp, (int)blockIdx.x - ndiag, lds);
}

// ---------------------------------------------------------------------------
// Triangular solves, one elimination-tree level at a time.
// Forward (_lulsolve!, SpkLUFactor.jl:269-323). A front's vector y (length
// L per right-hand side, in the vector arena) starts as the permuted rhs on
// its W own columns and zero on its R update rows, and receives the
// children's update vectors through the relative-index maps (the
// multifrontal form of the reference's scatter rhs[lindx] += temp). Then,
// block by block (the factor's blocks of pieces): a one-workgroup chain
// solves the block's pieces in LDS (interchanges, unit-lower block solve,
// rows below inside the block) and a many-workgroup GEMV updates every row
// below the block, y[c1:L] -= L[c1:L, c0:c1] y[c0:c1].
// Backward (_luusolve!, :325-377), blocks in reverse: t = U[c0:c1, c1:L]
// x[I[c1:L]] (one wave per row), then the chain: y = x - t, per piece in
// reverse the in-block U12 part and the upper block solve.
// Small fronts (L <= kSmallLmax) do everything in one workgroup.
// The serial block solves run on wave 0 only: LDS operations of one wave
// complete in issue order, so no barrier is needed between their steps.
// ---------------------------------------------------------------------------
template <typename T, int MAXW>
__device__ void stage_block(T* D, const T* lp, int nq, int jl) {
  for (int idx = threadIdx.x; idx < nq * nq; idx += kWG) {
    int i = idx % nq, k = idx / nq;
    D[i + k * MAXW] = lp[i + (int64_t)k * jl];
  }
}

// sum_k L[r, pc0 + k] * y[k] for k < nq: the column bases lcol[pc0 + k] are
// wave-uniform; loads are issued eight at a time.
template <typename T>
__device__ inline T dot_lrow(const DevSym& s, const T* __restrict__ lnz, int64_t pc0, int nq,
                             int r, const T* y) {
  T acc = T(0);
  int k = 0;
  for (; k + 8 <= nq; k += 8) {
    T lv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) lv[j] = lnz[s.lcol[pc0 + k + j] + r];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc += lv[j] * y[k + j];
  }
  for (; k < nq; ++k) acc += lnz[s.lcol[pc0 + k] + r] * y[k];
  return acc;
}


// Unit-lower solve of one piece's diagonal block on yb (LDS) by wave 0,
// after the piece's interchanges (glaswp! then gtrsm! 'l','l','n','u').
template <typename T>
__device__ void wave_lower_solve(T* yb, const T* D, int ld, const int* piv, int nq) {
  const int lane = threadIdx.x % 64;
  if (lane == 0)
    for (int k = 0; k < nq; ++k) {
      int ip = piv[k];
      if (ip != k) {
        T t = yb[k];
        yb[k] = yb[ip];
        yb[ip] = t;
      }
    }
  wave_lds_sync();
  for (int k = 0; k < nq; ++k) {
    const T yk = yb[k];
    if (yk != T(0))
      for (int i = k + 1 + lane; i < nq; i += 64) yb[i] -= yk * D[i + k * ld];
    wave_lds_sync();
  }
}

// Upper solve (gtrsm! 'l','u','n','n') of one piece's block by wave 0.
template <typename T>
__device__ void wave_upper_solve(T* yb, const T* D, int ld, int nq) {
  const int lane = threadIdx.x % 64;
  for (int k = nq - 1; k >= 0; --k) {
    T yk = yb[k];
    if (yk != T(0)) {
      yk /= D[k + k * ld];
      wave_lds_sync();
      if (lane == 0) yb[k] = yk;
      for (int i = lane; i < k; i += 64) yb[i] -= yk * D[i + k * ld];
    }
    wave_lds_sync();
  }
}

// Most right-hand sides the small-front solves take per staging of a panel.
constexpr int kSmallRhsBlock = 16;

// Small fronts, one wave per front (fpw <= 4 fronts per workgroup, as many
// as the LDS holds), everything in LDS; lane loops stride 64 and a wave
// barrier orders the steps.
// Forward: y starts as the own rows of x and zero update rows, receives the
// children's update vectors in a fixed child order, then per piece: the
// piece's whole panel (rows from its first row to L) is staged once, the
// interchanges and the unit-lower block solve, and every row below takes its
// dot product with the block's entries from the staged panel (the order of
// dot_lrow). Only the update rows go back to the vector arena (the parent
// reads nothing else), the own rows to x. Dynamic LDS per wave: the staged
// panel (MAXW x maxL, maxL the level's longest small front) and yW[maxL].
template <typename T, int MAXW>
__global__ void __launch_bounds__(kWG)
fwd_small_kernel(DevSym s, const int32_t* __restrict__ list, int cnt, int fpw, int qb,
                 const T* __restrict__ lnz, const int64_t* __restrict__ ipvt,
                 T* __restrict__ rhs, int64_t ldr, int nrhs, T* __restrict__ vec, int maxL,
                 int maxq) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  ipvt += vs * s.n;
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  extern __shared__ unsigned char smem_raw[];
  __shared__ int spiv_all[4][MAXW];
  const int lane = threadIdx.x % 64, wave = wave_id();
  const int item = blockIdx.x * fpw + wave;
  if (wave >= fpw || item >= cnt) return;
  // the staged panel: rows of stride maxL, at most maxq columns (the level's
  // widest piece); then the qb solve vectors
  T* D = reinterpret_cast<T*>(smem_raw) + (int64_t)wave * (maxq + qb) * maxL;
  T* Y = D + (int64_t)maxq * maxL;
  int* spiv = spiv_all[wave];
  const int f = list[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  // right-hand sides in blocks of qb (y of the block's right-hand side qq
  // at Y + qq maxL), each piece's panel staged once per block; the lanes
  // take (right-hand side, row) pairs, and every right-hand side receives
  // the same operations in the same order as when solved alone
  for (int q0 = 0; q0 < nrhs; q0 += qb) {
    const int nq_b = min(qb, nrhs - q0);
    for (int idx = lane; idx < nq_b * L; idx += 64) {
      const int qq = idx / L, r = idx - qq * L;
      Y[qq * maxL + r] = (r < W) ? rhs[(int64_t)(q0 + qq) * ldr + fc0 + r] : T(0);
    }
    wave_lds_sync();
    for (int ci = s.fchp[f]; ci < s.fchp[f + 1]; ++ci) {
      const int c = s.fch[ci];
      const int Lc = s.fL[c], Wc = s.fW[c], Rc = Lc - Wc;
      const int32_t* rel = s.relmap + s.frel[c];
      const T* uc = vec + s.fvec[c] * nrhs + (int64_t)q0 * Lc + Wc;
      for (int idx = lane; idx < nq_b * Rc; idx += 64) {
        const int qq = idx / Rc, t = idx - qq * Rc;
        Y[qq * maxL + rel[t]] += uc[(int64_t)qq * Lc + t];
      }
      wave_lds_sync();
    }
    for (int p = sp0; p < sp0 + np; ++p) {
      int64_t pc0, nq64;
      piece_cols(s, p, pc0, nq64);
      const int nq = (int)nq64;
      const int cq = (int)(pc0 - fc0);
      const int jl = L - cq;  // the panel's rows
      const T* lp = lnz + (s.xlnz[pc0] - 1);
      for (int idx = lane; idx < jl * nq; idx += 64) {
        const int i = idx % jl, kk = idx / jl;
        D[i + kk * maxL] = lp[i + (int64_t)kk * jl];
      }
      for (int i = lane; i < nq; i += 64) spiv[i] = (int)ipvt[pc0 + i] - 1;
      wave_lds_sync();
      // the piece's interchanges (glaswp!), one lane per right-hand side
      if (lane < nq_b) {
        T* yb = Y + lane * maxL + cq;
        for (int k = 0; k < nq; ++k) {
          const int ip = spiv[k];
          if (ip != k) {
            const T t = yb[k];
            yb[k] = yb[ip];
            yb[ip] = t;
          }
        }
      }
      wave_lds_sync();
      // unit-lower block solve (gtrsm! 'l','l','n','u'): step k updates the
      // entries below k of every right-hand side
      for (int k = 0; k + 1 < nq; ++k) {
        const int m = nq - 1 - k;
        for (int idx = lane; idx < nq_b * m; idx += 64) {
          const int qq = idx / m, i = k + 1 + (idx - qq * m);
          T* yb = Y + qq * maxL + cq;
          const T yk = yb[k];
          if (yk != T(0)) yb[i] -= yk * D[i + k * maxL];
        }
        wave_lds_sync();
      }
      // the rows below the block: dot products with the staged panel
      const int nrb = L - cq - nq;
      for (int idx = lane; idx < nq_b * nrb; idx += 64) {
        const int qq = idx / nrb, r = cq + nq + (idx - qq * nrb);
        T* yW = Y + qq * maxL;
        const T* dr = D + (r - cq);
        T acc = T(0);
        for (int kk = 0; kk < nq; ++kk) acc += dr[kk * maxL] * yW[cq + kk];
        yW[r] -= acc;
      }
      wave_lds_sync();
    }
    T* yg = vec + s.fvec[f] * nrhs + (int64_t)q0 * L;
    for (int idx = lane; idx < nq_b * L; idx += 64) {
      const int qq = idx / L, r = idx - qq * L;
      const T v = Y[qq * maxL + r];
      if (r < W)
        rhs[(int64_t)(q0 + qq) * ldr + fc0 + r] = v;
      else
        yg[(int64_t)qq * L + r] = v;
    }
    wave_lds_sync();
  }
}

// Small fronts, backward, one wave per front: y gathered from x through the
// index list, then per piece in reverse its U rows staged once in LDS, the
// U12 products (fixed reduction order) and the upper block solve. Dynamic
// LDS per wave: the staged rows (MAXW x maxL) and yW[maxL].
template <typename T, int MAXW>
__global__ void __launch_bounds__(kWG)
bwd_small_kernel(DevSym s, const int32_t* __restrict__ list, int cnt, int fpw, int qb,
                 const T* __restrict__ lnz, const T* __restrict__ unz, T* __restrict__ rhs,
                 int64_t ldr, int nrhs, int maxL, int maxq) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  unz += vs * s.vs_unz;
  rhs += vs * ldr * nrhs;
  extern __shared__ unsigned char smem_raw[];
  __shared__ T syk_all[4][kSmallRhsBlock];  // step k's entry of each right-hand side
  const int lane = threadIdx.x % 64, wave = wave_id();
  const int item = blockIdx.x * fpw + wave;
  if (wave >= fpw || item >= cnt) return;
  // the staged U rows: at most maxq rows (the level's widest piece) of
  // stride maxq, maxL columns; then the qb solve vectors
  T* D = reinterpret_cast<T*>(smem_raw) + (int64_t)wave * (maxq + qb) * maxL;
  T* Y = D + (int64_t)maxq * maxL;
  const int f = list[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  const int32_t* I = s.lindx + (s.xlindx[sp0] - 1);
  // right-hand sides in blocks of qb (y of the block's right-hand side qq
  // at Y + qq maxL), each piece's U rows staged once per block; lanes take
  // (right-hand side, row) pairs, every right-hand side receiving the same
  // operations in the same order as when solved alone
  T* syk = syk_all[wave];
  for (int q0 = 0; q0 < nrhs; q0 += qb) {
    const int nq_b = min(qb, nrhs - q0);
    for (int idx = lane; idx < nq_b * L; idx += 64) {
      const int qq = idx / L, r = idx - qq * L;
      Y[qq * maxL + r] = rhs[(int64_t)(q0 + qq) * ldr + I[r] - 1];
    }
    wave_lds_sync();
    for (int p = sp0 + np - 1; p >= sp0; --p) {
      int64_t pc0, nq64;
      piece_cols(s, p, pc0, nq64);
      const int nq = (int)nq64;
      const int cq = (int)(pc0 - fc0);
      // the piece's U rows staged once: D(c, j - cq) = U(c, j) for j from
      // the piece's first column to L (the block from the packed LU in lnz,
      // the rest from unz)
      const int jl = L - cq;
      const T* lp = lnz + (s.xlnz[pc0] - 1);
      for (int idx = lane; idx < nq * jl; idx += 64) {
        const int c = idx % nq, j = idx / nq;
        D[c + j * maxq] = (j < nq) ? lp[c + (int64_t)j * jl] : unz[s.ucol[pc0 + c] + cq + j];
      }
      wave_lds_sync();
      // U12 products: 16 lanes per (right-hand side, row) pair, lane g
      // summing the columns g, g + 16, ... in order, then a fixed tree over
      // the 16 partial sums
      {
        const int grp = lane / 16, gl = lane % 16;
        for (int pr = grp; pr < nq_b * nq; pr += 4) {
          const int qq = pr / nq, c = pr - qq * nq;
          T* yW = Y + qq * maxL;
          T acc = T(0);
          for (int j = cq + nq + gl; j < L; j += 16) acc += D[c + (j - cq) * maxq] * yW[j];
          acc += __shfl_xor(acc, 8, 16);
          acc += __shfl_xor(acc, 4, 16);
          acc += __shfl_xor(acc, 2, 16);
          acc += __shfl_xor(acc, 1, 16);
          if (gl == 0) yW[cq + c] -= acc;
        }
      }
      wave_lds_sync();
      // upper block solve (gtrsm! 'l','u','n','n'): step k divides entry k
      // of every right-hand side by U(k, k) and updates the entries above it
      for (int k = nq - 1; k >= 0; --k) {
        if (lane < nq_b) syk[lane] = Y[lane * maxL + cq + k];
        wave_lds_sync();
        for (int idx = lane; idx < nq_b * (k + 1); idx += 64) {
          const int qq = idx / (k + 1), i = idx - qq * (k + 1);
          T yk = syk[qq];
          if (yk != T(0)) {
            yk /= D[k + k * maxq];
            T* yb = Y + qq * maxL + cq;
            if (i == k)
              yb[k] = yk;
            else
              yb[i] -= yk * D[i + k * maxq];
          }
        }
        wave_lds_sync();
      }
    }
    for (int idx = lane; idx < nq_b * W; idx += 64) {
      const int qq = idx / W, c = idx - qq * W;
      rhs[(int64_t)(q0 + qq) * ldr + fc0 + c] = Y[qq * maxL + c];
    }
    wave_lds_sync();
  }
}

// Large fronts, forward: y = [rhs on own columns; 0] + children, one thread
// per (front, row) item of the level (the factor's column items: a front's
// rows and columns share its index list). The children's entries come from
// the item's (child, row) list in ascending child order, so each entry is
// the same sum in the same order as a per-front gather; a level's fronts
// spread over many workgroups.
template <typename T>
__global__ void __launch_bounds__(kWG)
fwd_assemble_items_kernel(DevSym s, const int32_t* __restrict__ it_front,
                          const int32_t* __restrict__ it_col,
                          const int64_t* __restrict__ it_ea,
                          const int32_t* __restrict__ ea_child,
                          const int32_t* __restrict__ ea_t, const T* __restrict__ rhs,
                          int64_t ldr, int nrhs, T* __restrict__ vec, int nitems) {
  const int64_t vs = blockIdx.y;  // value set
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  const int item = blockIdx.x * kWG + threadIdx.x;
  if (item >= nitems) return;
  const int f = it_front[item], r = it_col[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int64_t e0 = it_ea[item], e1 = it_ea[item + 1];
  for (int q = 0; q < nrhs; ++q) {
    T y = (r < W) ? rhs[(int64_t)q * ldr + fc0 + r] : T(0);
    for (int64_t e = e0; e < e1; ++e) {
      const int c = ea_child[e], t = ea_t[e];
      y += vec[s.fvec[c] * nrhs + (int64_t)q * s.fL[c] + s.fW[c] + t];
    }
    vec[s.fvec[f] * nrhs + (int64_t)q * L + r] = y;
  }
}

// Backward: the update rows of each large front's solve vector gathered
// from the solution by the front's index list, one thread per (front, row)
// item of the level.
template <typename T>
__global__ void __launch_bounds__(kWG)
bwd_gather_items_kernel(DevSym s, const int32_t* __restrict__ it_front,
                        const int32_t* __restrict__ it_col, const T* __restrict__ rhs,
                        int64_t ldr, int nrhs, T* __restrict__ vec, int nitems) {
  const int64_t vs = blockIdx.y;  // value set
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  const int item = blockIdx.x * kWG + threadIdx.x;
  if (item >= nitems) return;
  const int f = it_front[item], j = it_col[item];
  const int L = s.fL[f], W = s.fW[f];
  if (j < W) return;
  const int g = s.lindx[s.xlindx[s.fsp0[f]] - 1 + j] - 1;
  for (int q = 0; q < nrhs; ++q)
    vec[s.fvec[f] * nrhs + (int64_t)q * L + j] = rhs[(int64_t)q * ldr + g];
}


// Per-block solve data, built once per factorisation and read by every
// solve (refinement solves many times). A block of bw columns is padded to
// bwp = 4 ng columns, ng = ceil(bw / 4), in ng panels of four columns.
//  - Lower part (8 ng (ng + 1) elements): panel g holds rows 4g .. bwp-1 of
//    columns 4g .. 4g+3 of the unit-lower factor, four values per row, at
//    lo_panel(g, ng); entries on or above the diagonal and padding are 0.
//    The rows of every piece carry that piece's interchanges also in the
//    columns of earlier pieces: the reference interchanges y's rows when it
//    reaches a piece, after earlier pieces have updated them; interchanging
//    y once at the start and the stored L rows instead does the same
//    operations on the same values.
//  - Upper part (8 ng (ng + 1) elements): panel g holds rows 0 .. 4g+3 of
//    columns 4g .. 4g+3 of U, at up_panel(g); entries below the diagonal are
//    0, padding columns are unit columns.
//  - The bwp reciprocals of U's diagonal (1 for padding).
// In all 16 ng (ng + 1) + 4 ng elements, from element 4 desc.aux. The
// integer part (kPrepInts per block, from desc.r0 kPrepInts) is the row each
// position takes after its piece's interchanges (composed, glaswp! order).
__device__ inline int lo_panel(int g, int ng) { return 4 * (4 * ng * g - 2 * g * (g - 1)); }
__device__ inline int up_panel(int g) { return 8 * g * (g + 1); }
constexpr int kPanelElems = 8 * 32 * 33;  // one part of a 128-column block

template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
  typedef double __attribute__((ext_vector_type(4))) type;
  typedef double __attribute__((ext_vector_type(2))) half;
};
template <>
struct Vec4<float> {
  typedef float __attribute__((ext_vector_type(4))) type;
  typedef float __attribute__((ext_vector_type(2))) half;
};

// One workgroup per block.
template <typename T>
__global__ void __launch_bounds__(kWG)
solve_prep_kernel(DevSym s, const SolveDesc* __restrict__ desc, const T* __restrict__ lnz,
                  const T* __restrict__ unz, const int64_t* __restrict__ ipvt,
                  T* __restrict__ prep, int32_t* __restrict__ prepi) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  unz += vs * s.vs_unz;
  ipvt += vs * s.n;
  prep += vs * s.vs_prep;
  prepi += vs * s.vs_prepi;
  __shared__ int spiv[128];
  __shared__ int sperm[128];
  __shared__ int pstart[128];
  __shared__ int pbeg[9];
  __shared__ int64_t lb[128], ub[128];
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const SolveDesc d = desc[blockIdx.x];
  const int64_t fc0 = s.fc0[d.front];
  const int c0 = d.c0, bw = d.c1 - d.c0;
  const int ng = (bw + 3) / 4, bwp = 4 * ng;
  const int64_t g0 = fc0 + c0;
  T* lo = prep + 4 * (int64_t)d.aux;
  T* up = lo + 8 * ng * (ng + 1);
  T* rinv = up + 8 * ng * (ng + 1);
  for (int i = tid; i < bw; i += kWG) {
    const int ps = piece_start_of_col(s, fc0, c0 + i) - c0;
    pstart[i] = ps;
    spiv[i] = ps + (int)ipvt[g0 + i] - 1;
    sperm[i] = i;
    lb[i] = s.lcol[g0 + i] + c0;
    ub[i] = s.ucol[g0 + i] + c0;
  }
  if (tid <= d.np) pbeg[tid] = (tid < d.np) ? (int)(s.xsuper[d.piece + tid] - 1 - g0) : bw;
  __syncthreads();
  if (wave == 0 && lane < d.np) {
    for (int k = pbeg[lane]; k < pbeg[lane + 1]; ++k) {
      const int kp = spiv[k];
      if (kp != k) {
        const int t = sperm[k];
        sperm[k] = sperm[kp];
        sperm[kp] = t;
      }
    }
  }
  __syncthreads();
  for (int g = wave; g < ng; g += 4) {
    T* P = lo + lo_panel(g, ng);
    const int ne = (bwp - 4 * g) * 4;
#pragma unroll 4
    for (int i = lane; i < ne; i += 64) {
      const int r = 4 * g + i / 4, c = 4 * g + i % 4;
      T v = T(0);
      if (r > c && r < bw && c < bw) v = lnz[lb[c] + (pstart[r] == pstart[c] ? r : sperm[r])];
      P[i] = v;
    }
  }
  for (int g = wave; g < ng; g += 4) {
    T* P = up + up_panel(g);
    const int ne = (4 * g + 4) * 4;
#pragma unroll 4
    for (int i = lane; i < ne; i += 64) {
      const int r = i / 4, c = 4 * g + i % 4;
      T v = T(0);
      if (r <= c) {
        if (c < bw)
          v = (r >= pstart[c]) ? lnz[lb[c] + r] : unz[ub[r] + c];
        else if (r == c)
          v = T(1);
      }
      P[i] = v;
    }
  }
  for (int c = tid; c < bwp; c += kWG) rinv[c] = (c < bw) ? T(1) / lnz[lb[c] + c] : T(1);
  int32_t* pi = prepi + (int64_t)d.r0 * kPrepInts;
  for (int i = tid; i < bw; i += kWG) pi[i] = sperm[i];
}

// Value i (0 <= i < 128, wave-uniform) of a vector held two entries per lane
// (entry i in lane i % 64 of y0 for i < 64, of y1 otherwise).
template <typename T>
__device__ inline T bcast2(T y0, T y1, int i) {
  return rdlane(i < 64 ? y0 : y1, i & 63);
}

// Entry src (0 <= src < 128, per lane) of the same two-per-lane vector.
template <typename T>
__device__ inline T gather2(T y0, T y1, int src) {
  const T a = __shfl(y0, src & 63, 64), b = __shfl(y1, src & 63, 64);
  return src < 64 ? a : b;
}

// v, computed where it stands for every lane (keeps the compiler from
// turning a select of it into a branch around its computation).
__device__ inline double pin(double v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ inline float pin(float v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The solve data of a block (nv four-element vectors) into LDS, all loads in
// flight at once; at most kPrepV vectors per thread.
constexpr int kPrepV = 9;
template <typename V>
__device__ inline void load_panels(const V* __restrict__ src, V* dst, int nv) {
  V buf[kPrepV];
#pragma unroll
  for (int u = 0; u < kPrepV; ++u) {
    const int i = threadIdx.x + u * kWG;
    if (i < nv) buf[u] = src[i];
  }
#pragma unroll
  for (int u = 0; u < kPrepV; ++u) {
    const int i = threadIdx.x + u * kWG;
    if (i < nv) dst[i] = buf[u];
  }
}

// Forward chain of one block (one workgroup per front of the step). The
// block's lower panels come from the solve data in one coalesced load. Wave
// 0 holds y in registers, two entries per lane (rows lane, lane + 64),
// takes the block's interchanges at once (cross-lane shuffles) and
// eliminates four columns per step: the four new entries from a small
// unit-lower solve every lane does alike, then every row takes its four
// products in column order — the same operations, in the same order, as
// the reference's column-by-column piece solves and in-block L21 updates.
// A row at or above the step's first column reads the panel's zero row
// (its own four products are then exact no-ops), so the three rows below
// the first come out as the small solve's entries and no lane branches.
// The next step's panel values are read from LDS while the current step
// computes; once the first 64 rows are final only the second half updates.
template <typename T>
struct FwdStep {
  T t10, t32;                        // L(k+1, k), L(k+3, k+2)
  typename Vec4<T>::half t2, t3;     // L(k+2, k:k+2), L(k+3, k:k+2)
  typename Vec4<T>::type a, b;       // the lane's two rows
};

template <typename T>
__device__ __attribute__((always_inline)) void fwd_chain_body(
    const DevSym& s, const SolveDesc& d, const T* __restrict__ prep,
    const int32_t* __restrict__ prepi, T* __restrict__ rhs, int64_t ldr, int nrhs,
    T* __restrict__ vec, unsigned char* smem_a32) {
  using V = typename Vec4<T>::type;
  V* smem_v4 = reinterpret_cast<V*>(smem_a32);
  const T* R = reinterpret_cast<const T*>(smem_a32);
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int f = d.front;
  const int L = d.L;
  const int64_t fc0 = d.fc0;
  const int c0 = d.c0, bw = d.c1 - d.c0;
  const int ng = (bw + 3) / 4, bwp = 4 * ng;
  const bool in0 = lane < bw, in1 = lane + 64 < bw;
  const int32_t* pi = prepi + (int64_t)d.r0 * kPrepInts;
  // every wave: the permutation and its first right-hand side (wave w
  // takes right-hand sides w, w + 4, ...), in flight with the panels
  int s0 = lane, s1 = lane + 64;
  T y0r = T(0), y1r = T(0);
  const T* yv = vec + d.fvec * nrhs + (int64_t)wave * L;
  if (wave < nrhs) {
    if (in0) {
      s0 = pi[lane];
      y0r = yv[c0 + lane];
    }
    if (in1) {
      s1 = pi[lane + 64];
      y1r = yv[c0 + 64 + lane];
    }
  }
  load_panels(reinterpret_cast<const V*>(prep + 4 * (int64_t)d.aux), smem_v4,
              2 * ng * (ng + 1));
  __syncthreads();
  if (wave >= nrhs) return;
  const int gA = min(ng, 16);
  // panel values of step g (rows clamped into the panel)
  auto load = [&](int g, FwdStep<T>& o, bool both) {
    const int k = 4 * g;
    const T* P = R + lo_panel(g, ng);
    using H = typename Vec4<T>::half;
    o.t10 = P[4];
    o.t2 = *(const H*)(P + 8);
    o.t3 = *(const H*)(P + 12);
    o.t32 = P[14];
    if (both) o.a = *(const V*)(P + 4 * (min(max(lane, k), bwp - 1) - k));
    o.b = *(const V*)(P + 4 * (min(max(lane + 64, k), bwp - 1) - k));
  };
  for (int q = wave; q < nrhs; q += 4) {
    T* yg = vec + d.fvec * nrhs + (int64_t)q * L;
    T* x = rhs + (int64_t)q * ldr;
    if (q > wave) {
      y0r = in0 ? yg[c0 + lane] : T(0);
      y1r = in1 ? yg[c0 + 64 + lane] : T(0);
    }
    T y0 = gather2(y0r, y1r, s0), y1 = gather2(y0r, y1r, s1);
    FwdStep<T> A, B;
    // pivots in rows 0..63: both halves update
    auto stepA = [&](int g, const FwdStep<T>& c, FwdStep<T>& n) {
      const int k = 4 * g;
      load(min(g + 1, ng - 1), n, true);
      const T b0 = rdlane(y0, k), b1 = rdlane(y0, k + 1);
      const T b2 = rdlane(y0, k + 2), b3 = rdlane(y0, k + 3);
      const T z0 = b0;
      const T z1 = b1 - c.t10 * z0;
      const T z2 = (b2 - c.t2.x * z0) - c.t2.y * z1;
      const T z3 = ((b3 - c.t3.x * z0) - c.t3.y * z1) - c.t32 * z2;
      y0 = (((y0 - c.a.x * z0) - c.a.y * z1) - c.a.z * z2) - c.a.w * z3;
      y1 = (((y1 - c.b.x * z0) - c.b.y * z1) - c.b.z * z2) - c.b.w * z3;
    };
    load(0, A, true);
    for (int g = 0; g < gA; g += 2) {
      stepA(g, A, B);
      if (g + 1 < gA) stepA(g + 1, B, A);
    }
    // pivots in rows 64..127: only the second half updates
    auto stepB = [&](int g, const FwdStep<T>& c, FwdStep<T>& n) {
      const int k = 4 * g - 64;
      load(min(g + 1, ng - 1), n, false);
      const T b0 = rdlane(y1, k), b1 = rdlane(y1, k + 1);
      const T b2 = rdlane(y1, k + 2), b3 = rdlane(y1, k + 3);
      const T z0 = b0;
      const T z1 = b1 - c.t10 * z0;
      const T z2 = (b2 - c.t2.x * z0) - c.t2.y * z1;
      const T z3 = ((b3 - c.t3.x * z0) - c.t3.y * z1) - c.t32 * z2;
      y1 = (((y1 - c.b.x * z0) - c.b.y * z1) - c.b.z * z2) - c.b.w * z3;
    };
    if (ng > 16) {
      load(16, A, false);
      for (int g = 16; g < ng; g += 2) {
        stepB(g, A, B);
        if (g + 1 < ng) stepB(g + 1, B, A);
      }
    }
    if (in0) {
      yg[c0 + lane] = y0;
      x[fc0 + c0 + lane] = y0;
    }
    if (in1) {
      yg[c0 + 64 + lane] = y1;
      x[fc0 + c0 + 64 + lane] = y1;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kWG)
fwd_chain_kernel(DevSym s, const SolveDesc* __restrict__ desc, const T* __restrict__ prep,
                 const int32_t* __restrict__ prepi, T* __restrict__ rhs, int64_t ldr, int nrhs,
                 T* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  prep += vs * s.vs_prep;
  prepi += vs * s.vs_prepi;
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  extern __shared__ __attribute__((aligned(32))) unsigned char smem_a32[];
  fwd_chain_body<T>(s, desc[blockIdx.x], prep, prepi, rhs, ldr, nrhs, vec, smem_a32);
}

// Forward: y[c1 + r0 + lane] -= L[., c0:c1] y[c0:c1]; wave w takes a
// quarter of the block's columns, partial sums added in a fixed order. The
// block's column offsets and y[c0:c1] are staged in LDS first, then every
// thread issues all of its L loads at once.
template <typename T>
__device__ __attribute__((always_inline)) void fwd_bgemv_body(
    const DevSym& s, const SolveDesc& d, const T* __restrict__ lnz, int nrhs,
    T* __restrict__ vec) {
  __shared__ T red[4][64];
  __shared__ int64_t lc[128];
  __shared__ T zs[128];
  const int tid = threadIdx.x, lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const int f = d.front;
  const int L = d.L;
  const int64_t fc0 = d.fc0;
  const int c0 = d.c0, c1 = d.c1, bw = c1 - c0;
  const int r = c1 + d.r0 + lane;
  const int cb = bw * wave / 4, ce = bw * (wave + 1) / 4;  // at most 32 columns
  if (tid < bw) lc[tid] = s.lcol[fc0 + c0 + tid];
  __syncthreads();
  // the block's entries of this tile's rows, read once for every rhs
  T lv[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) lv[j] = (r < L && cb + j < ce) ? lnz[lc[cb + j] + r] : T(0);
  for (int q = 0; q < nrhs; ++q) {
    T* yg = vec + d.fvec * nrhs + (int64_t)q * L;
    if (tid < bw) zs[tid] = yg[c0 + tid];
    // this tile's rows of y (no other workgroup touches them)
    const T yr = (wave == 0 && r < L) ? yg[r] : T(0);
    __syncthreads();
    T acc = T(0);
    if (r < L) {
#pragma unroll
      for (int j = 0; j < 32; ++j)
        if (cb + j < ce) acc += lv[j] * zs[cb + j];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && r < L) yg[r] = yr - (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
    __syncthreads();
  }
}


template <typename T>
__global__ void __launch_bounds__(kWG)
fwd_bgemv_kernel(DevSym s, const SolveDesc* __restrict__ desc, const T* __restrict__ lnz,
                 int nrhs, T* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  lnz += vs * s.vs_lnz;
  vec += vs * s.vs_vec * nrhs;
  fwd_bgemv_body<T>(s, desc[blockIdx.x], lnz, nrhs, vec);
}

// Backward: t[c] = U[c, c1:L] x[c1:L] for row c = c0 + r0 of the block (a
// workgroup takes one or a few rows in turn), stored in y[c]. x[c1:L] is read from the front's vector:
// its update rows from bwd_gather_items_kernel, its own columns as the backward
// chains of the later blocks left them. Every thread takes columns c1 + tid
// + 256 m, sixteen loads in flight; per-thread sums are added over the wave,
// then the four waves in a fixed order.
template <typename T>
__device__ __attribute__((always_inline)) void bwd_bgemv_body(
    const DevSym& s, const SolveDesc& d, const T* __restrict__ unz, int nrhs,
    T* __restrict__ vec) {
  constexpr int QB = 16;  // right-hand sides sharing one read of the row
  constexpr int CH = 16;  // columns per thread and pass
  __shared__ T red[QB][4];
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int f = d.front;
  const int c = d.c0 + d.r0;
  const int L = d.L;
  const int64_t fc0 = d.fc0;
  const T* u = unz + d.ucol;  // the row's U entries (its unz offset rides in the descriptor)
  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    const int nq = min(QB, nrhs - q0);
    T acc[QB];
#pragma unroll
    for (int qq = 0; qq < QB; ++qq) acc[qq] = T(0);
    for (int j0 = d.c1 + tid; j0 < L; j0 += CH * kWG) {
      T uv[CH];
#pragma unroll
      for (int t = 0; t < CH; ++t) {
        const int j = j0 + kWG * t;
        uv[t] = (j < L) ? u[j] : T(0);
      }
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) {
        if (qq < nq) {
          const T* yg = vec + d.fvec * nrhs + (int64_t)(q0 + qq) * L;
          T xv[CH];
#pragma unroll
          for (int t = 0; t < CH; ++t) {
            const int j = j0 + kWG * t;
            xv[t] = (j < L) ? yg[j] : T(0);
          }
#pragma unroll
          for (int t = 0; t < CH; ++t) acc[qq] += uv[t] * xv[t];
        }
      }
    }
#pragma unroll
    for (int qq = 0; qq < QB; ++qq) {
      if (qq < nq) {
        T a = acc[qq];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) red[qq][wave] = a;
      }
    }
    __syncthreads();
    if (tid < nq) {
      T* yg = vec + d.fvec * nrhs + (int64_t)(q0 + tid) * L;
      yg[c] = ((red[tid][0] + red[tid][1]) + red[tid][2]) + red[tid][3];
    }
    __syncthreads();
  }
}

// The same for R rows of one block of one front at once: every load of x
// serves the R rows. Each (row, right-hand side) sum takes bwd_bgemv_body's
// operations in its order (a thread's columns c1 + tid + 256 m in ascending
// m, the wave tree, the four waves in order), so the results are the same.
template <typename T, int R>
__device__ __attribute__((always_inline)) void bwd_bgemv_rows_body(
    const SolveDesc* __restrict__ dr, const T* __restrict__ unz, int nrhs, T* __restrict__ vec) {
  constexpr int QB = 8;  // right-hand sides per pass
  constexpr int CH = 8;  // columns per thread and chunk
  __shared__ T red[R][QB][4];
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const SolveDesc& d = dr[0];
  const int L = d.L, c1 = d.c1;
  T* const yv = vec + d.fvec * nrhs;
  const T* u[R];
#pragma unroll
  for (int i = 0; i < R; ++i) u[i] = unz + dr[i].ucol;
  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    const int nq = min(QB, nrhs - q0);
    T acc[R][QB];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) acc[i][qq] = T(0);
    for (int j0 = c1 + tid; j0 < L; j0 += CH * kWG) {
      T uv[R][CH];
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int t = 0; t < CH; ++t) {
          const int j = j0 + kWG * t;
          uv[i][t] = (j < L) ? u[i][j] : T(0);
        }
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) {
        if (qq < nq) {
          const T* yg = yv + (int64_t)(q0 + qq) * L;
          T xv[CH];
#pragma unroll
          for (int t = 0; t < CH; ++t) {
            const int j = j0 + kWG * t;
            xv[t] = (j < L) ? yg[j] : T(0);
          }
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int t = 0; t < CH; ++t) acc[i][qq] += uv[i][t] * xv[t];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) {
        if (qq < nq) {
          T a = acc[i][qq];
          for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
          if (lane == 0) red[i][qq][wave] = a;
        }
      }
    __syncthreads();
    if (tid < R * nq) {
      const int i = tid / nq, qq = tid - i * nq;
      yv[(int64_t)(q0 + qq) * L + dr[i].c0 + dr[i].r0] =
          ((red[i][qq][0] + red[i][qq][1]) + red[i][qq][2]) + red[i][qq][3];
    }
    __syncthreads();
  }
}

// rows: consecutive row descriptors per workgroup (1 or 4); four rows of one
// block of one front share their loads of x, any other group is taken one
// row after another (the results do not depend on the grouping)
template <typename T>
__global__ void __launch_bounds__(kWG)
bwd_bgemv_kernel(DevSym s, const SolveDesc* __restrict__ desc, int ndesc, int rows,
                 const T* __restrict__ unz, int nrhs, T* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  unz += vs * s.vs_unz;
  vec += vs * s.vs_vec * nrhs;
  const int base = (int)blockIdx.x * rows;
  if (rows == 4 && base + 3 < ndesc) {
    const SolveDesc* dr = desc + base;
    if (dr[1].front == dr[0].front && dr[2].front == dr[0].front && dr[3].front == dr[0].front &&
        dr[3].c0 == dr[0].c0) {
      bwd_bgemv_rows_body<T, 4>(dr, unz, nrhs, vec);
      return;
    }
  }
  for (int i = 0; i < rows; ++i) {
    const int di = base + i;
    if (di >= ndesc) break;
    bwd_bgemv_body<T>(s, desc[di], unz, nrhs, vec);
  }
}

// b / u from r = 1 / u (correctly rounded): q0 = b r, then one correction
// with the exact remainder b - q0 u (an FMA) gives the correctly rounded
// quotient b / u, as the division itself (Markstein's theorem; no overflow
// or underflow assumed), in three dependent operations.
__device__ inline double div_r(double b, double u, double r) {
  const double q0 = b * r;
  return __builtin_fma(__builtin_fma(-q0, u, b), r, q0);
}
__device__ inline float div_r(float b, float u, float r) {
  const float q0 = b * r;
  return __builtin_fmaf(__builtin_fmaf(-q0, u, b), r, q0);
}

// Backward chain of one block. The block's upper panels and diagonal
// reciprocals come from the solve data in one coalesced load. With t =
// U[c0:c1, c1:L] x[c1:L] already in the vector arena (bwd_bgemv_kernel),
// wave 0 holds y = x - t in registers and back-substitutes, last column
// first, four columns per step: the four new entries x_k = (y_k - ...) /
// U(k, k) every lane computes alike, then each earlier row takes its four
// products in column order — the in-block U12 products and the upper block
// solves of the reference, by columns. The next step's panel values are
// read from LDS while the current step computes.
template <typename T>
struct BwdStep {
  typename Vec4<T>::type d0, ri, a, b;  // U(k, k:k+4), 1 / U(j, j), the lane's rows
  typename Vec4<T>::half d1, d2;        // U(k+1, k+2:k+4), U(k+2, k+2:k+4)
  T d11, d33;                           // U(k+1, k+1), U(k+3, k+3)
};

template <typename T>
__device__ __attribute__((always_inline)) void bwd_chain_body(
    const DevSym& s, const SolveDesc& d, const T* __restrict__ prep, T* __restrict__ rhs,
    int64_t ldr, int nrhs, T* __restrict__ vec, unsigned char* smem_a32) {
  using V = typename Vec4<T>::type;
  V* smem_v4 = reinterpret_cast<V*>(smem_a32);
  const T* R = reinterpret_cast<const T*>(smem_a32);
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const int f = d.front;
  const int L = d.L;
  const int64_t fc0 = d.fc0;
  const int c0 = d.c0, c1 = d.c1, bw = c1 - c0;
  const int ng = (bw + 3) / 4;
  const bool in0 = lane < bw, in1 = lane + 64 < bw, sub = c1 < L;
  T x0r = T(0), x1r = T(0), t0r = T(0), t1r = T(0);
  // wave w takes right-hand sides w, w + 4, ...; its first one in flight
  // with the panels
  if (wave < nrhs) {
    const T* x = rhs + (int64_t)wave * ldr + fc0 + c0;
    const T* yg = vec + d.fvec * nrhs + (int64_t)wave * L + c0;
    if (in0) {
      x0r = x[lane];
      if (sub) t0r = yg[lane];
    }
    if (in1) {
      x1r = x[lane + 64];
      if (sub) t1r = yg[lane + 64];
    }
  }
  const int npan = 8 * ng * (ng + 1);
  load_panels(reinterpret_cast<const V*>(prep + 4 * (int64_t)d.aux + npan), smem_v4,
              2 * ng * (ng + 1) + ng);
  __syncthreads();
  if (wave >= nrhs) return;
  const T* Rinv = R + npan;
  const int gA = min(ng, 16);
  auto load = [&](int g, BwdStep<T>& o, bool both) {
    const int k = 4 * g;
    const T* P = R + up_panel(g);
    using H = typename Vec4<T>::half;
    o.d0 = *(const V*)(P + 4 * k);
    o.d11 = P[4 * k + 5];
    o.d1 = *(const H*)(P + 4 * k + 6);
    o.d2 = *(const H*)(P + 4 * k + 10);
    o.d33 = P[4 * k + 15];
    o.ri = *(const V*)(Rinv + k);
    o.a = *(const V*)(P + 4 * min(lane, k + 3));
    if (both) o.b = *(const V*)(P + 4 * min(lane + 64, k + 3));
  };
  for (int q = wave; q < nrhs; q += 4) {
    T* yg = vec + d.fvec * nrhs + (int64_t)q * L;
    T* x = rhs + (int64_t)q * ldr;
    if (q > wave) {
      x0r = in0 ? x[fc0 + c0 + lane] : T(0);
      x1r = in1 ? x[fc0 + c0 + 64 + lane] : T(0);
      t0r = (in0 && sub) ? yg[c0 + lane] : T(0);
      t1r = (in1 && sub) ? yg[c0 + 64 + lane] : T(0);
    }
    T y0 = in0 ? x0r - t0r : T(0);
    T y1 = in1 ? x1r - t1r : T(0);
    BwdStep<T> A, B;
    // the step's four new entries from its pivot values b0..b3
    auto solve4 = [&](const BwdStep<T>& c, T b0, T b1, T b2, T b3, T& x0, T& x1, T& x2,
                      T& x3) {
      x3 = div_r(b3, c.d33, c.ri.w);
      x2 = div_r(b2 - c.d2.y * x3, c.d2.x, c.ri.z);
      x1 = div_r((b1 - c.d1.y * x3) - c.d1.x * x2, c.d11, c.ri.y);
      x0 = div_r(((b0 - c.d0.w * x3) - c.d0.z * x2) - c.d0.y * x1, c.d0.x, c.ri.x);
    };
    // row update of y at offset dl from the step's first column: rows above
    // take the four products, the step's rows their new entries
    auto upd = [&](T y, const V& a, int dl, T x0, T x1, T x2, T x3) {
      T r = pin((((y - a.w * x3) - a.z * x2) - a.y * x1) - a.x * x0);
      r = (dl == 0) ? x0 : r;
      r = (dl == 1) ? x1 : r;
      r = (dl == 2) ? x2 : r;
      r = (dl == 3) ? x3 : r;
      r = (dl >= 4) ? y : r;
      return r;
    };
    // pivots in rows 64..127: every row of the first half updates
    auto stepB = [&](int g, const BwdStep<T>& c, BwdStep<T>& n) {
      const int k = 4 * g;
      load(max(g - 1, 0), n, true);
      T x0, x1, x2, x3;
      solve4(c, rdlane(y1, k - 64), rdlane(y1, k - 63), rdlane(y1, k - 62), rdlane(y1, k - 61),
             x0, x1, x2, x3);
      y0 = (((y0 - c.a.w * x3) - c.a.z * x2) - c.a.y * x1) - c.a.x * x0;
      y1 = upd(y1, c.b, lane + 64 - k, x0, x1, x2, x3);
    };
    if (ng > 16) {
      load(ng - 1, A, true);
      for (int g = ng - 1; g >= 16; g -= 2) {
        stepB(g, A, B);
        if (g - 1 >= 16) stepB(g - 1, B, A);
      }
    }
    // pivots in rows 0..63: only the first half updates
    auto stepA = [&](int g, const BwdStep<T>& c, BwdStep<T>& n) {
      const int k = 4 * g;
      load(max(g - 1, 0), n, false);
      T x0, x1, x2, x3;
      solve4(c, rdlane(y0, k), rdlane(y0, k + 1), rdlane(y0, k + 2), rdlane(y0, k + 3), x0, x1,
             x2, x3);
      y0 = upd(y0, c.a, lane - k, x0, x1, x2, x3);
    };
    load(gA - 1, A, false);
    for (int g = gA - 1; g >= 0; g -= 2) {
      stepA(g, A, B);
      if (g - 1 >= 0) stepA(g - 1, B, A);
    }
    if (in0) {
      x[fc0 + c0 + lane] = y0;
      yg[c0 + lane] = y0;
    }
    if (in1) {
      x[fc0 + c0 + 64 + lane] = y1;
      yg[c0 + 64 + lane] = y1;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kWG)
bwd_chain_kernel(DevSym s, const SolveDesc* __restrict__ desc, const T* __restrict__ prep,
                 T* __restrict__ rhs, int64_t ldr, int nrhs, T* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  prep += vs * s.vs_prep;
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  extern __shared__ __attribute__((aligned(32))) unsigned char smem_a32[];
  bwd_chain_body<T>(s, desc[blockIdx.x], prep, rhs, ldr, nrhs, vec, smem_a32);
}

// ---------------------------------------------------------------------------
// Float solves with explicit inverses of the diagonal blocks. The chains are
// the solves' serial critical path (one column group after another); with
// ML = L11^-1 P and MU = U11^-1 formed once per factorisation, every block
// step becomes GEMV work that all workgroups of the step do in parallel, one
// launch per step instead of a chain launch and a GEMV launch. (The double
// factor keeps substitution, the reference's operation order.)
// ---------------------------------------------------------------------------

// ML and MU of one block (one workgroup per block), each bwp x bwp
// row-major and zero outside bw x bw, from the block's panels (read from
// global memory, every lane the same entry): column j is formed by
// substitution on e_j in LDS (X[r][j]), the lower inverse first (its columns
// then permuted by the block's interchanges), then the upper one. 64 KB of
// LDS: two workgroups per CU.
constexpr int kInvPrepLds = 128 * 128 * 4;  // float
__global__ void __launch_bounds__(kWG)
solve_inv_prep_kernel(const SolveDesc* __restrict__ desc, const float* __restrict__ prep,
                      const int32_t* __restrict__ prepi, float* __restrict__ inv,
                      int64_t vs_prep, int64_t vs_prepi, int64_t vs_inv) {
  const int64_t vs = blockIdx.y;  // value set
  prep += vs * vs_prep;
  prepi += vs * vs_prepi;
  inv += vs * vs_inv;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_inv[];
  float* X = reinterpret_cast<float*>(smem_inv);  // X[r * 128 + j]
  __shared__ int iperm[128];
  const SolveDesc d = desc[blockIdx.x];
  const int bw = d.c1 - d.c0, ng = (bw + 3) / 4, bwp = 4 * ng;
  const int tid = threadIdx.x;
  const int npan = 8 * ng * (ng + 1);
  const float* lo = prep + 4 * (int64_t)d.aux;
  const float* up = lo + npan;
  const float* rinv = up + npan;
  const int32_t* pi = prepi + (int64_t)d.r0 * kPrepInts;
  float* ML = inv + d.inv;
  float* MU = ML + (int64_t)bwp * bwp;
  // Column j of each inverse is the solve of e_j: thread (j, h) = (tid %
  // 128, tid / 128) takes column j's rows of parity h, and all threads take
  // the eliminated column c in step, so that the panel entries are read as
  // broadcasts. Column j's entries receive the terms of c = j, j + 1, ...
  // (ML) or c = j, j - 1, ... (MU) in that order; the steps with c before j
  // (ML) or after it (MU) meet a zero entry of column j and change nothing.
  const int j = tid & 127, h = tid >> 7;
  const bool cj = j < bw;
  for (int i = tid; i < bw; i += kWG) iperm[pi[i]] = i;
  for (int i = tid; i < 128 * 128; i += kWG) X[i] = (i / 128 == i % 128) ? 1.0f : 0.0f;
  __syncthreads();
  for (int c = 0; c + 1 < bw; ++c) {
    if (cj && c >= j) {
      const float xc = X[c * 128 + j];
      const float* Pc = lo + lo_panel(c / 4, ng) + (c % 4) - 4 * 4 * (c / 4);
      for (int r = c + 1 + h; r < bw; r += 2) X[r * 128 + j] -= Pc[4 * r] * xc;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < bwp * bwp; idx += kWG) {
    const int i = idx / bwp, jj = idx % bwp;
    ML[idx] = (i < bw && jj < bw) ? X[i * 128 + iperm[jj]] : 0.0f;
  }
  __syncthreads();
  for (int i = tid; i < 128 * 128; i += kWG) X[i] = (i / 128 == i % 128) ? 1.0f : 0.0f;
  __syncthreads();
  for (int c = bw - 1; c >= 0; --c) {
    if (cj && c <= j) {
      const float xc = X[c * 128 + j] * rinv[c];
      if (h == 0) MU[(int64_t)c * bwp + j] = xc;  // row c of column j is final
      const float* Pc = up + up_panel(c / 4) + (c % 4);
      for (int r = h; r < c; r += 2) X[r * 128 + j] -= Pc[4 * r] * xc;
    }
    __syncthreads();
  }
  // the rest of MU: zeros below the diagonal and in the padding
  for (int idx = tid; idx < bwp * bwp; idx += kWG) {
    const int i = idx / bwp, jj = idx % bwp;
    if (!(i < bw && jj < bw && i <= jj)) MU[idx] = 0.0f;
  }
}

// Row i of a block's inverse, columns 64h .. 64h+63, as 16 four-vectors
// (zero outside the block). The row's entries outside columns [jlo, jhi)
// are zeros of the inverse: four-vectors wholly outside are not read.
__device__ __attribute__((always_inline)) inline void load_inv_row(
    const float* __restrict__ M, int bw, int bwp, int i, int h, int jlo, int jhi,
    f32x4 (&mv)[16]) {
  const f32x4* R = reinterpret_cast<const f32x4*>(M + (int64_t)i * bwp);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int j = 64 * h + 4 * k;
    mv[k] = (i < bw && j < bwp && j < jhi && j + 4 > jlo) ? R[j / 4]
                                                          : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  }
}

// v = M yb for a block of up to 128 entries: thread (i = tid % 128, h = tid /
// 128) sums its 64 columns in order, the two halves are added in order; yb
// (128 entries, zero past the block) in LDS, v to LDS.
__device__ __attribute__((always_inline)) inline void inv_apply(const f32x4 (&mv)[16],
                                                                const float* yb, float (*zp)[128],
                                                                float* v) {
  const int tid = threadIdx.x, i = tid & 127, h = tid >> 7;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int j = 64 * h + 4 * k;
    acc += mv[k].x * yb[j];
    acc += mv[k].y * yb[j + 1];
    acc += mv[k].z * yb[j + 2];
    acc += mv[k].w * yb[j + 3];
  }
  zp[h][i] = acc;
  __syncthreads();
  if (tid < 128) v[tid] = zp[0][tid] + zp[1][tid];
  __syncthreads();
}

// Forward block step: every workgroup forms z = ML y[x0:x1] of its block;
// the one marked write stores z as the block's solution entries (rhs only:
// the front's vector keeps y, which the step's other workgroups read); a
// tile workgroup then updates its 64 rows, y[x1 + r0 + lane] -= L[., x0:x1]
// z, with fwd_bgemv_body's operations.
__global__ void __launch_bounds__(kWG)
fwd_inv_kernel(DevSym s, const InvItem* __restrict__ items, const float* __restrict__ inv,
               const float* __restrict__ lnz, float* __restrict__ rhs, int64_t ldr, int nrhs,
               float* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  inv += vs * s.vs_inv;
  lnz += vs * s.vs_lnz;
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  __shared__ float yb[128];
  __shared__ float zp[2][128];
  __shared__ float zs[128];
  __shared__ float red[4][64];
  const int tid = threadIdx.x, lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const InvItem* const ip = items + blockIdx.x;
  const InvItem it = *ip;
  const int L = it.L, c0 = it.x0, bw = it.x1 - it.x0, bwp = 4 * ((bw + 3) / 4);
  // row i of ML = L^-1 P is zero from the end of its piece on (the
  // interchanges stay inside the pieces)
  int pend = bw;
#pragma unroll
  for (int q = 1; q < 8; ++q) {
    const int ps = ip->pst[q];
    if (ps > (tid & 127) && ps < pend) pend = ps;
  }
  f32x4 mv[16];
  load_inv_row(inv + it.inv, bw, bwp, tid & 127, tid >> 7, 0, pend, mv);
  const bool tile = it.r0 >= 0;
  const int r = it.x1 + it.r0 + lane;
  const int cb = bw * wave / 4, ce = bw * (wave + 1) / 4;
  // the tile's L entries, in flight with the inverse: block column c's lnz
  // offset from the item's piece table (a piece's columns follow each other
  // at its row count; the column index is uniform in the wave)
  float lv[32];
  {
    int q = 0;
    while (q < 7 && ip->pst[q + 1] <= cb) ++q;
    int64_t str = L - c0 - ip->pst[q];
    int64_t lcc = ip->pbs[q] + (int64_t)(cb - ip->pst[q]) * str;
    int nxt = q < 7 ? ip->pst[q + 1] : (1 << 30);
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const int c = cb + j;
      if (c == nxt) {
        ++q;
        str = L - c0 - ip->pst[q];
        lcc = ip->pbs[q];
        nxt = q < 7 ? ip->pst[q + 1] : (1 << 30);
      }
      lv[j] = (tile && r < L && c < ce) ? lnz[lcc + r] : 0.0f;
      lcc += str;
    }
  }
  float* const yf = vec + it.fvec * nrhs;
  // the first right-hand side's y
  if (tid < 128) yb[tid] = (tid < bw) ? yf[c0 + tid] : 0.0f;
  float yr = (tile && wave == 0 && r < L) ? yf[r] : 0.0f;
  __syncthreads();
  for (int q = 0; q < nrhs; ++q) {
    float* yg = yf + (int64_t)q * L;
    if (q > 0) {
      if (tid < 128) yb[tid] = (tid < bw) ? yg[c0 + tid] : 0.0f;
      yr = (tile && wave == 0 && r < L) ? yg[r] : 0.0f;
      __syncthreads();
    }
    inv_apply(mv, yb, zp, zs);
    if ((it.flags & 1) && tid < bw) rhs[(int64_t)q * ldr + it.fc0 + c0 + tid] = zs[tid];
    float acc = 0.0f;
    if (tile && r < L) {
#pragma unroll
      for (int j = 0; j < 32; ++j)
        if (cb + j < ce) acc += lv[j] * zs[cb + j];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (tile && wave == 0 && r < L)
      yg[r] = yr - (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
    __syncthreads();
  }
}

// Backward block step. Before it, the right-hand side holds y = z - t for
// the rows whose t is done. A row workgroup first forms x of the block
// after its row's block, x = MU y[x0:x1], then its row's t = U[c, rc1:L]
// x[rc1:L] (x[x0:x1] from LDS, of later blocks from the front's vector, the
// update rows as gathered) with bwd_bgemv_body's operations, and leaves y_c
// = z_c - t_c in the right-hand side. The workgroup marked write stores x of
// that block in the front's vector (read by later steps only). A finishing
// workgroup (one per front) forms x of the front's first block into the
// right-hand side and copies the other blocks' x there.
__global__ void __launch_bounds__(kWG)
bwd_inv_kernel(DevSym s, const InvItem* __restrict__ items, const float* __restrict__ inv,
               const float* __restrict__ unz, float* __restrict__ rhs, int64_t ldr, int nrhs,
               float* __restrict__ vec) {
  const int64_t vs = blockIdx.y;  // value set
  inv += vs * s.vs_inv;
  unz += vs * s.vs_unz;
  rhs += vs * ldr * nrhs;
  vec += vs * s.vs_vec * nrhs;
  constexpr int QB = 4;   // right-hand sides sharing one read of the row
  constexpr int CH = 16;  // columns per thread and pass
  __shared__ float yb[128];
  __shared__ float zp[2][128];
  __shared__ float xs[QB][128];
  __shared__ float red[QB][4];
  const int tid = threadIdx.x, lane = tid % 64, wave = wave_id();
  const InvItem it = items[blockIdx.x];
  const bool row = it.c >= 0, fin = (it.flags & 2) != 0;
  const int L = it.L;
  const int xlo = it.x0, xhi = it.x1, xbw = it.x1 - it.x0;  // x columns formed here
  const bool hasx = xbw > 0;
  const int bwp = 4 * ((xbw + 3) / 4);
  f32x4 mv[16];
  if (hasx)  // MU is upper triangular
    load_inv_row(inv + it.inv + (int64_t)bwp * bwp, xbw, bwp, tid & 127, tid >> 7, tid & 127,
                 xbw, mv);
  float* const yf = vec + it.fvec * nrhs;
  for (int q0 = 0; q0 < nrhs; q0 += QB) {
    const int nq = min(QB, nrhs - q0);
    if (hasx) {
      for (int qq = 0; qq < nq; ++qq) {
        const float* y = rhs + (int64_t)(q0 + qq) * ldr + it.fc0 + xlo;
        if (tid < 128) yb[tid] = (tid < xbw) ? y[tid] : 0.0f;
        __syncthreads();
        inv_apply(mv, yb, zp, xs[qq]);
      }
      if (it.flags & 3) {
        for (int qq = 0; qq < nq; ++qq)
          if (tid < xbw) {
            if (fin)
              rhs[(int64_t)(q0 + qq) * ldr + it.fc0 + xlo + tid] = xs[qq][tid];
            else
              yf[(int64_t)(q0 + qq) * L + xlo + tid] = xs[qq][tid];
          }
      }
      if (fin) {
        for (int qq = 0; qq < nq; ++qq)
          for (int j = xhi + tid; j < it.W; j += kWG)
            rhs[(int64_t)(q0 + qq) * ldr + it.fc0 + j] = yf[(int64_t)(q0 + qq) * L + j];
      }
    }
    if (row) {
      const float* u = unz + it.ucol;
      float acc[QB];
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) acc[qq] = 0.0f;
      for (int j0 = it.rc1 + tid; j0 < L; j0 += CH * kWG) {
        float uv[CH];
#pragma unroll
        for (int t = 0; t < CH; ++t) {
          const int j = j0 + kWG * t;
          uv[t] = (j < L) ? u[j] : 0.0f;
        }
#pragma unroll
        for (int qq = 0; qq < QB; ++qq) {
          if (qq < nq) {
            const float* yg = yf + (int64_t)(q0 + qq) * L;
            float xv[CH];
#pragma unroll
            for (int t = 0; t < CH; ++t) {
              const int j = j0 + kWG * t;
              xv[t] = (j >= L) ? 0.0f : (j < xhi && j >= xlo) ? xs[qq][j - xlo] : yg[j];
            }
#pragma unroll
            for (int t = 0; t < CH; ++t) acc[qq] += uv[t] * xv[t];
          }
        }
      }
#pragma unroll
      for (int qq = 0; qq < QB; ++qq) {
        if (qq < nq) {
          float a = acc[qq];
          for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
          if (lane == 0) red[qq][wave] = a;
        }
      }
      __syncthreads();
      if (tid < nq) {
        float* y = rhs + (int64_t)(q0 + tid) * ldr + it.fc0 + it.c;
        *y = *y - (((red[tid][0] + red[tid][1]) + red[tid][2]) + red[tid][3]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Permutation, residual and refinement helpers (double unless noted).
// pb[i] = b[perm[i]-1] (rhs = b[rperm], SpkSparseBase.jl:407) and
// x[j] = pb[invp[j]-1] (solution = rhs[rinvp], :413).
// ---------------------------------------------------------------------------
// value set blockIdx.y: in and out start ivs and ovs elements further
template <typename TI, typename TO>
__global__ void permute_kernel(const TI* __restrict__ in, int64_t ldi,
                               TO* __restrict__ out, int64_t ldo,
                               const int64_t* __restrict__ idx, int64_t n,
                               int nrhs, int64_t ivs, int64_t ovs) {
  in += (int64_t)blockIdx.y * ivs;
  out += (int64_t)blockIdx.y * ovs;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t src = idx[i] - 1;
  for (int q = 0; q < nrhs; ++q) out[i + q * ldo] = (TO)in[src + q * ldi];
}

// r = b - A x with A in CSR (original numbering): G consecutive lanes per
// row, lane g summing the row's entries g, g + G, ... in order, the G
// partial sums then added pairwise (G = 1: one row per thread, the entries in
// order). Value set blockIdx.y: its values nnz further, its rows of x, b and
// r n further (the sets' systems stacked as one block-diagonal system).
template <int G>
__global__ void __launch_bounds__(kWG)
residual_kernel(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
                const double* __restrict__ val, const double* __restrict__ x, int64_t ldx,
                const double* __restrict__ b, int64_t ldb, double* __restrict__ r, int64_t ldr,
                int64_t n, int nrhs, int64_t nnz) {
  const int64_t vs = blockIdx.y;
  val += vs * nnz;
  x += vs * n;
  b += vs * n;
  r += vs * n;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int g = (int)(threadIdx.x % G);
  if (i >= n) return;  // a row's G lanes leave together
  const int64_t k0 = rowptr[i], k1 = rowptr[i + 1];
  for (int q = 0; q < nrhs; ++q) {
    const double* xq = x + (int64_t)q * ldx;
    double acc = 0.0;
    for (int64_t k = k0 + g; k < k1; k += G) acc = fma(val[k], xq[colind[k]], acc);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
    if (g == 0) r[i + q * ldr] = b[i + q * ldb] - acc;
  }
}

__device__ inline double abs_nan_inf(double x) {
  const double a = fabs(x);
  return (a != a) ? __builtin_huge_val() : a;
}

// Row sums |A| for ||A||_inf (a NaN entry makes its row sum +inf); value set
// blockIdx.y as in residual_kernel.
__global__ void rowabs_kernel(const int64_t* __restrict__ rowptr,
                              const double* __restrict__ val,
                              double* __restrict__ out, int64_t n, int64_t nnz) {
  val += (int64_t)blockIdx.y * nnz;
  out += (int64_t)blockIdx.y * n;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += abs_nan_inf(val[k]);
  out[i] = acc;
}

// Max |v| over row chunks [clo[c], clo[c+1]) of each column, folded into the
// per-(segment, column) maxima out[cseg[c] + q * nseg] by an integer
// atomicMax on the bit patterns (a non-negative double orders as its bits,
// +inf above every finite value); out starts at +0.0. NaN entries count as
// +inf. One block per (chunk, column).
__global__ void segmax_chunk_kernel(const double* __restrict__ v, int64_t ld,
                                    const int64_t* __restrict__ clo,
                                    const int32_t* __restrict__ cseg, int nchunks, int nseg,
                                    unsigned long long* __restrict__ out) {
  __shared__ double red[kWG];
  const int c = blockIdx.x % nchunks, q = blockIdx.x / nchunks;
  const double* vq = v + (int64_t)q * ld;
  const int64_t lo = clo[c], hi = clo[c + 1];
  double m = 0.0;
  int64_t r = lo + threadIdx.x;
  for (; r + 7 * kWG < hi; r += 8 * kWG) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = abs_nan_inf(vq[r + u * kWG]);
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmax(m, a[u]);
  }
  for (; r < hi; r += kWG) m = fmax(m, abs_nan_inf(vq[r]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = kWG / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMax(out + cseg[c] + (int64_t)q * nseg, (unsigned long long)__double_as_longlong(red[0]));
}

// x += d (d in the factor precision, already in original numbering).
template <typename T>
__global__ void axpy_kernel(double* __restrict__ x, int64_t ldx,
                            const T* __restrict__ d, int64_t ldd, int64_t n,
                            int nrhs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nrhs; ++q) x[i + q * ldx] += (double)d[i + q * ldd];
}

// ---------------------------------------------------------------------------
// The 1-norm estimate of A^-1 (Hager's method, the reference's commented
// sparsespdainverseonenorm, SpkSparseSpdSolver.jl:346-431): its vector steps.
// ---------------------------------------------------------------------------
// mode 0: x_i = 1/n; mode 1: the unit vector e_j; mode 2: x_i = (-1)^i (1 +
// i/(n-1)) (0-based i: the reference's altsgn vector).
__global__ void condest_fill_kernel(double* __restrict__ x, int64_t n, int mode, int64_t j) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v;
  if (mode == 0)
    v = 1.0 / (double)n;
  else if (mode == 1)
    v = (i == j) ? 1.0 : 0.0;
  else
    v = ((i & 1) ? -1.0 : 1.0) * (1.0 + (double)i / (double)(n - 1));
  x[i] = v;
}

// Per-block partial results over [b chunk, (b+1) chunk): the sum of |x|
// (a fixed tree order), the largest |x| and its first index.
__global__ void condest_reduce_kernel(const double* __restrict__ x, int64_t n, int64_t chunk,
                                      double* __restrict__ psum, double* __restrict__ pmax,
                                      int64_t* __restrict__ pidx) {
  __shared__ double ss[kWG], sm[kWG];
  __shared__ int64_t si[kWG];
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(n, lo + chunk);
  double sum = 0.0, mx = -1.0;
  int64_t mi = lo;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kWG) {
    const double a = fabs(x[i]);
    sum += a;
    if (a > mx) {
      mx = a;
      mi = i;
    }
  }
  ss[threadIdx.x] = sum;
  sm[threadIdx.x] = mx;
  si[threadIdx.x] = mi;
  __syncthreads();
  for (int off = kWG / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      ss[threadIdx.x] += ss[threadIdx.x + off];
      const double m2 = sm[threadIdx.x + off];
      const int64_t i2 = si[threadIdx.x + off];
      if (m2 > sm[threadIdx.x] || (m2 == sm[threadIdx.x] && i2 < si[threadIdx.x])) {
        sm[threadIdx.x] = m2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum[blockIdx.x] = ss[0];
    pmax[blockIdx.x] = sm[0];
    pidx[blockIdx.x] = si[0];
  }
}

// Signs s_i = sign(1, x_i) (+1 for x_i >= 0): compare mode counts the
// entries whose sign differs from isgn; otherwise x and isgn become the
// signs.
__global__ void condest_sign_kernel(double* __restrict__ x, int32_t* __restrict__ isgn,
                                    int64_t n, int compare,
                                    unsigned int* __restrict__ changed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sg = (x[i] < 0.0) ? -1 : 1;
  if (compare) {
    if (sg != isgn[i]) atomicAdd(changed, 1u);
  } else {
    x[i] = (double)sg;
    isgn[i] = sg;
  }
}

// Factor conversions for download.
template <typename TI, typename TO>
__global__ void convert_kernel(const TI* __restrict__ in, TO* __restrict__ out,
                               int64_t len) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < len; i += stride) out[i] = (TO)in[i];
}

}  // namespace spk
