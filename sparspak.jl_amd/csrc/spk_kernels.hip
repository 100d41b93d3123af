// Kernels of the numeric phase. See spk_kernels.h for the storage model.
#include "spk_kernels.h"

namespace spk {

__device__ inline int64_t ld_i64(const int64_t* p, int64_t i) { return p[i]; }

// Column range of piece p (0-based supernode): first column (0-based) and
// width.
__device__ inline void piece_cols(const DevSym& s, int64_t p, int64_t& c0,
                                  int64_t& w) {
  c0 = s.xsuper[p] - 1;
  w = s.xsuper[p + 1] - s.xsuper[p];
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

template <typename T>
__global__ void scatter_a_kernel(const double* __restrict__ nzval,
                                 const int64_t* __restrict__ amap,
                                 int64_t nnz, T* __restrict__ lnz,
                                 T* __restrict__ unz) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int64_t d = amap[k];
  T v = (T)nzval[k];
  if (d >= 0)
    lnz[d] += v;
  else if (d <= -2)
    unz[-d - 2] += v;
}

// SPD input: every stored lower entry is mirrored into the U slot of its
// transpose so that the LU front kernels see the full symmetric matrix.
template <typename T>
__global__ void scatter_a2_kernel(const double* __restrict__ nzval,
                                  const int64_t* __restrict__ amap2,
                                  int64_t nnz, T* __restrict__ lnz,
                                  T* __restrict__ unz) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int64_t d = amap2[k];
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
template <typename T>
__device__ void lds_partial_lu(T* F, int L, int cq, int nq, int pivot,
                               int* piv_out, int* zero_out) {
  __shared__ int s_kp;
  __shared__ T s_inv;
  const int tid = threadIdx.x;
  for (int k = 0; k < nq; ++k) {
    const int col = cq + k;
    if (tid == 0) {
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
      s_kp = kp;
      T pv = F[kp + col * L];
      if (pv == T(0)) *zero_out = 1;
      s_inv = T(1) / pv;
      piv_out[k] = kp - cq;
    }
    __syncthreads();
    const int kp = s_kp;
    const bool nonzero = (F[kp + col * L] != T(0));
    if (kp != col && nonzero) {
      for (int c = cq + tid; c < L; c += kWG) {
        T t = F[col + c * L];
        F[col + c * L] = F[kp + c * L];
        F[kp + c * L] = t;
      }
    }
    __syncthreads();
    const T inv = s_inv;
    // Rows inside the diagonal block are scaled only for a nonzero pivot
    // (ggetrf!); rows below it always (gtrsm! multiplies by 1/U(j,j)).
    for (int r = col + 1 + tid; r < L; r += kWG)
      if (r >= cq + nq || nonzero) F[r + col * L] *= inv;
    __syncthreads();
    const int m = L - col - 1;
    const int cnt = m * m;
    for (int idx = tid; idx < cnt; idx += kWG) {
      int r = col + 1 + idx % m;
      int c = col + 1 + idx / m;
      F[r + c * L] -= F[r + col * L] * F[col + c * L];
    }
    __syncthreads();
  }
}

// One workgroup per small front (L <= the LDS limit of the launch):
// gather the pieces' panels and the children's update matrices into LDS,
// factor, write the panels back to lnz/unz/ipvt and the update matrix to the
// arena.
template <typename T>
__global__ void __launch_bounds__(kWG)
small_front_kernel(DevSym s, const int32_t* __restrict__ list,
                   T* __restrict__ lnz, T* __restrict__ unz,
                   int64_t* __restrict__ ipvt, T* __restrict__ arena,
                   unsigned long long* __restrict__ first_fail) {
  extern __shared__ unsigned char smem_raw[];
  T* F = reinterpret_cast<T*>(smem_raw);
  __shared__ int s_piv[kMaxPiece];
  __shared__ int s_zero;
  const int tid = threadIdx.x;
  const int f = list[blockIdx.x];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  for (int i = tid; i < L * L; i += kWG) F[i] = T(0);
  __syncthreads();
  // pieces' panels
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    const int jl = L - cq;
    const T* lp = lnz + (s.xlnz[pc0] - 1);
    for (int idx = tid; idx < jl * nq; idx += kWG) {
      int r = idx % jl, c = idx / jl;
      if (!s.sym || r >= c) F[(cq + r) + (cq + c) * L] = lp[idx];
    }
    if (!s.sym) {
      const int ul = jl - (int)nq;
      const T* up = unz + (s.xunz[pc0] - 1);
      for (int idx = tid; idx < ul * nq; idx += kWG) {
        int r = idx % ul, c = idx / ul;
        F[(cq + c) + (cq + (int)nq + r) * L] = up[idx];
      }
    }
  }
  __syncthreads();
  if (s.sym) {  // upper rows of the own columns from the lower triangle
    for (int idx = tid; idx < W * L; idx += kWG) {
      int r = idx % W, j = idx / W;
      if (j > r) F[r + j * L] = F[j + r * L];
    }
    __syncthreads();
  }
  // children's update matrices (one child at a time: rows of different
  // children may coincide)
  for (int ci = s.fchp[f]; ci < s.fchp[f + 1]; ++ci) {
    const int c = s.fch[ci];
    const int Rc = s.fL[c] - s.fW[c];
    const int32_t* rel = s.relmap + s.frel[c];
    const T* cb = arena + s.fcb[c];
    const int ld = s.fcbld[c];
    for (int idx = tid; idx < Rc * Rc; idx += kWG) {
      int tp = idx % Rc, t = idx / Rc;
      F[rel[tp] + rel[t] * L] += cb[tp + (int64_t)t * ld];
    }
    __syncthreads();
  }
  // factor piece by piece
  if (tid == 0) s_zero = 0;
  __syncthreads();
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    lds_partial_lu<T>(F, L, cq, (int)nq, s.pivot, s_piv, &s_zero);
    for (int k = tid; k < nq; k += kWG) ipvt[pc0 + k] = s_piv[k] + 1;
    if (tid == 0 && s_zero) {
      atomicMin(first_fail, (unsigned long long)(p + 1));
      s_zero = 0;
    }
    __syncthreads();
  }
  // write back panels
  for (int p = sp0; p < sp0 + np; ++p) {
    int64_t pc0, nq;
    piece_cols(s, p, pc0, nq);
    const int cq = (int)(pc0 - fc0);
    const int jl = L - cq;
    T* lp = lnz + (s.xlnz[pc0] - 1);
    for (int idx = tid; idx < jl * nq; idx += kWG) {
      int r = idx % jl, c = idx / jl;
      lp[idx] = F[(cq + r) + (cq + c) * L];
    }
    const int ul = jl - (int)nq;
    T* up = unz + (s.xunz[pc0] - 1);
    for (int idx = tid; idx < ul * nq; idx += kWG) {
      int r = idx % ul, c = idx / ul;
      up[idx] = F[(cq + c) + (cq + (int)nq + r) * L];
    }
  }
  // update matrix
  const int R = L - W;
  T* cb = arena + s.fcb[f];
  for (int idx = tid; idx < R * R; idx += kWG) {
    int tp = idx % R, t = idx / R;
    cb[idx] = F[(W + tp) + (W + t) * L];
  }
}

// ---------------------------------------------------------------------------
// Large fronts: the L x L work matrix lives in the arena.
// ---------------------------------------------------------------------------

// One workgroup per (front, column j): load column j from lnz/unz (zero in
// the update block), then add the children's update-matrix columns that map
// to j, one after the other.
template <typename T>
__global__ void __launch_bounds__(kWG)
large_init_kernel(DevSym s, const int32_t* __restrict__ it_front,
                  const int32_t* __restrict__ it_col,
                  const int64_t* __restrict__ it_ea,
                  const int32_t* __restrict__ ea_child,
                  const int32_t* __restrict__ ea_t, const T* __restrict__ lnz,
                  const T* __restrict__ unz, T* __restrict__ arena) {
  const int item = blockIdx.x;
  const int f = it_front[item];
  const int j = it_col[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  T* F = arena + s.foff[f] + (int64_t)j * L;
  for (int r = threadIdx.x; r < L; r += kWG) {
    bool is_u = false;
    if (s.sym) {  // lower triangle only: U entry (r, j) is L entry (j, r)
      int rr = max(r, j), cc = min(r, j);
      int64_t o = (cc < W) ? slot_of(s, fc0, W, rr, cc, is_u) : -1;
      F[r] = (o < 0) ? T(0) : lnz[o];
    } else {
      int64_t o = slot_of(s, fc0, W, r, j, is_u);
      F[r] = (o < 0) ? T(0) : (is_u ? unz[o] : lnz[o]);
    }
  }
  __syncthreads();
  for (int64_t e = it_ea[item]; e < it_ea[item + 1]; ++e) {
    const int c = ea_child[e];
    const int t = ea_t[e];
    const int Rc = s.fL[c] - s.fW[c];
    const int32_t* rel = s.relmap + s.frel[c];
    const T* col = arena + s.fcb[c] + (int64_t)t * s.fcbld[c];
    for (int tp = threadIdx.x; tp < Rc; tp += kWG) F[rel[tp]] += col[tp];
    __syncthreads();
  }
}

// Inverse of large_init_kernel's first half: panel entries back to lnz/unz.
template <typename T>
__global__ void __launch_bounds__(kWG)
large_writeback_kernel(DevSym s, const int32_t* __restrict__ it_front,
                       const int32_t* __restrict__ it_col,
                       T* __restrict__ lnz, T* __restrict__ unz,
                       const T* __restrict__ arena) {
  const int item = blockIdx.x;
  const int f = it_front[item];
  const int j = it_col[item];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const T* F = arena + s.foff[f] + (int64_t)j * L;
  const int rend = (j < W) ? L : W;
  for (int r = threadIdx.x; r < rend; r += kWG) {
    bool is_u = false;
    int64_t o = slot_of(s, fc0, W, r, j, is_u);
    if (o < 0) continue;
    if (is_u)
      unz[o] = F[r];
    else
      lnz[o] = F[r];
  }
}

__device__ inline int find_desc(const StepDesc* d, int nd, int b) {
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

// Diagonal block of piece `piece` of a large front, factored redundantly by
// every workgroup of the panel step (nq <= MAXW). D is nq x nq in LDS with
// leading dimension MAXW.
template <typename T, int MAXW>
__device__ void lds_getrf_block(T* D, int nq, int pivot, int* piv, int* zero) {
  __shared__ int s_kp;
  __shared__ T s_inv;
  const int tid = threadIdx.x;
  const int ld = MAXW;
  for (int k = 0; k < nq; ++k) {
    if (tid == 0) {
      int kp = k;
      if (pivot) {
        T amax = fabs(D[k + k * ld]);
        for (int i = k + 1; i < nq; ++i) {
          T a = fabs(D[i + k * ld]);
          if (a > amax) {
            amax = a;
            kp = i;
          }
        }
      }
      s_kp = kp;
      T pv = D[kp + k * ld];
      if (pv == T(0)) *zero = 1;
      s_inv = T(1) / pv;
      piv[k] = kp;
    }
    __syncthreads();
    const int kp = s_kp;
    const bool nonzero = D[kp + k * ld] != T(0);
    if (kp != k && nonzero)
      for (int c = tid; c < nq; c += kWG) {
        T t = D[k + c * ld];
        D[k + c * ld] = D[kp + c * ld];
        D[kp + c * ld] = t;
      }
    __syncthreads();
    if (nonzero)
      for (int r = k + 1 + tid; r < nq; r += kWG) D[r + k * ld] *= s_inv;
    __syncthreads();
    const int m = nq - k - 1;
    for (int idx = tid; idx < m * m; idx += kWG) {
      int r = k + 1 + idx % m, c = k + 1 + idx / m;
      D[r + c * ld] -= D[r + k * ld] * D[k + c * ld];
    }
    __syncthreads();
  }
}

// Panel step of piece `piece` for every large front at this step: factor the
// diagonal block (each workgroup redundantly, in LDS), then L21 = A21 U11^-1
// (row blocks) and U12 = L11^-1 P A12 (column blocks). The factored block
// goes to a staging slot and is copied into the front by the update step,
// so that no workgroup here reads a half-written block.
template <typename T, int MAXW>
__global__ void __launch_bounds__(kWG)
large_panel_kernel(DevSym s, const StepDesc* __restrict__ desc, int ndesc,
                   T* __restrict__ arena, T* __restrict__ dstage,
                   int64_t* __restrict__ ipvt,
                   unsigned long long* __restrict__ first_fail) {
  __shared__ T D[MAXW * MAXW];
  __shared__ T X[MAXW * kUcols];
  __shared__ int piv[MAXW];
  __shared__ int s_zero;
  const int tid = threadIdx.x;
  const int di = find_desc(desc, ndesc, blockIdx.x);
  const StepDesc d = desc[di];
  const int b = blockIdx.x - d.wg_begin;
  const int f = d.front;
  const int L = s.fL[f];
  const int64_t fc0 = s.fc0[f];
  int64_t pc0, nq64;
  piece_cols(s, d.piece, pc0, nq64);
  const int nq = (int)nq64;
  const int cq = (int)(pc0 - fc0);
  const int m = L - cq - nq;  // rows below / columns right of the block
  T* F = arena + s.foff[f];
  if (tid == 0) s_zero = 0;
  for (int idx = tid; idx < nq * nq; idx += kWG) {
    int r = idx % nq, c = idx / nq;
    D[r + c * MAXW] = F[(cq + r) + (int64_t)(cq + c) * L];
  }
  __syncthreads();
  lds_getrf_block<T, MAXW>(D, nq, s.pivot, piv, &s_zero);
  if (b == 0) {
    T* st = dstage + (int64_t)d.slot * MAXW * MAXW;
    for (int idx = tid; idx < nq * nq; idx += kWG) {
      int r = idx % nq, c = idx / nq;
      st[r + c * nq] = D[r + c * MAXW];
    }
    for (int k = tid; k < nq; k += kWG) ipvt[pc0 + k] = piv[k] + 1;
    if (tid == 0 && s_zero) atomicMin(first_fail, (unsigned long long)(d.piece + 1));
  }
  if (m <= 0) return;
  const int nl = d.nl;
  if (b < nl) {
    // L21 rows: x U11 = a, column by column (gtrsm! 'r','u','n','n' order);
    // the solved entries are read back from the row itself.
    const int r = cq + nq + b * kWG + tid;
    if (r < L) {
      T* row = F + r + (int64_t)cq * L;
      for (int j = 0; j < nq; ++j) {
        T v = row[(int64_t)j * L];
        for (int k = 0; k < j; ++k) v -= D[k + j * MAXW] * row[(int64_t)k * L];
        row[(int64_t)j * L] = v * (T(1) / D[j + j * MAXW]);
      }
    }
  } else {
    // U12 columns: apply the block's interchanges, then unit-lower solve.
    const int c0 = cq + nq + (b - nl) * kUcols;
    const int nc = min(kUcols, L - c0);
    for (int idx = tid; idx < nq * nc; idx += kWG) {
      int r = idx % nq, c = idx / nq;
      X[r + c * MAXW] = F[(cq + r) + (int64_t)(c0 + c) * L];
    }
    __syncthreads();
    if (tid < nc) {
      T* xc = X + tid * MAXW;
      for (int k = 0; k < nq; ++k) {
        int kp = piv[k];
        if (kp != k) {
          T t = xc[k];
          xc[k] = xc[kp];
          xc[kp] = t;
        }
      }
      for (int k = 0; k < nq; ++k) {
        T xk = xc[k];
        for (int i = k + 1; i < nq; ++i) xc[i] -= D[i + k * MAXW] * xk;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < nq * nc; idx += kWG) {
      int r = idx % nq, c = idx / nq;
      F[(cq + r) + (int64_t)(c0 + c) * L] = X[r + c * MAXW];
    }
  }
}

// Trailing update A22 -= L21 U12 on kTile x kTile tiles, plus one extra
// workgroup per front that copies the staged diagonal block into place.
template <typename T, int MAXW>
__global__ void __launch_bounds__(kWG)
large_update_kernel(DevSym s, const StepDesc* __restrict__ desc, int ndesc,
                    T* __restrict__ arena, const T* __restrict__ dstage) {
  constexpr int KC = 32;  // k chunk staged in LDS
  __shared__ T As[KC][kTile];
  __shared__ T Bs[KC][kTile];
  const int tid = threadIdx.x;
  const int di = find_desc(desc, ndesc, blockIdx.x);
  const StepDesc d = desc[di];
  const int b = blockIdx.x - d.wg_begin;
  const int f = d.front;
  const int L = s.fL[f];
  const int64_t fc0 = s.fc0[f];
  int64_t pc0, nq64;
  piece_cols(s, d.piece, pc0, nq64);
  const int nq = (int)nq64;
  const int cq = (int)(pc0 - fc0);
  T* F = arena + s.foff[f];
  const int nt = d.nl;  // tiles per dimension
  if (b == nt * nt) {
    const T* st = dstage + (int64_t)d.slot * MAXW * MAXW;
    for (int idx = tid; idx < nq * nq; idx += kWG) {
      int r = idx % nq, c = idx / nq;
      F[(cq + r) + (int64_t)(cq + c) * L] = st[idx];
    }
    return;
  }
  const int base = cq + nq;
  const int ti = b % nt, tj = b / nt;
  const int r0 = base + ti * kTile, c0 = base + tj * kTile;
  const int ri = (tid % 16) * 4, cj = (tid / 16) * 4;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) acc[a][bb] = T(0);
  for (int k0 = 0; k0 < nq; k0 += KC) {
    const int kc = min(KC, nq - k0);
    for (int idx = tid; idx < kc * kTile; idx += kWG) {
      int i = idx % kTile, k = idx / kTile;
      int r = r0 + i;
      As[k][i] = (r < L) ? F[r + (int64_t)(cq + k0 + k) * L] : T(0);
    }
    for (int idx = tid; idx < kc * kTile; idx += kWG) {
      int k = idx % kc, j = idx / kc;
      int c = c0 + j;
      Bs[k][j] = (c < L) ? F[(cq + k0 + k) + (int64_t)c * L] : T(0);
    }
    __syncthreads();
    for (int k = 0; k < kc; ++k) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[k][ri + a];
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) bv[bb] = Bs[k][cj + bb];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) acc[a][bb] += av[a] * bv[bb];
    }
    __syncthreads();
  }
#pragma unroll
  for (int bb = 0; bb < 4; ++bb) {
    int c = c0 + cj + bb;
    if (c >= L) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      int r = r0 + ri + a;
      if (r < L) F[r + (int64_t)c * L] -= acc[a][bb];
    }
  }
}

// ---------------------------------------------------------------------------
// Triangular solves, one workgroup per front, one launch per level.
// Forward (_lulsolve!, SpkLUFactor.jl:269-323): the front's vector y (length
// L, per right-hand side) starts as the permuted rhs on its own columns and
// zero elsewhere, receives the children's update vectors through the
// relative-index maps (the multifrontal form of the reference's scatter
// rhs[lindx] += temp), and is solved piece by piece: interchange, unit-lower
// block solve, y[below] -= L21 y[block]. Backward (_luusolve!, :325-377):
// gather the solution on the front's index list, then per piece in reverse
// order y[block] -= U12 y[right] and the upper block solve.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kWG)
fwd_solve_kernel(DevSym s, const int32_t* __restrict__ list,
                 const T* __restrict__ lnz, const int64_t* __restrict__ ipvt,
                 T* __restrict__ rhs, int64_t ldr, int nrhs,
                 T* __restrict__ vec) {
  const int tid = threadIdx.x;
  const int f = list[blockIdx.x];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  for (int q = 0; q < nrhs; ++q) {
    T* y = vec + s.fvec[f] * nrhs + (int64_t)q * L;
    T* x = rhs + (int64_t)q * ldr;
    for (int r = tid; r < L; r += kWG) y[r] = (r < W) ? x[fc0 + r] : T(0);
    __syncthreads();
    for (int ci = s.fchp[f]; ci < s.fchp[f + 1]; ++ci) {
      const int c = s.fch[ci];
      const int Lc = s.fL[c], Wc = s.fW[c];
      const int32_t* rel = s.relmap + s.frel[c];
      const T* uc = vec + s.fvec[c] * nrhs + (int64_t)q * Lc + Wc;
      for (int t = tid; t < Lc - Wc; t += kWG) y[rel[t]] += uc[t];
      __syncthreads();
    }
    for (int p = sp0; p < sp0 + np; ++p) {
      int64_t pc0, nq;
      piece_cols(s, p, pc0, nq);
      const int cq = (int)(pc0 - fc0);
      const int jl = L - cq;
      const T* lp = lnz + (s.xlnz[pc0] - 1);
      if (tid == 0) {
        for (int k = 0; k < nq; ++k) {
          int ip = (int)ipvt[pc0 + k] - 1;
          if (ip != k) {
            T t = y[cq + k];
            y[cq + k] = y[cq + ip];
            y[cq + ip] = t;
          }
        }
        for (int k = 0; k < nq; ++k) {
          T yk = y[cq + k];
          if (yk != T(0))
            for (int i = k + 1; i < nq; ++i) y[cq + i] -= yk * lp[i + (int64_t)k * jl];
        }
      }
      __syncthreads();
      for (int r = (int)nq + tid; r < jl; r += kWG) {
        T acc = T(0);
        for (int k = 0; k < nq; ++k) acc += lp[r + (int64_t)k * jl] * y[cq + k];
        y[cq + r] -= acc;
      }
      __syncthreads();
    }
    for (int r = tid; r < W; r += kWG) x[fc0 + r] = y[r];
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kWG)
bwd_solve_kernel(DevSym s, const int32_t* __restrict__ list,
                 const T* __restrict__ lnz, const T* __restrict__ unz,
                 T* __restrict__ rhs, int64_t ldr, int nrhs,
                 T* __restrict__ vec) {
  __shared__ T red[kWG];
  const int tid = threadIdx.x;
  const int f = list[blockIdx.x];
  const int L = s.fL[f], W = s.fW[f];
  const int64_t fc0 = s.fc0[f];
  const int sp0 = s.fsp0[f], np = s.fnp[f];
  const int32_t* I = s.lindx + (s.xlindx[sp0] - 1);
  for (int q = 0; q < nrhs; ++q) {
    T* y = vec + s.fvec[f] * nrhs + (int64_t)q * L;
    T* x = rhs + (int64_t)q * ldr;
    for (int r = tid; r < L; r += kWG) y[r] = x[I[r] - 1];
    __syncthreads();
    for (int p = sp0 + np - 1; p >= sp0; --p) {
      int64_t pc0, nq;
      piece_cols(s, p, pc0, nq);
      const int cq = (int)(pc0 - fc0);
      const int jl = L - cq;
      const int ul = jl - (int)nq;
      const T* lp = lnz + (s.xlnz[pc0] - 1);
      const T* up = unz + (s.xunz[pc0] - 1);
      // y[cq+c] -= sum_r U12t[r,c] y[cq+nq+r]: one wave per column c.
      const int wave = tid / 64, lane = tid % 64;
      for (int c = wave; c < nq; c += kWG / 64) {
        T acc = T(0);
        for (int r = lane; r < ul; r += 64) acc += up[r + (int64_t)c * ul] * y[cq + nq + r];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) red[c] = acc;
      }
      __syncthreads();
      if (tid == 0) {
        for (int c = 0; c < nq; ++c) y[cq + c] -= red[c];
        for (int k = (int)nq - 1; k >= 0; --k) {
          T yk = y[cq + k];
          if (yk != T(0)) {
            yk /= lp[k + (int64_t)k * jl];
            y[cq + k] = yk;
            for (int i = 0; i < k; ++i) y[cq + i] -= yk * lp[i + (int64_t)k * jl];
          }
        }
      }
      __syncthreads();
    }
    for (int r = tid; r < W; r += kWG) x[fc0 + r] = y[r];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Permutation, residual and refinement helpers (double unless noted).
// pb[i] = b[perm[i]-1] (rhs = b[rperm], SpkSparseBase.jl:407) and
// x[j] = pb[invp[j]-1] (solution = rhs[rinvp], :413).
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ void permute_kernel(const TI* __restrict__ in, int64_t ldi,
                               TO* __restrict__ out, int64_t ldo,
                               const int64_t* __restrict__ idx, int64_t n,
                               int nrhs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t src = idx[i] - 1;
  for (int q = 0; q < nrhs; ++q) out[i + q * ldo] = (TO)in[src + q * ldi];
}

// r = b - A x with A in CSR (original numbering), one row per thread.
__global__ void residual_kernel(const int64_t* __restrict__ rowptr,
                                const int32_t* __restrict__ colind,
                                const double* __restrict__ val,
                                const double* __restrict__ x, int64_t ldx,
                                const double* __restrict__ b, int64_t ldb,
                                double* __restrict__ r, int64_t ldr, int64_t n,
                                int nrhs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nrhs; ++q) {
    double acc = 0.0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      acc = fma(val[k], x[colind[k] + q * ldx], acc);
    r[i + q * ldr] = b[i + q * ldb] - acc;
  }
}

// Row sums |A| for ||A||_inf.
__global__ void rowabs_kernel(const int64_t* __restrict__ rowptr,
                              const double* __restrict__ val,
                              double* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) acc += fabs(val[k]);
  out[i] = acc;
}

// Column-wise max |v| over n rows of nrhs columns; out[q] (one block per q).
__global__ void colmaxabs_kernel(const double* __restrict__ v, int64_t ld,
                                 int64_t n, double* __restrict__ out) {
  __shared__ double red[kWG];
  const int q = blockIdx.x;
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kWG) m = fmax(m, fabs(v[i + q * ld]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = kWG / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = red[0];
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

// Factor conversions for download.
template <typename TI, typename TO>
__global__ void convert_kernel(const TI* __restrict__ in, TO* __restrict__ out,
                               int64_t len) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < len; i += stride) out[i] = (TO)in[i];
}

}  // namespace spk
