// ORACLE — test infrastructure only. CPU restatement of the numeric phase
// of Sparspak.jl (v0.3.13), used by tests/, by the package smoke check and by
// bench.py's cpu_baseline leg as the checker. Nothing in the product links or
// calls this library.
//
// What it restates (1-based index values, caller-owned arrays, in-place):
//   orc_inmatrix_csc_*   SparseCSCInterface.jl:101-169  (_inmatrix!, CSC)
//   orc_lufactor_*       SpkLUFactor.jl:60-255          (_lufactor!)
//   orc_lulsolve_*       SpkLUFactor.jl:269-323         (_lulsolve!)
//   orc_luusolve_*       SpkLUFactor.jl:325-377         (_luusolve!)
//   orc_inmatrix_spd_csc_* SpkSparseSpdBase.jl:234-311  (_inmatrix!, SPD,
//                        lower triangle only; CSC instead of Problem input)
//   orc_ldltfactor_ref_* SpkLDLtFactor.jl:58-246 with the reference's own
//                        _pchole!/_matrixdiagmm! (defective for width >= 2;
//                        kept to document that defect)
//   orc_ldltfactor_*     the same left-looking scheme with a textbook
//                        diagonal-block LDL^T and the leading dimension of
//                        _matrixdiagmm! corrected (the project's SPD oracle)
//   orc_ldltsolve_*      SpkLDLtFactor.jl:266-293       (_ldltsolve!)
// Dense kernels follow the loop order of GenericBlasLapackFragments.jl
// (ggemm! :108-212, ggetrf! :56-100, gtrsm! :264-480, ggemv! :219-255,
// glaswp! :486-495) and the index kernels of SpkSpdMMOps.jl:41-175, so the
// floating-point result is deterministic. getrf reports a zero pivot the way
// LAPACK's ?getrf does (info = first zero pivot), which is what the
// reference sees for Float64/Float32 (SpkSpdMMOps.jl:253-268).
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <vector>

namespace {
using i64 = int64_t;

template <typename T> inline T* B(T* p) { return p - 1; }

// ---- dense fragments (1-based, column-major, leading dimension ld) --------
#define IDX(ld, i, j) (((j) - 1) * (ld) + (i))

// ggemm!('n','t'): C = alpha*A*B^T + beta*C   (GenericBlasLapackFragments.jl:177-194)
template <typename T>
void gemm_nt(i64 m, i64 n, i64 k, T alpha, const T* a, i64 lda, const T* b,
             i64 ldb, T beta, T* c, i64 ldc) {
  const T* A = B(a);
  const T* Bm = B(b);
  T* C = B(c);
  if (m == 0 || n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
  if (alpha == T(0)) {
    if (beta == T(0))
      for (i64 i = 1; i <= n * m; ++i) C[i] = T(0);
    else
      for (i64 i = 1; i <= n * m; ++i) C[i] *= beta * C[i];
    return;
  }
  for (i64 j = 1; j <= n; ++j) {
    if (beta == T(0)) {
      for (i64 i = 1; i <= m; ++i) C[IDX(ldc, i, j)] = T(0);
    } else if (beta != T(1)) {
      for (i64 i = 1; i <= m; ++i) C[IDX(ldc, i, j)] = beta * C[IDX(ldc, i, j)];
    }
    for (i64 l = 1; l <= k; ++l) {
      T temp = alpha * Bm[IDX(ldb, j, l)];
      for (i64 i = 1; i <= m; ++i) C[IDX(ldc, i, j)] += temp * A[IDX(lda, i, l)];
    }
  }
}

// ggetrf! (GenericBlasLapackFragments.jl:56-100); returns the LAPACK info.
template <typename T>
i64 getrf(i64 m, i64 n, T* a, i64 lda, i64* ipiv0) {
  T* A = B(a);
  i64* ipiv = B(ipiv0);
  i64 info = 0;
  i64 minmn = m < n ? m : n;
  for (i64 k = 1; k <= minmn; ++k) {
    i64 kp = k;
    if (k < m) {
      T amax = std::fabs(A[IDX(lda, k, k)]);
      for (i64 i = k + 1; i <= m; ++i) {
        T absi = std::fabs(A[IDX(lda, i, k)]);
        if (absi > amax) {
          kp = i;
          amax = absi;
        }
      }
    }
    ipiv[k] = kp;
    if (A[IDX(lda, kp, k)] != T(0)) {
      if (k != kp)
        for (i64 i = 1; i <= n; ++i) {
          T tmp = A[IDX(lda, k, i)];
          A[IDX(lda, k, i)] = A[IDX(lda, kp, i)];
          A[IDX(lda, kp, i)] = tmp;
        }
      T akkinv = T(1) / A[IDX(lda, k, k)];
      for (i64 i = k + 1; i <= m; ++i) A[IDX(lda, i, k)] *= akkinv;
    } else if (info == 0) {
      info = k;
    }
    for (i64 j = k + 1; j <= n; ++j)
      for (i64 i = k + 1; i <= m; ++i)
        A[IDX(lda, i, j)] -= A[IDX(lda, i, k)] * A[IDX(lda, k, j)];
  }
  return info;
}

// gtrsm! variants used by the LU and LDL^T paths (alpha = 1 everywhere).
// side='r', uplo='u', trans='n', diag='n': B := B*inv(A)   (:376-395)
template <typename T>
void trsm_run(i64 m, i64 n, const T* a, i64 lda, T* b, i64 ldb) {
  const T* A = B(a);
  T* Bm = B(b);
  if (m == 0 || n == 0) return;
  for (i64 j = 1; j <= n; ++j) {
    for (i64 k = 1; k <= j - 1; ++k)
      if (A[IDX(lda, k, j)] != T(0))
        for (i64 i = 1; i <= m; ++i)
          Bm[IDX(ldb, i, j)] -= A[IDX(lda, k, j)] * Bm[IDX(ldb, i, k)];
    T temp = T(1) / A[IDX(lda, j, j)];
    for (i64 i = 1; i <= m; ++i) Bm[IDX(ldb, i, j)] *= temp;
  }
}
// side='r', uplo='l', trans='t', diag='u': B := B*inv(A^T)   (:459-477)
template <typename T>
void trsm_rltu(i64 m, i64 n, const T* a, i64 lda, T* b, i64 ldb) {
  const T* A = B(a);
  T* Bm = B(b);
  if (m == 0 || n == 0) return;
  for (i64 k = 1; k <= n; ++k) {
    for (i64 j = k + 1; j <= n; ++j)
      if (A[IDX(lda, j, k)] != T(0)) {
        T temp = A[IDX(lda, j, k)];
        for (i64 i = 1; i <= m; ++i) Bm[IDX(ldb, i, j)] -= temp * Bm[IDX(ldb, i, k)];
      }
  }
}
// side='l', uplo='l', trans='n', diag='u': B := inv(A)*B    (:330-344)
template <typename T>
void trsm_llnu(i64 m, i64 n, const T* a, i64 lda, T* b, i64 ldb) {
  const T* A = B(a);
  T* Bm = B(b);
  if (m == 0 || n == 0) return;
  for (i64 j = 1; j <= n; ++j)
    for (i64 k = 1; k <= m; ++k)
      if (Bm[IDX(ldb, k, j)] != T(0))
        for (i64 i = k + 1; i <= m; ++i)
          Bm[IDX(ldb, i, j)] -= Bm[IDX(ldb, k, j)] * A[IDX(lda, i, k)];
}
// side='l', uplo='u', trans='n', diag='n': B := inv(A)*B    (:312-329)
template <typename T>
void trsm_lunn(i64 m, i64 n, const T* a, i64 lda, T* b, i64 ldb) {
  const T* A = B(a);
  T* Bm = B(b);
  if (m == 0 || n == 0) return;
  for (i64 j = 1; j <= n; ++j)
    for (i64 k = m; k >= 1; --k)
      if (Bm[IDX(ldb, k, j)] != T(0)) {
        Bm[IDX(ldb, k, j)] /= A[IDX(lda, k, k)];
        for (i64 i = 1; i <= k - 1; ++i)
          Bm[IDX(ldb, i, j)] -= Bm[IDX(ldb, k, j)] * A[IDX(lda, i, k)];
      }
}

// ggemv! (GenericBlasLapackFragments.jl:219-255)
template <typename T>
void gemv_n(i64 m, i64 n, T alpha, const T* a, i64 lda, const T* x, T beta,
            T* y) {
  if (m == 0 || n == 0) return;
  for (i64 i = 0; i < m; ++i) y[i] *= beta;
  i64 ii0 = 0;
  for (i64 j = 0; j < n; ++j) {
    i64 ii = ii0;
    T alphax = alpha * x[j];
    for (i64 i = 0; i < m; ++i) {
      y[i] += alphax * a[ii];
      ii += 1;
    }
    ii0 += lda;
  }
}
template <typename T>
void gemv_t(i64 m, i64 n, T alpha, const T* a, i64 lda, const T* x, T beta,
            T* y) {
  if (m == 0 || n == 0) return;
  i64 ii0 = 0;
  for (i64 j = 0; j < n; ++j) {
    y[j] *= beta;
    T temp = T(0);
    i64 ii = ii0;
    for (i64 i = 0; i < m; ++i) {
      temp += a[ii] * x[i];
      ii += 1;
    }
    y[j] += alpha * temp;
    ii0 += lda;
  }
}

// ---- index kernels (SpkSpdMMOps.jl) ----------------------------------------
// _assmb! :41-49
template <typename T>
void assmb(i64 tlen, i64 nj, const T* temp, const i64* relcol0,
           const i64* relind0, const i64* xlnz0, T* lnz0, i64 jlen) {
  const i64* relcol = B(relcol0);
  const i64* relind = B(relind0);
  const i64* xlnz = B(xlnz0);
  T* lnz = B(lnz0);
  for (i64 j = 1; j <= nj; ++j) {
    i64 lbot = xlnz[jlen - relcol[j] + 1] - 1;
    for (i64 k = 1; k <= tlen; ++k) lnz[lbot - relind[k]] += temp[(j - 1) * tlen + k - 1];
  }
}
// _ldindx! :70-79
void ldindx(i64 jlen, const i64* lindx0, i64* indmap0) {
  const i64* lindx = B(lindx0);
  i64* indmap = B(indmap0);
  i64 kk = jlen - 1;
  for (i64 j = 1; j <= jlen; ++j) {
    indmap[lindx[j]] = kk;
    kk -= 1;
  }
}
// _igathr! :94-102
void igathr(i64 klen, const i64* lindx0, const i64* indmap0, i64* relind0) {
  const i64* lindx = B(lindx0);
  const i64* indmap = B(indmap0);
  i64* relind = B(relind0);
  for (i64 j = 1; j <= klen; ++j) relind[j] = indmap[lindx[j]];
}
// _mmpyi! :125-143
template <typename T>
void mmpyi(i64 m, i64 q, const i64* zindxr0, const i64* zindxc0, const T* x0,
           const T* y0, const i64* iz0, T* z0, const i64* relind0, T diag) {
  const i64* zindxr = B(zindxr0);
  const i64* zindxc = B(zindxc0);
  const T* x = B(x0);
  const T* y = B(y0);
  const i64* iz = B(iz0);
  T* z = B(z0);
  const i64* relind = B(relind0);
  for (i64 k = 1; k <= q; ++k) {
    T t = y[k] * diag;
    i64 zlast = iz[zindxc[k] + 1] - 1;
    for (i64 j = 1; j <= m; ++j) {
      i64 zi = relind[zindxr[j]];
      z[zlast - zi] -= t * x[j];
    }
  }
}
// _luswap! :168-175
template <typename T>
void luswap(i64 m, i64 n, T* a, i64 lda, const i64* ipvt0) {
  const i64* ipvt = B(ipvt0);
  for (i64 k = 1; k <= n; ++k) {
    i64 i = ipvt[k];
    T* va = a + (k - 1) * lda;
    T* vb = a + (i - 1) * lda;
    for (i64 j = 0; j < m; ++j) {
      T t = vb[j];
      vb[j] = va[j];
      va[j] = t;
    }
  }
}

// ---- _lufactor! (SpkLUFactor.jl:60-255) -------------------------------------
// Returns the reference's iflag (overwritten per supernode as at :230-233);
// *first_fail receives the first supernode whose diagonal block had a zero
// pivot (0 if none), which the reference does not keep.
template <typename T>
i64 lufactor(i64 n, i64 nsuper, const i64* xsuper0, const i64* snode0,
             const i64* xlindx0, const i64* lindx0, const i64* xlnz0, T* lnz0,
             const i64* xunz0, T* unz0, i64* ipvt0, i64* first_fail) {
  const i64* xsuper = B(xsuper0);
  const i64* snode = B(snode0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  const i64* xunz = B(xunz0);
  T* lnz = B(lnz0);
  T* unz = B(unz0);
  i64* ipvt = B(ipvt0);
  const T ONE = T(1);
  std::vector<i64> link_((size_t)nsuper + 1, 0), lngth_((size_t)nsuper + 1, 0);
  i64* link = link_.data();
  i64* lngth = lngth_.data();
  i64 iflag = 0, tmpsiz = 0;
  if (first_fail) *first_fail = 0;
  for (i64 i = 1; i <= nsuper; ++i) {
    link[i] = 0;
    lngth[i] = xlindx[i + 1] - xlindx[i];
    i64 width = xsuper[i + 1] - xsuper[i];
    i64 need = width * (xlnz[xsuper[i] + 1] - xlnz[xsuper[i]]);
    if (need > tmpsiz) tmpsiz = need;
  }
  std::vector<i64> map((size_t)n + 1, 0), relind((size_t)n + 1, 0);
  std::vector<T> temp((size_t)(tmpsiz > 0 ? tmpsiz : 1), T(0));
  i64 nxt = 0;
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 fj = xsuper[jsup], lj = xsuper[jsup + 1] - 1, nj = lj - fj + 1;
    i64 jlen = xlnz[fj + 1] - xlnz[fj];
    i64 jxpnt = xlindx[jsup], jlpnt = xlnz[fj], jupnt = xunz[fj];
    ldindx(jlen, &lindx[jxpnt], map.data() + 1);
    for (;;) {
      i64 ksup = link[jsup];
      if (ksup != 0) {
        link[jsup] = link[ksup];
        link[ksup] = 0;
      }
      if (ksup == 0) break;
      i64 fk = xsuper[ksup], lk = xsuper[ksup + 1] - 1, nk = lk - fk + 1;
      i64 ksuplen = xlnz[fk + 1] - xlnz[fk];
      i64 klen = lngth[ksup];
      i64 kxpnt = xlindx[ksup + 1] - klen;
      i64 klpnt = xlnz[fk + 1] - klen;
      i64 kupnt = xunz[fk + 1] - klen;
      i64 nups;
      if (klen == jlen) {
        gemm_nt<T>(jlen, nj, nk, -ONE, &lnz[klpnt], ksuplen, &unz[kupnt],
                   ksuplen - nk, ONE, &lnz[jlpnt], jlen);
        if (jlen > nj)
          gemm_nt<T>(jlen - nj, nj, nk, -ONE, &unz[kupnt + nj], ksuplen - nk,
                     &lnz[klpnt], ksuplen, ONE, &unz[jupnt], jlen - nj);
        nups = nj;
        if (klen > nj) nxt = lindx[jxpnt + nj];
      } else {
        nups = klen;
        for (i64 i = 0; i <= klen - 1; ++i) {
          nxt = lindx[kxpnt + i];
          if (nxt > lj) {
            nups = i;
            break;
          }
        }
        if (nk == 1) {
          mmpyi<T>(klen, nups, &lindx[kxpnt], &lindx[kxpnt], &lnz[klpnt],
                   &unz[kupnt], xlnz0, lnz0, map.data() + 1, ONE);
          mmpyi<T>(klen - nups, nups, &lindx[kxpnt + nups], &lindx[kxpnt],
                   &unz[kupnt + nups], &lnz[klpnt], xunz0, unz0,
                   map.data() + 1, ONE);
        } else {
          i64 kfirst = lindx[kxpnt];
          i64 klast = lindx[kxpnt + klen - 1];
          i64 inddif = map[kfirst] - map[klast];
          if (inddif < klen) {
            i64 ilpnt = xlnz[kfirst] + (kfirst - fj);
            gemm_nt<T>(klen, nups, nk, -ONE, &lnz[klpnt], ksuplen, &unz[kupnt],
                       ksuplen - nk, ONE, &lnz[ilpnt], jlen);
            i64 iupnt = xunz[kfirst];
            if (klen > nups)
              gemm_nt<T>(klen - nups, nups, nk, -ONE, &unz[kupnt + nups],
                         ksuplen - nk, &lnz[klpnt], ksuplen, ONE, &unz[iupnt],
                         jlen - nj);
          } else {
            i64 store = klen * nups;
            if (store > tmpsiz) iflag = -2;
            igathr(klen, &lindx[kxpnt], map.data() + 1, relind.data() + 1);
            gemm_nt<T>(klen, nups, nk, -ONE, &lnz[klpnt], ksuplen, &unz[kupnt],
                       ksuplen - nk, T(0), temp.data(), klen);
            assmb<T>(klen, nups, temp.data(), relind.data() + 1,
                     relind.data() + 1, &xlnz[fj], lnz0, jlen);
            if (klen > nups) {
              gemm_nt<T>(klen - nups, nups, nk, -ONE, &unz[kupnt + nups],
                         ksuplen - nk, &lnz[klpnt], ksuplen, T(0), temp.data(),
                         klen - nups);
              assmb<T>(klen - nups, nups, temp.data(), relind.data() + 1,
                       relind.data() + 1 + nups, &xunz[fj], unz0, jlen);
            }
          }
        }
      }
      if (klen > nups) {
        i64 nxtsup = snode[nxt];
        link[ksup] = link[nxtsup];
        link[nxtsup] = ksup;
        lngth[ksup] = klen - nups;
      } else {
        lngth[ksup] = 0;
      }
    }
    iflag = getrf<T>(nj, nj, &lnz[jlpnt], jlen, &ipvt[fj]);
    if (iflag != 0) {
      iflag = -1;
      if (first_fail && *first_fail == 0) *first_fail = jsup;
    }
    trsm_run<T>(jlen - nj, nj, &lnz[jlpnt], jlen, &lnz[jlpnt + nj], jlen);
    if (jlen > nj) {
      luswap<T>(jlen - nj, nj, &unz[jupnt], jlen - nj, &ipvt[fj]);
      trsm_rltu<T>(jlen - nj, nj, &lnz[jlpnt], jlen, &unz[jupnt], jlen - nj);
    }
    if (jlen > nj) {
      nxt = lindx[jxpnt + nj];
      i64 nxtsup = snode[nxt];
      link[jsup] = link[nxtsup];
      link[nxtsup] = jsup;
      lngth[jsup] = jlen - nj;
    } else {
      lngth[jsup] = 0;
    }
  }
  return iflag;
}

// ---- solves (SpkLUFactor.jl:269-377) ----------------------------------------
template <typename T>
i64 lulsolve(i64 nsuper, const i64* xsuper0, const i64* xlindx0,
             const i64* lindx0, const i64* xlnz0, const T* lnz0,
             const i64* ipiv0, T* rhs0) {
  const i64* xsuper = B(xsuper0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  const T* lnz = B(lnz0);
  const i64* ipiv = B(ipiv0);
  T* rhs = B(rhs0);
  if (nsuper <= 0) return -1;
  i64 maxl = 0;
  for (i64 j = 1; j <= nsuper; ++j) {
    i64 l = xlindx[j + 1] - xlindx[j];
    if (l > maxl) maxl = l;
  }
  std::vector<T> temp((size_t)maxl + 1, T(0));
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 fj = xsuper[jsup], nj = xsuper[jsup + 1] - fj;
    i64 jlen = xlnz[fj + 1] - xlnz[fj];
    i64 jxpnt = xlindx[jsup], jlpnt = xlnz[fj];
    // glaswp!
    for (i64 i = 1; i <= nj; ++i) {
      i64 ip = ipiv[fj + i - 1];
      if (ip != i) {
        T t = rhs[fj + i - 1];
        rhs[fj + i - 1] = rhs[fj + ip - 1];
        rhs[fj + ip - 1] = t;
      }
    }
    trsm_llnu<T>(nj, 1, &lnz[jlpnt], jlen, &rhs[fj], nj);
    gemv_n<T>(jlen - nj, nj, T(-1), &lnz[jlpnt + nj], jlen, &rhs[fj], T(0),
              temp.data());
    i64 jj = jxpnt + nj - 1;
    for (i64 j = 1; j <= jlen - nj; ++j) {
      jj += 1;
      i64 isub = lindx[jj];
      rhs[isub] += temp[j - 1];
      temp[j - 1] = T(0);
    }
  }
  return 0;
}

template <typename T>
i64 luusolve(i64 n, i64 nsuper, const i64* xsuper0, const i64* xlindx0,
             const i64* lindx0, const i64* xlnz0, const T* lnz0,
             const i64* xunz0, const T* unz0, T* rhs0) {
  const i64* xsuper = B(xsuper0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  const i64* xunz = B(xunz0);
  const T* lnz = B(lnz0);
  const T* unz = B(unz0);
  T* rhs = B(rhs0);
  if (nsuper <= 0) return -1;
  i64 maxl = 0;
  for (i64 j = 1; j <= nsuper; ++j) {
    i64 l = xlindx[j + 1] - xlindx[j];
    if (l > maxl) maxl = l;
  }
  std::vector<T> temp((size_t)maxl + 1, T(0));
  for (i64 jsup = nsuper; jsup >= 1; --jsup) {
    i64 fj = xsuper[jsup], nj = xsuper[jsup + 1] - fj;
    i64 jlen = xlnz[fj + 1] - xlnz[fj];
    i64 jxpnt = xlindx[jsup], jlpnt = xlnz[fj], jupnt = xunz[fj];
    i64 jj = jxpnt + nj - 1;
    for (i64 j = 1; j <= jlen - nj; ++j) {
      jj += 1;
      i64 isub = lindx[jj];
      if (isub > n) return -2;
      temp[j - 1] = rhs[isub];
    }
    if (jlen > nj)
      gemv_t<T>(jlen - nj, nj, T(-1), &unz[jupnt], jlen - nj, temp.data(),
                T(1), &rhs[fj]);
    trsm_lunn<T>(nj, 1, &lnz[jlpnt], jlen, &rhs[fj], nj);
  }
  return 0;
}

// ---- _inmatrix! for CSC input (SparseCSCInterface.jl:101-169) --------------
// Zeroes lnz/unz/ipiv, then adds every A entry into its slot. Returns 0, or
// -1 on the reference's "No space for matrix element" error.
template <typename T>
i64 inmatrix_csc(i64 n, const i64* colptr0, const i64* rowval0,
                 const double* nzval0, const i64* rinvp0, const i64* snode0,
                 const i64* xsuper0, const i64* xlindx0, const i64* lindx0,
                 const i64* xlnz0, T* lnz0, const i64* xunz0, T* unz0,
                 i64* ipiv0) {
  const i64* colptr = B(colptr0);
  const i64* rowval = B(rowval0);
  const double* nzval = B(nzval0);
  const i64* rinvp = B(rinvp0);
  const i64* snode = B(snode0);
  const i64* xsuper = B(xsuper0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  const i64* xunz = B(xunz0);
  T* lnz = B(lnz0);
  T* unz = B(unz0);
  i64* ipiv = B(ipiv0);
  if (n == 0) return -1;
  const i64 nsuper = snode[n];
  for (i64 k = 1; k <= xlnz[n + 1] - 1; ++k) lnz[k] = T(0);
  for (i64 k = 1; k <= xunz[n + 1] - 1; ++k) unz[k] = T(0);
  for (i64 k = 1; k <= n; ++k) ipiv[k] = 0;
  (void)nsuper;
  for (i64 i = 1; i <= n; ++i) {
    for (i64 iptr = colptr[i]; iptr <= colptr[i + 1] - 1; ++iptr) {
      i64 inew = rinvp[rowval[iptr]];
      i64 jnew = rinvp[i];
      T value = (T)nzval[iptr];
      if (inew >= xsuper[snode[jnew]]) {
        i64 jsup = snode[jnew];
        i64 fstsub = xlindx[jsup], lstsub = xlindx[jsup + 1] - 1;
        i64 nnzloc = 0;
        for (i64 nxtsub = fstsub; nxtsub <= lstsub; ++nxtsub) {
          i64 irow = lindx[nxtsub];
          if (irow > inew) return -1;
          if (irow == inew) {
            lnz[xlnz[jnew] + nnzloc] += value;
            break;
          }
          nnzloc += 1;
        }
      } else {
        i64 jsup = snode[inew];
        i64 fstcol = xsuper[jsup], lstcol = xsuper[jsup + 1] - 1;
        i64 width = lstcol - fstcol + 1;
        i64 lstsub = xlindx[jsup + 1] - 1, fstsub = xlindx[jsup] + width;
        i64 nnzloc = 0;
        for (i64 nxtsub = fstsub; nxtsub <= lstsub; ++nxtsub) {
          i64 irow = lindx[nxtsub];
          if (irow > jnew) return -1;
          if (irow == jnew) {
            unz[xunz[inew] + nnzloc] += value;
            break;
          }
          nnzloc += 1;
        }
      }
    }
  }
  return 0;
}

// ---- SPD path --------------------------------------------------------------
// SpkSparseSpdBase.jl:234-311, reading CSC instead of a Problem: only
// entries with row >= column (original numbering) are used.
template <typename T>
i64 inmatrix_spd_csc(i64 n, const i64* colptr0, const i64* rowval0,
                     const double* nzval0, const i64* rinvp0,
                     const i64* snode0, const i64* xsuper0, const i64* xlindx0,
                     const i64* lindx0, const i64* xlnz0, T* lnz0) {
  const i64* colptr = B(colptr0);
  const i64* rowval = B(rowval0);
  const double* nzval = B(nzval0);
  const i64* rinvp = B(rinvp0);
  const i64* snode = B(snode0);
  const i64* xsuper = B(xsuper0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  T* lnz = B(lnz0);
  if (n == 0) return -1;
  for (i64 k = 1; k <= xlnz[n + 1] - 1; ++k) lnz[k] = T(0);
  for (i64 i = 1; i <= n; ++i) {
    for (i64 ptr = colptr[i]; ptr <= colptr[i + 1] - 1; ++ptr) {
      if (rowval[ptr] >= i) {
        i64 inew = rinvp[rowval[ptr]];
        T value = (T)nzval[ptr];
        i64 jnew = rinvp[i];
        if (inew < jnew) {
          i64 t = inew;
          inew = jnew;
          jnew = t;
        }
        i64 jsup = snode[jnew], fstcol = xsuper[jsup];
        i64 fstsub = xlindx[jsup] + jnew - fstcol;
        i64 lstsub = xlindx[jsup + 1] - 1;
        for (i64 nxtsub = fstsub; nxtsub <= lstsub; ++nxtsub) {
          i64 irow = lindx[nxtsub];
          if (irow > inew) return -1;
          if (irow == inew) {
            i64 supcol = jnew - xsuper[snode[jnew]];
            i64 nnzloc = xlnz[jnew] + (nxtsub - fstsub) + supcol;
            lnz[nnzloc] += value;
            break;
          }
        }
      }
    }
  }
  return 0;
}

// _matrixdiagmm! (SpkLDLtFactor.jl:315-325): the reference reads lnz with
// leading dimension klen (its defect); `ld` selects klen (faithful) or
// ksuplen (corrected).
template <typename T>
void matrixdiagmm(const T* lnz, i64 klen, i64 nk, i64 ld, const T* diag,
                  T* temp) {
  for (i64 j = 1; j <= nk; ++j) {
    T d = diag[j - 1];
    for (i64 r = 1; r <= klen; ++r)
      temp[(j - 1) * klen + r - 1] = d * lnz[(j - 1) * ld + r - 1];
  }
}

// Diagonal block of the LDL^T: the reference's _pchole! (faithful) or a
// textbook right-looking LDL^T of the nj x nj block (corrected), followed by
// L21 = A21 * L11^{-T} * D^{-1} as at SpkLDLtFactor.jl:367-375.
template <typename T>
void pchole(T* a, i64 nj, i64 lda, bool faithful) {
  T* A = B(a);
  if (faithful) {
    for (i64 c = 1; c <= nj; ++c)
      for (i64 i = 1; i <= c - 1; ++i) {
        for (i64 r = c; r <= nj; ++r)
          A[IDX(lda, r, c)] += -A[IDX(lda, c, i)] * A[IDX(lda, i, i)] * A[IDX(lda, r, i)];
        T diag = A[IDX(lda, c, c)];
        for (i64 r = c + 1; r <= nj; ++r) A[IDX(lda, r, c)] /= diag;
      }
  } else {
    for (i64 k = 1; k <= nj; ++k) {
      T d = A[IDX(lda, k, k)];
      for (i64 j = k + 1; j <= nj; ++j) {
        T ljk = A[IDX(lda, j, k)] / d;
        for (i64 r = j; r <= nj; ++r) A[IDX(lda, r, j)] -= A[IDX(lda, r, k)] * ljk;
      }
      for (i64 r = k + 1; r <= nj; ++r) A[IDX(lda, r, k)] /= d;
    }
  }
  trsm_rltu<T>(lda - nj, nj, a, lda, a + nj, lda);
  for (i64 c = 1; c <= nj; ++c) {
    T diag = A[IDX(lda, c, c)];
    for (i64 r = nj + 1; r <= lda; ++r) A[IDX(lda, r, c)] /= diag;
  }
}

// _ldltfactor! (SpkLDLtFactor.jl:58-246)
template <typename T>
i64 ldltfactor(i64 n, i64 nsuper, const i64* xsuper0, const i64* snode0,
               const i64* xlindx0, const i64* lindx0, const i64* xlnz0,
               T* lnz0, bool faithful) {
  const i64* xsuper = B(xsuper0);
  const i64* snode = B(snode0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  T* lnz = B(lnz0);
  const T ONE = T(1);
  std::vector<i64> link((size_t)nsuper + 1, 0), lngth((size_t)nsuper + 1, 0);
  i64 iflag = 0, tmpsiz = 0, maxwidth = 0;
  for (i64 i = 1; i <= nsuper; ++i) {
    link[i] = 0;
    lngth[i] = xlindx[i + 1] - xlindx[i];
    i64 width = xsuper[i + 1] - xsuper[i];
    i64 need = width * (xlnz[xsuper[i] + 1] - xlnz[xsuper[i]]);
    if (need > tmpsiz) tmpsiz = need;
    if (width > maxwidth) maxwidth = width;
  }
  std::vector<i64> map((size_t)n + 1, 0), relind((size_t)n + 1, 0);
  std::vector<T> diag((size_t)maxwidth + 1), temp((size_t)tmpsiz + 1),
      temp2((size_t)(maxwidth * maxwidth) + 1);
  i64 nxt = 0;
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 fj = xsuper[jsup], lj = xsuper[jsup + 1] - 1, nj = lj - fj + 1;
    i64 jlen = xlnz[fj + 1] - xlnz[fj];
    i64 jxpnt = xlindx[jsup], jlpnt = xlnz[fj];
    ldindx(jlen, &lindx[jxpnt], map.data() + 1);
    for (;;) {
      i64 ksup = link[jsup];
      if (ksup != 0) {
        link[jsup] = link[ksup];
        link[ksup] = 0;
      }
      if (ksup == 0) break;
      i64 fk = xsuper[ksup], lk = xsuper[ksup + 1] - 1, nk = lk - fk + 1;
      i64 ksuplen = xlnz[fk + 1] - xlnz[fk];
      i64 klen = lngth[ksup];
      i64 kxpnt = xlindx[ksup + 1] - klen;
      i64 klpnt = xlnz[fk + 1] - klen;
      // _loaddiag! (:328-333)
      for (i64 j = 1; j <= nk; ++j) diag[j - 1] = lnz[xlnz[fk] + (j - 1) * ksuplen + j - 1];
      const i64 ldd = faithful ? -1 : ksuplen;
      i64 nups;
      if (klen == jlen) {
        matrixdiagmm<T>(&lnz[klpnt], nj, nk, ldd < 0 ? nj : ldd, diag.data(), temp2.data());
        gemm_nt<T>(jlen, nj, nk, -ONE, &lnz[klpnt], ksuplen, temp2.data(), nj,
                   ONE, &lnz[jlpnt], jlen);
        nups = nj;
        if (klen > nj) nxt = lindx[jxpnt + nj];
      } else {
        nups = klen;
        for (i64 i = 0; i <= klen - 1; ++i) {
          nxt = lindx[kxpnt + i];
          if (nxt > lj) {
            nups = i;
            break;
          }
        }
        if (nk == 1) {
          mmpyi<T>(klen, nups, &lindx[kxpnt], &lindx[kxpnt], &lnz[klpnt],
                   &lnz[klpnt], xlnz0, lnz0, map.data() + 1, lnz[xlnz[fk]]);
        } else {
          i64 kfirst = lindx[kxpnt];
          i64 klast = lindx[kxpnt + klen - 1];
          i64 inddif = map[kfirst] - map[klast];
          if (inddif < klen) {
            i64 ilpnt = xlnz[kfirst] + (kfirst - fj);
            i64 width = nups;
            matrixdiagmm<T>(&lnz[klpnt], width, nk, ldd < 0 ? width : ldd, diag.data(), temp2.data());
            gemm_nt<T>(klen, nups, nk, -ONE, &lnz[klpnt], ksuplen, temp2.data(),
                       width, ONE, &lnz[ilpnt], jlen);
          } else {
            i64 store = klen * nups;
            if (store > tmpsiz) return -2;
            igathr(klen, &lindx[kxpnt], map.data() + 1, relind.data() + 1);
            i64 width = nups;
            matrixdiagmm<T>(&lnz[klpnt], width, nk, ldd < 0 ? width : ldd, diag.data(), temp2.data());
            gemm_nt<T>(klen, nups, nk, -ONE, &lnz[klpnt], ksuplen, temp2.data(),
                       width, T(0), temp.data(), klen);
            assmb<T>(klen, nups, temp.data(), relind.data() + 1,
                     relind.data() + 1, &xlnz[fj], lnz0, jlen);
          }
        }
      }
      if (klen > nups) {
        i64 nxtsup = snode[nxt];
        link[ksup] = link[nxtsup];
        link[nxtsup] = ksup;
        lngth[ksup] = klen - nups;
      } else {
        lngth[ksup] = 0;
      }
    }
    pchole<T>(&lnz[xlnz[fj]], nj, jlen, faithful);
    if (jlen > nj) {
      nxt = lindx[jxpnt + nj];
      i64 nxtsup = snode[nxt];
      link[jsup] = link[nxtsup];
      link[nxtsup] = jsup;
      lngth[jsup] = jlen - nj;
    } else {
      lngth[jsup] = 0;
    }
  }
  return iflag;
}

// _ldltsolve! (SpkLDLtFactor.jl:266-293)
template <typename T>
i64 ldltsolve(i64 nsuper, const i64* xsuper0, const i64* xlindx0,
              const i64* lindx0, const i64* xlnz0, const T* lnz0, T* rhs0) {
  const i64* xsuper = B(xsuper0);
  const i64* xlindx = B(xlindx0);
  const i64* lindx = B(lindx0);
  const i64* xlnz = B(xlnz0);
  const T* lnz = B(lnz0);
  T* rhs = B(rhs0);
  for (i64 jsup = 1; jsup <= nsuper; ++jsup) {
    i64 fjcol = xsuper[jsup], ljcol = xsuper[jsup + 1] - 1;
    i64 fsub = xlindx[jsup], lsub = xlindx[jsup + 1] - 1;
    for (i64 jcol = fjcol; jcol <= ljcol; ++jcol) {
      i64 ofst = jcol - fjcol;
      i64 nzf = xlnz[jcol] + 1 + ofst;
      fsub = fsub + 1;
      T t = rhs[jcol];
      // rhs[lindx[fsub:lsub]] = rhs[lindx[fsub:lsub]] - t*lnz[nzf:nzl]
      // (right-hand side evaluated before the assignment; indices distinct)
      for (i64 q = 0; q <= lsub - fsub; ++q)
        rhs[lindx[fsub + q]] = rhs[lindx[fsub + q]] - t * lnz[nzf + q];
    }
  }
  for (i64 jsup = nsuper; jsup >= 1; --jsup) {
    i64 fjcol = xsuper[jsup], ljcol = xsuper[jsup + 1] - 1;
    i64 fsub = xlindx[jsup] + ljcol - fjcol;
    for (i64 jcol = ljcol; jcol >= fjcol; --jcol) {
      i64 ofst = jcol - fjcol;
      i64 nzf = xlnz[jcol] + ofst;
      i64 m = xlnz[jcol + 1] - nzf - 1;
      T t = rhs[jcol] / lnz[nzf];
      for (i64 k = 1; k <= m; ++k) t = t - lnz[nzf + k] * rhs[lindx[fsub + k]];
      rhs[jcol] = t;
      fsub = fsub - 1;
    }
  }
  return 0;
}

}  // namespace

#define ORC_EXPORT(SUF, T)                                                      \
  extern "C" int64_t orc_lufactor_##SUF(                                       \
      int64_t n, int64_t nsuper, const int64_t* xsuper, const int64_t* snode,  \
      const int64_t* xlindx, const int64_t* lindx, const int64_t* xlnz,        \
      T* lnz, const int64_t* xunz, T* unz, int64_t* ipvt,                      \
      int64_t* first_fail) {                                                   \
    return lufactor<T>(n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz,     \
                       xunz, unz, ipvt, first_fail);                           \
  }                                                                            \
  extern "C" int64_t orc_lulsolve_##SUF(                                       \
      int64_t nsuper, const int64_t* xsuper, const int64_t* xlindx,            \
      const int64_t* lindx, const int64_t* xlnz, const T* lnz,                 \
      const int64_t* ipiv, T* rhs) {                                           \
    return lulsolve<T>(nsuper, xsuper, xlindx, lindx, xlnz, lnz, ipiv, rhs);  \
  }                                                                            \
  extern "C" int64_t orc_luusolve_##SUF(                                       \
      int64_t n, int64_t nsuper, const int64_t* xsuper, const int64_t* xlindx, \
      const int64_t* lindx, const int64_t* xlnz, const T* lnz,                 \
      const int64_t* xunz, const T* unz, T* rhs) {                             \
    return luusolve<T>(n, nsuper, xsuper, xlindx, lindx, xlnz, lnz, xunz, unz, \
                       rhs);                                                   \
  }                                                                            \
  extern "C" int64_t orc_inmatrix_csc_##SUF(                                   \
      int64_t n, const int64_t* colptr, const int64_t* rowval,                 \
      const double* nzval, const int64_t* rinvp, const int64_t* snode,         \
      const int64_t* xsuper, const int64_t* xlindx, const int64_t* lindx,      \
      const int64_t* xlnz, T* lnz, const int64_t* xunz, T* unz,                \
      int64_t* ipiv) {                                                         \
    return inmatrix_csc<T>(n, colptr, rowval, nzval, rinvp, snode, xsuper,     \
                           xlindx, lindx, xlnz, lnz, xunz, unz, ipiv);         \
  }                                                                            \
  extern "C" int64_t orc_inmatrix_spd_csc_##SUF(                               \
      int64_t n, const int64_t* colptr, const int64_t* rowval,                 \
      const double* nzval, const int64_t* rinvp, const int64_t* snode,         \
      const int64_t* xsuper, const int64_t* xlindx, const int64_t* lindx,      \
      const int64_t* xlnz, T* lnz) {                                           \
    return inmatrix_spd_csc<T>(n, colptr, rowval, nzval, rinvp, snode, xsuper, \
                               xlindx, lindx, xlnz, lnz);                      \
  }                                                                            \
  extern "C" int64_t orc_ldltfactor_##SUF(                                     \
      int64_t n, int64_t nsuper, const int64_t* xsuper, const int64_t* snode,  \
      const int64_t* xlindx, const int64_t* lindx, const int64_t* xlnz,        \
      T* lnz) {                                                                \
    return ldltfactor<T>(n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz,   \
                         false);                                               \
  }                                                                            \
  extern "C" int64_t orc_ldltfactor_ref_##SUF(                                 \
      int64_t n, int64_t nsuper, const int64_t* xsuper, const int64_t* snode,  \
      const int64_t* xlindx, const int64_t* lindx, const int64_t* xlnz,        \
      T* lnz) {                                                                \
    return ldltfactor<T>(n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz,   \
                         true);                                                \
  }                                                                            \
  extern "C" int64_t orc_ldltsolve_##SUF(                                      \
      int64_t nsuper, const int64_t* xsuper, const int64_t* xlindx,            \
      const int64_t* lindx, const int64_t* xlnz, const T* lnz, T* rhs) {       \
    return ldltsolve<T>(nsuper, xsuper, xlindx, lindx, xlnz, lnz, rhs);        \
  }

ORC_EXPORT(f64, double)
ORC_EXPORT(f32, float)
