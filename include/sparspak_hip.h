/*
 * sparspak_hip.h — C ABI of libsparspak_hip.so, the MI355X (gfx950) numeric
 * phase of the Sparspak.jl sparse direct solver plus its host symbolic phase.
 *
 * Conventions (the reference's own): every array is caller-owned and is not
 * retained after the call returns; integer arrays are int64 and hold 1-based
 * index VALUES exactly as Sparspak.jl stores them; numeric arrays are updated
 * in place. Functions return 0 on success and a negative code on failure:
 *   -1     zero pivot in a diagonal block (SpkLUFactor.jl:230-233 meaning),
 *          or a structural error named by the function's comment
 *   -2     workspace / size limit
 *   -3     sequence error (a phase called before the one it depends on)
 *   -4     bad argument
 *   -5     refinement did not reach the requested backward error within
 *          max_iter corrections (the solution is returned; its backward
 *          error is in spkhip_stats)
 *   -6     the solution or residual is not finite (NaN or Inf): never
 *          certified; backward_error is +inf
 *   <= -1000  HIP runtime error (-1000 - hipError_t)
 * No C++ exception crosses this boundary. One handle serialises its calls on
 * its own HIP stream; different handles may be used from different threads.
 *
 * Device buffers of the caller (the _device entry points): before reading
 * them, the handle's stream waits for all work the caller has issued on its
 * stream — the legacy default stream (what PyTorch's default stream is on
 * ROCm) unless spkhip_set_caller_stream names another — so a producer kernel
 * launched just before the call needs no host synchronisation. Every entry
 * point returns only after its own work has completed.
 */
#ifndef SPARSPAK_HIP_H
#define SPARSPAK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================
 * Host symbolic phase (bit-exact restatements; run on the CPU).
 * ====================================================================== */

/* SpkGraph.jl:298-316 isstructuresymmetric -> 1 or 0. */
int64_t spk_isstructuresymmetric(int64_t nv, const int64_t* xadj,
                                 const int64_t* adj);
/* SpkGraph.jl:94-268 (counting pass): edge count after symmetrisation. */
int64_t spk_symmetrized_nedges(int64_t nv, const int64_t* xadj,
                               const int64_t* adj);
/* SpkGraph.jl:94-289 makestructuresymmetric + sortgraph!. Returns 1 if the
 * graph was already symmetric (copied unchanged), 0 if symmetrised. */
int64_t spk_makestructuresymmetric(int64_t nv, const int64_t* xadj,
                                   const int64_t* adj, int64_t* hxadj,
                                   int64_t* hadj);
/* SpkMMD.jl:38-209 mmd!/_generalmmd (delta = 0): perm (new->old), invp. */
int64_t spk_mmd(int64_t n, const int64_t* xadj, const int64_t* adj,
                int64_t* perm, int64_t* invp);
/* SpkETree.jl:62-88 _findetree!. */
void spk_etree(int64_t n, const int64_t* xadj, const int64_t* adj,
               const int64_t* perm, const int64_t* invp, int64_t* parent);
/* SpkETree.jl:106-139 _getpostorder! (weight == NULL: plain postorder). */
void spk_postorder(int64_t n, int64_t* parent, int64_t* perm, int64_t* invp,
                   int64_t* weight);
/* SpkSymFct.jl:37-59 _findcolumncounts! -> nnz(L) including the diagonal. */
int64_t spk_colcounts(int64_t n, const int64_t* xadj, const int64_t* adj,
                      const int64_t* perm, const int64_t* invp,
                      const int64_t* parent, int64_t* colcnt);
/* SpkSymFct.jl:415-464 _findsupernodes! -> nsuper (xsuper needs n+1). */
int64_t spk_supernodes(int64_t n, const int64_t* parent, const int64_t* colcnt,
                       int64_t maxsize, int64_t* xsuper, int64_t* snode,
                       int64_t* nofsub);
/* SpkSparseBase.jl:253-284 findnonzeroindexs (LU panel pointers). */
void spk_nonzero_indexs(int64_t n, const int64_t* colcnt, int64_t nsuper,
                        const int64_t* xsuper, int64_t* xlnz, int64_t* xunz);
/* SpkSparseSpdBase.jl:359-375 _findnonzeroindexs! (SPD panel pointers). */
void spk_nonzero_indexs_spd(int64_t n, const int64_t* colcnt, int64_t nsuper,
                            const int64_t* xsuper, int64_t* xlnz);
/* SpkSymFct.jl:105-246 _symbolicfact! (-1: "Inconsistency in data
 * structure"). */
int64_t spk_symbfact(int64_t n, const int64_t* xadj, const int64_t* adj,
                     const int64_t* perm, const int64_t* invp,
                     const int64_t* colcnt, int64_t nsuper,
                     const int64_t* xsuper, const int64_t* snode,
                     int64_t nofsub, int64_t* xlindx, int64_t* lindx);
/* SpkSymFct.jl:372-380 _findsupernodetree! (0 marks a root). */
void spk_supernode_tree(int64_t nsuper, const int64_t* xsuper,
                        const int64_t* snode, const int64_t* xlindx,
                        const int64_t* lindx, int64_t* sparent);
/* The front (fundamental supernode) schedule spkhip_symbolic_upload builds,
 * on the host, with the relative-index maps it uploads. These maps replace
 * the per-pair maps of _ldindx!/_igathr! (SpkSpdMMOps.jl:70-102): for the
 * t-th update row of front f (t = 0 .. L_f - W_f - 1, the rows of its index
 * list after its W_f own columns), relmap[frel[f] - 1 + t] is the 0-based
 * position of that row in the parent front's index list (of length L_p),
 * i.e. L_p - 1 - relind with relind = map[row] = distance from the bottom of
 * the list (_ldindx!). Outputs (1-based values): fsp0[f] first supernode of
 * front f, fnp[f] its supernodes, fparent[f] the parent front (0: root),
 * frel[f] (nfront + 1 entries) offsets into relmap. Array sizes: nsuper
 * (nsuper + 1 for frel) and nsub = xlindx[nsuper] - 1 for relmap. Returns
 * nfront, -4 when the index lists do not nest. No device is used. */
int64_t spk_front_maps(int64_t n, int64_t nsuper, const int64_t* xsuper,
                       const int64_t* snode, const int64_t* xlindx,
                       const int64_t* lindx, int64_t* fsp0, int64_t* fnp,
                       int64_t* fparent, int64_t* frel, int64_t* relmap);

/* The map from each entry of A (CSC, original numbering, nnz = colptr[n]-1)
 * to its slot, as spkhip_symbolic_upload builds it for _inmatrix!
 * (SpkSparseBase.jl:302-372, SparseCSCInterface.jl:112-163, SPD
 * SpkSparseSpdBase.jl:274-282): amap[k] >= 0 is the 0-based lnz slot,
 * amap[k] <= -2 the unz slot -(amap[k] + 2), -1 an entry not stored (the
 * index list ran out, or an SPD entry above the diagonal). Returns -1 ("No
 * space for matrix element": a larger index met first), else 0. nthreads:
 * host threads (0: SPKHIP_HOST_THREADS or the hardware's, at most 16); any
 * count gives the same map. No device is used. */
int64_t spk_amap(int64_t n, int64_t kind, int64_t nsuper, const int64_t* xsuper,
                 const int64_t* snode, const int64_t* xlindx, const int64_t* lindx,
                 const int64_t* xlnz, const int64_t* xunz, const int64_t* invp,
                 const int64_t* colptr, const int64_t* rowval, int64_t nthreads,
                 int64_t* amap);

/* ======================================================================
 * Device numeric phase.
 * ====================================================================== */

typedef struct spkhip_handle* spkhip_t;

/* Factor kinds. */
#define SPKHIP_KIND_LU 0  /* SparseSolver: LU, pivoting inside diagonal blocks */
#define SPKHIP_KIND_SPD 1 /* SparseSpdSolver: LDL^T, no pivoting */

/* Factor precisions. */
#define SPKHIP_FP64 0     /* factor and solve in double */
#define SPKHIP_FP32 1     /* factor and solve in float (refine in double) */
#define SPKHIP_FP32_BF16 2 /* float factor whose trailing updates in fronts of
                              more than 256 own columns take bf16 operands
                              (float accumulation); solves as SPKHIP_FP32 */
#define SPKHIP_FP32_FP16 3 /* as SPKHIP_FP32_BF16 with fp16 operands under
                              diagonal scaling: column k of L21 is multiplied
                              and row k of U12 divided by the power of two
                              nearest below |U(k,k)| (the product is
                              unchanged), which keeps both in fp16 range for
                              diagonally scaled problems */
#define SPKHIP_FP32_BF16X3 4 /* as SPKHIP_FP32_BF16 with split-bf16 operands
                              (x = hi + lo, three bf16 products per pair,
                              about 16 significant bits) */

typedef struct spkhip_stats {
  int64_t n, nsuper, nfront, nlevel;
  int64_t nnz_lnz, nnz_unz;       /* |lnz|, |unz| of the reference layout */
  int64_t first_fail_supernode;   /* 1-based; 0 = no zero pivot */
  int64_t refine_iters;
  double factor_flops;            /* F_LU or F_Chol from the column counts */
  double solve_flops;             /* per right-hand side */
  double front_bytes;             /* device front/update arena, bytes */
  double t_upload_ms, t_inmatrix_ms, t_factor_ms, t_solve_ms, t_refine_ms;
  double backward_error;          /* last normwise backward error (refine) */
} spkhip_stats_t;

/* Handle lifetime. device = HIP device ordinal. */
int64_t spkhip_create(spkhip_t* h, int64_t device);
int64_t spkhip_destroy(spkhip_t h);

/* Uploads the symbolic factor produced by the host phase and builds the
 * device schedule (front tree, levels, relative-index maps, A -> factor
 * offset map). Replaces the storage set-up at the end of
 * SpkSparseBase.jl:193-251 (kind LU) / SpkSparseSpdBase.jl:178-232 (SPD).
 * colptr/rowval: the CSC pattern of A in the ORIGINAL numbering
 * (SparseCSCInterface.jl:101-169 reads the same arrays); nnz = colptr[n]-1.
 * Returns -1 ("No space for matrix element") if an entry of A falls outside
 * the symbolic structure. */
int64_t spkhip_symbolic_upload(spkhip_t h, int64_t kind, int64_t n,
                               int64_t nsuper, const int64_t* xsuper,
                               const int64_t* snode, const int64_t* xlindx,
                               const int64_t* lindx, const int64_t* xlnz,
                               const int64_t* xunz, const int64_t* perm,
                               const int64_t* invp, const int64_t* colptr,
                               const int64_t* rowval);

/* _inmatrix! (SpkSparseBase.jl:302-372, SparseCSCInterface.jl:101-169):
 * zero lnz/unz/ipiv on the device and add every value of A (CSC order,
 * original numbering, nnz values) into its slot. nzval is a host pointer;
 * the _device variant takes a device pointer (resident values). */
int64_t spkhip_inmatrix(spkhip_t h, const double* nzval);
int64_t spkhip_inmatrix_device(spkhip_t h, const double* d_nzval);

/* _factor! -> _lufactor! (SpkSparseBase.jl:378-391, SpkLUFactor.jl:60-255)
 * or _ldltfactor! (SpkSparseSpdBase.jl:313-332). precision: SPKHIP_FP64,
 * SPKHIP_FP32, SPKHIP_FP32_BF16, SPKHIP_FP32_FP16 or SPKHIP_FP32_BF16X3.
 * Returns -1 if any diagonal block met a
 * zero pivot (the first failing supernode is in stats). */
int64_t spkhip_factor(spkhip_t h, int64_t precision);

/* _triangularsolve! (SpkSparseBase.jl:400-416) for nrhs right-hand sides
 * stored column-major in B (host, leading dimension ldb, original
 * numbering); overwritten by the solution. Uses the last factor. */
int64_t spkhip_solve(spkhip_t h, int64_t nrhs, double* b, int64_t ldb);
/* Same with device-resident B (double, column-major). */
int64_t spkhip_solve_device(spkhip_t h, int64_t nrhs, double* d_b,
                            int64_t ldb);
/* No counterpart in the reference. spkhip_solve_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): this call makes no
 * hipMalloc/hipFree and no other allocation. Same kernels and results as
 * spkhip_solve_device (bit-identical), launched directly (never replayed
 * from a captured graph). Its scratch (the permuted right-hand sides and
 * the per-front solve vectors) is a 64 MiB array the library holds on each
 * GPU it runs on until the process ends; it needs
 *   round_up(n * nrhs * s, 256) + vec * nrhs * s  bytes
 * (s = 8 for an SPKHIP_FP64 factor, 4 otherwise; vec = the handle's solve
 * vector length, about the number of row indices of the fronts, rounded up
 * to 64; both terms times the number of value sets) and returns
 * -2 when that exceeds 64 MiB (or when the factor's solve data was never
 * sized, which spkhip_factor does). Calls on the same GPU share that array:
 * calls that overlap in time on one GPU (other threads, other handles) race
 * and must be serialised by the caller. */
int64_t spkhip_solve_device_noalloc(spkhip_t h, int64_t nrhs, double* d_b,
                                    int64_t ldb);

/* Solve with iterative refinement: the residual r = b - A x and the update
 * x += d run in double on the device, the correction solve uses the last
 * factor (double or float). Stops when the normwise backward error
 * ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol for every
 * column, or after max_iter corrections. Host B (as spkhip_solve). */
int64_t spkhip_solve_refine(spkhip_t h, int64_t nrhs, double* b, int64_t ldb,
                            double tol, int64_t max_iter);
int64_t spkhip_solve_refine_device(spkhip_t h, int64_t nrhs, double* d_b,
                                   int64_t ldb, double tol, int64_t max_iter);
/* 1-norm condition estimate of an SPD factor: the reference's commented
 * sparsespdainverseonenorm / sparsespdcondest (SpkSparseSpdSolver.jl:346-459),
 * Hager's method with at most five iterations, on device solves with the
 * current factor (double or float). cond1 = ||A||_1 * est(||A^-1||_1), with
 * ||A||_1 the largest row sum of |a_ij| of the matrix given to
 * spkhip_symbolic_upload (the full symmetric matrix); ainv_norm1 the
 * estimate of ||A^-1||_1, a lower bound. -4 for an LU handle: the method
 * needs solves with A^T, which equal the solves with A only for a symmetric
 * factor; -3 before a factor. */
int64_t spkhip_condest(spkhip_t h, double* cond1, double* ainv_norm1);
/* Independent systems stacked in one factor (a block-diagonal batch): the
 * refinement's backward error is then taken per system, over the rows
 * [offsets[i], offsets[i+1]) (0-based, original numbering, offsets[0] = 0,
 * offsets[nseg] = n). Default: one segment. No counterpart in the reference
 * (it factors one system at a time). */
int64_t spkhip_set_segments(spkhip_t h, int64_t nseg, const int64_t* offsets);
/* Value sets: nvs matrices of the uploaded pattern (same perm, same
 * symbolic factor, different values) factored and solved together, each
 * launch covering every set. After this call, and until the next symbolic
 * upload (which returns to one set):
 *   spkhip_inmatrix[_device]: nzval holds nvs * nnz values, set v's at v*nnz;
 *   spkhip_factor: factors every set (a zero pivot in any set returns -1;
 *     stats name the lowest failing supernode over the sets);
 *   spkhip_solve*, spkhip_solve_refine*: B has n * nvs rows (ldb >= n*nvs),
 *     set v's system on rows [v*n, (v+1)*n); the backward error is taken per
 *     set and per segment of spkhip_set_segments;
 *   spkhip_download*: set v's lnz at v*nnz_lnz, unz at v*nnz_unz, ipvt at
 *     v*n of the host arrays.
 * Every set's result equals, bit for bit, that of a handle holding that set
 * alone. Values and factor held before the call are dropped (-3 until the
 * next spkhip_inmatrix). spkhip_condest takes one set only (-4). nvs in
 * [1, 65535]. The reference factors one matrix at a time; its
 * refactorisation with new values of one pattern (sparspaklu!,
 * SparseCSCInterface.jl:247-268) is what this batches. */
int64_t spkhip_set_value_sets(spkhip_t h, int64_t nvs);
/* The stream (a hipStream_t, NULL = the legacy default stream) on which the
 * caller produces the device buffers it passes to the _device entry points;
 * they order themselves after that stream's earlier work. No counterpart in
 * the reference (it has no device). */
int64_t spkhip_set_caller_stream(spkhip_t h, void* stream);

/* Copies the factor back in the reference layout: lnz (|lnz|), unz (|unz|),
 * ipvt (n, local 1..nj pivots). Double variant converts a float factor. */
int64_t spkhip_download(spkhip_t h, double* lnz, double* unz, int64_t* ipvt);
int64_t spkhip_download_f32(spkhip_t h, float* lnz, float* unz, int64_t* ipvt);

int64_t spkhip_stats(spkhip_t h, spkhip_stats_t* st);

/* Stateless drop-ins with the reference's argument lists (upload, compute on
 * the device, download). lnz/unz hold the _inmatrix! output on entry. */
/* SpkLUFactor.jl:60 _lufactor! */
int64_t spkhip_lufactor_f64(int64_t n, int64_t nsuper, const int64_t* xsuper,
                            const int64_t* snode, const int64_t* xlindx,
                            const int64_t* lindx, const int64_t* xlnz,
                            double* lnz, const int64_t* xunz, double* unz,
                            int64_t* ipvt);
int64_t spkhip_lufactor_f32(int64_t n, int64_t nsuper, const int64_t* xsuper,
                            const int64_t* snode, const int64_t* xlindx,
                            const int64_t* lindx, const int64_t* xlnz,
                            float* lnz, const int64_t* xunz, float* unz,
                            int64_t* ipvt);
/* SpkLUFactor.jl:269 _lulsolve! */
int64_t spkhip_lulsolve_f64(int64_t nsuper, const int64_t* xsuper,
                            const int64_t* xlindx, const int64_t* lindx,
                            const int64_t* xlnz, const double* lnz,
                            const int64_t* ipiv, double* rhs);
int64_t spkhip_lulsolve_f32(int64_t nsuper, const int64_t* xsuper,
                            const int64_t* xlindx, const int64_t* lindx,
                            const int64_t* xlnz, const float* lnz,
                            const int64_t* ipiv, float* rhs);
/* SpkLUFactor.jl:325 _luusolve! */
int64_t spkhip_luusolve_f64(int64_t n, int64_t nsuper, const int64_t* xsuper,
                            const int64_t* xlindx, const int64_t* lindx,
                            const int64_t* xlnz, const double* lnz,
                            const int64_t* xunz, const double* unz,
                            double* rhs);
int64_t spkhip_luusolve_f32(int64_t n, int64_t nsuper, const int64_t* xsuper,
                            const int64_t* xlindx, const int64_t* lindx,
                            const int64_t* xlnz, const float* lnz,
                            const int64_t* xunz, const float* unz, float* rhs);
/* SpkLDLtFactor.jl:58 _ldltfactor! (corrected numerics, same layout) */
int64_t spkhip_ldltfactor_f64(int64_t n, int64_t nsuper, const int64_t* xsuper,
                              const int64_t* snode, const int64_t* xlindx,
                              const int64_t* lindx, const int64_t* xlnz,
                              double* lnz);
int64_t spkhip_ldltfactor_f32(int64_t n, int64_t nsuper, const int64_t* xsuper,
                              const int64_t* snode, const int64_t* xlindx,
                              const int64_t* lindx, const int64_t* xlnz,
                              float* lnz);
/* SpkLDLtFactor.jl:266 _ldltsolve! */
int64_t spkhip_ldltsolve_f64(int64_t nsuper, const int64_t* xsuper,
                             const int64_t* xlindx, const int64_t* lindx,
                             const int64_t* xlnz, const double* lnz,
                             double* rhs);
int64_t spkhip_ldltsolve_f32(int64_t nsuper, const int64_t* xsuper,
                             const int64_t* xlindx, const int64_t* lindx,
                             const int64_t* xlnz, const float* lnz,
                             float* rhs);

/* Per-kernel-class profile: with profiling on, every launch of the numeric
 * phase is bracketed by HIP events on the handle's stream and its
 * algorithmic work (flops and bytes the algorithm must move) is added to its
 * class. No counterpart in the reference (it has no device); used by the
 * benchmark's roofline report. */
#define SPKHIP_K_ASSEMBLE 0   /* zero lnz/unz + scatter A (inmatrix)     */
#define SPKHIP_K_SMALL 1      /* whole small front in LDS                */
#define SPKHIP_K_INIT 2       /* large front: load panels + extend-add   */
#define SPKHIP_K_DIAG 3       /* large front: diagonal region of a block */
#define SPKHIP_K_UPDATE 4     /* large front: block trailing update      */
#define SPKHIP_K_WRITEBACK 5  /* large front: panels back to lnz/unz     */
#define SPKHIP_K_FWD 6        /* forward solve level                     */
#define SPKHIP_K_BWD 7        /* backward solve level                    */
#define SPKHIP_K_RESIDUAL 8   /* r = b - A x (refinement)                */
#define SPKHIP_K_OFFDIAG 9    /* large front: L rows below, U rows right */
#define SPKHIP_K_FWD_GEMV 10  /* forward solve: rows below a block       */
#define SPKHIP_K_BWD_GEMV 11  /* backward solve: U rows right of a block */
#define SPKHIP_K_DIAG_BORDER 12 /* diagonal regions beside the previous  */
                                /* block's update within its group       */
#define SPKHIP_KCLASS_COUNT 13
typedef struct spkhip_kstat {
  int64_t launches;
  double ms, flops, bytes;
} spkhip_kstat_t;
int64_t spkhip_profile_enable(spkhip_t h, int64_t on);
int64_t spkhip_profile_get(spkhip_t h, int64_t kclass, spkhip_kstat_t* out);

/* Plan knobs of the numeric phase and the autotuner that picks them. No
 * counterpart in the reference (its dense kernels have no tiles to choose).
 *   small_lmax      fronts (fundamental supernodes) with at most this many
 *                   rows are factored whole in LDS, larger ones block by
 *                   block on the matrix cores: the small-versus-MFMA split
 *                   (0..64, default 64)
 *   block_cols      widest block of consecutive pieces of a front, the
 *                   tile width of the diagonal, off-diagonal and update
 *                   kernels (32..128, default 64; SPD blocks at most 96)
 *   upd_group       blocks per wide trailing update, i.e. its depth in
 *                   blocks (1..64, default 8)
 *   small_wave_min  small fronts in a level from which one wave (instead of
 *                   one workgroup) takes a front (>= 1, default 16384)
 *   factor_ms       autotune: the best measured factor time; ignored by set
 * Every setting in range gives a correct factor. set rebuilds the launch
 * plan when a knob changes (the factor must then be recomputed: -3 from the
 * solves until it is). autotune needs values (inmatrix); it times
 * spkhip_factor at `precision` over a coordinate search of the knobs (one
 * untimed factor per plan, then the minimum of `reps` timed ones, reps 1..10),
 * keeps the fastest, and leaves the handle factored with it. */
typedef struct spkhip_tuning {
  int64_t small_lmax, block_cols, upd_group, small_wave_min;
  double factor_ms;
} spkhip_tuning_t;
int64_t spkhip_get_tuning(spkhip_t h, spkhip_tuning_t* out);
int64_t spkhip_set_tuning(spkhip_t h, const spkhip_tuning_t* t);
int64_t spkhip_autotune(spkhip_t h, int64_t precision, int64_t reps, spkhip_tuning_t* out);

/* Library/device information. */
int64_t spkhip_device_count(void);
const char* spkhip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SPARSPAK_HIP_H */
