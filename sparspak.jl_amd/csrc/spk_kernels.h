// Device kernels of the numeric phase (gfx950, wave64).
//
// Storage model. The reference keeps the factor as per-supernode panels
// (SpkSparseBase.jl:253-284): lnz holds a column-major jlen x nj panel per
// supernode (rows = the supernode's index list, the top nj x nj block is the
// LAPACK-packed L11\U11), unz holds U12^T as a (jlen-nj) x nj panel. The
// reference splits every fundamental supernode into pieces of <= 30 (LU) or
// <= 60 (SPD) columns; pivoting is confined to each piece's diagonal block.
//
// On the device the unit of work is a FRONT: one fundamental supernode,
// i.e. a maximal chain of pieces whose index lists nest (piece k+1's list is
// piece k's list without its own columns). A front with index list I (length
// L) and W own columns is factored as a dense L x L column-major matrix: its
// first W columns/rows are the pieces' panels, the trailing R x R block
// (R = L - W) receives the Schur complement (update matrix) that the
// front's parent adds into itself through a relative-index map. Fronts are
// processed one elimination-tree level at a time, all fronts of a level in
// the same launches. Results are written back into the reference's lnz /
// unz / ipvt layout, which the solves read.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace spk {

constexpr int kWG = 256;      // threads per workgroup (4 waves)
constexpr int kSmallLmax = 64; // largest front the LDS small-front kernels take

// Device view of the symbolic factor and the front schedule.
struct DevSym {
  int64_t n, nsuper, nfront;
  const int64_t* xsuper;  // nsuper+1, 1-based values
  const int64_t* snode;   // n, 1-based values
  const int64_t* xlnz;    // n+1, 1-based values
  const int64_t* xunz;    // n+1, 1-based values
  const int64_t* xlindx;  // nsuper+1, 1-based values
  const int32_t* lindx;   // nsub, 1-based values
  // per front
  const int32_t* fL;      // index-list length
  const int32_t* fW;      // own columns
  const int32_t* fc0;     // first column (0-based)
  const int32_t* fsp0;    // first piece (0-based supernode)
  const int32_t* fnp;     // number of pieces
  const int64_t* foff;    // arena offset of the L x L work matrix (large
                          // fronts) — unused for small fronts
  const int64_t* fcb;     // arena offset of the update matrix
  const int32_t* fcbld;   // its leading dimension
  const int64_t* frel;    // offset of the front's relative-index map
  const int32_t* relmap;  // position in the parent's index list
  const int32_t* fchp;    // children CSR pointer (nfront+1)
  const int32_t* fch;     // children
  const int64_t* fvec;    // offset in the solve vector arena (L per rhs)
  const int64_t* lcol;    // per column g: lnz offset of front row 0, i.e.
                          // L(front row r, g) = lnz[lcol[g] + r]
  const int64_t* ucol;    // per row g: U(g, front column j) = unz[ucol[g] + j]
  int pivot;              // 1: pivot inside diagonal blocks (LU); 0: none
  int sym;                // 1: U entries are read from the transposed lnz slot
                          //    (SPD input: lower triangle only)
  // value sets: one pattern, several sets of values, set v = blockIdx.y of
  // every launch; element strides between consecutive sets' data (ipvt: n;
  // right-hand sides: ldr * nrhs; solve vectors: vs_vec per right-hand side)
  int64_t vs_lnz, vs_unz, vs_arena, vs_prep, vs_prepi, vs_inv, vs_vec;
};

// Per-launch descriptor of the blocked large-front kernels (one per front
// taking part in the launch).
struct BlkDesc {
  int32_t front;
  int32_t piece;     // first piece (0-based supernode) of the block
  int32_t np;        // pieces in the block
  int32_t c0, c1;    // block columns [c0, c1), front-local
  int32_t wg_begin;  // prefix sum of workgroups over the launch
  int32_t nl;        // off-diagonal kernel: workgroups taking rows below
  int32_t pad;
  int32_t fk0, fk1;  // diagonal kernel: previous block's columns folded in
                     // first (fk1 > fk0), its update running alongside
  // the front's fields and the widths of the block's pieces (at most eight),
  // so that a kernel reads no symbolic array before its tiles
  int32_t L;         // the front's rows
  int32_t pad2;
  int64_t fc0;       // the front's first column
  int64_t foff;      // the front's offset in the arena
  int32_t pw[8];     // widths of pieces 0 .. np-1
};

// Per-launch descriptor of the trailing-update kernel: one rectangle of one
// front, F[r0:r0+nr, j0:j0+nc] -= F[r0:r0+nr, k0:k1] F[k0:k1, j0:j0+nc],
// in 128 x 128 tiles.
struct GemmDesc {
  int32_t front;
  int32_t k0, k1;    // inner dimension: front columns / rows [k0, k1)
  int32_t r0, nr;    // output rows
  int32_t j0, nc;    // output columns
  int32_t wg_begin;  // prefix sum of workgroups over the launch
  int32_t ntr;       // row tiles
  int32_t big;       // the front has more than 256 own columns
  int32_t mirror;    // symmetric front: entries whose row lies past their
                     // tile's last column also write their transposes
                     // (no tile above the diagonal is computed)
};

// The large-front assembly's items, resolved on the host so that its kernels
// read no symbolic array before the column: one per (front f, column j)...
struct InitItem {
  int64_t fcol;  // arena offset of column j of the front's work matrix
  int64_t lcb;   // own column: lnz offset of front row 0 (as DevSym::lcol);
                 // -1: an update column
  int64_t fc0;   // the front's first column
  int32_t L, W;  // the front's rows and own columns
  int32_t r1;    // first row of the column's L part (its piece's first
                 // column); W for an update column
  int32_t j;
};
// ... and one per (child, update column) pair added into a column.
struct EaItem {
  int64_t col;  // arena offset of the child's update-matrix column
  int64_t rel;  // offset of the child's relative-index map
  int32_t Rc;   // the child's update rows
  int32_t pad;
};

// Integers of a block's solve data: the composed row permutation.
constexpr int kPrepInts = 128;

// Per-workgroup descriptor of the blocked solve kernels.
struct SolveDesc {
  int32_t front;
  int32_t piece;   // first piece of the block
  int32_t np;      // pieces in the block
  int32_t c0, c1;  // block columns [c0, c1), front-local
  int32_t r0;      // gemv: first row of the workgroup's tile (relative);
                   // chain and prep: the block's index in the solve data
  int32_t aux;     // chain and prep: the block's solve data starts at
                   // element 4 * aux
  int32_t L, W;    // the front's index-list length and own columns
  int64_t fc0;     // the front's first column
  int64_t fvec;    // the front's offset in the solve vector arena
  int64_t ucol;    // backward GEMV: s.ucol of its row (unz offset)
  int64_t inv;     // solve data (chain descriptors): element offset of the
                   // block's two explicit inverses (float factors)
};

// One workgroup of an explicit-inverse block step (fwd_inv_kernel,
// bwd_inv_kernel), with every field it needs (no dependent descriptor load).
// The block whose new entries it forms from that block's inverse: columns
// [x0, x1) (x1 == x0: none), inverse at element inv. Forward: the tile's
// first row below the block, c1 + r0 (r0 < 0: no tile). Backward: the row
// c of its GEMV (c < 0: none), its U entries from ucol, summed from column
// rc1. write: it stores the entries it formed (forward: the right-hand
// side; backward: the front's vector); fin: backward, it stores them in the
// right-hand side and copies the front's other columns [x1, W) there.
// Forward items also carry the block's pieces: piece q starts at block
// column pst[q] (pst[np..7] past the block), its first column's lnz offset
// in front coordinates is pbs[q] (as DevSym::lcol), and its columns follow
// each other at the stride L - (x0 + pst[q]), its row count.
struct InvItem {
  int64_t fc0, fvec, inv, ucol;
  int32_t L, W, x0, x1, r0, c, rc1, flags;  // flags: 1 write, 2 fin
  int32_t pst[8];
  int64_t pbs[8];
};

}  // namespace spk
