// kernels.h -- launchers of the gfx950 kernels of the pressure-solve path.
// All launches go on the given stream; all indices are 0-based int32.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace ufvm {
namespace k {

constexpr int kBlock = 256;
constexpr int kFoldBlocks = 512;   // grid cap of the grid-stride FCG kernels
constexpr int kFcgBlock = 1024;    // threads per block of the FCG kernels
constexpr int kSweepBlock = 1024;  // threads of the one-workgroup level-scheduled sweeps
inline int nblocks(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

// The GPU hands workgroup b to XCD b % 8 (each XCD has its own L2).  This
// maps it to the b-th of a contiguous share of the grid per XCD instead, so
// that rows next to each other -- and the neighbour values they read -- meet
// in one L2.  A bijection of [0, G) for any G.
constexpr int kXcds = 8;
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t G) {
  const int32_t q = G / kXcds, r = G % kXcds, x = b % kXcds, i = b / kXcds;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Rows of the FP64 operator product (k_spmv_dot and its batched form): block
// b of nb takes spmv_trips(n, nb) consecutive kFcgBlock-row chunks from chunk
// xcd_block(b, nb) * trips, so each XCD takes one contiguous share of the
// rows and the neighbour values that share gathers stay in its own L2 (with
// the rows interleaved over the XCDs, every L2 fetched most of the vector).
__host__ __device__ __forceinline__ int32_t spmv_trips(int32_t n, int32_t nb) {
  const int32_t chunks = (n + kFcgBlock - 1) / kFcgBlock;
  return (chunks + nb - 1) / nb;
}

// Column indices of a sliced-ELL copy: one int32 per entry (col), or, when
// every 64-row chunk's columns lie within 65536 of the chunk's smallest, a
// uint16 offset per entry (d16) from a base per chunk (base).  Exactly one
// form is set; padding entries point at a valid row with value 0.
struct SellCols {
  const int32_t *col = nullptr;
  const uint16_t *d16 = nullptr;
  const int32_t *base = nullptr;
};

// Device view of chunk q's column indices: col[e] is the column of entry e
// (an index into the level's entry stream that lies in chunk q).
template <bool PK>
struct ColSrc;
template <>
struct ColSrc<false> {
  const int32_t *c;
  __device__ __forceinline__ ColSrc(const SellCols &s, int32_t) : c(s.col) {}
  __device__ __forceinline__ int32_t operator[](int32_t e) const { return c[e]; }
  __device__ __forceinline__ int32_t nt(int32_t e) const { return __builtin_nontemporal_load(c + e); }
};
template <>
struct ColSrc<true> {
  const uint16_t *d;
  int32_t b;
  __device__ __forceinline__ ColSrc(const SellCols &s, int32_t q) : d(s.d16), b(s.base[q]) {}
  __device__ __forceinline__ int32_t operator[](int32_t e) const { return b + (int32_t)d[e]; }
  __device__ __forceinline__ int32_t nt(int32_t e) const {
    return b + (int32_t)__builtin_nontemporal_load(d + e);
  }
};
// entries per run of the sliced-ELL row loops (loads issued together)
constexpr int kSellRun = 8;

// f(std::bool_constant<packed>): launchers pick the kernel of cols' form
template <typename F>
inline void with_cols(const SellCols &cols, F &&f) {
  if (cols.d16)
    f(std::true_type{});
  else
    f(std::false_type{});
}

// ---- reference-order FP64 kernels (bit-exact with the reference arithmetic)
// r = bc - ac*x - sum_k anb_k x_col(k) (cconn order); partial[blk] = sum |r|
// device-side stopping test of a smoothing call: phase 0 (initial residual)
// or 1 (after a sweep) on st[0..3] = initial, final, stop, NaN seen
void alg_test(hipStream_t s, const double *partial, int32_t np, int32_t n, double rrf, int phase,
              double *st);
void residual_ref(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col,
                  const double *ac, const double *anb, const double *bc, const double *x,
                  double *r, double *partial, const double *stop = nullptr);
// out[0] = sum(partial[0..np)) in a fixed order
void reduce_final(hipStream_t s, const double *partial, int32_t np, double *out);
// Gauss-Seidel update of the listed rows (one level of a lexicographic sweep)
void gs_rows(hipStream_t s, const int32_t *rows, int32_t count, const int32_t *ptr,
             const int32_t *col, const double *ac, const double *anb, const double *b,
             double *x, const double *stop = nullptr);
// A whole level-scheduled sweep in one workgroup (levels lp[q] .. lp[q+1]
// of rows, device arrays, a barrier between levels): for schedules of many
// small levels, one launch instead of one per level, same arithmetic
void gs_sweep_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    const int32_t *ptr, const int32_t *col, const double *ac, const double *anb,
                    const double *b, double *x, const double *stop = nullptr);
// The same sweep pipelined (see kernels.hip) for schedules whose levels hold
// at most kPipeBlock rows, over a schedule-ordered ELL copy of the level:
// slot u of position s at u * n + s (w slots; ecol static, past a row's
// length column 0; elen the row lengths), eanb / eac refreshed from
// the CSR values by pipe_values before the sweeps; the level pointers in LDS
// (gs_sweep_pipe_lds bytes).  Same arithmetic.
constexpr int kPipeBlock = 256;
size_t gs_sweep_pipe_lds(int32_t nlev);
void pipe_values(hipStream_t s, const int32_t *rows, int32_t n, int32_t w, const int32_t *ptr,
                 const double *anb, const double *ac, double *eanb, double *eac);
void gs_sweep_pipe(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev, int32_t n,
                   int32_t w, const int32_t *ecol, const int32_t *elen, const double *eanb,
                   const double *eac, const double *b, double *x, const double *stop = nullptr);
// The DILU forward / backward substitutions pipelined the same way, over ELL
// copies in each schedule's order (forward: the lower-neighbour lists;
// backward: the entries with column above the row); eval = anb[esrc]
// refreshed by pipe_gather; y (n values) holds dinv_i * rc_i of the forward
// sweep.  Same arithmetic as dilu_fwd_row / dilu_bwd_row.
void pipe_gather(hipStream_t s, int64_t m, const int32_t *esrc, const double *anb, double *eval);
void dilu_pipe(hipStream_t s, bool backward, const int32_t *rows, const int32_t *lp, int32_t nlev,
               int32_t n, int32_t w, const int32_t *ecol, const int32_t *elen, const double *eval,
               const double *dinv, double *rc, double *y, double *x,
               const double *stop = nullptr);
// The DILU factor pipelined the same way over an ELL copy of the lower lists
// (forward-schedule order, the skipped neighbour left out); eji / eij the
// slots' a_ji / a_ij.  Same arithmetic as dilu_factor_row.
void dilu_factor_pipe(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                      int32_t n, int32_t w, const int32_t *ecol, const int32_t *elen,
                      const double *eji, const double *eij, const double *ac, double *dinv);
// The SOR forward and backward sweep of a small level held whole in LDS and
// run by one wave (see kernels.hip): rows of at most w <= 16 entries, n below
// 65,536, sor_staged_lds bytes of LDS.  Returns false if the launch failed
// (then nothing ran).  Same arithmetic.
size_t sor_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk);
// A whole SOR smoothing call with the device-side stopping test (initial
// residual and test, up to n_iter sweeps each with its residual and test; st
// and r as the separate kernels leave them) on such a level, in one launch
// of sor_smooth_staged_lds bytes of LDS.  Returns false if the launch failed.
size_t sor_smooth_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk);
bool sor_smooth_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                       const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                       const int32_t *ptr, const int32_t *col, const double *ac,
                       const double *anb, const double *b, double *x, double *r, int32_t n_iter,
                       double rrf, double *st);
bool sor_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                const int32_t *ptr, const int32_t *col, const double *ac, const double *anb,
                const double *b, double *x, const double *stop = nullptr);
// The DILU substitutions of a small level held whole in LDS and run by one
// wave (see kernels.hip): lower lists and entries above the diagonal of at
// most w <= 16 per row, n below 65,536, dilu_staged_lds bytes of LDS.
// Returns false if the launch failed (then nothing ran).  Same arithmetic.
size_t dilu_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk);
bool dilu_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                 const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                 const int32_t *ptr, const int32_t *col, const double *anb, const int32_t *lptr,
                 const int32_t *lnb, const int32_t *lpji, const double *dinv, double *rc,
                 double *x, const double *stop = nullptr);
// The DILU factor of such a level in LDS, one wave (dilu_factor_row's
// arithmetic; dilu_factor_staged_lds bytes).  Returns false if the launch
// failed (then nothing ran).
size_t dilu_factor_staged_lds(int32_t n, int32_t w, int32_t nf);
bool dilu_factor_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                        int32_t n, int32_t w, const double *ac, const double *anb,
                        const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                        const int32_t *lpij, int32_t skip_row, double *dinv);
// the same with the level's unknowns (x, rc, dinv) in LDS: n <= kLdsRows
// (48 KB of doubles, inside the 64 KB any workgroup may allocate)
constexpr int32_t kLdsRows = 6144;
void sor_lds(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf, const int32_t *br,
             const int32_t *blp, int32_t nbk, int32_t n, const int32_t *ptr, const int32_t *col,
             const double *ac, const double *anb, const double *b, double *x,
             const double *stop = nullptr);
void dilu_lds(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf, const int32_t *br,
              const int32_t *blp, int32_t nbk, int32_t n, const int32_t *ptr, const int32_t *col,
              const double *anb, const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
              const double *dinv, double *rc, double *x, const double *stop = nullptr);
void dilu_factor_lds(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                     int32_t n, const double *ac, const double *anb, const int32_t *lptr,
                     const int32_t *lnb, const int32_t *lpji, const int32_t *lpij,
                     int32_t skip_row, double *dinv);
void dilu_factor_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                       const double *ac, const double *anb, const int32_t *lptr,
                       const int32_t *lnb, const int32_t *lpji, const int32_t *lpij,
                       int32_t skip_row, double *dinv);
template <typename V>
void dilu_fwd_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    const V *anb, const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                    const V *dinv, V *rc, const double *stop = nullptr);
template <typename V>
void dilu_bwd_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    int32_t n, const int32_t *ptr, const int32_t *col, const V *anb,
                    const V *dinv, V *rc, V *x, const double *stop = nullptr);
// DILU in colour order on a sliced-ELL level whose colours are contiguous
// row ranges: forward pass of the colour rows [r0, r1) (rc from b - A x and
// the earlier colours; zero: x = 0) and backward pass (z from the later
// colours, x += z; zero: x = z).  ILU(0): lval = the factor's off-diagonals
// (the lower sum; nullptr for DILU, where they are A's), and the backward
// pass gets the factor as val.
template <typename V>
void mcilu_fwd_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off,
                    const SellCols &cols, const V *val, const V *lval, const V *ac, const V *b, const V *x,
                    const V *dinv, V *rc, bool zero);
template <typename V>
void mcilu_bwd_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off,
                    const SellCols &cols, const V *val, const V *dinv, V *rc, V *x, bool zero);
// ILU(0) factorisation of the rows [r0, r1) of one colour class (the factor's
// off-diagonals overwrite F, which holds A's on entry; update lists Ilu0Lists)
void ilu0_factor_rows(hipStream_t s, int32_t r0, int32_t r1, const double *ac,
                      const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                      const int32_t *hptr, const int32_t *hsrc, const int32_t *hdst, double *F,
                      double *dinv);
// DILU factorisation of the listed rows (pull form of cfdFactorizeILU)
void dilu_factor_rows(hipStream_t s, const int32_t *rows, int32_t count, const double *ac,
                      const double *anb, const int32_t *lptr, const int32_t *lnb,
                      const int32_t *lpji, const int32_t *lpij, int32_t skip_row, double *dinv);
// DILU substitution sweeps over one level of the lexicographic schedule
// (double: the reference arithmetic; float: the FP32 precision-form smoother)
template <typename V>
void dilu_fwd_rows(hipStream_t s, const int32_t *rows, int32_t count, const V *anb,
                   const int32_t *lptr, const int32_t *lnb, const int32_t *lpji, const V *dinv,
                   V *rc, const double *stop = nullptr);
template <typename V>
void dilu_bwd_rows(hipStream_t s, const int32_t *rows, int32_t count, int32_t n,
                   const int32_t *ptr, const int32_t *col, const V *anb, const V *dinv, V *rc,
                   V *x, const double *stop = nullptr);
// bc_c[p] = sum over children (increasing fine index) of r
void restrict_ref(hipStream_t s, int32_t nc, const int32_t *cptr, const int32_t *child,
                  const double *r, double *bc_c);
// x[i] += xc[parent[i]]
void prolong_ref(hipStream_t s, int32_t n, const int32_t *parent, const double *xc, double *x);
// the reference level as colour-class sliced-ELL chunks: r = b - A x of
// every row (per-block |r| sums) and Gauss-Seidel over the chunks [q0, q1)
void residual_csell(hipStream_t s, int64_t nslot, const int32_t *row, const int32_t *off,
                    const int32_t *col, const double *val, const double *ac, const double *bc,
                    const double *x, double *r, double *partial, const double *stop = nullptr);
void gs_csell(hipStream_t s, int32_t q0, int32_t q1, const int32_t *row, const int32_t *off,
              const int32_t *col, const double *val, const double *ac, const double *b,
              double *x, const double *stop = nullptr);
// Galerkin coarse operator in the reference accumulation order, from the
// gather lists of galerkin_lists (nout = nc + coarse nnz values)
void coarse_lhs(hipStream_t s, int32_t nout, int32_t nc, const int32_t *ptr, const int32_t *src,
                const double *fac, const double *fanb, double *cac, double *canb);

// ---- multicolour / precision-form kernels (V = float or double)
// sliced-ELL (64-row chunks, column-major inside a chunk) variants of the
// colour pass, the residual and the damped Jacobi sweep
template <typename V>
void mcgs_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off, const SellCols &cols,
               const V *val, const V *ac, const V *b, V *x, int32_t zero_above);
// rows below zero_rows get r = 0 without computing it.  wide: every entry of a
// row loaded in one round (levels whose rows are longer than a run)
template <typename V>
void residual_sell(hipStream_t s, int32_t n, int32_t zero_rows, const int32_t *off,
                   const SellCols &cols, const V *val, const V *ac, const V *b, const V *x, V *r,
                   bool wide = false);
// parent/xc not null: the sweep starts from x + xc[parent] (fused prolongation)
template <typename V>
void jacobi_sell(hipStream_t s, int32_t n, const int32_t *off, const SellCols &cols,
                 const V *val, const V *ac, const V *b, const V *x, const int32_t *parent, const V *xc, V *out,
                 V omega, bool wide = false);
// The finest level with FP16 off-diagonals scaled by their row's diagonal
// (s_ij = a_ij / a_ii; FP32 arithmetic): colour pass x_i = b_i / a_ii -
// sum s_ij x_j, residual r_i = b_i + a_ii (-x_i - sum s_ij x_j); scale_half
// fills the FP16 values from the FP64 copy (row by row over the chunks)
void mcgs_sell_h(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off, const SellCols &cols,
                 const _Float16 *val, const float *ac, const float *b, float *x,
                 int32_t zero_above);
void residual_sell_h(hipStream_t s, int32_t n, int32_t zero_rows, const int32_t *off,
                     const SellCols &cols, const _Float16 *val, const float *ac, const float *b,
                     const float *x, float *r);
void scale_half(hipStream_t s, int32_t n, const int32_t *off, const double *val, const double *ac,
                _Float16 *out16);
// block-local multicolour GS over blocks of at most kBRows rows (see kernels.hip)
constexpr int kBRows = 1024;
template <typename V>
void bmcgs(hipStream_t s, int32_t nblk, const int32_t *bptr, const int32_t *boff, int32_t ncol,
           const int32_t *ptr, const int32_t *col, const V *ac, const V *anb, const V *b,
           const V *x_in, const int32_t *parent, const V *xc, V *x_out, int mode, int nsweep);
// restriction over sliced-ELL child lists (off per 64 coarse rows, padding
// -1), bc = scale * sum, alone or fused with a Jacobi sweep from zero on the
// coarse level
template <typename V>
void restrict_jacobi(hipStream_t s, int32_t nc, const int32_t *off, const int32_t *child,
                     const V *r, const V *acc, V *bc, V *xc, V omega, V scale);
template <typename V>
void residual_fast(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col, const V *ac,
                   const V *anb, const V *b, const V *x, V *r);
template <typename V>
void restrict_fast(hipStream_t s, int32_t nc, const int32_t *off, const int32_t *child,
                   const V *r, V *bc, V scale);
template <typename V>
// x[i] += xc[parent[i]] for rows r0 <= i < n
void prolong_fast(hipStream_t s, int32_t r0, int32_t n, const int32_t *parent, const V *xc,
                  V *x);
template <typename V>
void dense_solve(hipStream_t s, int32_t n, const double *ainv, const V *b, V *x);

template <typename V>
void fill(hipStream_t s, int64_t n, V *x, V v);
template <typename A, typename B>
void convert(hipStream_t s, int64_t n, const A *a, B *b);
void permute_gather(hipStream_t s, int32_t n, const int32_t *idx, const double *src, double *dst);
void permute_scatter(hipStream_t s, int32_t n, const int32_t *idx, const double *src,
                     double *dst);

// dense Gauss-Jordan inverse in place (row-major n x n), no pivoting, in
// blocks of 64 columns; scratch holds gj_scratch(n) doubles
size_t gj_scratch(int32_t n);
void gj_inverse(hipStream_t s, int32_t n, double *a, double *scratch);
// a = csr (+ shift*11^T) as dense row-major
void csr_to_dense(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col,
                  const double *ac, const double *anb, double shift, double *a);

// ---- FPCG (FP64, level-1 order of the multicolour permutation).  Each call
// is a streaming kernel writing per-block partial sums plus a one-block kernel
// reducing them and updating the device scalars scal[] (layout in kernels.hip).
// out[i] = idx[i] >= 0 ? (V) src[idx[i]] : 0
template <typename V>
void gather_vals(hipStream_t s, int64_t n, const int32_t *idx, const double *src, V *out);
// ring[0] = *src, ring[1] = -1 (a device scalar handed to host-visible memory)
void ring_put(hipStream_t s, const double *src, double *ring);
// FCG(1) iteration kernels (scalar slots of scal[] listed in kernels.hip).
// w = A p (sliced-ELL finest level); partial[0, g) = per-block p.w
void spmv_dot(hipStream_t s, int32_t n, const int32_t *off, const SellCols &cols,
              const double *ac, const double *val, const double *p, double *w, double *partial);
// Strides between the right-hand sides of a batch of k solves (the FCG
// kernels below run the k solves as grid rows blockIdx.y, each exactly as a
// single solve): FP64 vectors, V-cycle-precision vectors, partials, device
// scalars and ring slots.  All zero for one solve.
struct BatchStrides {
  int64_t ld = 0, ldv = 0, ps = 0, ss = 0, rs = 0;
};
// partial[c*g + b] = per-block sums of z, z.w, z.r, w, r
template <typename V>
void fcg_dots(hipStream_t s, int32_t n, const V *z, const double *w, const double *r,
              double *partial, int k = 1, BatchStrides bs = {});
// folds fcg_dots' partials (m, rho, beta into scal) ; p = (z - m) - beta p
template <typename V>
void fcg_update_p(hipStream_t s, int32_t n, const V *z, const double *partial, int singular,
                  int first, double *scal, double *p, int k = 1, BatchStrides bs = {});
// blocks of the FCG streaming kernels (each writes one partial per block)
int fold_blocks(int32_t n);

// ---- K right-hand sides of one system at once (kernels_batch.hip): vectors
// RHS-major (vector r at offset r * ld), k in {1, 2, 4, 8}; each vector gets
// the arithmetic of the single-vector kernel above
template <typename V>
void mcgs_sell_b(hipStream_t s, int k, int32_t r0, int32_t r1, const int32_t *off,
                 const SellCols &cols, const V *val, const V *ac, const V *b, V *x, int64_t ld,
                 int32_t zero_above);
template <typename V>
void residual_sell_b(hipStream_t s, int k, int32_t n, int32_t zero_rows, const int32_t *off,
                     const SellCols &cols, const V *val, const V *ac, const V *b, const V *x,
                     V *r, int64_t ld);
// parent / xc null: no prolongation fused
template <typename V>
void jacobi_sell_b(hipStream_t s, int k, int32_t n, const int32_t *off, const SellCols &cols,
                   const V *val, const V *ac, const V *b, const V *x, int64_t ld,
                   const int32_t *parent, const V *xc, int64_t ldc, V *out, V omega);
// acc null: plain restriction, else fused with the first Jacobi sweep (xc)
template <typename V>
void restrict_b(hipStream_t s, int k, int32_t nc, const int32_t *off, const int32_t *child,
                const V *r, int64_t ldf, const V *acc, V *bc, V *xc, int64_t ldc, V omega,
                V scale);
template <typename V>
void prolong_b(hipStream_t s, int k, int32_t r0, int32_t n, const int32_t *parent, const V *xc,
               int64_t ldc, V *x, int64_t ld);
template <typename V>
void dense_solve_b(hipStream_t s, int k, int32_t n, const double *ainv, const V *b, int64_t ld,
                   V *x);
// w_r = A p_r and the p_r.w_r block partials at partial + r * pstride
void spmv_dot_b(hipStream_t s, int k, int32_t n, const int32_t *off, const SellCols &cols,
                const double *ac, const double *val, const double *p, double *w, int64_t ld,
                double *partial, int64_t pstride);
// folds spmv_dot's partials (delta, alpha into scal) ; x += alpha p ;
// r -= alpha w ; rv = (V) r ; then rr = r.r into scal[10], the convergence
// flag scal[11] once rr <= scal[13] (no update at all once it is set), and
// {rr, iteration} into ring slot (scal[14] & 1) when ring is not null.  Rows
// i < n0 also get z0[i] = rv[i] / acv[i] (the next V-cycle's first colour pass
// from zero).
template <typename V>
void fcg_update_xr(hipStream_t s, int32_t n, double *scal, const double *p, const double *w,
                   double *x, double *r, V *rv, double *partial, double *ring, int32_t n0,
                   const V *acv, V *z0, int k = 1, BatchStrides bs = {});
// arm a solve: scal[11] = 0 (flag), scal[13] = thr, scal[14] = 0 (iteration)
void fcg_arm(hipStream_t s, double *scal, double thr, int done = 0);
// reduce nblk partial groups of width `width` into out[0..width)
void reduce_groups(hipStream_t s, const double *partial, int32_t nblk, int32_t width,
                   double *out);
// r.r and |b| helpers
void dot_partial(hipStream_t s, int32_t n, const double *a, const double *b, double *partial);
void sum_partial(hipStream_t s, int32_t n, const double *a, double *partial);
void axpy_const(hipStream_t s, int32_t n, const double *scal_slot, double mul, double *x);

// ---- grouped launches (serving form).  One launch covers the same phase of
// g solves of different systems: blockIdx.y selects the member, whose
// arguments are entry y of a device table; blocks past the member's own
// block count nb return, and so does every block of a member whose solve
// has converged (*done != 0, the solve's flag scal[11]).  Each member keeps
// the rows, block decomposition and reduction order of its own single-solve
// launch, so its result keeps the bits of its own solve.
enum GOp : int32_t {
  G_MCGS,          // colour class [r0, n) Gauss-Seidel pass; aux = zero_above; out = x
  G_RESID,         // r = b - A x, rows < aux set to 0; out = r
  G_JACOBI,        // damped Jacobi (omega); parent != null: x + xc[parent] as input
  G_RESTRICT,      // bc = scale * sum of children (off/child: child lists, x: fine r);
                   // ac != null: also out2 = omega bc / ac (first Jacobi sweep)
  G_PROLONG,       // x[i] += xc[parent[i]] for rows [r0, n); out = x
  G_DENSE,         // x = Ainv b (FP64 inverse, one wavefront per row); out = x
};
template <typename V>
struct GArgs {
  int32_t n, r0, aux, nb;
  int32_t wide;                    // G_RESID / G_JACOBI: rows loaded in one round
  const int32_t *off, *child, *parent;
  SellCols cols;
  const V *val, *ac, *b, *x, *xc;
  V *out, *out2;
  const double *ainv;
  V omega, scale;
  const double *done;
};
enum GFcgOp : int32_t { G_DOTS, G_UPDATE_P, G_UPDATE_P_FIRST, G_SPMV, G_UPDATE_XR, G_RR };
template <typename V>
struct GFcgArgs {
  int32_t n, nb, n0, singular;   // rows, fold_blocks(n), colour-0 rows, constant nullspace
  const V *z, *acv;              // preconditioned residual; finest diagonal (V-cycle copy)
  V *rv, *z0;                    // V-cycle right-hand side and iterate of the finest level
  const int32_t *off;            // FP64 finest operator (sliced-ELL)
  SellCols cols;
  const double *ac, *val;
  double *p, *w, *x, *r, *partial, *scal, *ring;
};
// blocks of a dense coarsest solve of n rows (one 64-lane wavefront per row)
inline int32_t dense_blocks(int32_t n) { return (n + kBlock / 64 - 1) / (kBlock / 64); }
// gx: the largest member block count (grid x), g: members (grid y)
template <typename V>
void group_row(hipStream_t s, GOp op, int32_t gx, int32_t g, const GArgs<V> *tab);
template <typename V>
void group_fcg(hipStream_t s, GFcgOp op, int32_t gx, int32_t g, const GFcgArgs<V> *tab);

}  // namespace k
}  // namespace ufvm
