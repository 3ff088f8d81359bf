/*
 * sadmrg.h -- C ABI of the MI355X two-site DMRG inner-loop engine.
 *
 * Plain C: opaque handles, plain pointers and sizes.  Every entry point
 * returns 0 on success or a nonzero status (SADMRG_ERR_*); the message of the
 * last failure on the calling thread is returned by sadmrg_last_error().
 * (The reference prints and calls exit() instead, e.g.
 * src/block_tensor_inst.cpp:2191-2196; there is no C interface in the
 * reference, whose boundary is its C++ class API: block_tensor.hpp:38-193,
 * dense.hpp:26-156, lanczos.hpp:55-91, dmrg.hpp:29-133.)
 *
 * Block tensors cross the boundary in one format:
 *   header  = [nbond, locspin, {dir, nmom, q[nmom], dim[nmom]} x nbond]
 *             q = 2J (SU(2)) or charge (U(1)), sorted ascending per bond
 *   present = one byte per block; blocks are numbered column-major over the
 *             bonds' sectors, bond 0 fastest (src/symmetry/su2.cpp:432-450)
 *   data    = the present blocks, concatenated in block order, each block
 *             column-major (bond 0 fastest, dense_inst.cpp:31-45); complex
 *             data is interleaved (re, im)
 * Device data is FP64 (double or complex<double>).
 */
#ifndef SADMRG_H
#define SADMRG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sadmrg_ctx sadmrg_ctx;       /* one GPU, one stream, caching allocator */
typedef struct sadmrg_cgc sadmrg_cgc;       /* SU(2) recoupling tables */
typedef struct sadmrg_tensor sadmrg_tensor; /* BlockTensor<Scalar,Sym> */
typedef struct sadmrg_heff sadmrg_heff;     /* H_eff matvec plan of one two-site step */
typedef struct sadmrg_dmrg sadmrg_dmrg;     /* DMRG<Scalar,Sym> driver */

enum {
  SADMRG_OK = 0,
  SADMRG_ERR_INVALID = 1,
  SADMRG_ERR_HIP = 2,
  SADMRG_ERR_UNSUPPORTED = 3,
  SADMRG_ERR_NUMERICS = 4,
  SADMRG_ERR_NO_DEVICE = 5
};
enum { SADMRG_U1 = 0, SADMRG_SU2 = 1 };
enum { SADMRG_REAL = 0, SADMRG_COMPLEX = 1 };
/* operand precision of the bulk H_eff products (outputs are FP64):
   SADMRG_BF16 is an ablation (bfloat16 operands, FP32 accumulation) */
enum { SADMRG_FP64 = 0, SADMRG_FP32 = 1, SADMRG_BF16 = 2 };

/* ---- runtime ------------------------------------------------------------ */
const char* sadmrg_last_error(void);
int sadmrg_version(int* major, int* minor);
int sadmrg_ctx_create(int device, sadmrg_ctx** out);
int sadmrg_ctx_destroy(sadmrg_ctx* ctx);
int sadmrg_ctx_sync(sadmrg_ctx* ctx);
int sadmrg_ctx_memory(sadmrg_ctx* ctx, long long* bytes_in_use, long long* bytes_cached);
/* Host-side counters of the context since its creation: device allocations
   that missed the allocator cache (hipMalloc calls) and stream
   synchronisations, each with the wall seconds spent in them (any pointer
   may be NULL).  No reference counterpart. */
int sadmrg_ctx_stats(sadmrg_ctx* ctx, long long* mallocs, double* malloc_seconds, long long* syncs,
                     double* sync_seconds);
/* the context's HIP stream (hipStream_t), for callers that time or order work on it */
int sadmrg_ctx_stream(sadmrg_ctx* ctx, void** stream);

/* ---- SU(2) tables (src/symmetry/su2.cpp:41-75, src/cgc_tables.cpp:102-357) */
/* <j1 m1; j2 m2 | j3 m3>, all arguments doubled (CGCTable::cg) */
int sadmrg_clebsch_gordan(int j1, int m1, int j2, int m2, int j3, int m3, double* out);
int sadmrg_cgc_create(int physpn, sadmrg_cgc** out);
int sadmrg_cgc_destroy(sadmrg_cgc* cgc);
/* fac_hamilt_vec[i][j][k] (cgc_tables.hpp:27) */
int sadmrg_cgc_hamilt_vec(sadmrg_cgc* cgc, int i, int j, int k, double* out);
/* fac_operator_onsite_{left(side 0),rght(1)}[j][m][l][i][p] (cgc_tables.hpp:30-31) */
int sadmrg_cgc_onsite(sadmrg_cgc* cgc, int side, int j, int m, int l, int i, int p, double* out);
/* fac_operator_transformation_{left,rght}[j][m][l][i][k][p] (cgc_tables.hpp:34-35) */
int sadmrg_cgc_transformation(sadmrg_cgc* cgc, int side, int j, int m, int l, int i, int k, int p, double* out);
/* fac_operator_pairup_{left,rght}[j][i][k][m][p] (cgc_tables.hpp:38-39) */
int sadmrg_cgc_pairup(sadmrg_cgc* cgc, int side, int j, int i, int k, int m, int p, double* out);

/* ---- block tensors (BlockTensor<Scalar,Sym>, block_tensor.hpp:38-193) ---- */
/* BlockTensor::set(StructType, data, cgc) (block_tensor.hpp:63); cgc2_kind
   selects the CGC of 2-leg SU(2) blocks: 0 identity, 1 singlet */
int sadmrg_tensor_create(sadmrg_ctx* ctx, int sym, int cplx, const int* header, int header_len,
                         const unsigned char* present, const double* data, int cgc2_kind, sadmrg_tensor** out);
/* empty (null) tensor, the state after BlockTensor::clean() (block_tensor.hpp:59) */
int sadmrg_tensor_null(sadmrg_ctx* ctx, int sym, int cplx, sadmrg_tensor** out);
int sadmrg_tensor_destroy(sadmrg_tensor* t);
int sadmrg_tensor_header(const sadmrg_tensor* t, int* header, int cap, int* len);
int sadmrg_tensor_nten(const sadmrg_tensor* t, int* nten);
int sadmrg_tensor_nelem(const sadmrg_tensor* t, long long* nelem);
/* present: nten bytes; data: nelem (x2 if complex) doubles (BlockTensor::get_elems, block_tensor.hpp:99) */
int sadmrg_tensor_download(const sadmrg_tensor* t, unsigned char* present, double* data);
/* BlockTensor::operator= (block_tensor.hpp:107) */
int sadmrg_tensor_copy(sadmrg_tensor* dst, const sadmrg_tensor* src);
int sadmrg_tensor_device_ptr(const sadmrg_tensor* t, void** ptr);

/* BlockTensor::hamiltonian_vector_multiplication (block_tensor.hpp:179) */
int sadmrg_hamiltonian_vector_multiplication(sadmrg_tensor* out, const sadmrg_tensor* vec, const sadmrg_tensor* op1,
                                             const sadmrg_tensor* op2, sadmrg_cgc* cgc);
/* BlockTensor::contract (block_tensor.hpp:142) */
int sadmrg_contract(sadmrg_tensor* out, const sadmrg_tensor* t1, int i1, const sadmrg_tensor* t2, int i2);
/* Environment updates (block_tensor.hpp:163-176).  kind: 0 overlap_initial,
   1 overlap_transformation(uu,vv,op1), 2 operator_initial(uu,vv,op1),
   3 operator_transformation(uu,vv,op1), 4 operator_pairup(uu,vv,op1,op2),
   5 adjoint of overlap_transformation: out = T*(op1) with the structure of
   op2, where T(X) = overlap_transformation(uu,vv,X) (no reference
   counterpart; real tensors). */
int sadmrg_env_update(sadmrg_tensor* out, int kind, const sadmrg_tensor* uu, const sadmrg_tensor* vv,
                      const sadmrg_tensor* op1, const sadmrg_tensor* op2, int flag, sadmrg_cgc* cgc);
/* BlockTensor::svd (block_tensor.hpp:135) with the global max_dcut as an
   argument; wout receives kept then discarded singular values (*nw of them). */
int sadmrg_svd(const sadmrg_tensor* vec, sadmrg_tensor* tu, double p1, sadmrg_tensor* tv, double p2, int max_dcut,
               double* wout, int wout_cap, int* nw);
/* BlockTensor::fuse(BondType, BondType) (block_tensor.hpp:129); bond = [dir, nmom, q..., dim...] */
int sadmrg_fuse(sadmrg_tensor* out, const int* bond1, const int* bond2);
/* BlockTensor::fuse_to_singlet (block_tensor.hpp:131) */
int sadmrg_fuse_to_singlet(sadmrg_tensor* out, const int* bond1, const int* bond2);
/* BlockTensor::make_spinor_start (block_tensor.hpp:156) */
int sadmrg_make_spinor_start(sadmrg_tensor* out, int physpn);
/* In-place transformations.  op: 0 left2right_vectran, 1 right2left_vectran
   (block_tensor.hpp:124-125), 2 take_conjugate(bond=a) (:145), 3
   makeup_input_vector (:187), 4 *= (a + i b) (:110-111), 5 normalize_vector
   (:119), 6 random_initialize_tensor (:189), 7 shift(a, b) (:122),
   8 operator=(double a) (:106). */
int sadmrg_tensor_transform(sadmrg_tensor* t, int op, double a, double b);
/* y += (re + i im) x  (operator+= / -=, block_tensor.hpp:108-109) */
int sadmrg_tensor_axpy(sadmrg_tensor* y, double re, double im, const sadmrg_tensor* x);
/* inner_prod_c (block_tensor.hpp:117): out2 = {re, im} of x^H y */
int sadmrg_tensor_dot(const sadmrg_tensor* x, const sadmrg_tensor* y, double* out2);
/* reseed the global generator of random_init (dense_inst.cpp:23, default 42) */
int sadmrg_rng_seed(unsigned long long seed);

/* ---- H_eff plan: DMRG::hamiltonian_vector_multiplication_idmrg
        (dmrg.hpp:114-115) as one grouped launch pair ------------------------ */
int sadmrg_heff_create(sadmrg_ctx* ctx, sadmrg_cgc* cgc, const sadmrg_tensor* hl, const sadmrg_tensor* hr, int nterm,
                       const sadmrg_tensor* const* o1, const sadmrg_tensor* const* o2,
                       const sadmrg_tensor* psi_layout, sadmrg_heff** out);
int sadmrg_heff_destroy(sadmrg_heff* h);
/* psi_out is (re)laid out like psi_layout */
int sadmrg_heff_apply(sadmrg_heff* h, const sadmrg_tensor* psi_in, sadmrg_tensor* psi_out, int precision);
/* device pointers in the layout of psi_layout; enqueued on the context stream */
int sadmrg_heff_apply_device(sadmrg_heff* h, const void* d_in, void* d_out, int precision);
/* one grouped launch of the pair (stage 1: X = psi^T O1, stage 2: output
   blocks); stage 0 runs both.  For per-kernel timing. */
int sadmrg_heff_apply_stage(sadmrg_heff* h, const void* d_in, void* d_out, int precision, int stage);
/* The same product with one launch per output block of each stage (no
   grouping across blocks, no split reductions): the per-block baseline of
   the grouped launch, as the reference issues one GEMM per block.  No
   reference counterpart. */
int sadmrg_heff_apply_unbatched(sadmrg_heff* h, const void* d_in, void* d_out, int precision);
/* Process-wide: every H_eff application, the Lanczos solves of the DMRG
   driver included, runs as sadmrg_heff_apply_unbatched does (one launch per
   output block of each stage), so whole sweeps can be timed against the
   grouped launch.  enable 0/1 (default 0), -1 leaves the setting;
   *previous receives the old value.  No reference counterpart. */
int sadmrg_heff_per_block(int enable, int* previous);
/* Process-wide: every H_eff stage runs launch plan `plan` (0 .. 15: tile
   configuration + 4 x split variant + 8 x launch kind, as in
   sadmrg_heff_stats; complex tables accept tile configurations 0 and 3 and
   ignore the rest) instead of the autotuned or default plan; -1 restores
   them, below -1 leaves the setting.  *previous receives the old value.  For
   tests and measurements of single plans.  No reference counterpart. */
int sadmrg_heff_launch_plan(int plan, int* previous);
/* Process-wide: complex H_eff plans run as real tables on planar operands
   (1: each complex block product becomes four real ones on the re and im
   planes, through the real grouped kernel with its 16-byte staging, split
   reductions, persistent launches and autotuner) or on the complex kernel
   over interleaved data (0, the default: one load of each operand panel
   serves its four MFMA products, faster on small blocks).  -1 leaves the
   setting; *previous receives the old value.  No reference counterpart. */
int sadmrg_heff_complex_planar(int enable, int* previous);
/* One plan over n independent plans (instances, e.g. many chains of
   different lengths): their block tables merged, so each H_eff stage of
   every instance runs in a single launch.  Apply it with
   sadmrg_heff_apply_device on the instances' vectors stored back to back in
   the order of parts (sadmrg_heff_stats reports the total length).  The
   parts must share the context and scalar type and carry no penalty
   projectors; they may be destroyed afterwards.  Tensor-level apply and
   Lanczos reject a merged plan.  No reference counterpart. */
int sadmrg_heff_merge(const sadmrg_heff* const* parts, int n, sadmrg_heff** out);
int sadmrg_heff_stage_flops(sadmrg_heff* h, int stage, double* flops);
/* info[8] = {psi blocks, stage-1 problems, stage-2 problems, stage-1 tiles,
   stage-2 tiles, stage-1 launch plan, stage-2 launch plan, output tiles with
   a split reduction}.  Launch plan = tile configuration (0: 64x64, 1:
   128x128, 2: 128x64, 3: 32x32 output tiles; complex tables 0 and 3) + 4 x
   split variant (0: about 3, 1: about 8 reduction parts per resident
   workgroup) + 8 x launch kind (1: persistent). */
int sadmrg_heff_stats(sadmrg_heff* h, double* flops, double* bytes, long long* n, int* info);
/* MFMA flops issued by the launch plans of the plan's last application for
   operands of `precision` (stage 1 or 2, 0: both): every output tile is
   computed whole over its problem's segments, K rounded up to the K step.
   The algorithmic flops over this figure is the share of MFMA work that is
   not edge-tile padding.  No reference counterpart. */
int sadmrg_heff_issued_flops(sadmrg_heff* h, int stage, int precision, double* issued);
/* tile-configuration autotuner of the H_eff stages (process-wide, on by
   default): enable 0/1 (-1 leaves the setting), clear forgets the tuned
   entries, *entries receives the table size.  No reference counterpart (the
   reference has one BLAS path). */
int sadmrg_autotune(int enable, int clear, int* entries);
/* 16-byte operand staging in the H_eff grouped GEMMs (process-wide, on by
   default; the H_eff plans lay their operands out for it): enable 0/1, -1
   leaves the setting; *previous receives the old value.  Off stages single
   elements (the comparison baseline).  No reference counterpart. */
int sadmrg_gemm_vec_staging(int enable, int* previous);
/* Persistent launches of the real grouped GEMMs (process-wide): a grid of
   about the resident workgroup count (one or two per CU) whose workgroups
   take output tiles from a device ticket counter until the table is
   exhausted, instead of one workgroup per tile placed by the hardware
   dispatcher; the last workgroup to leave resets the counter, so launches
   replay from HIP graphs unchanged.  Results are identical.  mode 0: never;
   1: every real launch; 2 (default): the H_eff autotuner times both launch
   kinds for each stage and keeps the faster (other tables launch one
   workgroup per tile); -1 leaves the setting.  *previous receives the old
   mode.  Replaces the reference's loop of per-block GEMMs over sector
   triples (dmrg_inst.cpp:320-347, block_tensor_inst.cpp:2249-2279) like the
   default launch; no reference counterpart of its own. */
int sadmrg_gemm_persistent(int mode, int* previous);
/* Split-reduction policy of the grouped GEMMs (process-wide, for plans built
   afterwards).  A table with fewer output tiles than the device has room for
   splits its reductions.  Balanced (1, the default): each problem's
   reduction is cut into parts of about equal depth, sized to a number of
   parts per resident workgroup that the autotuner picks per table.  Uniform
   (0, the comparison baseline): every problem is cut into the same number of
   parts.  -1 leaves the setting; *previous receives the old value.  No
   reference counterpart. */
int sadmrg_gemm_balanced_split(int enable, int* previous);
/* Truncation route of sweep-step SVDs (BlockTensor::svd with (p1, p2) =
   (1, 0) or (0, 1), block_tensor_inst.cpp:501-637, real data): sectors whose
   larger side has at least gram_min rows are decomposed through the Gram
   matrix when that provably leaves the kept set unchanged, by gesdd
   otherwise.  value > 0 sets gram_min, 0 turns the Gram route off, -1 leaves
   the setting; *previous receives the old value.  Process-wide, default 256.
   No reference counterpart (the reference always calls gesvd). */
int sadmrg_svd_gram_min(int value, int* previous);
/* Small-sector SVD of the truncation: in sweep steps (one side absorbs S),
   real sectors with min(m, n) <= max_order whose working set fits 64 KB of
   LDS are decomposed by a
   one-sided Jacobi kernel (one workgroup per sector, all of a step's small
   sectors in one launch; relatively accurate singular values), in place of
   rocSOLVER gesdd, which replaces the reference's gesvd
   (dense_inst.cpp:486) for the others.  max_order 0 turns it off, -1 leaves
   the setting (process-wide, default 64, at most 88); *previous receives the
   old value; counts[2] = {sectors decomposed by it, sectors handed to gesdd
   after 60 sweeps without convergence} since the last sadmrg_svd_stats
   reset.  No reference counterpart. */
int sadmrg_svd_jacobi(int max_order, int* previous, long long* counts);
/* Batched symmetric eigensolver of the truncation's Gram route: for each of
   `count` real symmetric matrices a[q] (n[q] x n[q], column-major, leading
   dimension lda[q], lower triangle read, contents destroyed; device
   pointers), all eigenvalues in ascending order into w[q] (device, n[q]
   doubles) and the eigenvectors of the k[q] largest into z[q] (device,
   n[q] x k[q], leading dimension ldz[q]; column c belongs to
   w[q][n[q] - 1 - c]; k[q] = 0 asks for none).  Tridiagonalisation of every
   matrix in one persistent launch, bisection, inverse iteration with
   Cholesky-QR re-orthonormalisation, back-transformation; n[q] <= 2048.
   The batched counterpart of ScalarTraits<double>::heev (scalar.hpp:109-114),
   which the reference calls one matrix at a time. */
int sadmrg_eigh_batched(sadmrg_ctx* ctx, int count, const int* n, double* const* a, const int* lda,
                        double* const* w, const int* k, double* const* z, const int* ldz);
/* Inverse iteration of sadmrg_eigh_batched over clusters of close
   eigenvalues (closer than 32 eps |T|): mode 0 (default) iterates the
   members in parallel lanes and orthonormalises each cluster afterwards,
   running a cluster again serially when two members came out nearly
   parallel; mode 1 runs every cluster serially in one lane (LAPACK dstein's
   order, Gram-Schmidt after every solve); -1 only reads.  *previous gets the
   mode before the call, *serial_redo the process-wide count of clusters run
   again serially.  No reference counterpart. */
int sadmrg_eigh_clusters(int mode, int* previous, long long* serial_redo);
/* sadmrg_eigh_batched with every k[q] = 0 (eigenvalues only), allocating
   nothing: no hipMalloc, hipFree or allocating library call.  Its working
   memory (tridiagonal factors, publication words, task tables) is a static
   device array of 2 MiB per device, looked up when the device's first
   context is created and held until the process ends; a batch needs 88
   bytes per matrix row plus 96 per matrix (sum of n[q] up to about 23000)
   and a larger one is rejected with SADMRG_ERR_INVALID before any work.
   Calls that overlap on one device (other contexts or threads) share the
   array and race.  No counterpart in the reference: for callers that cannot
   allocate (graph-capturing or tight serving loops). */
int sadmrg_eigvalsh_batched_noalloc(sadmrg_ctx* ctx, int count, const int* n, double* const* a, const int* lda,
                                    double* const* w);
/* Eigensolver of the Gram route inside BlockTensor::svd (process-wide):
   1 (default) the hand-written batched solver above; 0 rocSOLVER's
   strided-batched syevd on matrices padded to the batch's largest order
   (the comparison).  previous (optional) receives the old setting.  No
   reference counterpart. */
int sadmrg_svd_eigensolver(int value, int* previous);
/* Truncation counters (process-wide): counts[4] = {svd calls, sectors
   decomposed through the Gram route, sectors decomposed by gesdd, Gram
   sectors redone by gesdd}; *seconds = wall time inside svd; reset != 0
   clears them after reading.  No reference counterpart. */
int sadmrg_svd_stats(long long* counts, double* seconds, int reset);
/* Lanczos step mode (process-wide, both on by default): device_steps runs
   the Krylov steps between convergence checks without host
   synchronisation (alpha, beta and the step index stay on the device);
   graph records one such step as a HIP graph once per solve and replays
   it.  The arithmetic is that of the host-driven step.  -1 leaves a
   setting; prev[2] (optional) receives the old values; counts[2]
   (optional) = {steps replayed from graphs, graphs captured} since the
   last reset.  No reference counterpart. */
int sadmrg_lanczos_mode(int device_steps, int graph, int* prev, long long* counts, int reset);
/* Lanczos::lanczos1 (lanczos.hpp:70) on the plan; psi is the start vector on
   entry and the Ritz vector on return.  counters[4] = {FP64 matvecs, FP32
   matvecs, Lanczos steps, refinement restarts}. */
int sadmrg_lanczos(sadmrg_heff* h, sadmrg_tensor* psi, int mlanc, int precision, int refine, double refine_tol,
                   double* energy, int* counters, double* residual);
/* Every option of the solve behind sadmrg_lanczos, for callers that set
   more than the basis size and precision.  mlanc: Krylov vectors (1..2047;
   the reference uses min(40, ndata + 1), dmrg_inst.cpp:267-270);
   check_every: convergence check after every check_every-th step (the
   reference's 4, dmrg_inst.cpp:1334; a value >= mlanc runs the whole
   basis); refine_krylov / max_restarts: FP32 Krylov steps per FP64
   residual-correction restart and the restart cap; device_steps / graph:
   the step mode of this solve (both also need the process-wide switches of
   sadmrg_lanczos_mode).  Bases of more than 64 vectors run host-driven
   steps.  Arguments out of range return SADMRG_ERR_INVALID.  No reference
   counterpart beyond Lanczos::lanczos1 (lanczos.hpp:70). */
typedef struct sadmrg_lanczos_options {
  int mlanc;
  int precision;
  int refine;
  double refine_tol;
  int refine_krylov;
  int max_restarts;
  int check_every;
  int device_steps;
  int graph;
} sadmrg_lanczos_options;
/* Fills *opt with the defaults of sadmrg_lanczos (mlanc 40, FP64, no
   refinement, check every 4th step, device steps and graphs on). */
int sadmrg_lanczos_default_options(sadmrg_lanczos_options* opt);
int sadmrg_lanczos_ex(sadmrg_heff* h, sadmrg_tensor* psi, const sadmrg_lanczos_options* opt, double* energy,
                      int* counters, double* residual);
/* Lanczos::lanczos1 (lanczos.hpp:70, dmrg_inst.cpp:1286-1348) of every
   instance of a merged plan (sadmrg_heff_merge, n_inst parts; a plain plan
   is one instance) at once.  Each Krylov step applies the merged plan once
   (both stages of every instance in one launch each); inner products, the
   recurrence, re-orthogonalisation, normalisation and Ritz vectors are
   segmented per instance, and every instance keeps its own alpha/beta,
   basis size min(mlanc, n_i + 1), breakdown stop and convergence test
   (after every check_every-th step; the check matvec is shared by the
   instances checking at that step).  refine: FP64 checks, then FP64
   residual-correction restarts of the instances with |r| > refine_tol |E|,
   batched, their Krylov steps in the bulk precision (FP32 operands for the
   mixed path; FP64 gives a refined reference); energies are FP64 Rayleigh
   quotients.  d_psi: the instances' start vectors back to back
   (the layout of sadmrg_heff_apply_device), overwritten by their Ritz
   vectors.  energies[n_inst]; steps, restarts, residuals (n_inst each,
   optional); counters[2] (optional) = merged FP64 / FP32 applications.
   mlanc in 2..62; device_steps and graph are ignored (host-driven steps).
   No reference counterpart beyond lanczos1 per instance. */
int sadmrg_lanczos_batched(sadmrg_heff* h, void* d_psi, int n_inst, const sadmrg_lanczos_options* opt,
                           double* energies, int* steps, int* restarts, double* residuals, int* counters);

/* ---- DMRG driver (DMRG<Scalar,Sym>, dmrg.hpp:29-133) --------------------- */
/* DMRG(xx, yy, bondd, physpn, totspin, 0) (dmrg.hpp:99) */
int sadmrg_dmrg_create(sadmrg_ctx* ctx, int sym, int cplx, int xx, int yy, int bondd, int physpn, int totspin,
                       sadmrg_dmrg** out);
/* DMRG(xx, yy, bondd, physpn, totspin, max_exci) (dmrg.hpp:99,
   dmrg_inst.cpp:63-110): the driver targets state max_exci and keeps
   stored states 0..max_exci-1 (orth) with their energies.  Every H_eff
   on two adjacent sites then carries the penalty -E_j |phi_j><phi_j|,
   phi_j = stored state j projected on the current basis
   (dmrg_inst.cpp:352-363).  max_exci <= 64; excited states need ns >= 6. */
int sadmrg_dmrg_create_exci(sadmrg_ctx* ctx, int sym, int cplx, int xx, int yy, int bondd, int physpn, int totspin,
                            int max_exci, sadmrg_dmrg** out);
int sadmrg_dmrg_destroy(sadmrg_dmrg* d);
int sadmrg_dmrg_ns(sadmrg_dmrg* d, int* ns);
/* max_exci of the driver */
int sadmrg_dmrg_exci(sadmrg_dmrg* d, int* exci);
/* stored state j < max_exci := src's MPS and energy (device copy).  The
   reference loads stored states only from files (read_mps with read2,
   dmrg_inst.cpp:1082-1141); no reference counterpart. */
int sadmrg_dmrg_set_lower_state(sadmrg_dmrg* d, int j, const sadmrg_dmrg* src);
/* gs_enr[j] of stored state j < max_exci (the penalty weight is -gs_enr[j]) */
int sadmrg_dmrg_lower_energy(sadmrg_dmrg* d, int j, double* e);
/* deep copy of src's couplings, MPS, environments and settings into dst
   (created with the same lattice, symmetry and scalar type); the reference
   copies a DMRG state only through its checkpoint files (dmrg_inst.cpp:947-1177) */
int sadmrg_dmrg_copy_state(sadmrg_dmrg* dst, const sadmrg_dmrg* src);
/* replaces hmap (ns x ns, row-major) and coup (dmrg_inst.cpp:117-176) */
int sadmrg_dmrg_set_couplings(sadmrg_dmrg* d, const int* hmap, const double* coup, int ncoup);
int sadmrg_dmrg_get_couplings(sadmrg_dmrg* d, int* hmap, double* coup, int cap);
/* Channel weights, U(1) drivers only (no reference counterpart: the
   reference's couplings are real, one per bond, dmrg.hpp:79).  A bond
   (j < k) with coupling index c contributes coup[c] * sum_q w[c][q]
   O_j(q) O_k(q) over the spin operators' channel labels q = -2, 0, +2,
   instead of coup[c] * S_j.S_k (all weights 1).  w holds 6 doubles per
   coupling index: (re, im) for q = -2, 0, +2; n = 6 * ncoup, or 0 to
   restore the isotropic exchange.  w(-2) must be the conjugate of w(+2)
   and w(0) real (Hermitian bond terms); complex weights need a complex
   driver.  XXZ with a z-axis Dzyaloshinskii-Moriya term D (S^x S^y -
   S^y S^x): w = (Jxy - iD, Jz, Jxy + iD); the sign convention of D does not
   change energies. */
int sadmrg_dmrg_set_channel_weights(sadmrg_dmrg* d, const double* w, int n);
/* the global max_dcut (main/main_dmrg_su2_real.cpp:32) */
int sadmrg_dmrg_set_max_dcut(sadmrg_dmrg* d, int max_dcut);
int sadmrg_dmrg_set_precision(sadmrg_dmrg* d, int precision, int refine, double refine_tol);
/* refinement restarts of the mixed path: FP32 Krylov steps per restart and
   the restart cap (defaults 6 and 6); no reference counterpart */
int sadmrg_dmrg_set_refinement(sadmrg_dmrg* d, int krylov_steps, int max_restarts);
/* Lanczos convergence check after every `every`-th Krylov step; the
   reference checks every 4th (dmrg_inst.cpp:1334), the default.  A shorter
   interval lets the mixed path stop with fewer FP32 matvecs. */
int sadmrg_dmrg_set_check_interval(sadmrg_dmrg* d, int every);
/* Mixed precision with refinement (no reference counterpart): which sweep
   steps compute the FP64 energy, i.e. the FP64 Rayleigh quotient and the
   residual-correction restarts.  1 (default): every step.  k > 1: every
   k-th step of a sweep and its last step.  0: the last step only.  The
   other steps make no FP64 matvec: they stop when the recurrence residual
   beta_m |y_m| of the FP32 Krylov Ritz pair is at most refine_tol |E| and
   record that Ritz value (flag 0 in sadmrg_dmrg_records).  iDMRG steps
   always compute the FP64 energy. */
int sadmrg_dmrg_set_fp64_energy_every(sadmrg_dmrg* d, int every);
/* 1: one line per sweep step on stderr (bond, energy, kept states, stage
   times), as the reference prints every step of DMRG::sweep to cout
   (dmrg_inst.cpp:881-897); 0 (default): silent. */
int sadmrg_dmrg_set_verbose(sadmrg_dmrg* d, int level);
/* Checkpoint files in the reference's format (dmrg_inst.cpp:947-1177), in
   directory dir (NULL or "" = the current directory, where the reference
   writes): uu-{ns}-{bondd}-{totspin}-{exci}-{i}.dat / .bin per site
   (save_mps2, dmrg.hpp:129), ww-{ns}-{bondd}-{totspin}-{exci}.bin (save_ww,
   dmrg.hpp:130), enr-{ns}-{bondd}-{totspin}-{exci}.dat (save_enr, appended,
   dmrg.hpp:125).  The own files carry the driver's max_exci as exci;
   save_mps names the files with the given exci as save_mps2 does. */
int sadmrg_dmrg_save_mps(sadmrg_dmrg* d, const char* dir, int exci);
int sadmrg_dmrg_save_ww(sadmrg_dmrg* d, const char* dir);
int sadmrg_dmrg_save_enr(sadmrg_dmrg* d, const char* dir);
/* read_mps(read1) (dmrg.hpp:127) when read2 < 0, read_mps(read1, read2)
   (dmrg.hpp:128: also read_ww and read_enr) otherwise; *pass = 1 when every
   site's files were read, as the reference's return value.  With
   max_exci > 0, read_mps(read1, read2) also loads stored states
   0..max_exci-1 and their energies from the files of bond dimension read2;
   missing files of a stored state are an error.  A malformed or short
   file, or a CGC block that is not the canonical one of its labels, is an
   error (the reference reads such files unchecked). */
int sadmrg_dmrg_read_mps(sadmrg_dmrg* d, const char* dir, int read1, int read2, int* pass);
int sadmrg_dmrg_read_ww(sadmrg_dmrg* d, const char* dir, int read1);   /* read_ww (dmrg.hpp:131) */
int sadmrg_dmrg_read_enr(sadmrg_dmrg* d, const char* dir, int read1);  /* read_enr (dmrg.hpp:126) */
/* sweep() ends with save_mps2(exci), save_ww(), save_enr() into dir, as the
   reference's sweep does (dmrg_inst.cpp:940-942); NULL (the default) turns
   the files off. */
int sadmrg_dmrg_set_checkpoint_dir(sadmrg_dmrg* d, const char* dir);
/* ww[exci]: singular values of the last truncation (kept, then discarded) or
   what read_ww loaded; *n receives the count, up to cap are copied. */
int sadmrg_dmrg_ww(sadmrg_dmrg* d, double* w, int cap, int* n);
int sadmrg_dmrg_idmrg(sadmrg_dmrg* d);         /* do_idmrg (dmrg.hpp:107) */
int sadmrg_dmrg_prepare_sweep(sadmrg_dmrg* d); /* prepare_sweep (dmrg.hpp:110) */
int sadmrg_dmrg_sweep(sadmrg_dmrg* d);         /* sweep (dmrg.hpp:111) */
int sadmrg_dmrg_energy(sadmrg_dmrg* d, double* e); /* get_enr (dmrg.hpp:102) */
/* per-step records since the last call: 14 doubles each {il, ir, energy,
   truncation error, kept, FP64 matvecs, FP32 matvecs, t_env, t_lanczos, t_svd,
   H_eff flops, reduced elements of the wave function, the last Ritz value of
   the FP32 Krylov recurrence (mixed precision; 0 otherwise), 1 if energy is
   an FP64 Rayleigh quotient (0: the FP32 Ritz value, see
   sadmrg_dmrg_set_fp64_energy_every)}; *n = number written; records are
   consumed */
int sadmrg_dmrg_records(sadmrg_dmrg* d, double* rec, int cap, int* n);
/* which: 0 uu[i], 1 hh[i], 2 opr[i][j]; for stored state j < max_exci:
   3 orth[j][i], 4 ovlp[j][i], 5 overlapvec[j] (i ignored) */
int sadmrg_dmrg_get_tensor(sadmrg_dmrg* d, int which, int i, int j, sadmrg_tensor* out);
int sadmrg_dmrg_set_tensor(sadmrg_dmrg* d, int which, int i, int j, const sadmrg_tensor* t);
/* step methods: 0 wavefunc_transformation(a,b), 1 prepare_site_operator_from_left(a,b),
   2 prepare_site_operator_from_right(a,b), 3 prepare_site_operator_from_left(a),
   4 prepare_site_operator_from_right(a), 5 prepare_sweep() (dmrg.hpp:90-121) */
int sadmrg_dmrg_call(sadmrg_dmrg* d, int code, int a, int b);
/* dE/dJ_c for every coupling index c (cap entries of grad; ns+1 covers every
   index of set_couplings' coup array) at the two-site wave function of the
   last Lanczos solve: bonds across the two sites are measured directly,
   bonds inside a block through the adjoint of the block-Hamiltonian
   recursion (env kind 5).  *energy = <psi|H|psi> from the same pieces.  No
   reference counterpart (the reference computes no gradients). */
int sadmrg_dmrg_coupling_gradient(sadmrg_dmrg* d, double* grad, int cap, double* energy);
/* prepare_input_vector (dmrg.hpp:121) */
int sadmrg_dmrg_input_vector(sadmrg_dmrg* d, int il, int ir, sadmrg_tensor* out);
/* hamiltonian_vector_multiplication_idmrg (dmrg.hpp:114) */
int sadmrg_dmrg_heff(sadmrg_dmrg* d, int il, int ir, const sadmrg_tensor* vin, sadmrg_tensor* vout, int precision);
/* The H_eff plan of step (il, ir) as the sweep's Lanczos solve builds it:
   the operand set of hamiltonian_vector_multiplication_idmrg
   (dmrg_inst.cpp:283-371) on the driver's current environments, for wave
   functions in the layout of `layout`.  Apply it with sadmrg_heff_apply*,
   release it with sadmrg_heff_destroy; it holds copies of the operands. */
int sadmrg_dmrg_heff_create(sadmrg_dmrg* d, int il, int ir, const sadmrg_tensor* layout, sadmrg_heff** out);
/* lanczos_solve_eigenvector_idmrg (dmrg.hpp:90); vec updated in place */
int sadmrg_dmrg_lanczos(sadmrg_dmrg* d, int il, int ir, sadmrg_tensor* vec, double* energy);
/* vec.svd(uu[a], p1, uu[b], p2, ww) as in DMRG::sweep (dmrg_inst.cpp:894) */
int sadmrg_dmrg_svd(sadmrg_dmrg* d, const sadmrg_tensor* vec, int a, int b, double p1, double p2, double* wout,
                    int cap, int* nw);

#ifdef __cplusplus
}
#endif
#endif /* SADMRG_H */
