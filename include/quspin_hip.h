/* libquspin_hip: C-ABI of the MI355X (gfx950) symmetry-reduced H.v path.
 *
 * Plain pointers and sizes only.  Every "device" pointer is HIP device
 * memory; "host" pointers are CPU memory.  `stream` is a hipStream_t (NULL =
 * default stream).  Functions return 0 (QSX_OK) or a negative status
 * (QSX_ERR_*); qsx_last_error() gives the message of the calling thread's
 * last failure.  Vectors are C-contiguous (Ns, nvec) arrays.
 *
 * Reference interfaces replaced (paths relative to the QuSpin 1.0.1 tree):
 *   basis cores   quspin_extensions ... _basis_general_core.*_basis_core_wrap,
 *                 called from src/quspin/basis/basis_general/base_general.py
 *   sparse tools  parallel_sparse_tools.matvec._oputils._matvec and
 *                 expm_multiply_parallel_wrapper._wrapper_* (rendered sources:
 *                 docs/modules/parallel_sparse_tools/matvec/matvec_core.html,
 *                 .../expm_multiply_parallel_core/expm_multiply_parallel_core.html)
 */
#ifndef QUSPIN_HIP_H
#define QUSPIN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QSX_OK = 0, QSX_ERR_ARG = -1, QSX_ERR_HIP = -2, QSX_ERR_OVERFLOW = -3, QSX_ERR_COMPLEX = -4,
       QSX_ERR_UNSUPPORTED = -5 };
/* vector / matrix element types */
enum { QSX_F32 = 0, QSX_F64 = 1, QSX_C64 = 2, QSX_C128 = 3 };
/* basis kinds */
enum { QSX_KIND_HCB = 0, QSX_KIND_SPINLESS_FERMION = 1, QSX_KIND_SPINFUL_FERMION = 2 };

typedef struct qsx_core qsx_core;             /* symmetry group + bit layout of one basis */
typedef struct qsx_oplist qsx_oplist;         /* compiled operator-string list */
typedef struct qsx_hv_ctx qsx_hv_ctx;         /* bound H.v (operator, basis, dtype, nvec, flags) */
typedef struct qsx_csr_builder qsx_csr_builder;

const char* qsx_last_error(void);
int qsx_version(void);

/* ---- basis core --------------------------------------------------------- */

/* Replaces hcb_basis_core_wrap(basis_dtype, N, maps, pers, qs)
 * (base_hcb.py:74-76), spinless_fermion_basis_core_wrap (fermion.py:169-171)
 * and spinful_fermion_basis_core_wrap(..., double_occupancy)
 * (fermion.py:543-545).  maps: host int32[n_maps][N] (entry -(j+1) = site j
 * with inversion), pers/qs: host int32[n_maps].  N = number of bits (2L for
 * spinful fermions), N <= 64, n_maps <= 32. */
int qsx_core_create(int kind, int N, const int32_t* maps, const int32_t* pers, const int32_t* qs, int n_maps,
                    int double_occupancy, qsx_core** out);
void qsx_core_destroy(qsx_core* core);
int qsx_core_group_order(const qsx_core* core);
int qsx_core_map_info(const qsx_core* core, int k, int* rev, int* ngroups, int* per);

/* Replaces _core.make_basis(basis, n, Np=, count=) (base_general.py:864-873).
 * Np: host int32 particle numbers (hcb: n_Np values; spinful: n_Np
 * (Nup, Ndown) pairs; n_Np = 0: no conservation).  basis_out == NULL counts
 * only.  Writes the representatives DESCENDING (uint32 for N <= 32, else
 * uint64) and n (n_bytes = 1, 2 or 4 per entry).  *Ns_out receives Ns.  When
 * capacity < Ns nothing is written and QSX_ERR_OVERFLOW is returned with
 * *Ns_out = the size needed (the reference's core signals this case by a
 * negative return; a binding maps one onto the other). */
int qsx_make_basis(qsx_core* core, const int32_t* Np, int n_Np, void* basis_out, void* n_out, int n_bytes,
                   int64_t capacity, int64_t* Ns_out, void* stream);

/* Replaces _core.make_basis_blocks(N_p, basis) (base_general.py:945-947):
 * device int64 begin/end[2^N_p] bounding the states with each N_p-bit prefix. */
int qsx_make_basis_blocks(qsx_core* core, int N_p, const void* basis, int64_t Ns, int64_t* begin, int64_t* end,
                          void* stream);

/* Replaces _core.representative(states, out, g_out=, sign_out=)
 * (base_general.py:1085-1117): orbit maxima; g_out int32[n][n_maps] and
 * sign_out int8[n] optional. */
int qsx_representative(qsx_core* core, const void* states, void* out, int32_t* g_out, int8_t* sign_out, int64_t n,
                       void* stream);

/* Replaces _core.normalization(states, out) (base_general.py:1158, 1178). */
int qsx_normalization(qsx_core* core, const void* states, void* out, int out_bytes, int64_t n, void* stream);

/* Sector rank of arbitrary states (ascending, -1 outside the sector of the
 * last make_basis) and the sector descriptor info[8] = {mode, Np, Nup, Ndown,
 * sector size, sector_ok, rank_ok, N}.  Internal index of the rank lookups;
 * no direct counterpart in the reference (its native make_basis_blocks
 * index plays this role). */
int qsx_sector_rank(qsx_core* core, const void* states, int64_t n, int64_t* out, int64_t* info, void* stream);

/* Lookup table of the sector (state rank -> basis index | element code):
 * size query and build.  Trades HBM for the orbit search of the H.v; no
 * counterpart in the reference.  nrm / n_bytes: the basis n array (device).
 * When lut_len > 0 the code field is an id into a coefficient table of
 * lut_len double2 (chi * sqrt(n) per distinct character chi(g) of the group
 * and distinct n of the basis), filled by qsx_table_build, so the H.v needs
 * no read of the target's n; otherwise it is the group element. */
int qsx_table_info(const qsx_core* core, int64_t Ns, const void* nrm, int n_bytes, int64_t* entries, int* entry_bytes,
                   int* idx_bits, int64_t* lut_len);
int qsx_table_build(qsx_core* core, const void* basis, int64_t Ns, const int64_t* begin, const int64_t* end, int N_p,
                    const void* nrm, int n_bytes, void* table, int entry_bytes, int idx_bits, void* lut,
                    int64_t lut_len, void* stream);

/* ---- operators ----------------------------------------------------------- */

/* Compile an operator list: opstrs = all operator strings concatenated, lens
 * their lengths, indx the sites concatenated, J (re, im) pairs, all host.
 * symmetric = 1 when the list commutes with the basis symmetries (enables
 * the one-launch gather form).  Site/character validation as in
 * base_general.py:299-311. */
int qsx_oplist_create(qsx_core* core, int n_terms, const char* opstrs, const int32_t* lens, const int32_t* indx,
                      const double* J, int symmetric, qsx_oplist** out);
void qsx_oplist_destroy(qsx_oplist* l);
int qsx_oplist_info(const qsx_oplist* l, int* n_off, int* n_diag, int* complex_me);

/* Replaces _core.op(row, col, ME, opstr, indx, J, basis, n, begin, end, N_p)
 * (base_general.py:324-336), argument order kept (each array followed by its
 * element size or length): one element per basis state (column); ME is
 * complex128, zero where the element vanishes; row/col int32 or int64
 * (idx_bytes 4 or 8). */
int qsx_op(qsx_core* core, void* row, void* col, int idx_bytes, void* ME, const char* opstr, const int32_t* indx,
           int n_op, double J_re, double J_im, const void* basis, const void* n, int n_bytes, int64_t Ns,
           const int64_t* begin, const int64_t* end, int N_p, void* stream);

/* Replaces _core.inplace_op(v_in, v_out, conjugated, transposed, opstr, indx,
 * J, basis, n, begin, end, N_p) (base_general.py:395-408), argument order
 * kept: v_out += J Op v_in (transposed / conjugated). */
int qsx_inplace_op(qsx_core* core, const void* v_in, void* v_out, int dtype, int64_t nvec, int conjugated,
                   int transposed, const char* opstr, const int32_t* indx, int n_op, double J_re, double J_im,
                   const void* basis, const void* n, int n_bytes, int64_t Ns, const int64_t* begin,
                   const int64_t* end, int N_p, void* stream);

/* Fused form of the loop over terms in base_general.py:392-408 plus the
 * diagonal product of quantum_LinearOperator._matvec
 * (quantum_LinearOperator_core.py:520-540), one launch:
 *   v_out = beta v_out + a_diag (D + diag_arr) v_in + a_off Op v_in. */
int qsx_hv(qsx_core* core, qsx_oplist* l, const void* basis, const void* n, int n_bytes, int64_t Ns,
           const int64_t* begin, const int64_t* end, int N_p, const void* v_in, void* v_out, int dtype, int64_t nvec,
           int transposed, int conjugated, int diag_conj, double a_off_re, double a_off_im, double a_diag_re,
           double a_diag_im, int beta, const void* diag_arr, int diag_arr_cplx, void* stream);

/* Bound form of qsx_hv for repeated products (Lanczos, expm): create once,
 * choose the lookup (table) and schedule (0 one lane per row, 1 compacted
 * queue, 2/3/4 = 2/4/8 lanes per row, 5/6 = batches of 4/8 terms, 7 =
 * split-chain blocks, 8/9 = direct-rank chunks of 4/8 terms with predicated
 * gathers, 10 = batches of 14 terms with the table, 11 = split-chain layout
 * in two passes through a buffer the context allocates, 12 = as 10 with the
 * workgroups of each XCD on one contiguous run of rows), apply. */
int qsx_hv_ctx_create(qsx_core* core, qsx_oplist* l, const void* basis, const void* n, int n_bytes, int64_t Ns,
                      const int64_t* begin, const int64_t* end, int N_p, int dtype, int64_t nvec, int transposed,
                      int conjugated, int diag_conj, const void* diag_arr, int diag_arr_cplx, qsx_hv_ctx** out);
/* Destroys the context and the launch sequences captured with it
 * (qsx_lanczos_run).  The set_table / set_schedule / set_diag calls below
 * likewise drop the captures made before them. */
void qsx_hv_ctx_destroy(qsx_hv_ctx* ctx);
/* lut: the coefficient table of qsx_table_build and its length (NULL and 0
 * when qsx_table_info gave lut_len 0) */
int qsx_hv_ctx_set_table(qsx_hv_ctx* ctx, const void* table, int entry_bytes, int idx_bits, const void* lut,
                         int lut_len);
int qsx_hv_ctx_set_schedule(qsx_hv_ctx* ctx, int sched);
/* Replace the context's diagonal (diagonal terms and extra diagonal) by a
 * precomputed per-row array (Ns doubles, or Ns (re, im) pairs when
 * diag_arr_cplx), as quantum_LinearOperator keeps `_diagonal`
 * (quantum_LinearOperator_core.py:137-156); the caller keeps it alive. */
int qsx_hv_ctx_set_diag(qsx_hv_ctx* ctx, const void* diag_arr, int diag_arr_cplx);
int qsx_hv_ctx_apply(qsx_hv_ctx* ctx, const void* x, void* y, double a_off_re, double a_off_im, double a_diag_re,
                     double a_diag_im, int beta, void* stream);

/* Column 1-norms colabs[j] = sum_i |A_ij + shift delta_ij| (Ns doubles) and
 * the diagonal dg[j] = A_jj (Ns (re, im) pairs) of the context's operator,
 * without assembling it: duplicates of one (i, j) are summed before |.|, as
 * in the canonical CSR whose _wrapper_csr_trace / _wrapper_csr_1_norm
 * (EMP.html:196-202) give ExpmMultiplyParallel's mu and ||A - mu I||_1.
 * QSX_ERR_UNSUPPORTED (nothing usable written) when a column holds more
 * than 64 distinct rows, or for a transposed context whose list does not
 * commute with the group: assemble the CSR then. */
int qsx_hv_ctx_column_norms(qsx_hv_ctx* ctx, double shift_re, double shift_im, double* colabs, double* dg,
                            void* stream);

/* Replaces the COO -> CSR assembly of basis._make_matrix (base.py:265-304)
 * for hamiltonian: canonical CSR (sorted columns, duplicates summed, zeros
 * dropped) of a*Op.  Two steps: create (returns nnz), finish (writes
 * indptr/indices/data in the requested types), destroy. */
int qsx_csr_builder_create(qsx_core* core, qsx_oplist* l, const void* basis, const void* n, int n_bytes, int64_t Ns,
                           const int64_t* begin, const int64_t* end, int N_p, int transposed, int conjugated,
                           double a_re, double a_im, qsx_csr_builder** out, int64_t* nnz, void* stream);
int qsx_csr_builder_complex(const qsx_csr_builder* b);
int qsx_csr_builder_finish(qsx_csr_builder* b, int dtype, int idx_bytes, void* indptr, void* indices, void* data,
                           void* stream);
void qsx_csr_builder_destroy(qsx_csr_builder* b);

/* ---- sectors and projections --------------------------------------------- */
/* The projector of a reduced basis, P[:, r] = c_r sum_g chi(g) sign(g, r)
 * |g.r>, c_r = 1/sqrt(n_r |G|) (get_proj, base_general.py:521-582, the
 * coefficient at :571-575).  "Full space": the 2^N states in descending order
 * (pcon = 0, Ns_full = 2^N) or the states of the particle sector recorded by
 * the last qsx_make_basis, descending (pcon = 1, Ns_full = its size). */

/* Replaces _core.op_shift_sector(v_in, v_out, opstr, indx, J, basis, n,
 * other_basis, other_n) (base_general.py:504-514), argument order kept (each
 * array followed by its element size or length): one term, v_out (Ns x nvec)
 * += P^H Op P_other v_in (other_Ns x nvec), where core describes the target
 * basis (basis, n, its prefix blocks) and other_core the source basis; both
 * must carry the same symmetry maps (characters may differ).  dtype
 * QSX_F64 or QSX_C128. */
int qsx_op_shift_sector(qsx_core* core, qsx_core* other_core, const void* v_in, void* v_out, int dtype,
                        int64_t nvec, const char* opstr, const int32_t* indx, int n_op, double J_re, double J_im,
                        const void* basis, const void* n, int n_bytes, int64_t Ns, const int64_t* begin,
                        const int64_t* end, int N_p, const void* other_basis, const void* other_n, int other_n_bytes,
                        int64_t other_Ns, void* stream);

/* Replaces _core.get_amp(states, out, n, mode) (base_general.py:1254): out[i]
 * = P[s_i, rep(s_i)] = sign conj(chi(g)) sqrt(n_rep / |G|) (g.s_i = rep,
 * 0 outside the sector), in out_dtype; mode 0 "representative", 1
 * "full_basis" (states, device, are replaced by their representatives). */
int qsx_get_amp(qsx_core* core, void* states, void* out, int out_dtype, int64_t n, int mode, void* stream);

/* Replace _core.project_from_dense(basis, n, v0, v_out, basis_pcon=) and
 * _core.project_to_dense(...) (base_general.py:787-789, 677-679): v_out
 * (Ns_full x nvec, zeros on entry) = P v0 (Ns x nvec), and v_out (Ns x nvec)
 * = P^H v0 (Ns_full x nvec). */
int qsx_project_from_dense(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns,
                           const void* v0, void* v_out, int dtype, int64_t nvec, int pcon, int64_t Ns_full,
                           void* stream);
int qsx_project_to_dense(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns, const void* v0,
                         void* v_out, int dtype, int64_t nvec, int pcon, int64_t Ns_full, void* stream);

/* Replaces _core.get_proj(basis, dtype, sign, c, indices, indptr,
 * basis_pcon=) (base_general.py:580-582): P as CSC arrays, rows ascending
 * within a column.  Call with indices == NULL for *nnz_out, then with
 * indptr (Ns + 1), indices and data (nnz) allocated (idx_bytes 4 or 8). */
int qsx_get_proj(qsx_core* core, const void* basis, const void* n, int n_bytes, int64_t Ns, int pcon, int64_t Ns_full,
                 int dtype, int idx_bytes, void* indptr, void* indices, void* data, int64_t* nnz_out, void* stream);

/* ---- sparse tools -------------------------------------------------------- */

/* Replaces _matvec(A, x, out=, overwrite_out=, a=) for csr A
 * (hamiltonian_core.py:607-612, 741-772; matvec_core.html): y = a A x, or
 * y += a A x when accumulate.  idx64 selects int64 indices; group = lanes
 * per row (0: automatic).  Matrix dtype = vector dtype or its real part. */
int qsx_csr_matvec(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype, int64_t Ns,
                   int64_t nnz, const void* x, void* y, int vec_dtype, int64_t nvec, double a_re, double a_im,
                   int accumulate, int group, void* stream);

/* Deterministic device reductions: op 0, out[0] + i out[1] = sum of n
 * (re, im) pairs x; op 1, out[0] = max of n doubles x. */
int qsx_reduce(int op, const double* x, int64_t n, double* out, void* stream);

/* Replace _wrapper_csr_trace(indptr, indices, data) and
 * _wrapper_csr_1_norm(indptr, indices, data, shift) (EMP.html:196-202);
 * results in device doubles (trace: re, im).  colsum: Ns device doubles. */
int qsx_csr_trace(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype, int64_t Ns,
                  double* out, void* stream);
int qsx_csr_1_norm(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype, int64_t Ns,
                   double shift_re, double shift_im, double* colsum, double* out, void* stream);
/* Same with an extra diagonal: ||A + shift I + xs diag(xdiag)||_1, the norm
 * LazyOperatorNormInfo needs (EMP.html:399-466) for H(t) = A + f(t) D with a
 * diagonal dynamic part D (xdiag, Ns device doubles). */
int qsx_csr_1_norm_diag(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                        int64_t Ns, double shift_re, double shift_im, const double* xdiag, double xs_re, double xs_im,
                        double* colsum, double* out, void* stream);
/* Column data for the norms above without a pass over A per shift:
 * colabs[j] = sum_i |A_ij| (Ns device doubles), dg[j] = A_jj (Ns (re, im)
 * pairs).  The pieces of _wrapper_csr_1_norm (EMP.html:200-202) that do not
 * depend on the shift. */
int qsx_csr_col_split(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype, int64_t Ns,
                      double* colabs, double* dg, void* stream);
/* The partition selection of ExpmMultiplyParallel (EMP.html:377-385, scipy
 * _fragment_3_1's theta branch) for a block of time steps of
 * H(t) = A + f(t) D, without leaving the device: es holds nsteps step
 * records (a, mu, dscale = f(t) set; m and s written), the norms
 * ||A + dscale D - mu I||_1 come from the column data (colabs, dg) and D
 * (Ns doubles or null); theta/mth list the nth theta_m values and their m in
 * the order scipy visits them, bound is the branch's bound on |a| * norm;
 * fail[k] = 1 marks a step above it.  nmax: nsteps words of scratch. */
int qsx_partition_exact(const double* colabs, const double* dg, const double* D, int64_t N, void* es, int nsteps,
                        const double* theta, const int* mth, int nth, double bound, unsigned long long* nmax,
                        int* fail, void* stream);

/* Device CSR algebra (hamiltonian_core.py:1686-1720 tocsr(time), and the
 * hamiltonian's __add__, __mul__, astype, transpose, conj, getH, which the
 * reference runs through scipy).  Inputs are canonical CSR (sorted, unique
 * column indices per row); so are the results.
 *
 * qsx_csr_sum_create: sum_k coef_k op_k(A_k) of K <= 32 matrices with Nr rows
 * (op_k = complex conjugate where conjs[k] = 1; coefs = K (re, im) pairs;
 * per matrix: index width idx64s[k], value dtype dtypes[k]); exact-zero sums
 * are dropped, as scipy's addition drops them.  Returns nnz; finish writes
 * indptr (Nr + 1), indices and values (nnz) in the given widths. */
typedef struct qsx_csr_sum qsx_csr_sum;
int qsx_csr_sum_create(int K, const void* const* indptrs, const void* const* indices, const void* const* datas,
                       const int* idx64s, const int* dtypes, const int* conjs, const double* coefs, int64_t Nr,
                       qsx_csr_sum** out, int64_t* nnz, void* stream);
int qsx_csr_sum_finish(qsx_csr_sum* h, int out_dtype, int idx_bytes, void* out_indptr, void* out_indices,
                       void* out_data, void* stream);
void qsx_csr_sum_destroy(qsx_csr_sum* h);
/* (Conjugate) transpose into preallocated arrays: indptr (Nc + 1), indices and
 * values (nnz).  Deterministic (radix sort on (column, row)). */
int qsx_csr_transpose(const void* indptr, const void* indices, const void* data, int idx64, int dtype, int64_t Nr,
                      int64_t Nc, int64_t nnz, int conj, int out_dtype, int idx_bytes, void* out_indptr,
                      void* out_indices, void* out_data, void* stream);

/* Replaces _wrapper_expm_multiply(indptr, indices, data, s, m_star, tol, mu,
 * a, v, work) and _wrapper_expm_multiply_batch (EMP.html:347-372): v <-
 * e^{a A} v in place, work = 2*Ns*nvec elements of the vector dtype. */
int qsx_expm_multiply_csr(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                          int64_t Ns, int64_t nnz, int s, int m_star, double tol, double mu_re, double mu_im,
                          double a_re, double a_im, void* v, void* work, int vec_dtype, int64_t nvec, int group,
                          void* stream);

/* Allocation-free form of qsx_expm_multiply_csr (same arguments, same
 * kernels, same results).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * Taylor loop's control block (truncation flag, per-term norms, per-block
 * partials) lives in a library-owned device array of
 * qsx_expm_multiply_csr_noalloc_max_bytes() bytes (4 MiB) per GPU, held until
 * the process ends; a partition whose block does not fit is rejected with
 * QSX_ERR_ARG.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race.  Each call clears the
 * block itself, with a kernel. */
int qsx_expm_multiply_csr_noalloc(const void* indptr, const void* indices, const void* data, int idx64,
                                  int mat_dtype, int64_t Ns, int64_t nnz, int s, int m_star, double tol, double mu_re,
                                  double mu_im, double a_re, double a_im, void* v, void* work, int vec_dtype,
                                  int64_t nvec, int group, void* stream);
int64_t qsx_expm_multiply_csr_noalloc_max_bytes(void);

/* The same Taylor loop on the matrix-free operator of an H.v context (no
 * CSR in memory); no counterpart in the reference. */
int qsx_expm_multiply_mf(void* hv_ctx, int64_t Ns, int s, int m_star, double tol, double mu_re, double mu_im,
                         double a_re, double a_im, void* v, void* work, int vec_dtype, int64_t nvec, void* stream);

/* Lanczos recurrence pieces (src/quspin/tools/lanczos/_lanczos_utils.py:
 * 173-204): r -= b_prev v; a = Re<q, r>; r -= a q; b = ||r|| with FP64
 * accumulation (a_out, b_out device doubles), q = r / b, vdot, nrm2, axpy,
 * and lin_comb_Q_T (:366-425). */
int qsx_lanczos_update(int dtype, const void* v, const double* b_prev, const void* q, void* r, int64_t n,
                       double* a_out, double* b_out, void* stream);
int qsx_lanczos_div(int dtype, const void* r, const double* b, void* q, int64_t n, void* stream);
/* The whole m-step recurrence of lanczos_full / lanczos_iter
 * (_lanczos_utils.py:158-227) on the device: op_kind 0 = H.v context
 * (hv_ctx), 1 = CSR (indptr/indices/data).  Q holds q0 in row 0 and receives
 * all m vectors when store_Q, else it is a ring of 3 vectors; ab = m x
 * (a re/im, b re/im) doubles.  use_graph captures the sequence once per
 * (operator, buffers, sizes) and replays it; a capture on an H.v context
 * lives until that context changes or is destroyed, and holds reduction
 * scratch of its own. */
int qsx_lanczos_run(int op_kind, void* hv_ctx, const void* indptr, const void* indices, const void* data, int idx64,
                    int mat_dtype, int64_t nnz, int group, int dtype, int64_t n, int m, void* Q, int store_Q, void* r,
                    double* ab, int use_graph, void* stream);
int qsx_vdot(int dtype, const void* x, const void* y, int64_t n, double* out, void* stream);
int qsx_nrm2(int dtype, const void* x, int64_t n, double* out, void* stream);
int qsx_axpy(int dtype, const void* x, void* y, int64_t n, double a_re, double a_im, const double* sdev, double sgn,
             int cplx_scale, void* stream);
int qsx_lincomb(int dtype, const void* Q, int64_t m, int64_t n, const double* c, void* out, void* stream);

/* Plain 16-byte-vector copy (reference point for HBM counters). */
int qsx_copy(const void* src, void* dst, int64_t bytes, void* stream);

/* ---- batched sectors x vectors ---------------------------------------------- */

/* Many symmetry sectors (one H.v context each, same dtype and site count, each
 * with a lookup table or rank lookup) applied to nvec vectors per sector in
 * ONE launch: vectors of all sectors are stacked row-wise, sector k from row
 * row_off[k] (qsx_hv_batch_rows).  One wavefront per row, lanes over the
 * vectors.  No single counterpart in the reference: it batches the
 * per-sector quantum_LinearOperator.dot / basis.inplace_Op calls
 * (quantum_LinearOperator_core.py:288-347, base_general.py:354-410). */
typedef struct qsx_hv_batch qsx_hv_batch;
int qsx_hv_batch_create(int nsec, qsx_hv_ctx* const* ctxs, int64_t nvec, int chunk_rows, qsx_hv_batch** out);
/* CSR form of a batch (same apply / expm entry points): sector k is a device
 * CSR matrix with Ns[k] rows (indptr int64 when idx64[k], int32 indices,
 * data of the vector dtype) -- the matrices hamiltonian.tocsr() builds
 * (hamiltonian_core.py:607-612; base.py:265-304), applied to all sectors at
 * once.  chunk_rows <= 63. */
int qsx_csr_batch_create(int nsec, const int64_t* Ns, const void* const* indptr, const int32_t* idx64,
                         const void* const* indices, const void* const* data, int dtype, int64_t nvec,
                         int chunk_rows, qsx_hv_batch** out);
void qsx_hv_batch_destroy(qsx_hv_batch* b);
int qsx_hv_batch_rows(const qsx_hv_batch* b, int64_t* row_off);
/* Rows per wavefront chunk; xcd_slices = 1 gives each XCD a contiguous slice
 * of the chunks, 0 keeps dispatch order (tuning only: results are equal). */
int qsx_hv_batch_set_schedule(qsx_hv_batch* b, int chunk_rows, int xcd_slices);
/* y = a * H x (+ y if beta) for every sector. */
int qsx_hv_batch_apply(qsx_hv_batch* b, const void* x, void* y, double a_re, double a_im, int beta, void* stream);

/* e^{a H_k} V_k for every sector k in lockstep: per-sector partition (s[k],
 * m[k]), tolerance tol[k] and shift mu[2k] (re, im) exactly as
 * expm_multiply_parallel(H_k, a).dot(V_k) would use them (EMP.html:267-374,
 * _wrapper_expm_multiply_batch; Al-Mohy & Higham Alg. 3.2, scipy
 * _expm_multiply.py:280-303).  v = stacked F, overwritten; work = 2 * rows *
 * nvec elements; nvec <= 64.  use_graph = 1 captures the whole launch
 * sequence as one hipGraph per (v, work, s_max, m_max) and replays it. */
int qsx_expm_multiply_batch(qsx_hv_batch* b, const int32_t* s, const int32_t* m, const double* tol,
                            const double* mu, double a_re, double a_im, void* v, void* work, int use_graph,
                            void* stream);

/* ---- time stepping with learned partitions ------------------------------------ */

/* Device record of one Taylor step of one sector: a, mu, factor of the extra
 * diagonal, tolerance, partition (s, m).  qsx_expm_record_bytes() gives its
 * size; field order: double2 a, double2 mu, double2 dscale, double tol,
 * int s, int m, int pad[2], then padding to a multiple of 16 bytes (80). */
int qsx_expm_record_bytes(void);

/* nsteps consecutive steps v <- e^{a_t (A + dscale_t diag(xdiag))} v
 * with A a device CSR (as qsx_expm_multiply_csr) and the records es[t]
 * (ExpmSec layout above); the same tail check into fail[t].  ctl_work:
 * qsx_evolve_csr_ctl_bytes(s_max, m_max) bytes of device scratch owned by the
 * caller.  vec_dtype is the dtype of the Taylor terms (work) and the SpMV;
 * state_dtype that of v, the accumulated sum: the same, or its FP64
 * counterpart (FP32 terms with FP64 accumulation), or -1 for the same.
 * use_graph captures the whole sequence once per set of operator arrays,
 * sizes and buffers and replays it on later calls with the same set; at
 * most 64 captures are kept (oldest dropped first).  qsx_graph_cache_clear
 * drops every capture. */
int64_t qsx_evolve_csr_ctl_bytes(int s_max, int m_max);
int qsx_evolve_csr_steps(const void* indptr, const void* indices, const void* data, int idx64, int mat_dtype,
                         int64_t Ns, int64_t nnz, const double* xdiag, int nsteps, const void* es, int s_max,
                         int m_max, void* v, void* work, int vec_dtype, int state_dtype, int64_t nvec, int* fail,
                         void* ctl_work, int64_t ctl_bytes, int group, int use_graph, void* stream);
void qsx_graph_cache_clear(void);

/* ---- species-factorised operator (spinful fermions, no spatial symmetry) ---- */
/* In the (N_up, N_down) sector of spinful_fermion_basis_general without
 * symmetry maps the state index is u * Nd + d (basis_general/fermion.py:466-474 sizes the sector as the
 * product C(N, N_up) C(N, N_down); up occupations are the high bits), and
 * an operator whose off-diagonal terms each act on one species ('|' strings of
 * hamiltonian's static list, hamiltonian_core.py:292-356) is
 *     H = H_up (x) 1 + 1 (x) H_down + diag(sdiag) + xs diag(xdiag).
 * The factors are device CSR matrices of the spinless sectors (the layout of
 * qsx_csr_matvec, int32 column indices, shared dtype mat_dtype); the arrays
 * stay owned by the caller.  The product is the same as hamiltonian.dot on
 * the CSR of H (hamiltonian_core.py:615-789) with one read and one write of
 * psi per row instead of the CSR stream. */
typedef struct qsx_kron qsx_kron;
int qsx_kron_create(const void* ip_u, const void* ind_u, const void* dat_u, int idx64_u, int64_t Nu, int64_t nnz_u,
                    const void* ip_d, const void* ind_d, const void* dat_d, int idx64_d, int64_t Nd, int64_t nnz_d,
                    const double* sdiag, int mat_dtype, qsx_kron** out);
/* also drops the step sequences captured with this operator */
void qsx_kron_destroy(qsx_kron* K);
/* Kernel form: 0 by size (default: one pass while psi fits in L2, else a
 * row pass over H_down and an L2-resident column-strip pass over H_up), 1 one
 * pass, 2 two passes.  The two-pass form allocates an N_u*N_d-element buffer
 * on its first product or step sequence (outside any capture). */
int qsx_kron_set_mode(qsx_kron* K, int mode);
/* y = a (H_up (x) 1 + 1 (x) H_down + diag(sdiag) + xs diag(xdiag)) x
 * (+ y if accumulate); sdiag (given at create: the static diagonal of the
 * terms that act on both species) and xdiag: Nu*Nd device doubles or NULL. */
int qsx_kron_matvec(const qsx_kron* K, const double* xdiag, double xs_re, double xs_im, const void* x, void* y,
                    int vec_dtype, double a_re, double a_im, int accumulate, void* stream);
/* ||H + diag(sdiag) - mu I + xs diag(xdiag)||_1 into out[0] (the norm LazyOperatorNormInfo
 * needs, EMP.html:399-466): offu / offd are the factors' off-diagonal column
 * 1-norms (device doubles), diag_u / diag_d their diagonals as complex
 * (re, im) device doubles; work: 1024 device doubles. */
int qsx_kron_1_norm(const qsx_kron* K, const double* offu, const double* diag_u, const double* offd,
                    const double* diag_d, const double* xdiag, double xs_re, double xs_im, double mu_re, double mu_im,
                    double* work, double* out, void* stream);
/* qsx_evolve_csr_steps on the factorised operator (one vector). */
int qsx_evolve_kron_steps(const qsx_kron* K, const double* xdiag, int nsteps, const void* es, int s_max, int m_max,
                          void* v, void* work, int vec_dtype, int state_dtype, int* fail, void* ctl_work,
                          int64_t ctl_bytes, int use_graph, void* stream);

/* ---- partition predictor (MLP 4 -> 64 -> 64 -> 2, tanh, FP64) ------------------ */

/* Replaces the per-step 1-norm estimate and _fragment_3_1 choice of
 * ExpmMultiplyParallel._calculate_partition (EMP.html:377-385; scipy
 * _expm_multiply.py:503-558) for time-dependent H(t).  params: W1[64][4],
 * b1[64], W2[64][64], b2[64], W3[2][64], b3[2]; norm[12] = input shift[4],
 * input scale[4], output shift[2], output scale[2].  Inputs per sample:
 * (t, dt, drive phase, ||H(t)||_1 bound); outputs (m, s). */
int qsx_mlp_param_count(void);
int qsx_mlp_forward(const double* params, const double* norm, const double* X, int64_t n, double* Y, void* stream);
/* Gradient of L = 1/(2n) sum |y_hat - t_hat|^2 (normalised outputs); work:
 * qsx_mlp_grad_work() device doubles; loss may be null. */
int64_t qsx_mlp_grad_work(void);
int qsx_mlp_grad(const double* params, const double* norm, const double* X, const double* T, int64_t n, double* grad,
                 double* loss, double* work, int64_t work_elems, void* stream);
int qsx_mlp_adam(double* params, const double* grad, double* m, double* v, int64_t step, double lr, double beta1,
                 double beta2, double eps, void* stream);
/* (m, s) of n steps into their records: m = ceil(y_m) + m_margin, s =
 * round(y_s), clamped to [1, m_cap] x [1, s_cap]. */
int qsx_mlp_partition(const double* params, const double* norm, const double* X, int64_t n, void* es, int m_margin,
                      int m_cap, int s_cap, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QUSPIN_HIP_H */
