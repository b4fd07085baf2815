// Internal declarations shared by the HIP translation units.
#pragma once

#include <string>

#include "qepw_common.h"
#include "qepw_plan.h"
#include "../../include/qepw.h"

namespace qepw {

constexpr int kMaxLdsBytes = 160 * 1024;

// qepw_fft.hip
int stick_group_for(int n3, int nst);
size_t apply_local_band_bytes(const Plan& p, int precision);
int apply_local_group(int count, const Plan* const* plans, int precision,
                      const void* const* vectors, const int64_t* ldv, const int* nb,
                      const void* const* potentials, const void* const* kinetics,
                      void* const* results, const int64_t* ldr, void* const* workspaces,
                      const size_t* ws_bytes, hipStream_t stream);
int apply_local(const Plan& p, int precision, const void* vectors, int64_t ldv, int nb,
                const void* potential, const void* kinetic, void* result, int64_t ldr,
                void* workspace, size_t ws_bytes, hipStream_t stream);
size_t density_workspace_bytes(const Plan& p, int precision, int nb);
int accumulate_density(const Plan& p, int precision, const void* vectors, int64_t ldv, int nb,
                       const double* weights, double scale, double* out, void* workspace,
                       size_t ws_bytes, hipStream_t stream);
unsigned team_slot_mask(int n, const std::vector<int>& xs, const std::vector<int>& ys);
size_t inverse_real_bytes(const Plan& p);
int inverse_real(const Plan& p, const void* grid_g, void* out_real, void* workspace,
                 size_t ws_bytes, hipStream_t stream);

// qepw_gemm.hip
size_t gemm_workspace_bytes(int precision, int op_a, int64_t m, int64_t n, int64_t k);
int gemm_set_rule(int precision, int op_a, const int64_t* ranges, int variant, int64_t splits);
void gemm_clear_rules();
int gemm_variant_count();
int gemm_plan(int precision, int op_a, int64_t m, int64_t n, int64_t k, int64_t* out);
int gemm(int precision, int op_a, int64_t m, int64_t n, int64_t k, const double* alpha,
         const void* a, int64_t lda, const void* b, int64_t ldb, const double* beta,
         void* c, int64_t ldc, void* workspace, size_t ws_bytes, hipStream_t stream);
int herm_product(int n, const void* x, const void* y, const void* d, void* c, int64_t ld,
                 double alpha, double beta, hipStream_t stream);

// qepw_vec.hip
int column_squared_norms(int precision, const void* v, int64_t ldv, int64_t nrows, int nb,
                         double* out, hipStream_t stream);
int column_inner_products(int precision, const void* l, int64_t ldl, const void* r,
                          int64_t ldr, int64_t nrows, int nb, double* out, hipStream_t stream);
int subtract_band_energies(int precision, void* r, int64_t ldr, const void* x, int64_t ldx,
                           const double* e, int64_t nrows, int nb, hipStream_t stream);
int qe_precondition(int precision, const void* r, int64_t ldr, const int* cols,
                    const double* e, const double* diag, void* out, int64_t ldo,
                    int64_t nrows, int nb, hipStream_t stream);
int normalize_selected_columns(int precision, const void* src, int64_t lds,
                               const double* sq, const int* sel, void* dst, int64_t ldd,
                               int64_t nrows, int nb, hipStream_t stream);
int restart_applied(int precision, const void* x, int64_t ldx, const void* r, int64_t ldr,
                    const double* e, void* hv, int64_t ldh, int64_t nrows, int nb,
                    hipStream_t stream);
int permute_rows(int src_precision, int dst_precision, int direction, const int* perm,
                 const void* src, int64_t lds, void* dst, int64_t ldd, int64_t nrows, int nb,
                 hipStream_t stream);
int row_real_dot(const void* a, int64_t lda, const void* b, int64_t ldb, int64_t nrows,
                 int ncols, double scale, double* acc, hipStream_t stream);
int copy_columns(int src_precision, int dst_precision, const void* src, int64_t lds,
                 const int* scol, void* dst, int64_t ldd, const int* dcol, const double* scale,
                 int64_t nrows, int ncols, hipStream_t stream);
int set_block(int precision, void* dst, int64_t ldd, int64_t nrows, int ncols, double diag,
              const double* d, hipStream_t stream);
int scale_rows(int precision, const void* src, int64_t lds, const double* s, int64_t sstride,
               const double* cfac, void* dst, int64_t ldd, int64_t nrows, int nb,
               hipStream_t stream);
int build_projectors(int precision, const void* beta, int64_t ldb, int nch, const double* q,
                     const double* tau, int natom, void* out, int64_t ldo, int64_t nrows,
                     hipStream_t stream);
int phase_blocks(int precision, const void* beta, int64_t ldb, int nch, const void* phases,
                 int64_t ldp, int natom, const int* perm, void* out, int64_t ldo, int64_t nrows,
                 hipStream_t stream);
int projector_forces(int precision, const void* dO, int64_t lddo, int64_t dstride,
                     const void* DO, int64_t lddO, const double* f, int nb, const int* off,
                     int natom, double* out, hipStream_t stream);

int column_axpby(int precision, const void* x, int64_t ldx, const double* a, void* y,
                 int64_t ldy, const double* b, int64_t nrows, int ncols, hipStream_t stream);
int column_diagonal_expectations(int precision, const void* v, int64_t ldv, const double* d,
                                 int64_t nrows, int nb, double* out, hipStream_t stream);
int rmm_precondition(int precision, const void* r, int64_t ldr, const int* cols,
                     const double* kinetic, const double* expect, void* out, int64_t ldo,
                     int64_t nrows, int ncols, hipStream_t stream);
int history_combine(int precision, const void* hist, int64_t ldh, int64_t hstride, int count,
                    const int* cols, const double* coef, void* out, int64_t ldo, int64_t nrows,
                    int ncols, hipStream_t stream);

int projector_basis(const double* q, int64_t npw, const double* table, int64_t ldt, int nq,
                    double dq, const int* chan_radial, const int* chan_l, const int* chan_m,
                    int nch, void* beta, int64_t ldb, hipStream_t stream);

// qepw_eig.hip
int hermitian_complete(void* a, int64_t lda, int first, int active, hipStream_t stream);
int shift_scale(const void* a, int64_t lda, int n, double sigma, void* x, int64_t ldx,
                double* bound, hipStream_t stream);
int hermitize(void* x, int64_t ldx, int n, hipStream_t stream);
int sign_check(const void* x, int64_t ldx, const void* x2, int64_t ld2, int n, double* out,
               hipStream_t stream);
int set_diagonal(void* a, int64_t lda, int n, const double* d, hipStream_t stream);
size_t sign_iteration_workspace(int n);
int sign_iteration(void* x, void* x2, void* y, int64_t ld, int n, const double* alphas,
                   int nalpha, int* in_x2, void* ws, size_t ws_bytes, hipStream_t stream);

// qepw_rr.hip
void set_host_lapack(void* zpotrf, void* ztrtri, void* zheevd, void* cheevd);
bool host_lapack_ready();
const std::string& rr_message();
size_t rr_workspace_bytes(int maxsub);
int rr_standardize(void* H, void* S, void* W, void* A, int maxsub, int first, int active,
                   int* not_pd, void* work, size_t work_bytes, hipStream_t stream);
int rr_spectral(const qepw_rr_spectral_args* a, double* values, qepw_rr_spectral_result* out);
int cholesky_qr(int precision, const void* block, int64_t ldb, int64_t npw, int nb,
                double condition, void* out, int64_t ldo, int* ok, void* work,
                size_t work_bytes, hipStream_t stream);

}  // namespace qepw
