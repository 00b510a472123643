// ppe.h -- pressure Poisson equation: assembly and PCG solve on the device.
#pragma once
#include <string>
#include "mpsr_internal.h"

namespace mpsr {

template <typename R> struct PpeSolver;

template <typename R>
PpeSolver<R>* ppe_create(const Dev<R>& d, const mpsr_params& P, const mpsr_ppe_options& opt, double n0_lap,
                         double lambda_lap, double n0_src);
template <typename R> void ppe_destroy(PpeSolver<R>* s);
// The PND pass of the PPE path (PND into out.n, and each Poisson row's active flag, neighbour
// bound, diagonal and Dirichlet flag); ppe_solve assembles from its results.
template <typename R> void ppe_start(PpeSolver<R>* s, const Dev<R>& d, hipStream_t st);
// Assembles A p = b from the current binning and the ppe_start results and solves it; writes the
// pressure to in.p / out.p and to the sorted pressure array used by the gradient kernel.
template <typename R>
mpsr_status ppe_solve(PpeSolver<R>* s, Dev<R>& d, const mpsr_ppe_options& opt, int profile, hipStream_t st);
// Iteration counts and residuals of the last solve (synchronises); error when it failed.
template <typename R> mpsr_status ppe_check(PpeSolver<R>* s, mpsr_timings* tm);
template <typename R> const char* ppe_why(PpeSolver<R>* s);
template <typename R> mpsr_status ppe_get(PpeSolver<R>* s, const Dev<R>& d, double* rhs, double* x, int* dir);
// The assembled matrix in slot order: row pointers [n+1], column slots, off-diagonal values,
// diagonal [n] (Dirichlet rows: no entries, diagonal 1).
template <typename R>
mpsr_status ppe_get_matrix(PpeSolver<R>* s, const Dev<R>& d, int* ptr, int* col, double* val, double* diag, int cap,
                           int* nnz);
template <typename R> void ppe_update_params(PpeSolver<R>* s, const mpsr_params& P);
// the neighbour list of the last ppe_start (valid until the next binning)
template <typename R> NbList ppe_neighbors(const PpeSolver<R>* s);
// SpMV launch timing of a step / step_many call starts afresh
template <typename R> void ppe_begin_call(PpeSolver<R>* s);
// the last solve's iterations, final FP64 relative residual and whether it rebuilt the coarse operator
template <typename R> void ppe_last(const PpeSolver<R>* s, int* iters, double* relres, int* rebuilt);
// the learned warm start / preconditioner network (parameters flat FP32)
template <typename R> mpsr_status ppe_mpnn_set(PpeSolver<R>* s, const float* w, hipStream_t st);
template <typename R> mpsr_status ppe_mpnn_get(PpeSolver<R>* s, float* w, hipStream_t st);
template <typename R>
mpsr_status ppe_mpnn_train(PpeSolver<R>* s, const Dev<R>& d, double lr, double beta, int update, float* grads,
                           double* parts, hipStream_t st);
template <typename R>
mpsr_status ppe_mpnn_rows(PpeSolver<R>* s, const Dev<R>& d, float* X, float* Y, double* scale, hipStream_t st);

// The PPE solves of the PPE buckets of a batched step (FP64 Jacobi-PCG, the single-domain solver with
// MPSR_PCG_FP64 and MPSR_PRECOND_JACOBI): solver j belongs to the bucket dev[j] of the group's Dev
// table.  A step runs ppe_batch_pnd and ppe_batch_head (assembly, first residual, start and a
// first chunk of iterations), then chunks of ppe_batch_iterate until ppe_batch_converged, then
// ppe_batch_tail (final residual, checks, write-back).
template <typename R> struct PpeBatch;
template <typename R> PpeBatch<R>* ppe_batch_create();
template <typename R> void ppe_batch_destroy(PpeBatch<R>* b);
template <typename R>
mpsr_status ppe_batch_build(PpeBatch<R>* b, PpeSolver<R>* const* solvers, const int* dev, const Dev<R>* host_devs,
                            const mpsr_ppe_options* opts, int n, int block);
template <typename R> void ppe_batch_pnd(PpeBatch<R>* b, const BatchLaunch<R>& bl, hipStream_t st);
template <typename R> void ppe_batch_head(PpeBatch<R>* b, const BatchLaunch<R>& bl, int iters, hipStream_t st);
template <typename R> void ppe_batch_iterate(PpeBatch<R>* b, int iters, hipStream_t st);
template <typename R> mpsr_status ppe_batch_converged(PpeBatch<R>* b, hipStream_t st, bool* all);
template <typename R> void ppe_batch_tail(PpeBatch<R>* b, const BatchLaunch<R>& bl, hipStream_t st);
// the solvers' host state after a batched call; *failed = the first solver whose last solve failed
// (why: the reason), -1 when none
template <typename R> mpsr_status ppe_batch_finish(PpeBatch<R>* b, int* failed, std::string* why);

}  // namespace mpsr
